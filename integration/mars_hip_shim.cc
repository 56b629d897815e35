// mars::cuda on top of libmars_hip (C ABI of include/mars_hip.h).
//
// A maintainer of the reference drops this file into its cuda/ directory in
// place of cuda/orthonormalize.cu, compiles it with the host compiler, defines
// MARS_USE_CUDA for marsalgo.cc (what the reference's USE_CUDA option does,
// CMakeLists.txt:110-113) and links libmars_hip.so.  The reference's call sites
// (marsalgo.cc:455-481 orthonormalize, 675-752 orthonormalize_batch, 375-379
// context_destroy) are then unchanged; its `cuda=True` runs on the MI355X.
//
// Each function below replaces the declaration of cuda/mars_cuda.h named in its
// comment.  Errors come back from the C ABI as status codes and are rethrown as
// std::runtime_error, as the reference's CUDA_CHECK did
// (orthonormalize.cu:35-53), so the binding's exception capture
// (marslib.cc:113-145) is unchanged.  The DGKS counter is a plain long at the C
// boundary; it is added to the caller's std::atomic<long> here.

#include "mars_cuda.h"
#include "mars_hip.h"

#include <stdexcept>
#include <string>

namespace mars {
namespace cuda {

struct Context {
    mars_hip_ctx *h;
};

static void check(int rc)
{
    if (rc != MARS_HIP_OK) throw std::runtime_error(std::string("mars_hip: ") + mars_hip_last_error());
}

// mars_cuda.h:43
Context *context_create(size_t n, size_t max_terms, double tol)
{
    Context *c = new Context{nullptr};
    const int rc = mars_hip_context_create(n, max_terms, tol, &c->h);
    if (rc != MARS_HIP_OK) {
        delete c;
        check(rc);
    }
    return c;
}

// mars_cuda.h:48
void context_destroy(Context *ctx)
{
    if (!ctx) return;
    mars_hip_context_destroy(ctx->h);
    delete ctx;
}

// mars_cuda.h:55
size_t context_batch_capacity(Context *ctx)
{
    size_t cap = 0;
    check(mars_hip_context_batch_capacity(ctx->h, &cap));
    return cap;
}

// mars_cuda.h:71
void context_sync_basis(Context *ctx, size_t m, const float *B, size_t ldB, const float *Bo, size_t ldBo)
{
    check(mars_hip_context_sync_basis(ctx->h, m, B, ldB, Bo, ldBo));
}

// mars_cuda.h:80
void context_set_target(Context *ctx, const float *y)
{
    check(mars_hip_context_set_target(ctx->h, y));
}

// mars_cuda.h:95
void context_sync_xcols(Context *ctx, size_t p_x, const float *X, size_t ldX, const float *s, const int *cols,
                        size_t ncols)
{
    check(mars_hip_context_sync_xcols(ctx->h, p_x, X, ldX, s, cols, ncols));
}

// mars_cuda.h:127
void orthonormalize(Context *ctx, size_t m, size_t p, const float *x, const int *mask, float *Bx, size_t ldBx,
                    double *ybx, std::atomic<long> *dgks_counter)
{
    long fired = 0;
    check(mars_hip_orthonormalize(ctx->h, m, p, x, mask, Bx, ldBx, ybx, dgks_counter ? &fired : nullptr));
    if (dgks_counter) *dgks_counter += fired;
}

// mars_cuda.h:152
void orthonormalize_batch(Context *ctx, size_t m, size_t P, const int *src_xcol, const int *src_basis, double *ybx,
                          std::atomic<long> *dgks_counter)
{
    long fired = 0;
    check(mars_hip_orthonormalize_batch(ctx->h, m, P, src_xcol, src_basis, ybx, dgks_counter ? &fired : nullptr));
    if (dgks_counter) *dgks_counter += fired;
}

}  // namespace cuda
}  // namespace mars
