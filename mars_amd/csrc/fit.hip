// Resident fit state on the device: constructor, append, step evaluation and
// the C ABI of libmars_hip (include/mars_hip.h).
#include "internal.h"
#include "kernels.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace mh {
namespace {

thread_local std::string g_last_error;

/* ---------------------------------------------------------------------- */
/* Constructor kernels (marsalgo.cc:266-370)                              */
/* ---------------------------------------------------------------------- */
enum { S_YN = 0, S_WN, S_YBO0, S_MEAN, S_YVAR, S_BADX, S_PROJ, S_PROJ2, S_VV, S_COUNT };

__global__ __launch_bounds__(256) void k_init1(const float *__restrict__ y, const float *__restrict__ w,
                                               size_t n, double *__restrict__ yd,
                                               double *__restrict__ sw, double *__restrict__ part)
{
    __shared__ double sh[256];
    double ay = 0.0, aw = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)kRedBlocks * 256) {
        double yi = (double)y[i];
        double si = sqrt((double)w[i]);
        if (!isfinite(yi)) yi = si = 0.0;
        const double d = yi * si;
        yd[i] = d;
        sw[i] = si;
        ay = fma(d, d, ay);
        aw = fma(si, si, aw);
    }
    ay = block_sum_256(ay, sh);
    aw = block_sum_256(aw, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = ay;
        part[kRedBlocks + blockIdx.x] = aw;
    }
}

__global__ __launch_bounds__(256) void k_init2(double *__restrict__ yd, const double *__restrict__ sw,
                                               size_t n, const double *__restrict__ part,
                                               float *__restrict__ y32, float *__restrict__ B,
                                               float *__restrict__ Bo, size_t ldb,
                                               double *__restrict__ part2, double *__restrict__ scal)
{
    __shared__ double sh[256];
    const double yn = sqrt(sum_partials(part, sh));
    const double wn = sqrt(sum_partials(part + kRedBlocks, sh));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[S_YN] = yn;
        scal[S_WN] = wn;
    }
    double ab = 0.0, am = 0.0;
    if (yn > 0.0 && wn > 0.0) {
        for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)kRedBlocks * 256) {
            const double d = yd[i] / yn;
            const double s = sw[i] / wn;
            yd[i] = d;
            const float yf = (float)d;
            const float b = (float)s;
            y32[i] = yf;
            B[i * ldb] = b;
            Bo[i * ldb] = b;
            ab = fma((double)b, (double)yf, ab);
            am += d;
        }
    }
    ab = block_sum_256(ab, sh);
    am = block_sum_256(am, sh);
    if (threadIdx.x == 0) {
        part2[blockIdx.x] = ab;
        part2[kRedBlocks + blockIdx.x] = am;
    }
}

__global__ __launch_bounds__(256) void k_init3(const double *__restrict__ yd, size_t n,
                                               const double *__restrict__ part2,
                                               double *__restrict__ part3, double *__restrict__ scal)
{
    __shared__ double sh[256];
    const double ybo0 = sum_partials(part2, sh);
    const double mean = sum_partials(part2 + kRedBlocks, sh) / (double)n;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[S_YBO0] = ybo0;
        scal[S_MEAN] = mean;
    }
    double av = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)kRedBlocks * 256) {
        const double dv = yd[i] - mean;
        av = fma(dv, dv, av);
    }
    av = block_sum_256(av, sh);
    if (threadIdx.x == 0) part3[blockIdx.x] = av;
}

__global__ __launch_bounds__(256) void k_init4(const double *__restrict__ part3, size_t n,
                                               double *__restrict__ scal, double *__restrict__ ybo)
{
    __shared__ double sh[256];
    const double v = sum_partials(part3, sh);
    if (threadIdx.x == 0) {
        scal[S_YVAR] = v / (double)n;
        ybo[0] = scal[S_YBO0];
    }
}

/* Column sums of squares, 64 blocks per column. */
__global__ __launch_bounds__(256) void k_colsq(const float *__restrict__ X, size_t ldx, size_t n,
                                               double *__restrict__ part)
{
    __shared__ double sh[256];
    const float *x = X + (size_t)blockIdx.y * ldx;
    double a = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += 64 * 256) {
        const double v = (double)x[i];
        a = fma(v, v, a);
    }
    a = block_sum_256(a, sh);
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * 64 + blockIdx.x] = a;
}

__global__ void k_colscale(const double *__restrict__ part, int px, size_t n,
                           float *__restrict__ s, double *__restrict__ scal)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= px) return;
    double sq = 0.0;
    for (int b = 0; b < 64; ++b) sq += part[(size_t)j * 64 + b];
    const double rms = sqrt(sq / (double)n);
    if (!isfinite(rms)) scal[S_BADX] = 1.0;
    s[j] = rms > 0.0 ? (float)(1.0 / rms) : 1.0f;
}

/* ---------------------------------------------------------------------- */
/* Append kernels (marsalgo.cc:760-875, kernels.cc:239-281)               */
/* ---------------------------------------------------------------------- */
__device__ inline float relu0(float a) { return a < 0.0f ? 0.0f : a; }

__global__ __launch_bounds__(256) void k_app_build(const float *__restrict__ X, size_t ldx,
                                                   const float *__restrict__ sc,
                                                   const float *__restrict__ B, size_t ldb,
                                                   int xcol, int bcol, int type, float h, size_t n,
                                                   double *__restrict__ v, double *__restrict__ part,
                                                   double *__restrict__ scal)
{
    __shared__ double sh[256];
    const float s = sc[xcol];
    const float *xp = X + (size_t)xcol * ldx;
    double a = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)kRedBlocks * 256) {
        const float sb = __fmul_rn(s, B[i * ldb + bcol]);
        float f;
        if (type == 0)
            f = __fmul_rn(sb, xp[i]);
        else if (type == 1)
            f = __fmul_rn(sb, relu0(__fsub_rn(xp[i], h)));
        else
            f = __fmul_rn(sb, relu0(__fsub_rn(h, xp[i])));
        const double d = (double)f;
        v[i] = d;
        a = fma(d, d, a);
    }
    a = block_sum_256(a, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = a;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[S_PROJ] = 0.0;
        scal[S_PROJ2] = 0.0;
    }
}

/*
 * One modified Gram-Schmidt step j of the residual v against Bo's column j:
 * apply the update of column j-1 (coefficient = sum of the previous step's
 * partials) and accumulate the partials of column j's coefficient; the final
 * step (j == m) accumulates |v|^2 instead.  On the first sweep, step 0 also
 * stores the raw-norm-normalized B column.  `gate` (retry sweep) makes the
 * launch a no-op unless the DGKS gate fired.
 */
__global__ __launch_bounds__(256) void k_app_mgs(const float *__restrict__ Bo, float *__restrict__ B,
                                                 size_t ldb, int mcol, size_t n,
                                                 double *__restrict__ v,
                                                 const double *__restrict__ prev,
                                                 double *__restrict__ next, int j, int m,
                                                 int first, const int *__restrict__ gate,
                                                 double *__restrict__ scal)
{
    __shared__ double sh[256];
    if (gate && !*gate) return;
    double c = 0.0;
    float rn = 1.0f;
    if (j > 0) c = sum_partials(prev, sh);
    if (j == 0 && first) rn = (float)sqrt(sum_partials(prev, sh));
    double a = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)kRedBlocks * 256) {
        double vi = v[i];
        if (j == 0 && first) B[i * ldb + mcol] = __fdiv_rn((float)vi, rn);
        if (j > 0) {
            vi = fma(-c, (double)Bo[i * ldb + (j - 1)], vi);
            v[i] = vi;
        }
        if (j < m)
            a = fma((double)Bo[i * ldb + j], vi, a);
        else
            a = fma(vi, vi, a);
    }
    a = block_sum_256(a, sh);
    if (threadIdx.x == 0) next[blockIdx.x] = a;
    if (j > 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        double *acc = first ? &scal[S_PROJ] : &scal[S_PROJ2];
        *acc = fma(c, c, *acc);
    }
}

__global__ __launch_bounds__(256) void k_app_gate(const double *__restrict__ part, double tol,
                                                  int *__restrict__ gate, double *__restrict__ scal,
                                                  long long *__restrict__ dgks, int retry_done)
{
    __shared__ double sh[256];
    const double vv = sum_partials(part, sh);
    if (threadIdx.x == 0) {
        if (!retry_done) {
            const int fire = (vv > tol && vv * kGate < scal[S_PROJ]) ? 1 : 0;
            gate[0] = fire;
            if (fire) *dgks += 1;
            scal[S_VV] = vv;
        } else if (gate[0]) {
            scal[S_VV] = vv;
        }
    }
}

__global__ __launch_bounds__(256) void k_app_store(float *__restrict__ Bo, size_t ldb, int mcol,
                                                   size_t n, const double *__restrict__ v,
                                                   const float *__restrict__ y, double tol,
                                                   const double *__restrict__ scal,
                                                   double *__restrict__ part)
{
    __shared__ double sh[256];
    const double w = sqrt(scal[S_VV]);
    const bool ok = w * w > tol;
    double a = 0.0;
    if (ok) {
        for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)kRedBlocks * 256) {
            const float b = (float)(v[i] / w);
            Bo[i * ldb + mcol] = b;
            a = fma((double)b, (double)y[i], a);
        }
    }
    a = block_sum_256(a, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = a;
}

__global__ __launch_bounds__(256) void k_app_final(const double *__restrict__ part, int m, size_t n,
                                                   double tol, const double *__restrict__ scal,
                                                   double *__restrict__ ybo, const int *__restrict__ gate,
                                                   AppendOut *__restrict__ out)
{
    __shared__ double sh[256];
    const double yb = sum_partials(part, sh);
    if (threadIdx.x != 0) return;
    const double w = sqrt(scal[S_VV]);
    AppendOut o;
    o.fired = gate[0];
    o.ybo_m = yb;
    o.status = 0;
    o.mse = -1.0;
    if (w * w > tol) {
        double sq = yb * yb;
        for (int k = 0; k < m; ++k) sq = fma(ybo[k], ybo[k], sq);
        const double mse = (1.0 - sq) / (double)n;
        if (mse >= -tol) {
            o.status = 1;
            o.mse = mse > 0.0 ? mse : 0.0;
            ybo[m] = yb;
        }
    }
    *out = o;
}

/* ---------------------------------------------------------------------- */
/* Dense (p_x, m) outputs and the argmax (mars.py:342-343)                */
/* ---------------------------------------------------------------------- */
__global__ void k_dense_init(double *__restrict__ lin, double *__restrict__ hin,
                             double *__restrict__ cut, int *__restrict__ knot, size_t total)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    lin[i] = 0.0;
    hin[i] = 0.0;
    cut[i] = __builtin_nan("");
    knot[i] = -1;
}

__global__ void k_scatter(const int *__restrict__ cxc, const int *__restrict__ cpar, int P, int m,
                          const double *__restrict__ clin, const double *__restrict__ chin,
                          const int *__restrict__ cknot, const int *__restrict__ crow,
                          const float *__restrict__ X, size_t ldx, int hinge,
                          double *__restrict__ lin, double *__restrict__ hin,
                          double *__restrict__ cut, int *__restrict__ knot)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P) return;
    const size_t d = (size_t)cxc[c] * m + cpar[c];
    lin[d] = clin[c];
    if (hinge && cknot[c] >= 0) {
        hin[d] = chin[c];
        cut[d] = (double)X[(size_t)cxc[c] * ldx + crow[c]];
        knot[d] = cknot[c];
    }
}

/* First maximum in row-major order, like numpy.argmax (values are >= 0). */
__global__ __launch_bounds__(1024) void k_argmax(const double *__restrict__ lin,
                                                 const double *__restrict__ hin,
                                                 const int *__restrict__ knot,
                                                 const double *__restrict__ cut, size_t total,
                                                 BestOut *__restrict__ out)
{
    __shared__ double sv[2][1024];
    __shared__ long long si[2][1024];
    const int t = threadIdx.x;
    const size_t chunk = (total + 1023) / 1024;
    const size_t b = t * chunk, e = b + chunk < total ? b + chunk : total;
    double v1 = -1.0, v2 = -1.0;
    long long i1 = -1, i2 = -1;
    for (size_t i = b; i < e; ++i) {
        if (lin[i] > v1) { v1 = lin[i]; i1 = (long long)i; }
        if (hin[i] > v2) { v2 = hin[i]; i2 = (long long)i; }
    }
    sv[0][t] = v1; si[0][t] = i1;
    sv[1][t] = v2; si[1][t] = i2;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) {
            for (int q = 0; q < 2; ++q) {
                const double a = sv[q][t], bb = sv[q][t + s];
                const long long ia = si[q][t], ib = si[q][t + s];
                // larger value wins; on equal values the earlier index wins
                if (bb > a || (bb == a && ib >= 0 && (ia < 0 || ib < ia))) {
                    sv[q][t] = bb;
                    si[q][t] = ib;
                }
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        BestOut o;
        o.v1 = sv[0][0] > 0.0 ? sv[0][0] : 0.0;
        o.c1 = sv[0][0] > 0.0 ? (int)si[0][0] : 0;
        o.v2 = sv[1][0] > 0.0 ? sv[1][0] : 0.0;
        o.c2 = sv[1][0] > 0.0 ? (int)si[1][0] : 0;
        o.knot2 = total ? knot[o.c2] : -1;
        o.row2 = 0;
        o.pad1 = o.pad2 = 0;
        *out = o;
    }
}

/* ---------------------------------------------------------------------- */
/* Host side                                                              */
/* ---------------------------------------------------------------------- */
void ctx_base_init(mars_hip_ctx *ctx, size_t n, size_t max_terms, double tol)
{
    require(n > 0, "empty dataset");
    require(max_terms >= 1, "max_terms must be >= 1");
    MH_CHECK(hipGetDevice(&ctx->device));
    MH_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    MH_CHECK(hipEventCreate(&ctx->ev0));
    MH_CHECK(hipEventCreate(&ctx->ev1));
    ctx->n = n;
    ctx->max_terms = max_terms;
    ctx->ldb = (max_terms + 3) / 4 * 4;
    ctx->tol = tol;
    ctx->B.ensure(n * ctx->ldb);
    ctx->Bo.ensure(n * ctx->ldb);
    MH_CHECK(hipMemsetAsync(ctx->B.p, 0, n * ctx->ldb * sizeof(float), ctx->stream));
    MH_CHECK(hipMemsetAsync(ctx->Bo.p, 0, n * ctx->ldb * sizeof(float), ctx->stream));
    ctx->y.ensure(n);
    ctx->ybo.ensure(max_terms);
    ctx->scal.ensure(S_COUNT);
    ctx->red.ensure(4 * kRedBlocks);
    ctx->dgks.ensure(1);
    ctx->fired_count.ensure(2);
    MH_CHECK(hipMemsetAsync(ctx->dgks.p, 0, sizeof(long long), ctx->stream));
    MH_CHECK(hipMemsetAsync(ctx->scal.p, 0, S_COUNT * sizeof(double), ctx->stream));
}

void ctx_free(mars_hip_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    DBuf<float> *fb[] = {&ctx->X, &ctx->s, &ctx->B, &ctx->Bo, &ctx->y, &ctx->delta, &ctx->rbuf};
    for (auto *b : fb) b->release();
    DBuf<double> *db[] = {&ctx->ybo, &ctx->tpart, &ctx->T, &ctx->tnorm, &ctx->spart, &ctx->ssum,
                          &ctx->ybx, &ctx->lin, &ctx->hin, &ctx->red, &ctx->vcol, &ctx->scal,
                          &ctx->dense, &ctx->hstate, &ctx->hbest, &ctx->mixed};
    for (auto *b : db) b->release();
    DBuf<int> *ib[] = {&ctx->order, &ctx->fired_list, &ctx->fired_count, &ctx->knot, &ctx->krow,
                       &ctx->hstate_i, &ctx->hbest_i, &ctx->cand.xc, &ctx->cand.par};
    for (auto *b : ib) b->release();
    ctx->dgks.release();
    ctx->aout.release();
    ctx->bout.release();
    for (auto &r : ctx->pend) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : ctx->evpool) (void)hipEventDestroy(e);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

void fit_create(mars_hip_ctx *ctx, const float *X, size_t ldX, const float *y, const float *w,
                size_t n, size_t px, size_t max_terms)
{
    require(X && y && w, "null input array");
    require(ldX >= n, "invalid leading dimension for X");
    require(px >= 1, "X has no columns");
    ctx_base_init(ctx, n, max_terms, (n * 0.02) * 2.220446049250313e-16);
    ctx->px = px;
    ctx->resident = true;
    ctx->X.ensure(n * px);
    ctx->s.ensure(px);
    hipStream_t st = ctx->stream;
    MH_CHECK(hipMemcpy2DAsync(ctx->X.p, n * sizeof(float), X, ldX * sizeof(float), n * sizeof(float),
                              px, hipMemcpyHostToDevice, st));
    DBuf<float> yw;
    yw.ensure(2 * n);
    MH_CHECK(hipMemcpyAsync(yw.p, y, n * sizeof(float), hipMemcpyHostToDevice, st));
    MH_CHECK(hipMemcpyAsync(yw.p + n, w, n * sizeof(float), hipMemcpyHostToDevice, st));
    DBuf<double> tmp;
    tmp.ensure(2 * n);
    double *yd = tmp.p, *sw = tmp.p + n;
    double *part = ctx->red.p, *part2 = ctx->red.p + 2 * kRedBlocks;
    hipLaunchKernelGGL(k_init1, dim3(kRedBlocks), dim3(256), 0, st, yw.p, yw.p + n, n, yd, sw, part);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_init2, dim3(kRedBlocks), dim3(256), 0, st, yd, sw, n, part, ctx->y.p,
                       ctx->B.p, ctx->Bo.p, ctx->ldb, part2, ctx->scal.p);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_init3, dim3(kRedBlocks), dim3(256), 0, st, yd, n, part2, part, ctx->scal.p);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_init4, dim3(1), dim3(256), 0, st, part, n, ctx->scal.p, ctx->ybo.p);
    MH_LAUNCH_CHECK();
    DBuf<double> cpart;
    cpart.ensure(px * 64);
    hipLaunchKernelGGL(k_colsq, dim3(64, px), dim3(256), 0, st, ctx->X.p, n, n, cpart.p);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_colscale, dim3((px + 255) / 256), dim3(256), 0, st, cpart.p, (int)px, n,
                       ctx->s.p, ctx->scal.p);
    MH_LAUNCH_CHECK();
    double sc[S_COUNT];
    MH_CHECK(hipMemcpyAsync(sc, ctx->scal.p, sizeof(sc), hipMemcpyDeviceToHost, st));
    MH_CHECK(hipStreamSynchronize(st));
    yw.release();
    tmp.release();
    cpart.release();
    require(sc[S_YN] > 0.0 && sc[S_WN] > 0.0, "target Y is all zero or NANs");
    require(sc[S_BADX] == 0.0, "not all columns in X are finite");
    ctx->yvar = sc[S_YVAR];
    ctx->ybo_h.assign(1, sc[S_YBO0]);
    ctx->m = 1;
    ctx->sorted.assign(px, 0);
}

void fit_append(mars_hip_ctx *ctx, char type, int xcol, int bcol, float h, double *mse_out)
{
    require(ctx->resident, "context was not created by mars_hip_fit_create");
    if (ctx->m >= ctx->max_terms) throw Error(MARS_HIP_EFULL, "basis matrix is full");
    if (bcol < 0 || (size_t)bcol >= ctx->m) {
        char msg[80];
        snprintf(msg, sizeof(msg), "invalid basis column number: %d", bcol);
        throw Error(MARS_HIP_EINVAL, msg);
    }
    require(xcol >= 0 && (size_t)xcol < ctx->px, "invalid X column index");
    int t;
    if (type == 'l') t = 0;
    else if (type == '+') t = 1;
    else if (type == '-') t = 2;
    else throw Error(MARS_HIP_EINVAL, "invalid basis type");

    const size_t n = ctx->n, m = ctx->m;
    hipStream_t st = ctx->stream;
    ctx->vcol.ensure(n);
    ctx->aout.ensure(1);
    ctx->knot.ensure(std::max<size_t>(ctx->knot.cap, 1));
    double *pa = ctx->red.p, *pb = ctx->red.p + kRedBlocks, *pc = ctx->red.p + 2 * kRedBlocks;
    double *pp[2] = {pa, pb};
    int *gate = ctx->fired_count.p;  // slot 0: DGKS gate of this append

    {
        KTimer kt(ctx, K_APPEND, 0.0, 8.0 * n * (3.0 * (m + 1) + 4));
        hipLaunchKernelGGL(k_app_build, dim3(kRedBlocks), dim3(256), 0, st, ctx->X.p, n, ctx->s.p,
                           ctx->B.p, ctx->ldb, xcol, bcol, t, h, n, ctx->vcol.p, pa, ctx->scal.p);
        MH_LAUNCH_CHECK();
        for (int pass = 0; pass < 2; ++pass) {
            for (size_t j = 0; j <= m; ++j) {
                hipLaunchKernelGGL(k_app_mgs, dim3(kRedBlocks), dim3(256), 0, st, ctx->Bo.p, ctx->B.p,
                                   ctx->ldb, (int)m, n, ctx->vcol.p, pp[j % 2], pp[(j + 1) % 2], (int)j,
                                   (int)m, pass == 0 ? 1 : 0, pass == 0 ? nullptr : gate, ctx->scal.p);
                MH_LAUNCH_CHECK();
            }
            hipLaunchKernelGGL(k_app_gate, dim3(1), dim3(256), 0, st, pp[(m + 1) % 2], ctx->tol, gate,
                               ctx->scal.p, ctx->dgks.p, pass);
            MH_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_app_store, dim3(kRedBlocks), dim3(256), 0, st, ctx->Bo.p, ctx->ldb,
                           (int)m, n, ctx->vcol.p, ctx->y.p, ctx->tol, ctx->scal.p, pc);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_app_final, dim3(1), dim3(256), 0, st, pc, (int)m, n, ctx->tol,
                           ctx->scal.p, ctx->ybo.p, gate, ctx->aout.p);
        MH_LAUNCH_CHECK();
    }
    AppendOut o;
    MH_CHECK(hipMemcpyAsync(&o, ctx->aout.p, sizeof(o), hipMemcpyDeviceToHost, st));
    MH_CHECK(hipStreamSynchronize(st));
    resolve_timing(ctx);
    if (o.fired) ctx->dgks_host += 1;
    if (o.status == 1) {
        ctx->ybo_h.push_back(o.ybo_m);
        ctx->m += 1;
        *mse_out = o.mse;
    } else {
        *mse_out = -1.0;
    }
}

double host_dsse(const mars_hip_ctx *ctx)
{
    double s = 0.0;
    for (double v : ctx->ybo_h) s = std::fma(v, v, s);
    return s;
}

void fit_eval(mars_hip_ctx *ctx, const unsigned char *mask, size_t rows, size_t stride,
              int min_span, int end_span, int linear_only, double *lin_out, double *hin_out,
              double *cut_out, mars_hip_best *best)
{
    require(ctx->resident, "context was not created by mars_hip_fit_create");
    require(mask != nullptr, "null mask");
    require(rows == ctx->px, "invalid dimension for mask");
    require(stride >= ctx->m, "invalid mask row stride");
    require(min_span >= 1, "min_span must be >= 1");
    require(end_span >= 0 && (size_t)end_span < ctx->n, "end_span must be in [0, n)");
    const size_t m = ctx->m, total = rows * m;
    hipStream_t st = ctx->stream;

    // candidates in row-major (xcol, parent) order
    std::vector<int> all_xc, all_par;
    for (size_t i = 0; i < rows; ++i)
        for (size_t j = 0; j < m; ++j)
            if (mask[i * stride + j]) {
                all_xc.push_back((int)i);
                all_par.push_back((int)j);
            }

    ctx->dense.ensure(3 * std::max<size_t>(total, 1));
    ctx->hstate_i.ensure(std::max<size_t>(total, 1));
    ctx->bout.ensure(1);
    double *dlin = ctx->dense.p, *dhin = dlin + total, *dcut = dhin + total;
    int *dknot = ctx->hstate_i.p;
    if (total) {
        hipLaunchKernelGGL(k_dense_init, dim3((total + 255) / 256), dim3(256), 0, st, dlin, dhin,
                           dcut, dknot, total);
        MH_LAUNCH_CHECK();
    }
    MH_CHECK(hipEventRecord(ctx->ev0, st));
    const XSrc xs{ctx->X.p, ctx->n, ctx->s.p};
    const size_t cap = block_capacity(ctx);
    size_t a = 0;
    while (a < all_xc.size()) {
        // whole X columns per block
        size_t b = a;
        while (b < all_xc.size()) {
            size_t e = b;
            while (e < all_xc.size() && all_xc[e] == all_xc[b]) ++e;
            if (b > a && e - a > cap) break;
            b = e;
        }
        ctx->cand.h_xc.assign(all_xc.begin() + a, all_xc.begin() + b);
        ctx->cand.h_par.assign(all_par.begin() + a, all_par.begin() + b);
        upload_cands(ctx);
        const int P = (int)(b - a);
        if (ctx->form == MARS_HIP_FORM_MIXED) {
            mixed_orthonormalize(ctx, xs, P);
            require(linear_only, "the mixed-precision form currently scores linear_only fits");
        } else {
            PlainPlan pl = plain_orthonormalize(ctx, xs, P, !linear_only);
            if (!linear_only) hinge_sweep(ctx, xs, pl, min_span, end_span);
        }
        hipLaunchKernelGGL(k_scatter, dim3((P + 255) / 256), dim3(256), 0, st, ctx->cand.xc.p,
                           ctx->cand.par.p, P, (int)m, ctx->lin.p, ctx->hin.p, ctx->knot.p,
                           ctx->krow.p, ctx->X.p, ctx->n, linear_only ? 0 : 1, dlin, dhin, dcut,
                           dknot);
        MH_LAUNCH_CHECK();
        a = b;
    }
    hipLaunchKernelGGL(k_argmax, dim3(1), dim3(1024), 0, st, dlin, dhin, dknot, dcut, total,
                       ctx->bout.p);
    MH_LAUNCH_CHECK();
    MH_CHECK(hipEventRecord(ctx->ev1, st));
    BestOut bo;
    MH_CHECK(hipMemcpyAsync(&bo, ctx->bout.p, sizeof(bo), hipMemcpyDeviceToHost, st));
    if (lin_out && total) MH_CHECK(hipMemcpyAsync(lin_out, dlin, total * 8, hipMemcpyDeviceToHost, st));
    if (hin_out && total) MH_CHECK(hipMemcpyAsync(hin_out, dhin, total * 8, hipMemcpyDeviceToHost, st));
    if (cut_out && total) MH_CHECK(hipMemcpyAsync(cut_out, dcut, total * 8, hipMemcpyDeviceToHost, st));
    double cut2 = NAN;
    if (total) MH_CHECK(hipMemcpyAsync(&cut2, dcut + bo.c2, 8, hipMemcpyDeviceToHost, st));
    MH_CHECK(hipStreamSynchronize(st));
    float ms = 0.f;
    MH_CHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->eval_ms += ms;
    resolve_timing(ctx);
    if (best) {
        best->sse0 = host_dsse(ctx);
        best->sse1 = bo.v1;
        best->xcol1 = m ? bo.c1 / (int)m : 0;
        best->bcol1 = m ? bo.c1 % (int)m : 0;
        best->sse2 = bo.v2;
        best->xcol2 = m ? bo.c2 / (int)m : 0;
        best->bcol2 = m ? bo.c2 % (int)m : 0;
        best->knot2 = bo.knot2;
        best->cut2 = cut2;
    }
}

/* ---------------------------------------------------------------------- */
/* Mirror mode (the reference's mars::cuda boundary)                       */
/* ---------------------------------------------------------------------- */
__global__ void k_put_bcol(float *__restrict__ dst, size_t ldb, int col, const float *__restrict__ src, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i * ldb + col] = src[i];
}

void mirror_sync_basis(mars_hip_ctx *ctx, size_t m, const float *B, size_t ldB, const float *Bo,
                       size_t ldBo)
{
    require(!ctx->resident, "resident contexts own their basis");
    require(m <= ctx->max_terms, "m exceeds max_terms");
    require(m == 0 || (B && Bo), "null basis pointer");
    const size_t n = ctx->n;
    hipStream_t st = ctx->stream;
    if (m > ctx->m_synced) {
        DBuf<float> col;
        col.ensure(n);
        std::vector<float> h(n);
        for (size_t j = ctx->m_synced; j < m; ++j) {
            MH_CHECK(hipMemcpyAsync(col.p, B + j * ldB, n * sizeof(float), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_put_bcol, dim3((n + 255) / 256), dim3(256), 0, st, ctx->B.p, ctx->ldb,
                               (int)j, col.p, n);
            MH_LAUNCH_CHECK();
            MH_CHECK(hipStreamSynchronize(st));
            for (size_t i = 0; i < n; ++i) h[i] = Bo[i * ldBo + j];
            MH_CHECK(hipMemcpyAsync(col.p, h.data(), n * sizeof(float), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_put_bcol, dim3((n + 255) / 256), dim3(256), 0, st, ctx->Bo.p, ctx->ldb,
                               (int)j, col.p, n);
            MH_LAUNCH_CHECK();
            MH_CHECK(hipStreamSynchronize(st));
        }
        col.release();
        ctx->m_synced = m;
    }
    ctx->m = m;
}

void mirror_set_target(mars_hip_ctx *ctx, const float *y)
{
    require(y != nullptr, "null target");
    if (ctx->y_key == y) return;
    MH_CHECK(hipMemcpyAsync(ctx->y.p, y, ctx->n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    MH_CHECK(hipStreamSynchronize(ctx->stream));
    ctx->y_key = y;
}

void mirror_sync_xcols(mars_hip_ctx *ctx, size_t px, const float *X, size_t ldX, const float *s,
                       const int *cols, size_t ncols)
{
    require(X && s, "null X or scale pointer");
    const size_t n = ctx->n;
    if (ctx->px == 0) {
        ctx->px = px;
        ctx->X.ensure(n * px + n);  // + one scratch column for single-column calls
        ctx->s.ensure(px + 1);
        ctx->xres.assign(px, 0);
    }
    require(px == ctx->px, "p_x changed between calls");
    MH_CHECK(hipMemcpyAsync(ctx->s.p, s, px * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    for (size_t k = 0; k < ncols; ++k) {
        const int c = cols[k];
        require(c >= 0 && (size_t)c < px, "X column out of range");
        if (ctx->xres[c]) continue;
        MH_CHECK(hipMemcpyAsync(ctx->X.p + (size_t)c * n, X + (size_t)c * ldX, n * sizeof(float),
                                hipMemcpyHostToDevice, ctx->stream));
        ctx->xres[c] = 1;
    }
    MH_CHECK(hipStreamSynchronize(ctx->stream));
}

__global__ void k_gather_cols(const float *__restrict__ rbuf, size_t ldr, int p, size_t n,
                              float *__restrict__ out, size_t ldo)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < n) out[(size_t)j * ldo + i] = rbuf[i * ldr + j];
}

void mirror_orthonormalize(mars_hip_ctx *ctx, size_t m, size_t p, const float *x, const int *mask,
                           float *Bx, size_t ldBx, double *ybx, long *fired)
{
    require(!ctx->resident, "use mars_hip_fit_eval on resident contexts");
    require(m == ctx->m && m <= ctx->m_synced, "basis not synced up to m");
    require(p <= ctx->max_terms, "p exceeds max_terms");
    require(x && mask, "null x or mask");
    require(Bx == nullptr || ldBx >= ctx->n, "invalid leading dimension for Bx");
    if (p == 0) return;
    const size_t n = ctx->n;
    hipStream_t st = ctx->stream;
    DBuf<float> xs1;
    xs1.ensure(n + 1);
    const float one = 1.0f;
    MH_CHECK(hipMemcpyAsync(xs1.p, x, n * sizeof(float), hipMemcpyHostToDevice, st));
    MH_CHECK(hipMemcpyAsync(xs1.p + n, &one, sizeof(float), hipMemcpyHostToDevice, st));
    ctx->cand.h_xc.assign(p, 0);
    ctx->cand.h_par.assign(mask, mask + p);
    for (int v : ctx->cand.h_par) require(v >= 0 && (size_t)v < m, "mask entry out of range");
    upload_cands(ctx);
    long long d0 = 0;
    MH_CHECK(hipMemcpyAsync(&d0, ctx->dgks.p, sizeof(d0), hipMemcpyDeviceToHost, st));
    const XSrc src{xs1.p, n, xs1.p + n};
    PlainPlan pl = plain_orthonormalize(ctx, src, (int)p, Bx != nullptr);
    if (Bx) {
        DBuf<float> out;
        out.ensure(n * p);
        hipLaunchKernelGGL(k_gather_cols, dim3((n + 255) / 256, p), dim3(256), 0, st, ctx->rbuf.p,
                           pl.ldr, (int)p, n, out.p, n);
        MH_LAUNCH_CHECK();
        MH_CHECK(hipMemcpy2DAsync(Bx, ldBx * sizeof(float), out.p, n * sizeof(float), n * sizeof(float),
                                  p, hipMemcpyDeviceToHost, st));
        MH_CHECK(hipStreamSynchronize(st));
        out.release();
    }
    if (ybx) MH_CHECK(hipMemcpyAsync(ybx, ctx->ybx.p, p * sizeof(double), hipMemcpyDeviceToHost, st));
    long long d1 = 0;
    MH_CHECK(hipMemcpyAsync(&d1, ctx->dgks.p, sizeof(d1), hipMemcpyDeviceToHost, st));
    MH_CHECK(hipStreamSynchronize(st));
    xs1.release();
    if (fired) *fired += (long)(d1 - d0);
}

void mirror_orthonormalize_batch(mars_hip_ctx *ctx, size_t m, size_t P, const int *src_xcol,
                                 const int *src_basis, double *ybx, long *fired)
{
    require(!ctx->resident, "use mars_hip_fit_eval on resident contexts");
    require(m == ctx->m && m <= ctx->m_synced, "basis not synced up to m");
    require(P <= block_capacity(ctx), "P exceeds the batch capacity");
    require(ctx->px > 0, "no X columns resident (call mars_hip_context_sync_xcols)");
    if (P == 0) return;
    require(src_xcol && src_basis && ybx, "null argument");
    ctx->cand.h_xc.assign(src_xcol, src_xcol + P);
    ctx->cand.h_par.assign(src_basis, src_basis + P);
    for (size_t j = 0; j < P; ++j) {
        require(src_xcol[j] >= 0 && (size_t)src_xcol[j] < ctx->px && ctx->xres[src_xcol[j]],
                "src_xcol not resident");
        require(src_basis[j] >= 0 && (size_t)src_basis[j] < m, "src_basis out of range");
    }
    upload_cands(ctx);
    hipStream_t st = ctx->stream;
    long long d0 = 0, d1 = 0;
    MH_CHECK(hipMemcpyAsync(&d0, ctx->dgks.p, sizeof(d0), hipMemcpyDeviceToHost, st));
    const XSrc src{ctx->X.p, ctx->n, ctx->s.p};
    plain_orthonormalize(ctx, src, (int)P, false);
    MH_CHECK(hipMemcpyAsync(ybx, ctx->ybx.p, P * sizeof(double), hipMemcpyDeviceToHost, st));
    MH_CHECK(hipMemcpyAsync(&d1, ctx->dgks.p, sizeof(d1), hipMemcpyDeviceToHost, st));
    MH_CHECK(hipStreamSynchronize(st));
    if (fired) *fired += (long)(d1 - d0);
}

template <typename F>
int guarded(F &&f)
{
    try {
        f();
        g_last_error.clear();
        return MARS_HIP_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return MARS_HIP_EDEVICE;
    } catch (...) {
        g_last_error = "unknown error";
        return MARS_HIP_EDEVICE;
    }
}

} // namespace
} // namespace mh

/* ---------------------------------------------------------------------- */
/* C ABI                                                                  */
/* ---------------------------------------------------------------------- */
using namespace mh;

extern "C" {

const char *mars_hip_last_error(void) { return g_last_error.c_str(); }

int mars_hip_context_create(size_t n, size_t max_terms, double tol, mars_hip_ctx **out)
{
    return guarded([&] {
        require(out != nullptr, "null output pointer");
        *out = nullptr;
        auto *ctx = new mars_hip_ctx;
        try {
            ctx_base_init(ctx, n, max_terms, tol);
            MH_CHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            ctx_free(ctx);
            throw;
        }
        *out = ctx;
    });
}

int mars_hip_context_destroy(mars_hip_ctx *ctx)
{
    return guarded([&] { ctx_free(ctx); });
}

int mars_hip_context_batch_capacity(mars_hip_ctx *ctx, size_t *capacity)
{
    return guarded([&] {
        require(ctx && capacity, "null argument");
        *capacity = block_capacity(ctx);
    });
}

int mars_hip_context_sync_basis(mars_hip_ctx *ctx, size_t m, const float *B, size_t ldB,
                                const float *Bo, size_t ldBo)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        mirror_sync_basis(ctx, m, B, ldB, Bo, ldBo);
    });
}

int mars_hip_context_set_target(mars_hip_ctx *ctx, const float *y)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        mirror_set_target(ctx, y);
    });
}

int mars_hip_context_sync_xcols(mars_hip_ctx *ctx, size_t p_x, const float *X, size_t ldX,
                                const float *s, const int *cols, size_t ncols)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        require(!ctx->resident, "resident contexts own their X");
        require(ldX >= ctx->n, "invalid leading dimension for X");
        mirror_sync_xcols(ctx, p_x, X, ldX, s, cols, ncols);
    });
}

int mars_hip_orthonormalize(mars_hip_ctx *ctx, size_t m, size_t p, const float *x,
                            const int *mask, float *Bx, size_t ldBx, double *ybx,
                            long *dgks_fired)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        mirror_orthonormalize(ctx, m, p, x, mask, Bx, ldBx, ybx, dgks_fired);
    });
}

int mars_hip_orthonormalize_batch(mars_hip_ctx *ctx, size_t m, size_t P, const int *src_xcol,
                                  const int *src_basis, double *ybx, long *dgks_fired)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        mirror_orthonormalize_batch(ctx, m, P, src_xcol, src_basis, ybx, dgks_fired);
    });
}

int mars_hip_fit_create(const float *X, size_t ldX, const float *y, const float *w, size_t n,
                        size_t p_x, size_t max_terms, int form, mars_hip_ctx **out)
{
    return guarded([&] {
        require(out != nullptr, "null output pointer");
        require(form == MARS_HIP_FORM_PLAIN || form == MARS_HIP_FORM_MIXED, "unknown precision form");
        *out = nullptr;
        auto *ctx = new mars_hip_ctx;
        try {
            fit_create(ctx, X, ldX, y, w, n, p_x, max_terms);
            ctx->form = form;
        } catch (...) {
            ctx_free(ctx);
            throw;
        }
        *out = ctx;
    });
}

int mars_hip_fit_info(mars_hip_ctx *ctx, int *nbasis, double *yvar, double *dsse)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        if (nbasis) *nbasis = (int)ctx->m;
        if (yvar) *yvar = ctx->yvar;
        if (dsse) *dsse = host_dsse(ctx);
    });
}

int mars_hip_fit_set_form(mars_hip_ctx *ctx, int form)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        require(form == MARS_HIP_FORM_PLAIN || form == MARS_HIP_FORM_MIXED, "unknown precision form");
        ctx->form = form;
    });
}

int mars_hip_fit_eval(mars_hip_ctx *ctx, const unsigned char *mask, size_t mask_rows,
                      size_t mask_stride, int min_span, int end_span, int linear_only,
                      double *linear, double *hinge, double *cut, mars_hip_best *best)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        fit_eval(ctx, mask, mask_rows, mask_stride, min_span, end_span, linear_only, linear,
                 hinge, cut, best);
    });
}

int mars_hip_fit_append(mars_hip_ctx *ctx, char type, int xcol, int bcol, float h, double *mse)
{
    return guarded([&] {
        require(ctx != nullptr && mse != nullptr, "null argument");
        fit_append(ctx, type, xcol, bcol, h, mse);
    });
}

int mars_hip_fit_basis(mars_hip_ctx *ctx, float *B, float *Bo, size_t *m_out)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        const size_t n = ctx->n, m = ctx->m, ldb = ctx->ldb;
        if (m_out) *m_out = m;
        std::vector<float> h(n * ldb);
        float *dst[2] = {B, Bo};
        const float *src[2] = {ctx->B.p, ctx->Bo.p};
        for (int q = 0; q < 2; ++q) {
            if (!dst[q]) continue;
            MH_CHECK(hipMemcpyAsync(h.data(), src[q], n * ldb * sizeof(float), hipMemcpyDeviceToHost,
                                    ctx->stream));
            MH_CHECK(hipStreamSynchronize(ctx->stream));
            for (size_t i = 0; i < n; ++i)
                for (size_t j = 0; j < m; ++j) dst[q][i * m + j] = h[i * ldb + j];
        }
    });
}

int mars_hip_timing(mars_hip_ctx *ctx, int enable)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        MH_CHECK(hipStreamSynchronize(ctx->stream));
        resolve_timing(ctx);
        ctx->timing = enable != 0;
    });
}

int mars_hip_kernel_stats(mars_hip_ctx *ctx, mars_hip_kstat *out, int capacity, int *count)
{
    return guarded([&] {
        require(ctx != nullptr && count != nullptr, "null argument");
        MH_CHECK(hipStreamSynchronize(ctx->stream));
        resolve_timing(ctx);
        int k = 0;
        for (int c = 0; c < K_NCLASS; ++c) {
            if (ctx->kacc[c].launches == 0) continue;
            if (out && k < capacity) {
                mars_hip_kstat &r = out[k];
                std::memset(r.name, 0, sizeof(r.name));
                std::strncpy(r.name, kClassName[c], sizeof(r.name) - 1);
                r.launches = ctx->kacc[c].launches;
                r.ms = ctx->kacc[c].ms;
                r.flops = ctx->kacc[c].flops;
                r.bytes = ctx->kacc[c].bytes;
            }
            ++k;
        }
        *count = k;
    });
}

int mars_hip_stats(mars_hip_ctx *ctx, long *dgks_fired, double *eval_ms)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        long long d = 0;
        MH_CHECK(hipMemcpyAsync(&d, ctx->dgks.p, sizeof(d), hipMemcpyDeviceToHost, ctx->stream));
        MH_CHECK(hipStreamSynchronize(ctx->stream));
        if (dgks_fired) *dgks_fired = (long)d;
        if (eval_ms) *eval_ms = ctx->eval_ms;
    });
}

} // extern "C"
