// Shared definitions of libmars_hip: context state, error plumbing, device helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mars_hip.h"

namespace mh {

/* Error carrying a MARS_HIP_* status. */
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void require(bool ok, const char *msg)
{
    if (!ok) throw Error(MARS_HIP_EINVAL, msg);
}

#define MH_CHECK(expr)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::mh::Error(MARS_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define MH_LAUNCH_CHECK() MH_CHECK(hipGetLastError())

/* DGKS gate: retry when residual energy * 9 < projected energy (kernels.h:17). */
constexpr double kGate = 9.0;

/* Device buffer that only grows. */
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t cap = 0;
    void ensure(size_t count)
    {
        if (count <= cap) return;
        if (p) MH_CHECK(hipFree(p));
        p = nullptr;
        cap = 0;
        MH_CHECK(hipMalloc((void **)&p, sizeof(T) * (count ? count : 1)));
        cap = count;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

/* Per-step candidate block: candidates c in [0, P), grouped by X column. */
struct Cands {
    DBuf<int> xc, par;   // X column and parent basis column per candidate
    std::vector<int> h_xc, h_par;
};

/* Result slot returned by the device append. */
struct AppendOut {
    double mse;
    double ybo_m;
    int status;  // 1 accepted, 0 degenerate
    int fired;
};

/* Kernel classes timed by mars_hip_timing. */
enum KClass { K_TGEMM = 0, K_PROJECT, K_FINALIZE, K_DGKS, K_SWEEP, K_SORT, K_APPEND, K_MIXED_T,
              K_MIXED_P, K_BATCH, K_NCLASS };
constexpr const char *kClassName[K_NCLASS] = {"tgemm_f64", "project_f64", "finalize", "dgks_retry",
                                             "hinge_sweep", "sort", "append", "tgemm_mixed",
                                             "project_mixed", "batch_step"};

struct KAcc {
    long launches = 0;
    double ms = 0, flops = 0, bytes = 0;
};

struct KRec {
    int cls;
    hipEvent_t a, b;
    double flops, bytes;
};

struct BestOut {
    double v1;
    int c1;
    int pad1;
    double v2;
    int c2;
    int knot2;
    int row2;
    int pad2;
};

} // namespace mh

struct mars_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    size_t n = 0;          // rows
    size_t max_terms = 0;  // basis capacity
    size_t ldb = 0;        // row stride of B and Bo (>= max_terms, multiple of 4)
    size_t px = 0;         // resident X columns
    double tol = 0;
    int form = MARS_HIP_FORM_PLAIN;

    // Resident fit state (row-major B / Bo, column-major raw X)
    mh::DBuf<float> X, s, B, Bo, y;
    mh::DBuf<double> ybo;
    std::vector<double> ybo_h;
    size_t m = 0;          // live basis count
    double yvar = 0;
    bool resident = false;

    // Mirror-mode bookkeeping
    size_t m_synced = 0;
    const float *y_key = nullptr;
    std::vector<char> xres;  // resident flag per X column

    // Hinge sort cache (per X column: stable descending order and f32 deltas)
    mh::DBuf<int> order;
    mh::DBuf<float> delta;
    std::vector<char> sorted;

    // Scoring workspace
    mh::Cands cand;
    mh::DBuf<double> tpart, T, tnorm, spart, ssum, ybx, lin, hin;
    mh::DBuf<float> rbuf;
    mh::DBuf<int> fired_list, fired_count, knot, krow;
    mh::DBuf<long long> dgks;       // device DGKS counter
    mh::DBuf<double> red;           // reduction scratch
    mh::DBuf<double> vcol;          // append: f64 residual column
    mh::DBuf<double> scal;          // append / init scalars
    mh::DBuf<mh::AppendOut> aout;
    mh::DBuf<mh::BestOut> bout;
    mh::DBuf<double> dense;         // (p_x, m) x 3 result arrays
    mh::DBuf<int> hstate_i;
    mh::DBuf<double> hstate;        // hinge chunk summaries / carries
    mh::DBuf<double> hbest;
    mh::DBuf<int> hbest_i;
    mh::DBuf<double> mixed;         // mixed-form scratch

    long dgks_host = 0;
    double eval_ms = 0;

    // optional per-kernel timing
    bool timing = false;
    std::vector<mh::KRec> pend;
    std::vector<hipEvent_t> evpool;
    mh::KAcc kacc[mh::K_NCLASS];
};

namespace mh {

/* Brackets the launches of one kernel class with events when timing is on. */
struct KTimer {
    mars_hip_ctx *ctx;
    KRec rec;
    bool on;
    KTimer(mars_hip_ctx *c, int cls, double flops, double bytes) : ctx(c), on(c->timing)
    {
        if (!on) return;
        rec.cls = cls;
        rec.flops = flops;
        rec.bytes = bytes;
        rec.a = take();
        rec.b = take();
        MH_CHECK(hipEventRecord(rec.a, ctx->stream));
    }
    ~KTimer() noexcept(false)
    {
        if (!on) return;
        MH_CHECK(hipEventRecord(rec.b, ctx->stream));
        ctx->pend.push_back(rec);
    }
    hipEvent_t take()
    {
        if (!ctx->evpool.empty()) {
            hipEvent_t e = ctx->evpool.back();
            ctx->evpool.pop_back();
            return e;
        }
        hipEvent_t e;
        MH_CHECK(hipEventCreate(&e));
        return e;
    }
};

/* Fold completed timing records into the per-class sums (stream must be idle). */
inline void resolve_timing(mars_hip_ctx *ctx)
{
    for (auto &r : ctx->pend) {
        float ms = 0.f;
        MH_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
        KAcc &a = ctx->kacc[r.cls];
        a.launches += 1;
        a.ms += ms;
        a.flops += r.flops;
        a.bytes += r.bytes;
        ctx->evpool.push_back(r.a);
        ctx->evpool.push_back(r.b);
    }
    ctx->pend.clear();
}

} // namespace mh
