// Shared definitions of libmars_hip: fit / context state, error plumbing, timing.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstddef>
#include <map>
#include <cstdint>
#include <algorithm>
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

/* Buffers replaced while queued kernels may still read them.  hipFree waits for the whole
 * device, so they are freed at the next ctx_sync, after the stream has drained. */
void retire_device(void *p, size_t bytes);
void retire_host(void *p);
void free_retired();

/* Device bytes held by this library's buffers in the process (retired blocks included until
 * they are freed), and the most held at any time since the last reset. */
void dev_account(long long delta);

/* Device buffer that only grows: the first allocation is exact, a regrowth takes at least
 * 2x (lists that grow every step reallocate O(log) times) and retires the old block. */
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t cap = 0;
    void ensure(size_t count)
    {
        if (count <= cap) return;
        const size_t want = cap ? std::max(count, 2 * cap) : count;
        if (p) retire_device(p, sizeof(T) * (cap ? cap : 1));
        p = nullptr;
        cap = 0;
        MH_CHECK(hipMalloc((void **)&p, sizeof(T) * (want ? want : 1)));
        dev_account((long long)(sizeof(T) * (want ? want : 1)));
        cap = want;
    }
    void release()
    {
        if (p) {
            (void)hipFree(p);
            dev_account(-(long long)(sizeof(T) * (cap ? cap : 1)));
        }
        p = nullptr;
        cap = 0;
    }
};

/* Pinned host block for readbacks: grows 2x, and a replaced block is retired (freed at the next
 * synchronisation), since a copy into it may still be in flight. */
struct PinBuf {
    char *p = nullptr;
    size_t cap = 0;
    template <typename T>
    T *get(size_t count)
    {
        const size_t bytes = std::max<size_t>(count * sizeof(T), 1);
        if (bytes > cap) {
            const size_t want = std::max(bytes, 2 * cap);
            if (p) retire_host(p);
            p = nullptr;
            cap = 0;
            MH_CHECK(hipHostMalloc((void **)&p, want, hipHostMallocDefault));
            cap = want;
        }
        return reinterpret_cast<T *>(p);
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

/* Device view of one fit, read by every batched kernel. */
struct FitDev {
    const float *X;    // raw X, column-major (ldx)
    const float *s;    // per-column 1/rms scale
    const float *B;    // basis, row-major (ldb); null in batch contexts
    const float *Bo;   // orthonormal basis, row-major (ldb); null in batch contexts
    const float *BoC;  // the same basis, column-major (ldx)
    const float *BC;   // basis B, column-major (ldx)
    const float *y;    // normalized target (f32)
    const double *ybo; // Bo^T y per live column
    float *rbuf;       // residual / normalized candidate columns: element (row, col) at
                       // row * rsr + col * rsc (row-major, rsr = ldr, for the hinge sweep and
                       // column downloads; column-major, rsc = ldx, for linear scoring)
    double *v;         // append: f64 residual of the new column (n)
    double *tc;        // mixed form: cached T, [xc][par][k] (ldb x ldb per X column)
    double *tcn;       // mixed form: cached |Bx|^2, Bx.y per (xc, par)
    long long n, ldx, ldb, rsr, rsc;
    int m, c0;         // live basis count, first candidate of this fit
    int rbyl;          // how a candidate's rbuf column is found: 0 column cand - c0, 1 (list passes)
                       // position in the fit's part of the list, 2 float offset roff[cand]
    int px;            // X columns
    int act;           // batch driver: the fit takes part in this step (candidates are generated)
    int pad_;
    const long long *roff;  // rbyl == 2: per candidate, float offset of its column in rbuf (-1: none)
    double tol;
};

/* One fit's resident state. */
struct FitState {
    size_t n = 0, px = 0;
    size_t ldx = 0;   // device column stride of X (n rounded up to 32: aligned vector loads)
    double tol = 0, yvar = 0;
    size_t m = 0;
    std::vector<double> ybo_h;
    DBuf<float> X, s, B, Bo, BoC, BC, y, rbuf;
    DBuf<double> ybo, v;
    DBuf<double> G;   // append: Gram rows of the stored Bo columns (ldb x ldb)
    size_t ldr = 0;
    // mixed form: T cache valid for basis columns and parents < mc
    DBuf<double> tc, tcn;
    size_t mc = 0;
    // hinge sort cache
    DBuf<int> order;
    DBuf<float> delta;
    std::vector<char> sorted;
    // mirror-mode bookkeeping
    size_t m_synced = 0;
    const float *y_key = nullptr;
    std::vector<char> xres;
    // a view (the sorted hinge path's per-X-column fit): its buffers alias storage owned elsewhere
    bool view = false;
    void release()
    {
        if (view) {
            detach();
            return;
        }
        X.release(); s.release(); B.release(); Bo.release(); BoC.release(); BC.release(); y.release();
        rbuf.release();
        ybo.release(); v.release(); G.release(); order.release(); delta.release(); tc.release(); tcn.release();
    }
    void detach()
    {
        for (DBuf<float> *b : {&X, &s, &B, &Bo, &BoC, &BC, &y, &rbuf, &delta}) *b = DBuf<float>{};
        for (DBuf<double> *b : {&ybo, &v, &G, &tc, &tcn}) *b = DBuf<double>{};
        order = DBuf<int>{};
    }
};

/* Work item of the batched kernels: candidates [c0, c0+nc) of fit `fit`, rows [r0, r1).
 * pad: in the orthonormalization passes, the list position where the fit's candidates begin
 * (compact residual columns are numbered from there); in the cache update, its tile code. */
struct Item {
    int fit, c0, nc, pad;
    long long r0, r1;
};

/* Writable basis storage of one fit (append); B and Bo are null without row-major copies. */
struct FitW {
    float *B, *Bo, *BoC, *BC;
    double *ybo;
    double *G;  // Gram matrix of the stored Bo columns, G[j * ldb + l] for l < j
};

/* Deferred small uploads: each is staged in pinned host memory at `off`; one kernel reads the
 * pieces from there over the host link and scatters them to their destinations. */
struct UpDesc {
    void *dst;
    unsigned long long off, bytes;
};
constexpr int kUpMax = 16;  // pending uploads per flush (the table travels as a kernel argument)
struct UpTable {
    UpDesc d[kUpMax];
    int n = 0;
};

/* Result slot returned by the device append. */
struct AppendOut {
    double mse;
    double ybo_m;
    int status;  // 1 accepted, 0 degenerate
    int fired;
};

struct AppendArg {
    int type, xcol, bcol, active;
    float h;
    int pad;
};

/* Per-fit winner of one step (device argmax). */
struct BestOut {
    double v1;
    int c1;
    int pad1;
    double v2;
    int c2;
    int knot2;
};

constexpr int kDStat = 8;  // device counters per context (mars_hip_ctx::dstat)

/* Kernel classes timed by mars_hip_timing. */
enum KClass { K_TGEMM = 0, K_PROJECT, K_FINALIZE, K_DGKS, K_SWEEP, K_SORT, K_APPEND, K_MIXED_A,
              K_MIXED_B, K_RECHECK, K_ARGMAX, K_GRAPH_PLAIN, K_TGEMM32, K_PROJECT32, K_HRECHECK, K_TGEMM_BF16,
              // host wall-clock sections of a batch step (include the waits they contain)
              K_H_CANDS, K_H_SCORE, K_H_ARGMAX, K_H_APPEND, K_H_STEP, K_NCLASS };
constexpr const char *kClassName[K_NCLASS] = {"tgemm_f64", "project_f64", "finalize", "dgks_retry",
                                             "hinge_sweep", "sort", "append", "mixed_update",
                                             "mixed_score", "mixed_recheck", "argmax", "plain_graph",
                                             "tgemm_f32", "project_f32", "hinge_recheck", "tgemm_bf16x3",
                                             "host_cands", "host_score", "host_argmax", "host_append",
                                             "host_step"};

struct KAcc {
    long launches = 0;
    double ms = 0, flops = 0, bytes = 0;
};

struct KRec {
    int cls;
    hipEvent_t a, b;
    double flops, bytes;
};

/* Candidate list of one scoring call: candidates of all fits, grouped by fit,
 * then by X column, then parent. */
struct Cands {
    std::vector<int> fit, xc, par;     // host
    std::vector<int> fit_c0, fit_P;    // per fit: first candidate, count
    DBuf<int> d_fit, d_xc, d_par;      // device copies
};

} // namespace mh

struct mars_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_rb = nullptr;      // marks a readback into rb_pin as complete
    hipEvent_t ev_app = nullptr;     // marks the append results' readback as complete
    mh::AppendOut *aout_pin = nullptr;  // pinned host copy of the append results (aout_cap)
    size_t aout_cap = 0;
    int *rb_pin = nullptr;           // pinned host readback buffer (rb_cap ints)
    mh::PinBuf sel_pin, best_pin;    // readbacks of the rescoring list and of the per-fit argmax
    mh::PinBuf aarg_pin;             // readback of append arguments built on the device
    size_t rb_cap = 0;
    // pinned staging for host-to-device copies: bump-allocated, reset after a stream sync
    char *stage = nullptr;
    size_t stage_cap = 0, stage_off = 0;
    char *stage_dev = nullptr;       // device address of the staging space (deferred uploads)
    mh::UpTable up_pend;             // deferred uploads not yet sent

    size_t max_terms = 0;  // basis capacity (all fits)
    size_t ldb = 0;        // row stride of B and Bo (>= max_terms, multiple of 4)
    int form = MARS_HIP_FORM_PLAIN;
    bool split_bf16 = false;         // precision form: BF16-split products in the hinge path's first T-GEMM
    bool resident = false;
    bool row_major = true;  // keep row-major B and Bo (hinge sweep, basis download); batch contexts do not
    std::vector<mh::FitState> fits;

    // scoring workspace (shared by all fits)
    mh::Cands cand;
    mh::Cands rcand;                 // mixed form: candidates rescored exactly
    mh::DBuf<double> rlin, rybx;
    mh::DBuf<int> rsel;              // their positions in the scored list
    mh::DBuf<int> rkeep;             // 1: the candidate was ungated in the cache approximation,
                                     // so its residual is stored for the final pass
    bool rbuf_cm = false;            // candidate columns in rbuf are column-major (linear scoring)
    mh::DBuf<float> rarena;          // residual columns of the DGKS retry / device rescoring (bounded)
    mh::DBuf<long long> roff;        // device rescoring: per list entry, its column's offset in rarena
    mh::DBuf<mh::FitDev> d_fits2;    // device rescoring: fit views with arena-addressed residuals
    mh::DBuf<int> dl_int;            // device rescoring: per-fit ranges, counts, compaction tiles
    mh::DBuf<unsigned long long> dl_u64;  // device rescoring: arena bump counter
    mh::DBuf<int> cgate, cdone;      // mixed form, per candidate: bulk gate tripped; rescored exactly
    mh::DBuf<double> clo, chi;       // mixed form, per candidate: interval of the plain form's score
    mh::DBuf<unsigned long long> dstat;  // device counters (kDStat): [1] rescored, [2] gated,
                                         // [3] window violations, [4] DGKS retries of the rescoring,
                                         // [5], [6] algorithmic flops and HBM bytes of the device-driven
                                         // exact rescoring
    long redo_steps = 0;             // batch steps scored again on the host-driven path
    long hinge_rounds = 0;           // precision-form hinge evals that needed a further rescoring round
    long window_misses = 0;          // ... candidates still approximate after the last round
    mh::DBuf<int> dredo;             // 1: the step's device-driven rescoring must be redone on the host
    // batch driver, candidates generated on the device: per fit and basis function its degree and
    // the bit set of X columns it contains (words of 32 bits, kw per basis function)
    mh::DBuf<int> bdeg;
    mh::DBuf<unsigned> bhas;
    int kw = 1;
    std::vector<int> act_h;          // the active flags the last upload_fits sent
    mh::DBuf<int> dcand;             // batch driver: device count of the step's candidates
    mars_hip_tuning tune{32768.0, 4096.0, 16384.0, 65536.0, 512, 0, 24, 0};  // launch shapes
    bool use_graphs = false;         // batch driver: replay captured graphs of the scoring launches
    std::map<std::vector<unsigned long long>, hipGraphExec_t> graphs;  // by run_graph key
    long graph_captures = 0, graph_launches = 0;
    char *scratch = nullptr;         // device address of this device's no-alloc append scratch
    int rsel_hint = 0;               // the previous rescoring list's length
    mh::DBuf<mh::FitDev> d_fits;
    mh::DBuf<mh::Item> d_items;
    mh::DBuf<int> d_cmap;            // per-candidate first item / item count per pass
    mh::DBuf<mh::Item> d_ritems;     // DGKS retry passes (separate: the first passes' lists stay live)
    mh::DBuf<int> d_rcmap;
    mh::DBuf<mh::Item> d_litems;     // final pass over the candidates with s > tol
    mh::DBuf<int> d_lcmap, live_list;
    mh::DBuf<double> tpart, T, T2, tnorm, spart, ssum, ybx, lin, hin, aux;
    mh::DBuf<int> flag, fired_list, fired_count, knot, krow, ival;
    mh::DBuf<long long> dgks;        // device DGKS counter
    mh::DBuf<double> red;            // reduction scratch
    mh::DBuf<double> vcol;           // append: f64 residual columns
    mh::DBuf<double> adot, agram, acoef;  // append: partials of Q^T v and of a Gram row, MGS coefficients
    mh::DBuf<double> scal;           // append / init scalars
    mh::DBuf<mh::AppendOut> aout;
    mh::DBuf<mh::AppendArg> aarg;
    mh::DBuf<mh::FitW> d_fitw;       // per-fit writable basis pointers
    mh::DBuf<mh::BestOut> bout;
    mh::DBuf<double> dense;          // (p_x, m) x 3 result arrays
    mh::DBuf<int> dknot;
    mh::DBuf<double> hstate;         // hinge chunk summaries / carries
    mh::DBuf<double> hbest;
    mh::DBuf<int> hbest_i;
    mh::DBuf<double> hbest_ub;       // hinge chunk upper bounds (precision form)
    mh::DBuf<double> hub;            // per candidate: linear + upper bound of its best hinge score
    mh::DBuf<double> hbest2;         // precision-form hinge eval: exact hinge scores of the rescored
    mh::DBuf<int> rknot;             // ... and their knot positions and rows
    // sorted hinge path: per X column xc, the block rep + xc * rep_stride holds that column's
    // descending order of x, y, B[:, 0..rep_mcap) and Bo[:, 0..rep_mcap), column-major (ld = ldx,
    // rows past n zero); its first rep_m[xc] basis columns are current
    mh::DBuf<float> rep;
    size_t rep_mcap = 0, rep_stride = 0;
    std::vector<int> rep_m;
    int rep_state = 0;               // 0 undecided, 1 in use, -1 the gather sweep instead
    bool keep_cm = false;            // plain_score keeps normalised columns column-major
    mh::DBuf<float> hr;              // sorted hinge path: normalised candidate column c at c * ldx
    mh::Cands vcand;                 // sorted hinge path: the list over the per-column fit views
    mh::DBuf<int> rep_todo;          // replica extension list
    mh::DBuf<float> rep_stage;       // an append's new basis columns, row-major, for the replica gather

    long dgks_host = 0;
    double eval_ms = 0;
    long scored = 0, rescored = 0;  // candidates scored / rescored in FP64 by the mixed form
    long gated = 0;                 // mixed form: rescored because the bulk DGKS gate tripped

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

/* Host wall-clock time of a section, when timing is on. */
struct HostTimer {
    mars_hip_ctx *ctx;
    int cls;
    std::chrono::steady_clock::time_point t0;
    HostTimer(mars_hip_ctx *c, int k) : ctx(c), cls(k), t0(std::chrono::steady_clock::now()) {}
    ~HostTimer() { stop(); }
    /* Record now (once); cancel() drops the section instead. */
    void stop()
    {
        if (!ctx) return;
        if (ctx->timing) {
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            ctx->kacc[cls].launches += 1;
            ctx->kacc[cls].ms += ms;
        }
        ctx = nullptr;
    }
    void cancel() { ctx = nullptr; }
};

/* Fold completed timing records into the per-class sums (stream must be idle). */
/* Fold the recorded timers whose work has completed (they complete in recording order; work
 * queued behind the last synchronisation stays pending). */
inline void resolve_timing(mars_hip_ctx *ctx)
{
    size_t done = 0;
    for (auto &r : ctx->pend) {
        if (hipEventQuery(r.b) != hipSuccess) break;
        ++done;
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
    ctx->pend.erase(ctx->pend.begin(), ctx->pend.begin() + done);
}

} // namespace mh
