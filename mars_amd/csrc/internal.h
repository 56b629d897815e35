// Host-side entry points shared between the libmars_hip translation units.
#pragma once

#include "common.h"

#include <exception>
#include <functional>
#include <map>

namespace mh {

/* Work list of one batched pass plus, per candidate, where its partial sums live. */
struct Pass {
    std::vector<Item> items;
    std::vector<int> cmap;  // per candidate (or list position): first item, item count, slot
};

/*
 * Score the candidate columns `cs` (all fits) with the reference's FP64
 * arithmetic and f32 rounding points (kernels.cc:283-417,
 * marsalgo.cc:500-505).  Writes per candidate c: ybx[c] and lin[c] = ybx^2
 * (linear delta-SSE); with keep_bxn, the normalized f32 column stays in its
 * fit's rbuf (row-major, column c - c0).  `xover`/`sover`, when set, replace
 * fit 0's X and scales (mirror-mode single-column calls).  `rescore` (a few
 * candidates per fit, the mixed form's exact rescoring): the Sum r^2 pass
 * keeps the residuals (column-major), so the DGKS retry and the final pass
 * read r instead of recomputing it from Bo; the final pass runs while the
 * DGKS flags come back, and the fired candidates get it again after retry.
 * `keep` (device, one int per candidate, rescoring only): 0 = that candidate's
 * residual is not stored, because it most likely ends with s <= tol and is
 * never read; one that ends with s > tol after all is recomputed (before its
 * DGKS retry, or by a final pass over the late candidates).
 */
void plain_score(mars_hip_ctx *ctx, Cands &cs, bool keep_bxn, double *lin, double *ybx,
                 const float *xover = nullptr, const float *sover = nullptr, bool rescore = false,
                 const int *keep = nullptr, bool f32 = false);

/* FP64 scoring with device-built lists and work items (no host synchronisation).  With `list`:
 * the candidates cs[list[i]], i < *dcount (bound: upper bound on the count), lin[list[i]] = score,
 * done[list[i]] = 1 (done may be null), keep[i]: store entry i's first residual.  Without: every
 * candidate of cs (bound = its count), lin[c] / ybx[c], keep null.  *redo = 1 when a fired entry
 * found no arena room (the step must then be scored again with host-built lists).  narrow:
 * 16-candidate tiles. */
void plain_score_device(mars_hip_ctx *ctx, const Cands &cs, const int *list, const int *dcount, int bound,
                        int *keep, double *lin, double *ybx, int *done, int *redo, bool narrow, int phase = 0);
/* phase 1: size the buffers and upload the fit views only; phase 2: the kernel launches only (no
 * allocation, upload or synchronisation: they can be captured into a graph); 0: both. */

/* Run `launches` (kernel launches on the context's stream, nothing else) through a graph
 * captured the first time `key` is seen and replayed afterwards; directly when graphs are off.
 * The key must name every kernel argument the launches take (pointers, sizes, variants). */
void run_graph(mars_hip_ctx *ctx, const std::vector<unsigned long long> &key, const std::function<void()> &launches);
/* Key words for run_graph: the addresses of the context's scoring buffers; the launch shapes. */
void graph_key_buffers(mars_hip_ctx *ctx, std::vector<unsigned long long> &key);
void tuning_key(const mars_hip_ctx *ctx, std::vector<unsigned long long> &key);
/* True when the per-fit counts favour 16-candidate tiles. */
bool narrow_tiles(const mars_hip_ctx *ctx, const std::vector<int> &cnt);

/* Mixed-precision scoring of ctx->cand (linear delta-SSE only) from each
 * fit's incrementally maintained T cache (FP32 MFMA, FP64 across row chunks),
 * then an exact FP64 rescoring of the DGKS-gated and near-winning
 * candidates.  Writes ctx->lin. */
void mixed_score(mars_hip_ctx *ctx, bool host_lists = false, const int *dP = nullptr, int Pb = 0);
/* Launch the mixed form's cache update for every fit's newest basis column; it needs no candidate
 * list, so a caller can build the list while the update runs (mixed_score then finds it done). */
void mixed_prepare(mars_hip_ctx *ctx);
/* Allocate the mixed form's T cache of every fit that has none yet (at creation for the mixed
 * form, else at its first use). */
void ensure_tcache(mars_hip_ctx *ctx);
/* Size the buffers indexed by candidate or work item for the context's largest step (every
 * fit's X columns times max_terms - 1 parents, capped at 4M candidates), so that a forward pass
 * does not grow them step by step; called when a resident context is created. */
void reserve_step_buffers(mars_hip_ctx *ctx);

/* Hinge sweep of fit 0's candidates `cs`, whose normalized columns sit in its rbuf (column c),
 * scores lin / ybx in, hinge score, knot position and knot row per candidate out.  With hub:
 * also, per candidate, linear + an upper bound of every knot's score when each covariate carries
 * a relative error up to eps (the precision form's screen).  dpar: device parents of cs (default cs.d_par).
 * sorted: the columns are ctx->hr's, in each X column's sort order, and the sweep streams the
 * sorted replicas instead of gathering rows. */
void hinge_sweep_list(mars_hip_ctx *ctx, Cands &cs, const double *lin, const double *ybx, double *hin, int *knot,
                      int *krow, int min_span, int end_span, double eps = 0.0, double *hub = nullptr,
                      const int *dpar = nullptr, bool sorted = false);
/* Hinge eval of fit 0's candidates `cs` (grouped by X column): projection (FP32 first pass when
 * f32) and sweep, lin / ybx / hin / knot / krow per candidate out, ctx->flag and ctx->ssum per
 * candidate as plain_score leaves them.  Runs on the per-column sorted replicas when they fit
 * the device, else as plain_score + hinge_sweep_list over row-major columns. */
void hinge_score(mars_hip_ctx *ctx, Cands &cs, double *lin, double *ybx, double *hin, int *knot, int *krow,
                 int min_span, int end_span, bool f32 = false, double eps = 0.0, double *hub = nullptr);

/* Upload the per-fit device views (current m, rbuf, first candidate of `cs`). */
void upload_fits(mars_hip_ctx *ctx, const Cands &cs, const float *xover = nullptr,
                 const float *sover = nullptr, const std::vector<size_t> *arena_off = nullptr);

/* Upload the candidate arrays of `cs`. */
void upload_cands(mars_hip_ctx *ctx, Cands &cs);

/* Fill cs.fit_c0 / fit_P from cs.fit (candidates grouped by fit). */
void index_cands(const mars_hip_ctx *ctx, Cands &cs);

/* Make sure each fit's residual buffer can hold its candidates of `cs`. */
void ensure_rbuf(mars_hip_ctx *ctx, const Cands &cs);

} // namespace mh

namespace mh {

/* fit.hip */
void ctx_base_init(mars_hip_ctx *ctx, size_t max_terms);
void ctx_free(mars_hip_ctx *ctx);
void fit_init(mars_hip_ctx *ctx, FitState &fs, const float *X, size_t ldX, const float *y, const float *w,
              size_t n, size_t px);
/* Append one basis function to every fit with args[f].active.  `before_readback`, when set,
 * is called after the append's kernels and the copy of its results are queued and before the
 * host waits for those results: work it queues overlaps the host's reading of them.
 * `device_args`: the arguments are already in ctx->aarg on the device (queued before this call);
 * `args` then only marks the fits that may be active (it sizes the launch), and on return holds
 * the device's arguments, read back with the results (with `best`, when given, the per-fit
 * argmax ctx->bout as well). */
void fit_append_batch(mars_hip_ctx *ctx, std::vector<AppendArg> &args, std::vector<AppendOut> &out,
                      const std::function<void()> &before_readback = {}, bool device_args = false,
                      std::vector<BestOut> *best = nullptr);
/* Per-fit argmax of `lin` over each fit's candidates (first maximum in candidate order), queued
 * on the stream: ctx->bout receives it, ctx->ival holds the fits' candidate offsets and counts. */
void fit_eval(mars_hip_ctx *ctx, const unsigned char *mask, size_t rows, size_t stride, int min_span,
              int end_span, int linear_only, double *lin_out, double *hin_out, double *cut_out,
              mars_hip_best *best);
void argmax_launch(mars_hip_ctx *ctx, const Cands &cs, const double *lin);
/* The same over ranges already on the device (ctx->ival: first candidate, count per fit). */
void argmax_launch_ranges(mars_hip_ctx *ctx, const double *lin, size_t bound);
double host_dsse(const FitState &fs);
/* MARS_HIP_FORM_* checks; set_form maps MIXED_BF16 to the mixed form with BF16-split products. */
bool valid_form(int form);
void set_form(mars_hip_ctx *ctx, int form);
/* Copy the context's device counters (kDStat entries) to `out` (synchronises the stream). */
void read_dstat(mars_hip_ctx *ctx, unsigned long long *out);
/* Device address of the no-alloc append's scratch on the current device (resolved once). */
char *scratch_ptr();
extern thread_local std::string g_last_error;

/* Asynchronous host-to-device copy on the context's stream.  Small copies go through pinned
 * staging, so the call does not wait for the stream (a pageable copy does); the source may be
 * reused as soon as it returns. */
void h2d(mars_hip_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Deferred small upload: staged now, sent by the next h2d_flush (or ctx_sync) together with the
 * other pending ones (one scatter kernel reading pinned memory instead of a copy each).  The caller
 * flushes before launching any kernel that reads `dst`, and does not reallocate `dst` first. */
void h2d_defer(mars_hip_ctx *ctx, void *dst, const void *src, size_t bytes);
void h2d_flush(mars_hip_ctx *ctx);
/* Held by every function that defers uploads: an exception leaving it drops what is still
 * pending, so no later flush writes into a buffer reallocated in the meantime. */
struct PendingScope {
    mars_hip_ctx *ctx;
    int ex;
    explicit PendingScope(mars_hip_ctx *c) : ctx(c), ex(std::uncaught_exceptions()) {}
    ~PendingScope()
    {
        if (std::uncaught_exceptions() > ex) ctx->up_pend.n = 0;
    }
};
/* hipStreamSynchronize on the context's stream; the staging space is free again afterwards. */
void ctx_sync(mars_hip_ctx *ctx);

/* Run f, turning exceptions into a status code and the thread's last error text. */
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

} // namespace mh
