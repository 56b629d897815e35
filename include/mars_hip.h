/*
 *  mars_hip.h -- C ABI of the MI355X (gfx950) forward-pass library libmars_hip.so.
 *
 *  Plain C types only (no HIP, torch or C++ types), so the reference's
 *  C++ host (marsalgo.cc / marslib.cc) or any FFI can bind it directly.
 *
 *  Two ways to drive it:
 *
 *  1. Mirror mode -- drop-in for the reference's GPU boundary
 *     (cuda/mars_cuda.h:43-154).  The caller owns the fit state on the host
 *     (B col-major, Bo row-major, the scaled candidate columns) and the
 *     library keeps device copies in sync, exactly as the reference's
 *     mars::cuda::* functions do.  Each mars_hip_* function below names the
 *     mars::cuda function it replaces.
 *
 *  2. Resident mode -- the whole fit state lives in HBM.  mars_hip_fit_create
 *     performs the MarsAlgo constructor on the device (marsalgo.cc:266-370),
 *     mars_hip_fit_eval scores every (X column, parent basis) candidate --
 *     orthonormalization, linear delta-SSE, hinge sweep, device argmax -- and
 *     mars_hip_fit_append adds a basis function on the device
 *     (marsalgo.cc:760-875).  Only the (p_x, m) result arrays and a few
 *     scalars cross PCIe per forward step.
 *
 *  Every function returns MARS_HIP_OK (0) or a negative status; the message
 *  of the last failure on the calling thread is mars_hip_last_error().
 *  Calls on one context are synchronous and not thread-safe (one host thread
 *  per context), like the reference (mars_cuda.h:26-33).
 */
#ifndef MARS_HIP_H
#define MARS_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MARS_HIP_OK 0
#define MARS_HIP_EINVAL (-1)   /* bad argument / shape / state            */
#define MARS_HIP_EDEVICE (-2)  /* HIP runtime error (allocation, launch)  */
#define MARS_HIP_EFULL (-3)    /* basis matrix is full                    */

/* Precision forms of the candidate scoring. */
#define MARS_HIP_FORM_PLAIN 0  /* all FP64 arithmetic on f32 storage, CPU rounding points */
#define MARS_HIP_FORM_MIXED 1  /* linear fits: FP32-MFMA incremental projection cache; hinge fits: FP32-MFMA
                                  first projection; then exact FP64 rescoring of every candidate that can win */
#define MARS_HIP_FORM_MIXED_BF16 2  /* the mixed form with the hinge path's first T-GEMM on BF16-split
                                       hi/lo products (three BF16 MFMAs per product, FP32 sums) */

typedef struct mars_hip_ctx mars_hip_ctx;

/* Winner of one forward step, as chosen on the device (mars.py:342-343). */
typedef struct mars_hip_best {
    double sse1;   /* best linear delta-SSE (first maximum in row-major (xcol, bcol) order) */
    int xcol1, bcol1;
    double sse2;   /* best hinge-pair delta-SSE */
    int xcol2, bcol2;
    int knot2;     /* sorted position of the winning knot (-1 if none) */
    double cut2;   /* raw X value at that knot (NaN if none) */
    double sse0;   /* delta-SSE of the current basis (sum of ybo^2) */
} mars_hip_best;

/* Text of the last error on this thread ("" if none). */
const char *mars_hip_last_error(void);

/* ---------------------------------------------------------------------- */
/*  Mirror mode: the reference's mars::cuda API                           */
/* ---------------------------------------------------------------------- */

/* replaces mars::cuda::context_create (cuda/mars_cuda.h:43) */
int mars_hip_context_create(size_t n, size_t max_terms, double tol, mars_hip_ctx **out);

/* replaces mars::cuda::context_destroy (cuda/mars_cuda.h:48); NULL is a no-op */
int mars_hip_context_destroy(mars_hip_ctx *ctx);

/* replaces mars::cuda::context_batch_capacity (cuda/mars_cuda.h:55) */
int mars_hip_context_batch_capacity(mars_hip_ctx *ctx, size_t *capacity);

/* replaces mars::cuda::context_sync_basis (cuda/mars_cuda.h:71):
 * B (n, *) col-major ld ldB; Bo (n, m) row-major ld ldBo; uploads new columns only. */
int mars_hip_context_sync_basis(mars_hip_ctx *ctx, size_t m, const float *B, size_t ldB,
                                const float *Bo, size_t ldBo);

/* replaces mars::cuda::context_set_target (cuda/mars_cuda.h:80): the normalized f32 target */
int mars_hip_context_set_target(mars_hip_ctx *ctx, const float *y);

/* replaces mars::cuda::context_sync_xcols (cuda/mars_cuda.h:95):
 * make columns `cols` of X (n, p_x) col-major ld ldX, scaled by s, resident. */
int mars_hip_context_sync_xcols(mars_hip_ctx *ctx, size_t p_x, const float *X, size_t ldX,
                                const float *s, const int *cols, size_t ncols);

/* replaces mars::cuda::orthonormalize (cuda/mars_cuda.h:127): x (n) scaled candidate,
 * mask (p) parent columns; Bx (n, p) col-major ld ldBx [out, may be NULL];
 * ybx (p) [out, may be NULL]; dgks_fired [in/out, may be NULL] += retries. */
int mars_hip_orthonormalize(mars_hip_ctx *ctx, size_t m, size_t p, const float *x,
                            const int *mask, float *Bx, size_t ldBx, double *ybx,
                            long *dgks_fired);

/* replaces mars::cuda::orthonormalize_batch (cuda/mars_cuda.h:152): P candidates
 * (src_xcol[j], src_basis[j]) over resident columns; ybx (P) [out]. */
int mars_hip_orthonormalize_batch(mars_hip_ctx *ctx, size_t m, size_t P, const int *src_xcol,
                                  const int *src_basis, double *ybx, long *dgks_fired);

/* ---------------------------------------------------------------------- */
/*  Resident mode: MarsAlgo on the device                                 */
/* ---------------------------------------------------------------------- */

/* Device counterpart of MarsAlgo::MarsAlgo (marsalgo.cc:266-370) and
 * new_algo (marslib.cc:12-37): X (n, p_x) col-major ld ldX, y (n), w (n).
 * X, y, w are copied; the caller need not keep them.  `form` is a
 * MARS_HIP_FORM_* value. */
int mars_hip_fit_create(const float *X, size_t ldX, const float *y, const float *w, size_t n,
                        size_t p_x, size_t max_terms, int form, mars_hip_ctx **out);

/* MarsAlgo::nbasis / yvar / dsse (marsalgo.cc:381-400) */
int mars_hip_fit_info(mars_hip_ctx *ctx, int *nbasis, double *yvar, double *dsse);

/* Select the precision form for subsequent evals. */
int mars_hip_fit_set_form(mars_hip_ctx *ctx, int form);

/* Score all candidates of one forward step -- the device counterpart of the
 * binding's eval (marslib.cc:41-148) / MarsAlgo::eval (marsalgo.cc:402-662)
 * for every X column at once, plus the argmax (mars.py:342-343).
 *   mask        : (mask_rows, m) bool, row stride mask_stride bytes, m = nbasis
 *   linear/hinge/cut : (mask_rows, m) row-major [out, each may be NULL]
 *   best        : device-chosen winners [out, may be NULL]            */
int mars_hip_fit_eval(mars_hip_ctx *ctx, const unsigned char *mask, size_t mask_rows,
                      size_t mask_stride, int min_span, int end_span, int linear_only,
                      double *linear, double *hinge, double *cut, mars_hip_best *best);

/* Device counterpart of MarsAlgo::append (marsalgo.cc:760-875): type 'l',
 * '+' or '-'.  *mse receives the mse after the append, or -1 when the new
 * column is numerically degenerate (the basis is then unchanged). */
int mars_hip_fit_append(mars_hip_ctx *ctx, char type, int xcol, int bcol, float h, double *mse);

/* mars_hip_fit_append without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device or pinned memory during a call
 * (code under stream capture, serving loops, harnesses that forbid it).  Same
 * arguments, checks, result and basis as mars_hip_fit_append (the same
 * kernels), on a single-fit context from mars_hip_fit_create.  Its scratch is
 * one fixed device array of 4 MiB per GPU, held until the process ends: it
 * supports the append's limits (basis columns < 512) when
 * 16 * ldb * min(256, ceil(n / 4096)) + 12288 <= 4 MiB, ldb = max_terms
 * rounded up to 4; larger fits get MARS_HIP_EINVAL.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (several host threads or
 * contexts at once) race and must be serialised by the caller.  Synchronous:
 * returns after the result has been copied to *mse. */
int mars_hip_fit_append_noalloc(mars_hip_ctx *ctx, char type, int xcol, int bcol, float h, double *mse);

/* Forward-pass options of mars.fit (mars.py:190-247) for mars_hip_fit_forward, resolved by the
 * caller as mars.fit resolves them: max_terms after the dof clamp (mars.py:241-243), min_span
 * after Friedman's rule (mars.py:225-228), end_span = max(int(n * tail_span), 1) (mars.py:247). */
typedef struct mars_hip_forward_opts {
    int max_terms;
    int max_degree;
    double penalty;
    int min_span;
    int end_span;
    int self_interact;
    int linear_only;
    double n_true;
    int r2_window;
    double r2_thresh;
    double max_runtime;  /* seconds */
} mars_hip_forward_opts;

/* One model row (mars.py:261-275, 366-375). */
typedef struct mars_hip_model_row {
    int type;       /* 'i', 'l', '+' or '-' */
    int basis;
    int input;
    int order;
    double hinge;   /* knot (NaN for 'i' and 'l') */
    double mse;     /* append mse (normalized units); 0 for the intercept */
    float r2;
    float r2_cv;
    float time;     /* seconds since the call started, at this term's step */
    int pad;
} mars_hip_model_row;

/* One forward step: the chosen candidate, its scores, the appends' mse values. */
typedef struct mars_hip_step {
    int kind;       /* 'l' (linear term) or 'h' (hinge pair) */
    int xcol, bcol;
    int knot;       /* sorted position of the knot (-1 for 'l') */
    int n_cand;     /* candidates scored */
    int n_mse;      /* appends attempted (1 or 2; fewer when the term budget ran out) */
    double cut;     /* knot value (NaN for 'l') */
    double sse0, sse1, sse2;
    double mse[2];  /* each append's mse (-1: degenerate column, not added) */
} mars_hip_step;

/* The whole forward pass of a fresh single-fit context (nbasis 1), driven in
 * C++: mars.fit's loop (mars.py:285-406) with Fast-MARS ranking off (its
 * default, max_basis = max_inputs = None) and no per-candidate filter -- per
 * step the degree / self-interaction mask, mars_hip_fit_eval's scoring and
 * argmax, the GCV choice between the linear term and the hinge pair, the
 * appends and the stopping rules (float32 model fields compared in float32, as
 * numpy 2 compares them).  rows: at least opts->max_terms entries; *nrows rows
 * written (the intercept first).  steps (may be NULL): up to steps_capacity
 * step records; *nsteps (may be NULL) receives the step count. */
int mars_hip_fit_forward(mars_hip_ctx *ctx, const mars_hip_forward_opts *opts, mars_hip_model_row *rows,
                         int *nrows, mars_hip_step *steps, int steps_capacity, int *nsteps);

/* Copy the live basis out: B and Bo as (n, m) row-major f32 (either may be NULL). */
int mars_hip_fit_basis(mars_hip_ctx *ctx, float *B, float *Bo, size_t *m_out);

/* Cumulative DGKS retries (orthonormalize + append) and device milliseconds
 * spent in the scoring kernels since the context was created. */
int mars_hip_stats(mars_hip_ctx *ctx, long *dgks_fired, double *eval_ms);

/* Device memory held by the library's buffers in this process (all contexts; blocks
 * replaced but not yet freed included), and the most held since the last reset.
 * reset_peak != 0 sets the peak to the current amount after reading it.  No
 * counterpart in the reference (its context sizes scratch once, mars_cuda.h:43):
 * lets a caller check the footprint of a workload. */
int mars_hip_device_bytes(size_t *live, size_t *peak, int reset_peak);

/* ---------------------------------------------------------------------- */
/*  Batching form: many independent linear_only fits advanced together    */
/* ---------------------------------------------------------------------- */

/* Forward-pass options of mars.fit (mars.py:190-214) that the batch driver honours. */
typedef struct mars_hip_batch_opts {
    int max_degree;     /* default 3 */
    double penalty;     /* default 3.0 */
    int self_interact;  /* default 0 */
    int r2_window;      /* default 32 */
    double r2_thresh;   /* default 3e-5 */
} mars_hip_batch_opts;

/* One model row (mars.py:261-275); type 'i' (intercept) or 'l'. */
typedef struct mars_hip_term {
    int type;
    int basis;
    int input;
    int order;
    double mse;     /* append mse (normalized units); 0 for the intercept */
    float r2;
    float r2_cv;
} mars_hip_term;

/* Create `nfits` resident fits: fit i has X[i] (n[i], p_x[i]) col-major ld ldX[i],
 * y[i], w[i] (w[i] may be NULL for unit weights). */
int mars_hip_batch_create(int nfits, const float *const *X, const size_t *ldX, const float *const *y,
                          const float *const *w, const size_t *n, const size_t *p_x, size_t max_terms,
                          int form, mars_hip_ctx **out);

/* Run every fit's linear_only forward pass to completion, all fits advancing
 * one term per step together (mars.py:285-406 restated in C++): one batched
 * scoring pass, per-fit device argmax and one batched append per step.
 * terms: (nfits, max_terms) rows; nterms[i]: rows of fit i; steps: step count. */
int mars_hip_batch_fit(mars_hip_ctx *ctx, const mars_hip_batch_opts *opts, mars_hip_term *terms, int *nterms,
                       int *steps);

/* Candidates scored since creation (every form); those the mixed form rescored
 * exactly in FP64, and how many of them because their DGKS gate tripped in the
 * bulk pass (both 0 for the plain form).  Any pointer may be NULL. */
int mars_hip_score_stats(mars_hip_ctx *ctx, long *scored, long *rescored, long *gated);

/* Checks and launch counts since creation: window_violations counts mixed-form
 * candidates left unrescored whose score interval reached their fit's best
 * exact score (0 while the interval bounds hold: then every step's winner is an
 * FP64 value); host_redo_steps counts batch steps whose device-driven scoring
 * ran out of residual room and were scored again with host-built lists;
 * graph_captures / graph_launches count the batch step's scoring graphs
 * captured and launched (MARS_HIP_GRAPH=0 turns them off).  No counterpart in
 * the reference.  Any pointer may be NULL. */
int mars_hip_check_stats(mars_hip_ctx *ctx, long *window_violations, long *host_redo_steps, long *graph_captures,
                         long *graph_launches);

/* Launch shapes of a context, tunable per workload (mars_amd/autotune.py writes
 * the table the Python layer applies).  None of them changes a result beyond
 * the summation order of FP64 sums.  No counterpart in the reference (its
 * tile and budget knobs are environment variables, orthonormalize.cu:476-501). */
typedef struct mars_hip_tuning {
    double pass_items;    /* work items per FP64 T-GEMM / projection pass (default 32768) */
    double retry_items;   /* work items per DGKS-retry pass (default 4096) */
    double update_waves;  /* waves per mixed-form cache update (default 16384) */
    double sweep_waves;   /* waves per hinge-sweep pass (default 65536) */
    long sweep_min_rows;  /* sorted rows per hinge-sweep chunk, at least (default 512) */
    long append_rows;     /* rows per append block; 0: 8192 from 64 fits on, else 4096 */
    int narrow_avg;       /* 16-candidate tiles when fits average at most this many (default 24) */
    int pad;
} mars_hip_tuning;

int mars_hip_default_tuning(mars_hip_tuning *out);
int mars_hip_get_tuning(mars_hip_ctx *ctx, mars_hip_tuning *out);
/* Validated: pass_items, retry_items, update_waves, sweep_waves in [256, 1e6];
 * sweep_min_rows in [32, 1 << 20]; append_rows 0 or a multiple of 256 in
 * [256, 1 << 20]; narrow_avg in [0, 4096]. */
int mars_hip_set_tuning(mars_hip_ctx *ctx, const mars_hip_tuning *t);

/* Per-kernel-class timing with HIP events on the context's stream (off by
 * default).  Each record: launches, summed device milliseconds and the
 * summed algorithmic flops / HBM bytes of those launches.  The mixed form's
 * exact rescoring with device-built lists ("mixed_recheck") reports the work
 * counted on the device since the context was created: 4 n m flops per
 * rescored candidate and per DGKS retry (one T-GEMM and one projection pass
 * each); bytes: each candidate's own columns over its passes and its fit's Bo
 * columns once per pass. */
typedef struct mars_hip_kstat {
    char name[24];
    long launches;
    double ms;
    double flops;
    double bytes;
} mars_hip_kstat;

int mars_hip_timing(mars_hip_ctx *ctx, int enable);
int mars_hip_kernel_stats(mars_hip_ctx *ctx, mars_hip_kstat *out, int capacity, int *count);

/* ---------------------------------------------------------------------- */
/* After the forward pass (stateless; host pointers, synchronous).         */
/* ---------------------------------------------------------------------- */

/* Basis matrix of a model: replaces mars.expand (mars.py:423-450).  X (n, p_x)
 * and B (n, M) column-major; term i has type[i] in 'i' (B_i = 1), 'l'
 * (B_basis * X_input), '+' (B_basis * max(X_input - hinge, 0)), '-'
 * (B_basis * max(hinge - X_input, 0)), with basis[i] < i for every non-'i'
 * term.  '+' / '-' form the hinge and its product in f64 and round once. */
int mars_hip_expand(const float *X, size_t ldX, size_t n, size_t p_x, size_t M, const char *type, const int *basis,
                    const int *input, const double *hinge, float *B, size_t ldB);
int mars_hip_expand_f64(const double *X, size_t ldX, size_t n, size_t p_x, size_t M, const char *type,
                        const int *basis, const int *input, const double *hinge, double *B, size_t ldB);

/* out (ma, mb) column-major = A^T B with f64 accumulation: replaces mars.gram
 * (mars.py:456-490).  A (n, ma), B (n, mb) column-major; B == NULL gives
 * A^T A (exactly symmetric).  f32 inputs: exact f64 products. */
int mars_hip_gram(const float *A, size_t ldA, size_t n, size_t ma, const float *B, size_t ldB, size_t mb,
                  double *out, size_t ldOut);
int mars_hip_gram_f64(const double *A, size_t ldA, size_t n, size_t ma, const double *B, size_t ldB, size_t mb,
                      double *out, size_t ldOut);

/* Backward stepwise selection and coefficient solve of the backward pass:
 * replaces the host part of mars.prune (mars.py:493-698: _init_inverse_state,
 * _backward_eliminate and the final lstsq).  Input: the raw Gram products of
 * the (weight-folded) basis, XX = B^T B (M x M, row stride ldXX), XY = B^T y,
 * YY = y^T y, as mars_hip_gram computes them on the device; mask (M) the
 * candidate terms (NULL: all).  Columns are standardised to a unit Gram
 * diagonal (zero-norm columns are dropped), `ridge` loads that diagonal, the
 * intercept (column 0) is never removed.  Output: beta (M) in the original
 * scale, 0 for pruned terms; chosen (M) [may be NULL] the kept subset.
 * Host-only (no device work), synchronous. */
int mars_hip_prune_gram(const double *XX, size_t ldXX, const double *XY, double YY, size_t M,
                        const unsigned char *mask, double penalty, double n_true, double ridge,
                        double *beta, unsigned char *chosen);

#ifdef __cplusplus
}
#endif

#endif /* MARS_HIP_H */
