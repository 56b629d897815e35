// Batching form: many independent linear_only MARS fits advanced together.
//
// The forward-pass driver of mars.fit (mars.py:285-406) restated in C++ for
// the linear_only case with Fast-MARS off (every input and every parent is a
// candidate each step).  One step = one batched scoring pass over every
// active fit's candidates (plain FP64 or mixed form), one per-fit device
// argmax, the GCV choice, and one batched device append; fits stop on their
// own (no progress, R^2 threshold, GCV patience window, term budget).
#include "internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>

namespace mh {
namespace {

/* Append arguments from the device argmax: fit f appends the linear term of its best candidate
 * (X column, parent) when it had candidates and that candidate's score is above 0 (mars.py:345-346);
 * otherwise its argument is inactive.  When the step's rescoring must be redone (*redo), every
 * argument is inactive and marked (pad = 1): nothing is appended.  One thread per fit. */
__global__ void k_batch_args(int F, const int *__restrict__ c0, const int *__restrict__ cnt,
                             const BestOut *__restrict__ best, const int *__restrict__ cxc,
                             const int *__restrict__ cpar, const int *__restrict__ redo,
                             AppendArg *__restrict__ args)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    AppendArg a{0, 0, 0, 0, 0.0f, 0};
    if (redo && *redo) {
        a.pad = 1;
    } else if (cnt[f] > 0 && best[f].v1 > 0.0) {
        const int g = c0[f] + best[f].c1;
        a.xcol = cxc[g];
        a.bcol = cpar[g];
        a.active = 1;
    }
    args[f] = a;
}

struct FitDriver {
    size_t max_terms = 0;            // after the dof clamp (mars.py:241-243)
    std::vector<std::vector<int>> vars;  // variables of each basis function
    std::vector<unsigned char> has;  // has[b * p_x + i]: variable i occurs in basis function b
    std::vector<mars_hip_term> rows;
    bool active = true;
};

void batch_fit(mars_hip_ctx *ctx, const mars_hip_batch_opts &o, mars_hip_term *terms, int *nterms, int *steps)
{
    require(ctx->resident, "not a resident context");
    const size_t F = ctx->fits.size();
    std::vector<FitDriver> D(F);
    auto dof = [&](double m) { return m + o.penalty * (m - 1); };
    for (size_t f = 0; f < F; ++f) {
        const FitState &fs = ctx->fits[f];
        const double nt = (double)fs.n;
        size_t mt = 0;
        for (size_t t = 1; t <= ctx->max_terms; ++t)
            if (dof((double)t) < nt) mt = t;
        require(mt >= 1, "too few rows for any term");
        D[f].max_terms = mt;
        D[f].vars.assign(1, std::vector<int>());
        D[f].has.assign(fs.px, 0);
        mars_hip_term t0{'i', 0, 0, 0, 0.0, 0.0f, 0.0f};
        D[f].rows.assign(1, t0);
        D[f].active = fs.m < mt;
    }
    int step = 0;
    std::vector<AppendArg> args(F);
    std::vector<BestOut> best;
    bool device_lists = true;  // the mixed form's rescoring lists are built on the device
    while (true) {
        HostTimer step_timer(ctx, K_H_STEP);
        // the mixed form's cache update runs on the device while the host builds the candidates
        bool live = false;
        for (size_t f = 0; f < F; ++f) live = live || (D[f].active && ctx->fits[f].m < D[f].max_terms);
        if (live && ctx->form == MARS_HIP_FORM_MIXED) mixed_prepare(ctx);
        HostTimer cands_timer(ctx, K_H_CANDS);
        Cands &cs = ctx->cand;
        cs.fit.clear();
        cs.xc.clear();
        cs.par.clear();
        bool any = false;
        for (size_t f = 0; f < F; ++f) {
            FitDriver &d = D[f];
            if (d.active && ctx->fits[f].m >= d.max_terms) d.active = false;
            if (!d.active) continue;
            any = true;
            const FitState &fs = ctx->fits[f];
            const size_t px = fs.px;
            for (size_t i = 0; i < px; ++i)
                for (size_t b = 0; b < fs.m; ++b)
                    if ((int)d.vars[b].size() < o.max_degree && (o.self_interact || !d.has[b * px + i])) {
                        cs.fit.push_back((int)f);
                        cs.xc.push_back((int)i);
                        cs.par.push_back((int)b);
                    }
        }
        if (!any) {
            step_timer.cancel();
            cands_timer.cancel();
            break;
        }
        ++step;
        index_cands(ctx, cs);
        cands_timer.stop();
        const size_t P = cs.fit.size();
        ctx->scored += (long)P;
        ctx->lin.ensure(std::max<size_t>(P, 1));
        ctx->ybx.ensure(std::max<size_t>(P, 1));
        if (P) {
            HostTimer t(ctx, K_H_SCORE);
            if (ctx->form == MARS_HIP_FORM_MIXED)
                mixed_score(ctx, !device_lists);
            else
                plain_score(ctx, cs, false, ctx->lin.p, ctx->ybx.p);
            // no synchronisation here: the argmax that follows on the same stream returns the result
        }
        // the argmax and the append arguments stay on the device: the append is queued at once,
        // and the host reads the winners with the append's results
        {
            HostTimer t(ctx, K_H_ARGMAX);
            argmax_launch(ctx, cs, ctx->lin.p);
            ctx->aarg.ensure(F);  // before the kernel writes it: the append must not reallocate it
            hipLaunchKernelGGL(k_batch_args, dim3((unsigned)((F + 255) / 256)), dim3(256), 0, ctx->stream, (int)F,
                               (const int *)ctx->ival.p, (const int *)ctx->ival.p + F,
                               (const BestOut *)ctx->bout.p, (const int *)cs.d_xc.p, (const int *)cs.d_par.p,
                               device_lists ? (const int *)ctx->dredo.p : nullptr, ctx->aarg.p);
            MH_LAUNCH_CHECK();
        }
        std::vector<double> sse0(F, 0.0), m0(F, 0.0);
        for (size_t f = 0; f < F; ++f) {
            // fits that may append: the active ones with candidates (the device drops a fit whose
            // best score is not above 0)
            args[f] = AppendArg{0, 0, 0, (D[f].active && cs.fit_P[f] > 0) ? 1 : 0, 0.0f, 0};
            if (args[f].active) {
                sse0[f] = host_dsse(ctx->fits[f]);
                m0[f] = (double)ctx->fits[f].m;
            }
        }
        std::vector<AppendOut> out;
        {
            HostTimer t(ctx, K_H_APPEND);
            // mixed form: the next step's cache update for the columns being appended is queued
            // behind the append, assuming each succeeds; a degenerate append (rare) marks its
            // column uncached again below, and the entries computed for it are never read
            std::function<void()> ahead;
            if (ctx->form == MARS_HIP_FORM_MIXED)
                ahead = [&] {
                    // (not for a fit this append fills: it takes no further step)
                    std::vector<char> bump(F, 0);
                    for (size_t f = 0; f < F; ++f)
                        bump[f] = args[f].active && ctx->fits[f].m + 1 < D[f].max_terms;
                    for (size_t f = 0; f < F; ++f)
                        if (bump[f]) ctx->fits[f].m += 1;
                    mixed_prepare(ctx);
                    for (size_t f = 0; f < F; ++f)
                        if (bump[f]) ctx->fits[f].m -= 1;
                };
            fit_append_batch(ctx, args, out, ahead, true, &best);
            for (size_t f = 0; f < F; ++f) {
                FitState &fs = ctx->fits[f];
                if (fs.mc > fs.m) fs.mc = fs.m;
            }
        }
        if (F > 0 && args[0].pad == 1) {
            // the device-driven rescoring ran out of residual room: nothing was appended; score
            // this step again with host-built lists
            ctx->redo_steps += 1;
            --step;
            device_lists = false;
            continue;
        }
        device_lists = true;
        for (size_t f = 0; f < F; ++f) {
            FitDriver &d = D[f];
            if (!d.active) continue;
            if (!args[f].active) {  // no candidates, or all input data is filtered or zero (mars.py:345-346)
                d.active = false;
                continue;
            }
            const FitState &fs = ctx->fits[f];
            {
                // the GCV choice (mars.py:349-357): the linear term, as the fit is linear_only
                const double s1 = best[f].v1, n = (double)fs.n, m = m0[f];
                auto gcv = [&](double e, double mm) { return e / std::pow(1.0 - dof(mm) / n, 2); };
                const double mse1 = gcv((1.0 - sse0[f] - s1) / n, m + 1);
                const double mse2 = gcv((1.0 - sse0[f] - 0.0) / n, m + 2);
                require(mse1 <= mse2, "linear_only batch fit chose a hinge pair");
            }
            const double mse = out[f].mse;
            if (mse < 0) {  // degenerate append: no progress (mars.py:393-394)
                d.active = false;
                continue;
            }
            const double n = (double)fs.n;
            const double m = (double)(d.rows.size());
            const double gm = mse / std::pow(1.0 - dof(m + 1) / n, 2);
            std::vector<int> v = d.vars[args[f].bcol];
            v.push_back(args[f].xcol);
            mars_hip_term t;
            t.type = 'l';
            t.basis = args[f].bcol;
            t.input = args[f].xcol;
            t.order = (int)v.size();
            t.mse = mse;
            t.r2 = (float)(1.0 - mse / fs.yvar);
            t.r2_cv = (float)(1.0 - gm / fs.yvar);
            d.vars.push_back(v);
            d.has.resize(d.has.size() + fs.px, 0);
            for (int x : v) d.has[(d.vars.size() - 1) * fs.px + x] = 1;
            d.rows.push_back(t);
            // stopping rules (mars.py:395-406)
            // numpy compares the float32 model fields against Python floats in float32
            if (t.r2 > (float)(1.0 - o.r2_thresh)) {
                d.active = false;
                continue;
            }
            const size_t R = d.rows.size();
            if ((int)R > o.r2_window) {
                float all = -INFINITY, last = -INFINITY;
                for (size_t r = 0; r < R; ++r) {
                    all = std::max(all, d.rows[r].r2_cv);
                    if (r + o.r2_window >= R) last = std::max(last, d.rows[r].r2_cv);
                }
                if (last < all - (float)o.r2_thresh) d.active = false;
            }
        }
    }
    ctx_sync(ctx);  // work queued ahead (an update no step consumed) finishes; every timer resolves
    resolve_timing(ctx);
    for (size_t f = 0; f < F; ++f) {
        nterms[f] = (int)D[f].rows.size();
        for (size_t r = 0; r < D[f].rows.size(); ++r) terms[f * ctx->max_terms + r] = D[f].rows[r];
    }
    if (steps) *steps = step;
}

} // namespace
} // namespace mh

using namespace mh;

extern "C" {

int mars_hip_batch_create(int nfits, const float *const *X, const size_t *ldX, const float *const *y,
                          const float *const *w, const size_t *n, const size_t *p_x, size_t max_terms, int form,
                          mars_hip_ctx **out)
{
    return guarded([&] {
        require(out != nullptr, "null output pointer");
        require(nfits >= 1, "need at least one fit");
        require(X && ldX && y && n && p_x, "null argument");
        require(form == MARS_HIP_FORM_PLAIN || form == MARS_HIP_FORM_MIXED, "unknown precision form");
        *out = nullptr;
        auto *ctx = new mars_hip_ctx;
        try {
            ctx_base_init(ctx, max_terms);
            ctx->resident = true;
            ctx->row_major = false;  // the batch scores linear terms only: column-major copies suffice
            ctx->form = form;
            ctx->fits.resize(nfits);
            for (int f = 0; f < nfits; ++f) {
                std::vector<float> ones;
                const float *wf = w ? w[f] : nullptr;
                if (!wf) {
                    ones.assign(n[f], 1.0f);
                    wf = ones.data();
                }
                fit_init(ctx, ctx->fits[f], X[f], ldX[f], y[f], wf, n[f], p_x[f]);
            }
            if (form == MARS_HIP_FORM_MIXED) ensure_tcache(ctx);
            reserve_step_buffers(ctx);
        } catch (...) {
            ctx_free(ctx);
            throw;
        }
        *out = ctx;
    });
}

int mars_hip_batch_fit(mars_hip_ctx *ctx, const mars_hip_batch_opts *opts, mars_hip_term *terms, int *nterms,
                       int *steps)
{
    return guarded([&] {
        require(ctx && opts && terms && nterms, "null argument");
        require(opts->max_degree >= 1, "max_degree must be >= 1");
        batch_fit(ctx, *opts, terms, nterms, steps);
    });
}

} // extern "C"
