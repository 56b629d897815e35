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
#include <cstdlib>
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

/* ---------------------------------------------------------------------- */
/* Candidates generated on the device (the batch step's lists)             */
/* ---------------------------------------------------------------------- */

/* Parent b of a fit is a candidate for X column xc when its degree is below max_degree and, unless
 * self-interaction is allowed, xc is not among its variables (mars.py:319-323). */
__device__ inline bool cand_ok(const int *deg, const unsigned *has, int kw, int b, int xc, int maxdeg, int self)
{
    return deg[b] < maxdeg && (self || !((has[(size_t)b * kw + (xc >> 5)] >> (xc & 31)) & 1u));
}

/* Candidates per active fit; one block per fit. */
__global__ __launch_bounds__(256) void k_cand_count(const FitDev *__restrict__ fits, const int *__restrict__ bdeg,
                                                    const unsigned *__restrict__ bhas, int kw, int ldb, int maxdeg,
                                                    int self, int *__restrict__ cnt)
{
    __shared__ int sh[4];
    const int f = blockIdx.x, t = threadIdx.x;
    const FitDev F = fits[f];
    const int *deg = bdeg + (size_t)f * ldb;
    const unsigned *has = bhas + (size_t)f * ldb * kw;
    const int tot = F.act ? F.px * F.m : 0;
    int c = 0;
    for (int i = t; i < tot; i += 256) c += cand_ok(deg, has, kw, i % F.m, i / F.m, maxdeg, self) ? 1 : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((t & 63) == 0) sh[t >> 6] = c;
    __syncthreads();
    if (t == 0) cnt[f] = sh[0] + sh[1] + sh[2] + sh[3];
}

/* first[f] = candidates of the fits before f, *total = all of them; one block. */
__global__ __launch_bounds__(1024) void k_cand_scan(int F, const int *__restrict__ cnt, int *__restrict__ first,
                                                    int *__restrict__ total)
{
    __shared__ int sh[1024];
    const int t = threadIdx.x, chunk = (F + 1023) / 1024, f0 = min(F, t * chunk), f1 = min(F, f0 + chunk);
    int mine = 0;
    for (int f = f0; f < f1; ++f) mine += cnt[f];
    sh[t] = mine;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = t >= o ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    int at = sh[t] - mine;
    for (int f = f0; f < f1; ++f) {
        first[f] = at;
        at += cnt[f];
    }
    if (t == 1023) *total = sh[1023];
}

/* The candidate list, grouped by fit, then X column, then parent (the host driver's order);
 * one block per fit, order kept by wave ballots. */
__global__ __launch_bounds__(256) void k_cand_write(const FitDev *__restrict__ fits, const int *__restrict__ bdeg,
                                                    const unsigned *__restrict__ bhas, int kw, int ldb, int maxdeg,
                                                    int self, const int *__restrict__ first, int *__restrict__ cfit,
                                                    int *__restrict__ cxc, int *__restrict__ cpar)
{
    __shared__ int sw[4];
    const int f = blockIdx.x, t = threadIdx.x, w = t >> 6, lane = t & 63;
    const FitDev F = fits[f];
    const int *deg = bdeg + (size_t)f * ldb;
    const unsigned *has = bhas + (size_t)f * ldb * kw;
    const int tot = F.act ? F.px * F.m : 0;
    const unsigned long long below = (1ull << lane) - 1ull;
    int run = first[f];
    for (int base = 0; base < tot; base += 256) {
        const int i = base + t;
        const int b = i < tot ? i % F.m : 0, xc = i < tot ? i / F.m : 0;
        const bool take = i < tot && cand_ok(deg, has, kw, b, xc, maxdeg, self);
        const unsigned long long m = __ballot(take);
        if (lane == 0) sw[w] = __popcll(m);
        __syncthreads();
        int at = run;
        for (int k = 0; k < w; ++k) at += sw[k];
        if (take) {
            const int pos = at + __popcll(m & below);
            cfit[pos] = f;
            cxc[pos] = xc;
            cpar[pos] = b;
        }
        run += sw[0] + sw[1] + sw[2] + sw[3];
        __syncthreads();
    }
}

/* The basis function fit f has just appended (out.status 1) sits at index fits[f].m (the count
 * before the append): its degree is its parent's plus one and its variables are the parent's
 * plus the X column (mars.py:359-384).  One thread per fit. */
__global__ void k_meta_update(int F, const FitDev *__restrict__ fits, const AppendArg *__restrict__ args,
                              const AppendOut *__restrict__ out, int *__restrict__ bdeg, unsigned *__restrict__ bhas,
                              int kw, int ldb)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F || !args[f].active || out[f].status != 1) return;
    const int b = fits[f].m, p = args[f].bcol, xc = args[f].xcol;
    if (b >= ldb) return;
    int *deg = bdeg + (size_t)f * ldb;
    unsigned *has = bhas + (size_t)f * ldb * kw;
    deg[b] = deg[p] + 1;
    for (int k = 0; k < kw; ++k) has[(size_t)b * kw + k] = has[(size_t)p * kw + k];
    has[(size_t)b * kw + (xc >> 5)] |= 1u << (xc & 31);
}

__global__ void k_meta_clear(int n, int *__restrict__ bdeg, int nw, unsigned *__restrict__ bhas)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < max(n, nw); i += gridDim.x * blockDim.x) {
        if (i < n) bdeg[i] = 0;
        if (i < nw) bhas[i] = 0u;
    }
}

/* stat[0] += *n (candidates scored) */
__global__ void k_count_add_u(const int *__restrict__ n, unsigned long long *__restrict__ stat)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) stat[0] += (unsigned long long)*n;
}

__global__ void k_zero_redo(int *__restrict__ redo)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *redo = 0;
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
    // device metadata of the basis functions: every fit starts from its intercept (degree 0, no
    // variables) as the host driver does
    {
        size_t kw = 1;
        for (const FitState &fs : ctx->fits) kw = std::max(kw, (fs.px + 31) / 32);
        ctx->kw = (int)kw;
        const size_t nd = F * ctx->ldb, nw = nd * kw;
        ctx->bdeg.ensure(nd);
        ctx->bhas.ensure(nw);
        hipLaunchKernelGGL(k_meta_clear, dim3((unsigned)std::min<size_t>(1024, (nw + 255) / 256)), dim3(256), 0,
                           ctx->stream, (int)nd, ctx->bdeg.p, (int)nw, ctx->bhas.p);
        MH_LAUNCH_CHECK();
        for (size_t f = 0; f < F; ++f)
            require(ctx->fits[f].m == 1, "a batch context fits once (its basis holds the intercept only)");
    }
    ctx->act_h.assign(F, 1);
    // the largest step's candidate count: with graphs, every step's buffers, grids and bounds are
    // sized for it, so one captured graph per basis-size class serves every step
    size_t Pmax = 1;
    for (size_t f = 0; f < F; ++f) Pmax += ctx->fits[f].px * (D[f].max_terms > 1 ? D[f].max_terms - 1 : 1);
    {
        const char *e = std::getenv("MARS_HIP_GRAPH");
        ctx->use_graphs = !(e && e[0] == '0');
    }
    int step = 0;
    std::vector<AppendArg> args(F);
    std::vector<BestOut> best;
    bool device_lists = true;  // candidate, rescoring and retry lists are built on the device
    while (true) {
        HostTimer step_timer(ctx, K_H_STEP);
        // the mixed form's cache update runs on the device while the host builds the candidates
        bool live = false;
        for (size_t f = 0; f < F; ++f) live = live || (D[f].active && ctx->fits[f].m < D[f].max_terms);
        if (live && ctx->form == MARS_HIP_FORM_MIXED) mixed_prepare(ctx);
        HostTimer cands_timer(ctx, K_H_CANDS);
        Cands &cs = ctx->cand;
        bool any = false;
        size_t Pb = 0;  // candidates of this step
        cs.fit_c0.assign(F, 0);
        cs.fit_P.assign(F, 0);
        for (size_t f = 0; f < F; ++f) {
            FitDriver &d = D[f];
            if (d.active && ctx->fits[f].m >= d.max_terms) d.active = false;
            ctx->act_h[f] = d.active ? 1 : 0;
            if (!d.active) continue;
            any = true;
            // parent b serves every X column not among its variables (all of them with self-interaction)
            const FitState &fs = ctx->fits[f];
            size_t c = 0;
            for (size_t b = 0; b < fs.m; ++b)
                if ((int)d.vars[b].size() < o.max_degree)
                    c += o.self_interact ? fs.px : fs.px - d.vars[b].size();
            cs.fit_c0[f] = (int)Pb;
            cs.fit_P[f] = (int)c;
            Pb += c;
        }
        if (!any) {
            step_timer.cancel();
            cands_timer.cancel();
            break;
        }
        ++step;
        const int *dP = nullptr;  // device count of the candidates (device lists)
        const bool graph = device_lists && ctx->use_graphs;
        const size_t Pl = graph ? Pmax : Pb;  // the bound the scoring is launched with
        if (device_lists) {
            // the lists are generated on the device from each fit's basis metadata
            const int nblk = (int)((Pl + 1023) / 1024);
            ctx->ival.ensure(2 * F + 2 + 2 * (size_t)nblk);  // ranges, then the rescoring's scratch
            ctx->dcand.ensure(2);
            cs.d_fit.ensure(Pl);
            cs.d_xc.ensure(Pl);
            cs.d_par.ensure(Pl);
            Cands none;
            none.fit_c0.assign(F, 0);
            upload_fits(ctx, none);  // m and the active flags
            h2d_flush(ctx);
            const FitDev *fv = (const FitDev *)ctx->d_fits.p;
            hipLaunchKernelGGL(k_cand_count, dim3((unsigned)F), dim3(256), 0, ctx->stream, fv, (const int *)ctx->bdeg.p,
                               (const unsigned *)ctx->bhas.p, ctx->kw, (int)ctx->ldb, o.max_degree,
                               o.self_interact ? 1 : 0, ctx->ival.p + F);
            MH_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_cand_scan, dim3(1), dim3(1024), 0, ctx->stream, (int)F, (const int *)ctx->ival.p + F,
                               ctx->ival.p, ctx->dcand.p);
            MH_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_cand_write, dim3((unsigned)F), dim3(256), 0, ctx->stream, fv, (const int *)ctx->bdeg.p,
                               (const unsigned *)ctx->bhas.p, ctx->kw, (int)ctx->ldb, o.max_degree,
                               o.self_interact ? 1 : 0, (const int *)ctx->ival.p, cs.d_fit.p, cs.d_xc.p, cs.d_par.p);
            MH_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_count_add_u, dim3(1), dim3(64), 0, ctx->stream, (const int *)ctx->dcand.p,
                               ctx->dstat.p);
            MH_LAUNCH_CHECK();
            dP = ctx->dcand.p;
            cs.fit.clear();  // (no host copies of the lists; fit_c0 / fit_P hold the counts)
        } else {
            cs.fit.clear();
            cs.xc.clear();
            cs.par.clear();
            for (size_t f = 0; f < F; ++f) {
                FitDriver &d = D[f];
                if (!d.active) continue;
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
            index_cands(ctx, cs);
            Pb = cs.fit.size();
            ctx->scored += (long)Pb;
        }
        cands_timer.stop();
        const size_t P = Pb;
        ctx->lin.ensure(std::max<size_t>(Pl, 1));
        ctx->ybx.ensure(std::max<size_t>(Pl, 1));
        if (P) {
            HostTimer t(ctx, K_H_SCORE);
            if (ctx->form == MARS_HIP_FORM_MIXED) {
                mixed_score(ctx, !device_lists, dP, (int)(device_lists ? Pl : P));
            } else if (device_lists) {
                // every list after the DGKS gate (retries, the final pass) is built on the device
                const bool narrow = narrow_tiles(ctx, cs.fit_P);
                plain_score_device(ctx, cs, nullptr, dP, (int)Pl, nullptr, ctx->lin.p, ctx->ybx.p, nullptr,
                                   ctx->dredo.p, narrow, 1);
                auto body = [&] {
                    hipLaunchKernelGGL(k_zero_redo, dim3(1), dim3(64), 0, ctx->stream, ctx->dredo.p);
                    MH_LAUNCH_CHECK();
                    plain_score_device(ctx, cs, nullptr, dP, (int)Pl, nullptr, ctx->lin.p, ctx->ybx.p, nullptr,
                                       ctx->dredo.p, narrow, 2);
                };
                if (graph) {
                    size_t mmax = 1;
                    double flops = 0, bytes = 0;
                    for (size_t f = 0; f < F; ++f) {
                        const FitState &fs = ctx->fits[f];
                        mmax = std::max(mmax, fs.m);
                        // T = Bo^T Bx and the projection: 4 n m flops per candidate (bench_ortho.cc:139)
                        flops += 4.0 * fs.n * fs.m * cs.fit_P[f];
                        bytes += ctx->act_h[f] ? 4.0 * fs.n * (2.0 * fs.m + fs.px + 1) : 0.0;
                    }
                    std::vector<unsigned long long> key = {2ull, (unsigned long long)((mmax + 15) / 16),
                                                           (unsigned long long)narrow, (unsigned long long)Pl,
                                                           (unsigned long long)F};
                    tuning_key(ctx, key);
                    graph_key_buffers(ctx, key);
                    KTimer kt(ctx, K_GRAPH_PLAIN, flops, bytes);
                    run_graph(ctx, key, body);
                } else {
                    body();
                }
            } else {
                plain_score(ctx, cs, false, ctx->lin.p, ctx->ybx.p);
            }
            // no synchronisation here: the argmax that follows on the same stream returns the result
        }
        // the argmax and the append arguments stay on the device: the append is queued at once,
        // and the host reads the winners with the append's results
        {
            HostTimer t(ctx, K_H_ARGMAX);
            if (device_lists)
                argmax_launch_ranges(ctx, ctx->lin.p, P);  // the ranges the generator wrote into ival
            else
                argmax_launch(ctx, cs, ctx->lin.p);
            ctx->aarg.ensure(F);  // before the kernel writes it: the append must not reallocate it
            hipLaunchKernelGGL(k_batch_args, dim3((unsigned)((F + 255) / 256)), dim3(256), 0, ctx->stream, (int)F,
                               (const int *)ctx->ival.p, (const int *)ctx->ival.p + F,
                               (const BestOut *)ctx->bout.p, (const int *)cs.d_xc.p, (const int *)cs.d_par.p,
                               device_lists ? (const int *)ctx->dredo.p : nullptr, ctx->aarg.p);
            // (with host-built lists nothing on the device can ask for a redo)
            MH_LAUNCH_CHECK();
        }
        std::vector<double> sse0(F, 0.0), m0(F, 0.0);
        for (size_t f = 0; f < F; ++f) {
            // fits that may append: the active ones with candidates (the device drops a fit whose
            // best score is not above 0)
            // (an active fit always has candidates: the intercept is a parent for every column)
            args[f] = AppendArg{0, 0, 0, (D[f].active && (device_lists || cs.fit_P[f] > 0)) ? 1 : 0, 0.0f, 0};
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
            std::function<void()> ahead = [&] {
                // the appended basis functions' metadata, from the device's own arguments and
                // results (the fit views still hold the counts before the append)
                if (device_lists) {
                    hipLaunchKernelGGL(k_meta_update, dim3((unsigned)((F + 255) / 256)), dim3(256), 0, ctx->stream,
                                       (int)F, (const FitDev *)ctx->d_fits.p, (const AppendArg *)ctx->aarg.p,
                                       (const AppendOut *)ctx->aout.p, ctx->bdeg.p, ctx->bhas.p, ctx->kw,
                                       (int)ctx->ldb);
                    MH_LAUNCH_CHECK();
                }
                if (ctx->form != MARS_HIP_FORM_MIXED) return;
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
        require(valid_form(form), "unknown precision form");
        *out = nullptr;
        auto *ctx = new mars_hip_ctx;
        try {
            ctx_base_init(ctx, max_terms);
            ctx->resident = true;
            ctx->row_major = false;  // the batch scores linear terms only: column-major copies suffice
            set_form(ctx, form);
            form = ctx->form;
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
