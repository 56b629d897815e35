// Hinge-knot sweep on gfx950 (marsalgo.cc:508-657) and the per-column sort
// order it walks (marsalgo.cc:33-39, 521-529).
//
// Sweep state per candidate (X column xc, parent b, normalized column bx) after
// sorted row i, over lanes q = [Bo row (m) ; bx ; y ; b]:
//     g_i = sum_{l<=i} b_l q_l                   f_i = sum_{l<i} b_l (x_l - x_i) q_l
//     k_i = sum_{l<i} b_l^2 (x_l - x_i)^2
// advanced row by row as f += d g, g += b q, k += d^2 g_b + 2 d f_b with
// d = x_{i-1} - x_i.  The lane m+1 pair (g, f) is the reference's (vb, w) and
// the lane m+2 pair is (b2, bd).  At grid rows i = head+1 + t*min_span < tail:
//     ff = |f[0..m]|^2,  fy = f[0..m-1].ybo + f[m] ybx,
//     sse = (fy - w)^2 / (k - ff + tol)  if k - ff > tol.
// Rows are split into chunks scanned in parallel: pass 1 runs each chunk from
// the zero state, a carry scan combines the chunk summaries
//     g' = g + U,  f' = f + D g + F,  k' = k + 2 D f_b + D^2 g_b + K
// (D = sum of the chunk's deltas), and pass 2 re-runs each chunk from its exact
// carry-in while evaluating the grid.  Equal-x runs: a grid row whose x equals
// the previous grid row's x has the same score on the CPU (d = 0 leaves every
// accumulator unchanged) and can never win the strict '>' there, so it is
// skipped here -- the first grid row of a run is the one reported.
#include "internal.h"
#include "kernels.h"

#include <algorithm>

namespace mh {
namespace {

/* ---------------------------------------------------------------------- */
/* Stable descending sort of one X column: LSD radix, 8-bit digits        */
/* ---------------------------------------------------------------------- */
constexpr int kSortTile = 4096;  // 256 threads x 16 elements

/* Ascending key of the descending order; -0.0 and +0.0 map to one key. */
__device__ inline unsigned desc_key(float v)
{
    unsigned u = __float_as_uint(v == 0.0f ? 0.0f : v);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);  // ascending order of v
    return ~u;                                       // descending order of v
}

__global__ __launch_bounds__(256) void k_sort_keys(const float *__restrict__ x, size_t n,
                                                   unsigned *__restrict__ key, int *__restrict__ val)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    key[i] = desc_key(x[i]);
    val[i] = (int)i;
}

__global__ __launch_bounds__(256) void k_sort_hist(const unsigned *__restrict__ key, size_t n,
                                                   int shift, int nblocks, unsigned *__restrict__ hist)
{
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t b0 = (size_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortTile / 256; ++r) {
        const size_t i = b0 + (size_t)r * 256 + threadIdx.x;
        if (i < n) atomicAdd(&h[(key[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

/* Exclusive scan of the digit-major histogram (single block). */
__global__ __launch_bounds__(1024) void k_sort_scan(unsigned *__restrict__ hist, int count)
{
    __shared__ unsigned sh[1024];
    const int t = threadIdx.x;
    const int chunk = (count + 1023) / 1024;
    const int b = t * chunk, e = min(count, b + chunk);
    unsigned s = 0;
    for (int i = b; i < e; ++i) s += hist[i];
    sh[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const unsigned v = t >= o ? sh[t - o] : 0u;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    unsigned run = sh[t] - s;
    for (int i = b; i < e; ++i) {
        const unsigned v = hist[i];
        hist[i] = run;
        run += v;
    }
}

/* Stable scatter: rank within the tile = earlier elements with the same digit. */
__global__ __launch_bounds__(256) void k_sort_scatter(const unsigned *__restrict__ kin,
                                                      const int *__restrict__ vin, size_t n,
                                                      int shift, int nblocks,
                                                      const unsigned *__restrict__ offs,
                                                      unsigned *__restrict__ kout,
                                                      int *__restrict__ vout)
{
    __shared__ unsigned base[256];
    __shared__ unsigned wcnt[4][256];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    base[t] = offs[(size_t)t * nblocks + blockIdx.x];
    const size_t b0 = (size_t)blockIdx.x * kSortTile;
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (int r = 0; r < kSortTile / 256; ++r) {
        for (int q = 0; q < 4; ++q) wcnt[q][t] = 0;
        __syncthreads();
        const size_t i = b0 + (size_t)r * 256 + t;
        const bool ok = i < n;
        const unsigned k = ok ? kin[i] : 0u;
        const unsigned dg = (k >> shift) & 255u;
        unsigned long long match = __ballot(ok);
        for (int bit = 0; bit < 8; ++bit) {
            const unsigned long long bb = __ballot(ok && ((dg >> bit) & 1u));
            match &= ((dg >> bit) & 1u) ? bb : ~bb;
        }
        const unsigned rank = (unsigned)__popcll(match & lt);
        if (ok && rank == 0) wcnt[w][dg] = (unsigned)__popcll(match);
        __syncthreads();
        if (ok) {
            unsigned pos = base[dg] + rank;
            for (int q = 0; q < w; ++q) pos += wcnt[q][dg];
            kout[pos] = k;
            vout[pos] = vin[i];
        }
        __syncthreads();
        base[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
        __syncthreads();
    }
}

/* d[i] = f32(xs[k[i-1]] - xs[k[i]]) on the scaled column (marsalgo.cc:526-529), d[0] = 0. */
__global__ __launch_bounds__(256) void k_deltas(const float *__restrict__ x, float s,
                                                const int *__restrict__ ord, size_t n,
                                                float *__restrict__ d)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (i == 0) {
        d[0] = 0.0f;
        return;
    }
    const float a = __fmul_rn(x[ord[i - 1]], s);
    const float b = __fmul_rn(x[ord[i]], s);
    d[i] = __fsub_rn(a, b);
}

/* ---------------------------------------------------------------------- */
/* Sweep                                                                  */
/* ---------------------------------------------------------------------- */
struct SweepArgs {
    const float *Bo, *B, *rbuf, *y, *X, *s;
    size_t ldb, ldr, n;
    const double *ybo, *ybx;
    const int *grp;      // per X-column group: xc, first candidate, count
    const int *cpar;     // parent per candidate
    const int *order;    // [px][n]
    const float *delta;  // [px][n]
    int m, L2, C;        // basis count, padded lane count (KPL*64), chunks
    size_t tail, head, span, clen;
    double tol;
    double *sum;         // [P][C][L2][2]   chunk summaries, then carry-ins (f, g)
    double *ksum;        // [P][C]          chunk k, then carry-in k
    double *dsum;        // [ngroups][C]    chunk delta sums
    double *best;        // [P][C]
    int *besti;          // [P][C]
};

constexpr int kTileRows = 64;

template <int KPL, bool EVAL>
__global__ __launch_bounds__(1024) void k_sweep(SweepArgs a)
{
    extern __shared__ float lds[];
    float *sbo = lds;                                   // [64][m]
    int *srow = (int *)(sbo + kTileRows * a.m);         // [64]
    float *sd = (float *)(srow + kTileRows);            // [64]
    float *sy = sd + kTileRows;                         // [64]

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int chunk = blockIdx.x, pg = blockIdx.y, gi = blockIdx.z;
    const int xc = a.grp[3 * gi], c0 = a.grp[3 * gi + 1], np = a.grp[3 * gi + 2];
    const int j = pg * 16 + wv;             // parent slot within the group
    if (pg * 16 >= np) return;              // whole block idle (uniform)
    const bool active = j < np;
    const int c = c0 + (active ? j : 0);
    const int par = a.cpar[c];
    const int m = a.m;
    const size_t r0 = (size_t)chunk * a.clen;
    const size_t r1 = min(a.tail, r0 + a.clen);
    const int *ord = a.order + (size_t)xc * a.n;
    const float *dl = a.delta + (size_t)xc * a.n;

    double f[KPL], g[KPL], Y[KPL];
    double k2 = 0.0;
    const int sb = (m + 2) / 64, tb = (m + 2) % 64;
    const int sy_ = (m + 1) / 64, ty_ = (m + 1) % 64;
#pragma unroll
    for (int s = 0; s < KPL; ++s) {
        const int k = lane + 64 * s;
        f[s] = g[s] = 0.0;
        Y[s] = k < m ? a.ybo[k] : (k == m ? a.ybx[c] : 0.0);
    }
    if (EVAL && active) {
        const double *src = a.sum + ((size_t)c * a.C + chunk) * a.L2 * 2;
#pragma unroll
        for (int s = 0; s < KPL; ++s) {
            f[s] = src[2 * (lane + 64 * s)];
            g[s] = src[2 * (lane + 64 * s) + 1];
        }
        k2 = a.ksum[(size_t)c * a.C + chunk];
    }

    // grid rows inside this chunk and the tie-run tracker
    const size_t g0 = a.head + 1;
    size_t next = g0;
    if (r0 > g0) next = g0 + ((r0 - g0 + a.span - 1) / a.span) * a.span;
    long long last_nz = -1;  // last sorted row <= current with a nonzero delta
    if (EVAL) {
        const size_t lo = r0 > a.span ? r0 - a.span : 0;
        for (size_t l = lo; l < r0; ++l)
            if (dl[l] != 0.0f) last_nz = (long long)l;
    }
    double bsse = 0.0;
    int bidx = -1;
    double dacc = 0.0;

    for (size_t t0 = r0; t0 < r1; t0 += kTileRows) {
        const int nr = (int)min((size_t)kTileRows, r1 - t0);
        __syncthreads();
        if ((int)threadIdx.x < nr) {
            const int row = ord[t0 + threadIdx.x];
            srow[threadIdx.x] = row;
            sd[threadIdx.x] = dl[t0 + threadIdx.x];
            sy[threadIdx.x] = a.y[row];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nr * m; e += blockDim.x) {
            const int rr = e / m, k = e - rr * m;
            sbo[rr * m + k] = a.Bo[(size_t)srow[rr] * a.ldb + k];
        }
        __syncthreads();
        if (!active) continue;
        // per-lane preload of this parent's b and bx for the tile rows
        float bt = 0.f, bxt = 0.f;
        if (lane < nr) {
            const size_t row = (size_t)srow[lane];
            bt = a.B[row * a.ldb + par];
            bxt = a.rbuf[row * a.ldr + c];
        }
        for (int rr = 0; rr < nr; ++rr) {
            const size_t i = t0 + rr;
            const double d = (double)sd[rr];
            const double b = (double)__shfl(bt, rr, 64);
            const double bx = (double)__shfl(bxt, rr, 64);
            const double yv = (double)sy[rr];
            dacc += d;
            // k uses the b-lane (f, g) before this row's update
            double fb = 0.0, gb = 0.0;
#pragma unroll
            for (int s = 0; s < KPL; ++s)
                if (s == sb) {
                    fb = f[s];
                    gb = g[s];
                }
            k2 = fma(d * d, gb, k2);
            k2 = fma(d * 2, fb, k2);
#pragma unroll
            for (int s = 0; s < KPL; ++s) {
                const int k = lane + 64 * s;
                double q;
                if (k < m) q = (double)sbo[rr * m + k];
                else if (k == m) q = bx;
                else if (k == m + 1) q = yv;
                else if (k == m + 2) q = b;
                else q = 0.0;
                f[s] = fma(d, g[s], f[s]);
                g[s] = fma(b, q, g[s]);
            }
            if (EVAL) {
                if (d != 0.0) last_nz = (long long)i;
                if (i == next) {
                    next += a.span;
                    const bool dup = (i >= g0 + a.span) && (last_nz <= (long long)(i - a.span));
                    double ff = 0.0, fy = 0.0;
#pragma unroll
                    for (int s = 0; s < KPL; ++s) {
                        const int k = lane + 64 * s;
                        if (k <= m) {
                            ff = fma(f[s], f[s], ff);
                            fy = fma(f[s], Y[s], fy);
                        }
                    }
                    ff = wave_sum(ff);
                    fy = wave_sum(fy);
                    double fw = 0.0;
#pragma unroll
                    for (int s = 0; s < KPL; ++s)
                        if (s == sy_) fw = f[s];
                    const double w = __shfl(fw, ty_, 64);
                    const double kk = __shfl(k2, tb, 64);
                    const double uw = fy - w;
                    const double den = kk - ff;
                    const double sse = den > a.tol ? (uw * uw) / (den + a.tol) : 0.0;
                    if (!dup && sse > bsse) {
                        bsse = sse;
                        bidx = (int)i;
                    }
                }
            }
        }
    }
    if (!active) return;
    const size_t slot = (size_t)c * a.C + chunk;
    if (EVAL) {
        if (lane == 0) {
            a.best[slot] = bsse;
            a.besti[slot] = bidx;
        }
    } else {
        double *dst = a.sum + slot * a.L2 * 2;
#pragma unroll
        for (int s = 0; s < KPL; ++s) {
            dst[2 * (lane + 64 * s)] = f[s];
            dst[2 * (lane + 64 * s) + 1] = g[s];
        }
        if (lane == tb) a.ksum[slot] = k2;
        if (lane == 0 && j == 0) a.dsum[(size_t)gi * a.C + chunk] = dacc;
    }
}

/* Turn chunk summaries into carry-ins, in chunk order.  One thread per (candidate, lane). */
__global__ void k_sweep_carry(SweepArgs a, const int *__restrict__ cgrp, int P)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = (int)(t / a.L2), k = (int)(t % a.L2);
    if (c >= P) return;
    const int gi = cgrp[c];
    const int m = a.m;
    double f = 0.0, g = 0.0, kk = 0.0;
    // the thread of lane m+2 also carries k and needs its own (f, g) = (bd, b2)
    for (int ch = 0; ch < a.C; ++ch) {
        double *s = a.sum + (((size_t)c * a.C + ch) * a.L2 + k) * 2;
        const double F = s[0], U = s[1];
        const double D = a.dsum[(size_t)gi * a.C + ch];
        s[0] = f;
        s[1] = g;
        if (k == m + 2) {
            double *ks = a.ksum + (size_t)c * a.C + ch;
            const double K = *ks;
            *ks = kk;
            kk = kk + 2.0 * D * f + D * D * g + K;
        }
        f = f + D * g + F;
        g = g + U;
    }
}

/* Best knot over chunks (first maximum in row order), hinge = linear + best. */
__global__ void k_sweep_final(SweepArgs a, int P, const int *__restrict__ cxc,
                              const double *__restrict__ lin, double *__restrict__ hin,
                              int *__restrict__ knot, int *__restrict__ krow)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P) return;
    double b = 0.0;
    int bi = -1;
    for (int ch = 0; ch < a.C; ++ch) {
        const double v = a.best[(size_t)c * a.C + ch];
        const int vi = a.besti[(size_t)c * a.C + ch];
        if (vi >= 0 && v > b) {
            b = v;
            bi = vi;
        }
    }
    knot[c] = bi;
    if (bi >= 0) {
        hin[c] = lin[c] + b;
        krow[c] = a.order[(size_t)cxc[c] * a.n + bi];
    } else {
        hin[c] = 0.0;
        krow[c] = -1;
    }
}

template <int KPL>
void launch_sweep(const SweepArgs &a, dim3 grid, size_t lds, hipStream_t st, bool eval)
{
    if (eval)
        hipLaunchKernelGGL((k_sweep<KPL, true>), grid, dim3(1024), lds, st, a);
    else
        hipLaunchKernelGGL((k_sweep<KPL, false>), grid, dim3(1024), lds, st, a);
    MH_LAUNCH_CHECK();
}

} // namespace

/* Make the stable descending order and f32 deltas of X column xc of fit 0 resident. */
static void ensure_sorted(mars_hip_ctx *ctx, int xc)
{
    FitState &fs = ctx->fits[0];
    if (fs.sorted.size() != fs.px) fs.sorted.assign(fs.px, 0);
    if (fs.sorted[xc]) return;
    const size_t n = fs.n;
    hipStream_t st = ctx->stream;
    fs.order.ensure(fs.px * n);
    fs.delta.ensure(fs.px * n);
    const int nblocks = (int)((n + kSortTile - 1) / kSortTile);
    DBuf<unsigned> keys, hist;
    DBuf<int> vals;
    keys.ensure(2 * n);
    vals.ensure(n);
    hist.ensure((size_t)256 * nblocks);
    int *out = fs.order.p + (size_t)xc * n;
    unsigned *ka = keys.p, *kb = keys.p + n;
    int *va = vals.p, *vb = out;
    KTimer kt(ctx, K_SORT, 0.0, 16.0 * 4 * n);
    hipLaunchKernelGGL(k_sort_keys, dim3((n + 255) / 256), dim3(256), 0, st, fs.X.p + (size_t)xc * fs.ldx,
                       n, ka, va);
    MH_LAUNCH_CHECK();
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = 8 * pass;
        hipLaunchKernelGGL(k_sort_hist, dim3(nblocks), dim3(256), 0, st, ka, n, shift, nblocks, hist.p);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(1024), 0, st, hist.p, 256 * nblocks);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sort_scatter, dim3(nblocks), dim3(256), 0, st, ka, va, n, shift, nblocks,
                           hist.p, kb, vb);
        MH_LAUNCH_CHECK();
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    // after 4 passes the values are back in `vals` (even pass count): copy into the cache
    MH_CHECK(hipMemcpyAsync(out, va, n * sizeof(int), hipMemcpyDeviceToDevice, st));
    float sc = 1.0f;
    MH_CHECK(hipMemcpyAsync(&sc, fs.s.p + xc, sizeof(float), hipMemcpyDeviceToHost, st));
    MH_CHECK(hipStreamSynchronize(st));
    hipLaunchKernelGGL(k_deltas, dim3((n + 255) / 256), dim3(256), 0, st, fs.X.p + (size_t)xc * fs.ldx,
                       sc, out, n, fs.delta.p + (size_t)xc * n);
    MH_LAUNCH_CHECK();
    MH_CHECK(hipStreamSynchronize(st));
    keys.release();
    vals.release();
    hist.release();
    fs.sorted[xc] = 1;
}

void hinge_sweep(mars_hip_ctx *ctx, int min_span, int end_span)
{
    FitState &fs = ctx->fits[0];
    Cands &cs = ctx->cand;
    const int P = (int)cs.fit.size();
    const size_t n = fs.n;
    const int m = (int)fs.m;
    hipStream_t st = ctx->stream;
    ctx->hin.ensure(P);
    ctx->knot.ensure(P);
    ctx->krow.ensure(P);

    // X-column groups of this block (candidates are contiguous per column)
    std::vector<int> grp, cgrp(P);
    int maxnp = 0;
    for (int c = 0; c < P;) {
        int e = c;
        while (e < P && cs.xc[e] == cs.xc[c]) ++e;
        for (int q = c; q < e; ++q) cgrp[q] = (int)(grp.size() / 3);
        grp.push_back(cs.xc[c]);
        grp.push_back(c);
        grp.push_back(e - c);
        maxnp = std::max(maxnp, e - c);
        ensure_sorted(ctx, cs.xc[c]);
        c = e;
    }
    const int ng = (int)(grp.size() / 3);
    const size_t tail = n - (size_t)end_span;
    const int pgroups = (maxnp + 15) / 16;
    const int want = std::max(1, (512 + ng * pgroups - 1) / (ng * pgroups));
    const int maxc = (int)std::max<size_t>(1, tail / 4096);
    const int C = std::min(want, maxc);
    const size_t clen = (tail + C - 1) / C;
    const int KPL = (m + 3 + 63) / 64;
    require(KPL <= 4, "hinge sweep supports up to 253 basis functions");
    const int L2 = KPL * 64;

    DBuf<int> dgrp, dcgrp;
    dgrp.ensure(grp.size());
    dcgrp.ensure(P);
    MH_CHECK(hipMemcpyAsync(dgrp.p, grp.data(), grp.size() * sizeof(int), hipMemcpyHostToDevice, st));
    MH_CHECK(hipMemcpyAsync(dcgrp.p, cgrp.data(), P * sizeof(int), hipMemcpyHostToDevice, st));
    ctx->hstate.ensure((size_t)P * C * L2 * 2 + (size_t)P * C + (size_t)ng * C);
    ctx->hbest.ensure((size_t)P * C);
    ctx->hbest_i.ensure((size_t)P * C);

    SweepArgs a;
    a.Bo = fs.Bo.p;
    a.B = fs.B.p;
    a.rbuf = fs.rbuf.p;
    a.y = fs.y.p;
    a.X = fs.X.p;
    a.s = fs.s.p;
    a.ldb = ctx->ldb;
    a.ldr = fs.ldr;
    a.n = n;
    a.ybo = fs.ybo.p;
    a.ybx = ctx->ybx.p;
    a.grp = dgrp.p;
    a.cpar = cs.d_par.p;
    a.order = fs.order.p;
    a.delta = fs.delta.p;
    a.m = m;
    a.L2 = L2;
    a.C = C;
    a.tail = tail;
    a.head = (size_t)end_span;
    a.span = (size_t)min_span;
    a.clen = clen;
    a.tol = fs.tol;
    a.sum = ctx->hstate.p;
    a.ksum = a.sum + (size_t)P * C * L2 * 2;
    a.dsum = a.ksum + (size_t)P * C;
    a.best = ctx->hbest.p;
    a.besti = ctx->hbest_i.p;

    const dim3 grid(C, pgroups, ng);
    // reference algorithmic work: n (m+1) (4 + 4/min_span) flops per candidate (marsalgo.cc:576-649)
    {
        KTimer kt(ctx, K_SWEEP, (double)P * (double)tail * (m + 1) * (4.0 + 4.0 / min_span),
                  (double)ng * tail * (4.0 * m + 12.0) + (double)P * tail * 8.0);
        const size_t lds = (size_t)kTileRows * m * sizeof(float) + 3 * kTileRows * sizeof(float);
        for (int pass = 0; pass < 2; ++pass) {
            const bool eval = pass == 1;
            if (eval) {
                const size_t threads = (size_t)P * L2;
                hipLaunchKernelGGL(k_sweep_carry, dim3((threads + 255) / 256), dim3(256), 0, st, a,
                                   dcgrp.p, P);
                MH_LAUNCH_CHECK();
            }
            switch (KPL) {
            case 1: launch_sweep<1>(a, grid, lds, st, eval); break;
            case 2: launch_sweep<2>(a, grid, lds, st, eval); break;
            case 3: launch_sweep<3>(a, grid, lds, st, eval); break;
            default: launch_sweep<4>(a, grid, lds, st, eval); break;
            }
        }
        hipLaunchKernelGGL(k_sweep_final, dim3((P + 255) / 256), dim3(256), 0, st, a, P, cs.d_xc.p,
                           ctx->lin.p, ctx->hin.p, ctx->knot.p, ctx->krow.p);
        MH_LAUNCH_CHECK();
    }
    MH_CHECK(hipStreamSynchronize(st));
    dgrp.release();
    dcgrp.release();
}

} // namespace mh
