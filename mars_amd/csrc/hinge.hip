// Hinge-knot sweep on gfx950 (marsalgo.cc:508-657) and the per-column sort
// order it walks (marsalgo.cc:33-39, 521-529).
//
// Sweep state per candidate (X column xc, parent b, normalized column bx) after
// sorted row i, over lanes q = [b ; y ; bx ; Bo row (m)]:
//     g_i = sum_{l<=i} b_l q_l                   f_i = sum_{l<i} b_l (x_l - x_i) q_l
//     k_i = sum_{l<i} b_l^2 (x_l - x_i)^2
// advanced row by row as f += d g, g += b q, k += d^2 g_b + 2 d f_b with
// d = x_{i-1} - x_i.  The y lane's pair (g, f) is the reference's (vb, w) and
// the b lane's pair is (b2, bd).  At grid rows i = head+1 + t*min_span < tail:
//     ff = |f[bx, Bo]|^2,  fy = f[Bo].ybo + f[bx] ybx,
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

constexpr int kTile = 32;   // sorted rows per tile
constexpr int kPitch = 33;  // LDS floats per tile column (odd: column-major tile without bank conflicts)
constexpr int kSlots = 8;   // candidates per block

/* LDS bytes of one k_sweep block for `nc` tile columns (two buffers of everything). */
__host__ __device__ constexpr size_t sweep_lds(int nc)
{
    return 2 * (size_t)kSlots * kTile * sizeof(double)   // b per slot and row (f64)
         + 2 * (size_t)kTile * 2 * sizeof(double)        // (d, d*d) per row
         + 2 * (size_t)kTile * sizeof(int)               // sorted-order row indices
         + 2 * (size_t)nc * kPitch * sizeof(float);      // the tile, column-major
}

/*
 * One block = W = 8/PK waves sweeping kSlots candidates (the parents c0.. of one
 * X column) over the sorted rows [r0, r1) of its chunk.  PK = 2 packs two
 * candidates of at most 29 basis functions into the halves of a wave (L = 32
 * lanes each).  Lane slot u = lane % L + L s (s < KPL) of a candidate holds
 * the running (f, g) of one covariate column: u = 0 the parent b, u = 1 y,
 * u = 2 the candidate's Bx column, u = 3 .. m+2 the orthonormal basis; the
 * b lane also carries k = |h|^2 (marsalgo.cc:576-649).
 *
 * Tiles of 32 sorted rows are gathered into LDS column-major, columns
 *   [0, m) Bo row | [m, m+8) Bx per slot | m+8 y | m+9 zero | [m+10, m+18) b per slot,
 * double-buffered: the global gathers of tile t+1 are in flight while the
 * unrolled recurrence runs over tile t.  Rows past the chunk end are zero
 * (d = b = 0 leaves every accumulator unchanged), so every tile runs 32 rows.
 * Knot rows come as a per-tile bit mask (grid position) plus a mask of nonzero
 * deltas for the tie-run test, both in scalar registers.
 */
template <int KPL, int PK, bool EVAL>
__global__ __launch_bounds__(512) void k_sweep(SweepArgs a)
{
    constexpr int L = 64 / PK;
    constexpr int W = kSlots / PK;                  // waves per block
    constexpr int NT = 64 * W;
    constexpr int RPW = kTile / W;                  // gathered rows per wave and tile
    constexpr int UM = PK == 2 ? 1 : KPL + 1;       // 64-column groups of the tile
    extern __shared__ double lds_d[];
    const int m = a.m;
    const int nc = m + 2 * kSlots + 2;
    double *sb = lds_d;                                       // [2][kSlots][32]
    double2 *sdd = (double2 *)(sb + 2 * kSlots * kTile);      // [2][32]
    int *srow = (int *)(sdd + 2 * kTile);                     // [2][32]
    float *sq = (float *)(srow + 2 * kTile);                  // [2][nc][33]

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int half = PK == 2 ? lane >> 5 : 0, kl = lane & (L - 1);
    const int chunk = blockIdx.x, pg = blockIdx.y, gi = blockIdx.z;
    const int xc = a.grp[3 * gi], c0 = a.grp[3 * gi + 1], np = a.grp[3 * gi + 2];
    const int sbase = pg * kSlots;                  // first parent slot of this block
    if (sbase >= np) return;                        // whole block idle (uniform)
    const int nact = min(kSlots, np - sbase);
    const int slot = wv * PK + half;                // this lane's candidate slot
    const bool active = slot < nact;
    const bool wave_active = wv * PK < nact;
    const int c = c0 + sbase + (active ? slot : 0);
    const size_t r0 = (size_t)chunk * a.clen;
    const size_t r1 = min(a.tail, r0 + a.clen);
    const int *ord = a.order + (size_t)xc * a.n;
    const float *dl = a.delta + (size_t)xc * a.n;
    const int T = r1 > r0 ? (int)((r1 - r0 + kTile - 1) / kTile) : 0;
    auto tile_rows = [&](int t) { return (int)min((size_t)kTile, r1 - (r0 + (size_t)t * kTile)); };

    // ---- gather sources of this thread's tile columns (fixed over tiles)
    const float *cp[UM];
    size_t cld[UM];
    int ccol[UM], cbs[UM];
#pragma unroll
    for (int u = 0; u < UM; ++u) {
        const int col = lane + 64 * u;
        cp[u] = nullptr;
        cld[u] = 0;
        ccol[u] = col < nc ? col : -1;
        cbs[u] = -1;
        if (col < m) {
            cp[u] = a.Bo + col;
            cld[u] = a.ldb;
        } else if (col < m + kSlots) {
            const int j = col - m;
            if (j < nact) {
                cp[u] = a.rbuf + (c0 + sbase + j);
                cld[u] = a.ldr;
            }
        } else if (col == m + kSlots) {
            cp[u] = a.y;
            cld[u] = 1;
        } else if (col >= m + kSlots + 2 && col < nc) {
            const int j = col - m - kSlots - 2;
            cbs[u] = j;
            if (j < nact) {
                cp[u] = a.B + a.cpar[c0 + sbase + j];
                cld[u] = a.ldb;
            }
        }
    }

    // ---- this lane's covariate columns and state
    int qoff[KPL];
    double f[KPL], g[KPL], Y[KPL];
    bool inff[KPL];
    double k2 = 0.0;
#pragma unroll
    for (int s = 0; s < KPL; ++s) {
        const int u = kl + L * s;
        int col;
        if (u == 0) col = m + kSlots + 2 + slot;
        else if (u == 1) col = m + kSlots;
        else if (u == 2) col = m + slot;
        else if (u <= m + 2) col = u - 3;
        else col = m + kSlots + 1;
        qoff[s] = col * kPitch;
        f[s] = g[s] = 0.0;
        Y[s] = (u == 2) ? a.ybx[c] : ((u >= 3 && u <= m + 2) ? a.ybo[u - 3] : 0.0);
        inff[s] = u >= 2 && u <= m + 2;
    }
    if (EVAL && active) {
        const double *src = a.sum + ((size_t)c * a.C + chunk) * a.L2 * 2;
#pragma unroll
        for (int s = 0; s < KPL; ++s) {
            f[s] = src[2 * (kl + L * s)];
            g[s] = src[2 * (kl + L * s) + 1];
        }
        k2 = a.ksum[(size_t)c * a.C + chunk];
    }

    // ---- grid rows and the tie-run tracker
    const size_t g0 = a.head + 1;
    int lastnz = -1;  // last sorted row before the current tile with a nonzero delta
    if (EVAL) {
        const size_t lo = r0 > a.span ? r0 - a.span : 0;
        int mine = -1;
        for (size_t l = lo + lane; l < r0; l += 64)
            if (dl[l] != 0.0f) mine = (int)l;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mine = max(mine, __shfl_xor(mine, o, 64));
        lastnz = mine;
    }
    double bsse = 0.0, dacc = 0.0;
    int bidx = -1;

    // tile masks of the wave: bit rr = row rr is a grid row / has a nonzero delta
    auto tile_masks = [&](int t, float dv, unsigned &km, unsigned &nzm) {
        const int nr = tile_rows(t);
        const size_t i = r0 + (size_t)t * kTile + lane;
        const bool in = lane < nr;
        const bool grid = in && i >= g0 && ((i - g0) % a.span) == 0;
        km = (unsigned)__ballot(grid);
        nzm = (unsigned)__ballot(in && dv != 0.0f);
    };

    // ---- prologue: row indices of tiles 0 and 1, tile 0 into buffer 0
    if (tid < kTile) {
        if (T > 0 && tid < tile_rows(0)) srow[tid] = ord[r0 + tid];
        if (T > 1 && tid < tile_rows(1)) srow[kTile + tid] = ord[r0 + kTile + tid];
    }
    __syncthreads();
    unsigned kmask = 0, nzmask = 0;
    if (T > 0) {
        const int nr = tile_rows(0);
#pragma unroll
        for (int i = 0; i < RPW; ++i) {
            const int rr = wv + W * i;
            const bool in = rr < nr;
            const size_t row = in ? (size_t)srow[rr] : 0;
#pragma unroll
            for (int u = 0; u < UM; ++u) {
                const float v = (in && cp[u]) ? cp[u][row * cld[u]] : 0.0f;
                if (ccol[u] >= 0) sq[ccol[u] * kPitch + rr] = v;
                if (cbs[u] >= 0) sb[cbs[u] * kTile + rr] = (double)v;
            }
        }
        const float dv = lane < nr ? dl[r0 + lane] : 0.0f;
        if (tid < kTile) {
            const double d = (double)dv;
            sdd[tid] = make_double2(d, d * d);
        }
        if (EVAL) tile_masks(0, dv, kmask, nzmask);
    }
    __syncthreads();

    for (int t = 0; t < T; ++t) {
        const int cur = t & 1, nxt = cur ^ 1;
        const size_t tb0 = r0 + (size_t)t * kTile;
        const int nr1 = t + 1 < T ? tile_rows(t + 1) : 0;
        const int nr2 = t + 2 < T ? tile_rows(t + 2) : 0;
        // ---- prefetch tile t+1 (its row indices are in srow[nxt]) and tile t+2's indices
        float gv[RPW][UM];
#pragma unroll
        for (int i = 0; i < RPW; ++i) {
            const int rr = wv + W * i;
            const bool in = rr < nr1;
            const size_t row = in ? (size_t)srow[nxt * kTile + rr] : 0;
#pragma unroll
            for (int u = 0; u < UM; ++u) gv[i][u] = (in && cp[u]) ? cp[u][row * cld[u]] : 0.0f;
        }
        const float nd = lane < nr1 ? dl[tb0 + kTile + lane] : 0.0f;
        int nrow = 0;
        if (tid < nr2) nrow = ord[tb0 + 2 * kTile + tid];

        // ---- the recurrence over tile t (marsalgo.cc:576-649)
        if (wave_active) {
            const float *tq = sq + (size_t)cur * nc * kPitch;
            const double *tbv = sb + (cur * kSlots + slot) * kTile;
            const double2 *td = sdd + cur * kTile;
#pragma unroll
            for (int rr = 0; rr < kTile; ++rr) {
                const double2 dd = td[rr];
                const double b = tbv[rr];
                // k uses the b lane's (f, g) before this row's update
                k2 = fma(dd.y, g[0], k2);
                k2 = fma(dd.x * 2, f[0], k2);
#pragma unroll
                for (int s = 0; s < KPL; ++s) {
                    const double q = (double)tq[qoff[s] + rr];
                    f[s] = fma(dd.x, g[s], f[s]);
                    g[s] = fma(b, q, g[s]);
                }
                if (!EVAL) dacc += dd.x;
                if (EVAL && ((kmask >> rr) & 1u)) {
                    const size_t i = tb0 + rr;
                    const unsigned nzb = nzmask & (0xffffffffu >> (31 - rr));
                    const long long last = nzb ? (long long)(tb0 + 31 - __clz(nzb)) : (long long)lastnz;
                    const bool dup = (i >= g0 + a.span) && (last <= (long long)(i - a.span));
                    if (!dup) {
                        double ff = 0.0, fy = 0.0;
#pragma unroll
                        for (int s = 0; s < KPL; ++s)
                            if (inff[s]) {
                                ff = fma(f[s], f[s], ff);
                                fy = fma(f[s], Y[s], fy);
                            }
#pragma unroll
                        for (int o = L / 2; o > 0; o >>= 1) {
                            ff = ff + __shfl_xor(ff, o, 64);
                            fy = fy + __shfl_xor(fy, o, 64);
                        }
                        const double w = __shfl(f[0], half * L + 1, 64);
                        const double kk = __shfl(k2, half * L, 64);
                        const double uw = fy - w;
                        const double den = kk - ff;
                        const double sse = den > a.tol ? (uw * uw) / (den + a.tol) : 0.0;
                        if (sse > bsse) {
                            bsse = sse;
                            bidx = (int)i;
                        }
                    }
                }
            }
        }
        if (EVAL && nzmask) lastnz = (int)(tb0 + 31 - __clz(nzmask));
        if (EVAL && t + 1 < T) tile_masks(t + 1, nd, kmask, nzmask);
        // buffer nxt last held tile t-1, which every wave finished before the
        // previous iteration's barrier: store tile t+1 into it
#pragma unroll
        for (int i = 0; i < RPW; ++i) {
            const int rr = wv + W * i;
#pragma unroll
            for (int u = 0; u < UM; ++u) {
                if (ccol[u] >= 0) sq[((size_t)nxt * nc + ccol[u]) * kPitch + rr] = gv[i][u];
                if (cbs[u] >= 0) sb[(nxt * kSlots + cbs[u]) * kTile + rr] = (double)gv[i][u];
            }
        }
        if (tid < kTile) {
            const double d = (double)nd;
            sdd[nxt * kTile + tid] = make_double2(d, d * d);
            if (tid < nr2) srow[cur * kTile + tid] = nrow;
        }
        __syncthreads();
    }
    if (!active) return;
    const size_t cs = (size_t)c * a.C + chunk;
    if (EVAL) {
        if (kl == 0) {
            a.best[cs] = bsse;
            a.besti[cs] = bidx;
        }
    } else {
        double *dst = a.sum + cs * a.L2 * 2;
#pragma unroll
        for (int s = 0; s < KPL; ++s) {
            dst[2 * (kl + L * s)] = f[s];
            dst[2 * (kl + L * s) + 1] = g[s];
        }
        if (kl == 0) a.ksum[cs] = k2;
        if (kl == 0 && pg == 0 && slot == 0) a.dsum[(size_t)gi * a.C + chunk] = dacc;
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
    // the thread of lane slot 0 (the parent b) also carries k and needs its own (f, g) = (bd, b2)
    for (int ch = 0; ch < a.C; ++ch) {
        double *s = a.sum + (((size_t)c * a.C + ch) * a.L2 + k) * 2;
        const double F = s[0], U = s[1];
        const double D = a.dsum[(size_t)gi * a.C + ch];
        s[0] = f;
        s[1] = g;
        if (k == 0) {
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
void launch_sweep(const SweepArgs &a, dim3 grid, size_t lds, hipStream_t st, bool eval, int pk)
{
    if (pk == 2) {
        if (eval)
            hipLaunchKernelGGL((k_sweep<1, 2, true>), grid, dim3(64 * kSlots / 2), lds, st, a);
        else
            hipLaunchKernelGGL((k_sweep<1, 2, false>), grid, dim3(64 * kSlots / 2), lds, st, a);
    } else {
        if (eval)
            hipLaunchKernelGGL((k_sweep<KPL, 1, true>), grid, dim3(64 * kSlots), lds, st, a);
        else
            hipLaunchKernelGGL((k_sweep<KPL, 1, false>), grid, dim3(64 * kSlots), lds, st, a);
    }
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
    require(m <= 253, "hinge sweep supports up to 253 basis functions");
    const int PK = m + 3 <= 32 ? 2 : 1;  // candidates per wave
    const int KPL = PK == 2 ? 1 : (m + 3 + 63) / 64;
    const int L2 = KPL * (64 / PK);      // state lanes per candidate
    const size_t lds = sweep_lds(m + 2 * kSlots + 2);
    {
        int dev = 0, lmax = 0;
        MH_CHECK(hipGetDevice(&dev));
        MH_CHECK(hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        require(lds <= (size_t)lmax, "hinge sweep: basis too large for the device's LDS");
    }
    // about 4096 blocks (16 per CU), chunks of at least 512 sorted rows
    const int pgroups = (maxnp + kSlots - 1) / kSlots;
    const int want = std::max(1, (4096 + ng * pgroups - 1) / (ng * pgroups));
    const int maxc = (int)std::max<size_t>(1, tail / 512);
    const int C = std::min(want, maxc);
    const size_t clen = (tail + C - 1) / C;

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
        for (int pass = 0; pass < 2; ++pass) {
            const bool eval = pass == 1;
            if (eval) {
                const size_t threads = (size_t)P * L2;
                hipLaunchKernelGGL(k_sweep_carry, dim3((threads + 255) / 256), dim3(256), 0, st, a,
                                   dcgrp.p, P);
                MH_LAUNCH_CHECK();
            }
            switch (KPL) {
            case 1: launch_sweep<1>(a, grid, lds, st, eval, PK); break;
            case 2: launch_sweep<2>(a, grid, lds, st, eval, PK); break;
            case 3: launch_sweep<3>(a, grid, lds, st, eval, PK); break;
            default: launch_sweep<4>(a, grid, lds, st, eval, PK); break;
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
