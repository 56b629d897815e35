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
#include <cstdlib>

namespace mh {
namespace {

/* ---------------------------------------------------------------------- */
/* Stable descending sort of X columns: LSD radix, 8-bit digits, a batch   */
/* of columns per launch (blockIdx.y = the column's place in the batch)    */
/* ---------------------------------------------------------------------- */
constexpr int kSortTile = 4096;  // 256 threads x 16 elements

/* Ascending key of the descending order; -0.0 and +0.0 map to one key. */
__device__ inline unsigned desc_key(float v)
{
    unsigned u = __float_as_uint(v == 0.0f ? 0.0f : v);
    u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);  // ascending order of v
    return ~u;                                       // descending order of v
}

/* Column y of the batch: keys and values of X column cols[y] at key + 2 n y, val + 2 n y. */
__global__ __launch_bounds__(256) void k_sort_keys(const float *__restrict__ X, size_t ldx, const int *__restrict__ cols,
                                                   size_t n, unsigned *__restrict__ key, int *__restrict__ val)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const size_t y = blockIdx.y;
    key[2 * n * y + i] = desc_key(X[(size_t)cols[y] * ldx + i]);
    val[2 * n * y + i] = (int)i;
}

/* Digit histogram per tile: hist[y][digit][tile]. */
__global__ __launch_bounds__(256) void k_sort_hist(const unsigned *__restrict__ key, size_t n, int half, int shift,
                                                   int nblocks, unsigned *__restrict__ hist)
{
    __shared__ unsigned h[256];
    const size_t y = blockIdx.y;
    key += 2 * n * y + (size_t)half * n;
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t b0 = (size_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortTile / 256; ++r) {
        const size_t i = b0 + (size_t)r * 256 + threadIdx.x;
        if (i < n) atomicAdd(&h[(key[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(y * 256 + threadIdx.x) * nblocks + blockIdx.x] = h[threadIdx.x];
}

/* Exclusive scan of each column's digit-major histogram (one block per column). */
__global__ __launch_bounds__(1024) void k_sort_scan(unsigned *__restrict__ hist, int count)
{
    __shared__ unsigned sh[1024];
    hist += (size_t)blockIdx.x * count;
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

/* Stable scatter from half `half` to the other: rank within the tile = earlier elements with the
 * same digit. */
__global__ __launch_bounds__(256) void k_sort_scatter(unsigned *__restrict__ key, int *__restrict__ val, size_t n,
                                                      int half, int shift, int nblocks,
                                                      const unsigned *__restrict__ offs)
{
    __shared__ unsigned base[256];
    __shared__ unsigned wcnt[4][256];
    const size_t y = blockIdx.y;
    const unsigned *kin = key + 2 * n * y + (size_t)half * n;
    const int *vin = val + 2 * n * y + (size_t)half * n;
    unsigned *kout = key + 2 * n * y + (size_t)(1 - half) * n;
    int *vout = val + 2 * n * y + (size_t)(1 - half) * n;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    base[t] = offs[(y * 256 + t) * nblocks + blockIdx.x];
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

/* The sorted rows of column cols[y] (half 0 of its values) into order[xc][.], and the deltas
 * d[i] = f32(xs[k[i-1]] - xs[k[i]]) on the scaled column (marsalgo.cc:526-529), d[0] = 0. */
__global__ __launch_bounds__(256) void k_sort_out(const int *__restrict__ val, const float *__restrict__ X, size_t ldx,
                                                  const float *__restrict__ scale, const int *__restrict__ cols,
                                                  size_t n, int *__restrict__ order, float *__restrict__ delta)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const size_t y = blockIdx.y;
    const int xc = cols[y];
    const int *v = val + 2 * n * y;
    const float *x = X + (size_t)xc * ldx;
    const float s = scale[xc];
    const int k = v[i];
    order[(size_t)xc * n + i] = k;
    float d = 0.0f;
    if (i > 0) d = __fsub_rn(__fmul_rn(x[v[i - 1]], s), __fmul_rn(x[k], s));
    delta[(size_t)xc * n + i] = d;
}

/* ---------------------------------------------------------------------- */
/* Sweep                                                                  */
/* ---------------------------------------------------------------------- */
struct SweepArgs {
    const float *Bo, *B, *rbuf, *y, *X, *s;
    size_t ldb, ldr, n;
    const double *ybo, *ybx;
    const int *grp;      // per X-column group: xc, first candidate, count
    const int *cgrp;     // group per candidate
    const int *cpar;     // parent per candidate
    const int *order;    // [px][n]
    const float *delta;  // [px][n]
    // sorted sweep: X column xc's replica block rep + xc * rstride holds [x | y | B (mcap) | Bo (mcap)]
    // in the column's order (ld = ldx); candidate c's normalised column, in the same order, is hr + c * ldx
    const float *rep, *hr, *zero;   // zero: a column of ldx zeros (runs of absent columns)
    size_t rstride, ldx;
    int mcap;
    int m, L2, C;        // basis count, covariate slots per candidate, chunks
    size_t tail, head, span, clen;
    double tol;
    double *sum;         // [P][C][L2][2]   chunk summaries, then carry-ins (f, g)
    double *ksum;        // [P][C]          chunk k, then carry-in k
    double *dsum;        // [ngroups][C]    chunk delta sums
    double *best;        // [P][C]
    int *besti;          // [P][C]
    double eps;          // > 0: also bound each knot's score from above (the precision form's screen)
    double *bestub;      // [P][C]  largest upper bound over the chunk's knots (eps > 0)
};

constexpr int kTR = 16;    // sorted rows per tile

/*
 * Lane layout of one candidate: G lanes, each holding J consecutive covariate
 * slots u = t J + s (t = lane in the group): u = 0 the parent b, u = 1 y,
 * u = 2 the candidate's Bx column, u = 3 .. m+2 the orthonormal basis, the
 * rest zero padding up to PT = G J slots.
 */
/* Slots per lane: 8 or 12, with 4 lanes per candidate up to 48 slots (blocks of 16 or 32
 * candidates), then 8, 16, 32 or 64 lanes; blocks of 8 candidates keep 8 lanes below 64 slots.
 * 512 slots take a whole wave per candidate. */
constexpr int sweep_j(int pt, int nc)
{
    return pt == 32 ? (nc >= 16 ? 8 : 4) : (pt == 48 ? (nc >= 16 ? 12 : 6) : (pt == 96 || pt == 192 ? 12 : 8));
}

template <int PT, int NC>
struct SweepShape {
    static constexpr int J = sweep_j(PT, NC);
    static constexpr int G = PT / J;                    // lanes per candidate
    static constexpr int CW = 64 / G;                   // candidates per wave
    static constexpr int NW = NC / CW;                  // waves per block
    static constexpr int RPS = 64 * NW / PT;            // gather: tile rows per block-wide load
    static_assert(G <= 64 && NW <= 16, "a candidate within one wave, at most 1024 threads per block");
};

/* v from the lane DPP control CTRL names (a permutation within each row of 16 lanes). */
template <int CTRL>
__device__ __forceinline__ double dpp_perm(double v)
{
    const unsigned long long x = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)x, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(x >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

/* Sum over aligned groups of G lanes; every lane of a group gets the same bits. */
template <int G>
__device__ __forceinline__ double group_sum(double v)
{
    v = v + dpp_perm<0xB1>(v);   // quad_perm [1,0,3,2]
    v = v + dpp_perm<0x4E>(v);   // quad_perm [2,3,0,1]
    if (G >= 8) v = v + dpp_perm<0x141>(v);  // row_half_mirror: lane i <-> 7-i
    if (G >= 16) v = v + dpp_perm<0x140>(v);  // row_mirror: lane i <-> 15-i
    if (G >= 32) v = v + __shfl_xor(v, 16, 64);
    if (G >= 64) v = v + __shfl_xor(v, 32, 64);
    return v;
}

/*
 * The sorted sweep's tile loads: a tile of TR sorted rows is PT slot columns (y and the Bo
 * columns of X column xc's replica; the others a zero column) and, per candidate, b (its
 * parent's B column) and Bx (its normalised column), each a float4 run of 4 rows per lane
 * (NT threads per block).  load() fetches tile tt into registers, store() puts them into one
 * LDS buffer as FP64: sq[row][slot] and sbx[row][candidate] = (b, Bx).
 */
template <int PT, int NC, int NT, int TR = kTR, int NS = 1>
struct SortedRuns {
    static constexpr int QC = TR / 4;                         // float4 runs per column and tile
    static constexpr int NQ = (PT * QC + NT - 1) / NT;        // slot runs per lane
    static constexpr int NQC = (2 * NC * QC + NT - 1) / NT;   // candidate runs per lane
    static constexpr bool QPART = NQ * NT > PT * QC;          // some lanes have no slot run
    static constexpr bool CPART = NQC * NT > 2 * NC * QC;
    const float *qs[NQ], *cq[NQC];
    int qcol[NQ], qq[NQ], cj[NQC], cqq[NQC];
    bool cisb[NQC];
    float4 qv[NS][NQ], cv[NS][NQC];   // NS register sets: loads of NS tiles in flight
    size_t r0, r1;

    /* Runs are addressed from the chunk start r0 (a multiple of 16); candidates c0 + sbase + j,
     * j < nact, of X column xc. */
    __device__ void init(const SweepArgs &a, int tid, int xc, int m, int c0, int sbase, int nact, size_t cr0,
                         size_t cr1)
    {
        r0 = cr0;
        r1 = cr1;
        const float *blk = a.rep + (size_t)xc * a.rstride;  // [x | y | B (mcap) | Bo (mcap)]
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const int e = tid + NT * i;
            const int col = e / QC;
            qcol[i] = e < PT * QC ? col : -1;
            qq[i] = e % QC;
            const float *p = a.zero;
            if (col == 1)
                p = blk + a.ldx;
            else if (col >= 3 && col < m + 3)
                p = blk + (size_t)(2 + a.mcap + col - 3) * a.ldx;
            qs[i] = p + r0 + 4 * qq[i];
        }
#pragma unroll
        for (int i = 0; i < NQC; ++i) {
            const int e = tid + NT * i;
            const bool ok = e < 2 * NC * QC;
            cisb[i] = e < NC * QC;
            const int rem = e % (NC * QC);
            cj[i] = ok ? rem / QC : -1;
            cqq[i] = rem % QC;
            const float *p = a.zero;
            if (ok && cj[i] < nact) {
                const int cc = c0 + sbase + cj[i];
                p = cisb[i] ? blk + (size_t)(2 + a.cpar[cc]) * a.ldx : a.hr + (size_t)cc * a.ldx;
            }
            cq[i] = p + r0 + 4 * cqq[i];
        }
    }
    /* Tile tt into register set S; rows at or past r1 (only in the chunk's last tile) read as zero. */
    template <int S = 0>
    __device__ void load(int tt)
    {
        const size_t off = (size_t)tt * TR;
#pragma unroll
        for (int i = 0; i < NQ; ++i) qv[S][i] = *(const float4 *)(qs[i] + off);
#pragma unroll
        for (int i = 0; i < NQC; ++i) cv[S][i] = *(const float4 *)(cq[i] + off);
        const size_t tb = r0 + off;
        if (tb + TR > r1) {   // (uniform)
            auto cut = [&](float4 v, int q) {
                const size_t r = tb + 4 * q;
                return float4{r < r1 ? v.x : 0.f, r + 1 < r1 ? v.y : 0.f, r + 2 < r1 ? v.z : 0.f,
                              r + 3 < r1 ? v.w : 0.f};
            };
#pragma unroll
            for (int i = 0; i < NQ; ++i) qv[S][i] = cut(qv[S][i], qq[i]);
#pragma unroll
            for (int i = 0; i < NQC; ++i) cv[S][i] = cut(cv[S][i], cqq[i]);
        }
    }
    template <int S = 0, int SQ = PT + 2>
    __device__ void store(double (*sq)[SQ], double2 (*sbx)[NC])
    {
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            if (QPART && qcol[i] < 0) continue;
            const int rr = 4 * qq[i];
            sq[rr][qcol[i]] = (double)qv[S][i].x;
            sq[rr + 1][qcol[i]] = (double)qv[S][i].y;
            sq[rr + 2][qcol[i]] = (double)qv[S][i].z;
            sq[rr + 3][qcol[i]] = (double)qv[S][i].w;
        }
#pragma unroll
        for (int i = 0; i < NQC; ++i) {
            if (CPART && cj[i] < 0) continue;
            const int rr = 4 * cqq[i];
            double *dst = cisb[i] ? &sbx[rr][cj[i]].x : &sbx[rr][cj[i]].y;
            dst[0] = (double)cv[S][i].x;
            dst[2 * NC] = (double)cv[S][i].y;
            dst[4 * NC] = (double)cv[S][i].z;
            dst[6 * NC] = (double)cv[S][i].w;
        }
    }
};

/*
 * One block = NC candidates (parents c0.. of one X column) swept over the
 * sorted rows [r0, r1) of its chunk (marsalgo.cc:576-649).  Tiles of 16 sorted
 * rows go to LDS in FP64: row-major [0 | y | 0 | Bo row | 0...] (PT values,
 * rows padded by 2), (b, Bx) per candidate, and the deltas d.
 * SORTED: the tiles are read from the X column's sorted replicas (y, B, Bo in
 * the column's order) and from the candidates' columns that the projection
 * produced in that order: every source is a contiguous run of 16 rows (float4
 * per lane; absent columns read a zero column).  Otherwise each tile row is
 * gathered from the row-major B, Bo and candidate columns through the sort
 * order.  The LDS buffers are doubled: while the unrolled recurrence runs over
 * tile t, the global loads of tile t+1 are in flight (sorted at 32 slots: of
 * tiles t+1 and t+2, in two register sets; the deltas of both in every form).
 * Rows past the chunk end are
 * zero (d = b = 0 leaves every accumulator unchanged), so every tile runs 16
 * rows; chunks start at multiples of 16, so only the last tile before the tail
 * is partial.  Grid rows and nonzero deltas of a tile are bit masks in scalar
 * registers.  At a grid row each lane sums its slots' f.Y (Y = -1 on the y
 * slot, so the sum is fy - w) and k - |f|^2 over the bx and basis slots, the
 * group reduces both with DPP, and the quotient is formed only where it can
 * exceed the chunk's best so far.
 */
template <int PT, int NC, bool EVAL, bool SORTED>
__global__ __launch_bounds__((64 * SweepShape<PT, NC>::NW)) void k_sweep(SweepArgs a)
{
    using S = SweepShape<PT, NC>;
    constexpr int J = S::J, G = S::G, CW = S::CW, NW = S::NW, RPS = S::RPS;
    constexpr int NT = 64 * NW;                            // threads per block
    static_assert(SORTED || (RPS > 0 && kTR % RPS == 0 && 64 * NW == RPS * PT), "tile rows");
    static_assert(J % 2 == 0, "slot pairs");
    constexpr int NB = SORTED ? 1 : kTR / RPS;             // gather: Bo-row loads per lane and tile
    constexpr int NPC = (2 * NC * kTR + NT - 1) / NT;      // gather: per-candidate loads per lane and tile
    constexpr bool PPART = NPC * NT > 2 * NC * kTR;        // (64 lanes per candidate: half the threads load)
    constexpr int SQ = PT + 2;                             // LDS row stride of the y / Bo tile
    __shared__ __attribute__((aligned(16))) double sq[2][kTR][SQ];
    __shared__ double2 sbx[2][kTR][NC];                    // (b, Bx) per candidate
    __shared__ double sd[2][kTR];                          // d
    __shared__ int srow[2][kTR];                           // gather: sorted rows' memory rows

    const int m = a.m;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int t = lane % G, grp = lane / G;
    const int chunk = blockIdx.x, pg = blockIdx.y, gi = blockIdx.z;
    const int xc = a.grp[3 * gi], c0 = a.grp[3 * gi + 1], np = a.grp[3 * gi + 2];
    const int sbase = pg * NC;                     // first parent slot of this block
    if (sbase >= np) return;                       // whole block idle (uniform)
    const int nact = min(NC, np - sbase);
    const int slot = wv * CW + grp;                // this lane's candidate slot
    const bool active = slot < nact;
    const bool wave_active = wv * CW < nact;
    const int c = c0 + sbase + (active ? slot : 0);
    const size_t r0 = (size_t)chunk * a.clen;
    const size_t r1 = min(a.tail, r0 + a.clen);
    const int *ord = a.order + (size_t)xc * a.n;
    const float *dl = a.delta + (size_t)xc * a.n;
    const int T = r1 > r0 ? (int)((r1 - r0 + kTR - 1) / kTR) : 0;
    auto tile_rows = [&](int tt) { return (int)min((size_t)kTR, r1 - (r0 + (size_t)tt * kTR)); };

    // ---- gather sources (fixed over tiles)
    const int pos = tid % PT;   // tile column
    const int ro = tid / PT;    // row within a block-wide load
    const float *bp = nullptr;
    size_t bld = 0;
    const float *pp[NPC];
    size_t pld[NPC];
    int prr[NPC], pj[NPC];
    bool pisb[NPC], pok[NPC];
    // sorted: the tile loads, two tiles ahead at 32 slots (wider tiles: one, which keeps their occupancy)
    constexpr int NSR = SORTED && PT <= 32 ? 2 : 1;
    SortedRuns<PT, NC, NT, kTR, NSR> R;
    if constexpr (SORTED) {
        R.init(a, tid, xc, m, c0, sbase, nact, r0, r1);
    } else {
        if (pos == 1) {
            bp = a.y;
            bld = 1;
        } else if (pos >= 3 && pos < m + 3) {
            bp = a.Bo + (pos - 3);
            bld = a.ldb;
        }
#pragma unroll
        for (int i = 0; i < NPC; ++i) {
            const int e = tid + NT * i;
            pok[i] = !PPART || e < 2 * NC * kTR;
            pisb[i] = e < NC * kTR;
            const int rem = e % (NC * kTR);
            prr[i] = rem / NC;
            pj[i] = rem % NC;
            pp[i] = nullptr;
            pld[i] = 0;
            if (pok[i] && pj[i] < nact) {
                if (pisb[i]) {
                    pp[i] = a.B + a.cpar[c0 + sbase + pj[i]];
                    pld[i] = a.ldb;
                } else {
                    pp[i] = a.rbuf + (c0 + sbase + pj[i]);
                    pld[i] = a.ldr;
                }
            }
        }
    }
    // ---- lane state
    double f[J], g[J], Y[J];
    const double isb = t == 0 ? 1.0 : 0.0;   // this lane holds the b, y and bx slots
    double k2 = 0.0;
#pragma unroll
    for (int s = 0; s < J; ++s) {
        const int u = t * J + s;
        f[s] = g[s] = 0.0;
        Y[s] = u == 1 ? -1.0 : (u == 2 ? a.ybx[c] : ((u >= 3 && u <= m + 2) ? a.ybo[u - 3] : 0.0));
    }
    if (EVAL && active) {
        const double *src = a.sum + ((size_t)c * a.C + chunk) * a.L2 * 2;
#pragma unroll
        for (int s = 0; s < J; ++s) {
            f[s] = src[2 * (t * J + s)];
            g[s] = src[2 * (t * J + s) + 1];
        }
        k2 = a.ksum[(size_t)c * a.C + chunk];
    }

    // ---- grid rows and the tie-run tracker
    const size_t g0 = a.head + 1;
    size_t next = g0;  // next grid row at or after the current tile
    if (r0 > g0) next = g0 + ((r0 - g0 + a.span - 1) / a.span) * a.span;
    int lastnz = -1;   // last sorted row before the current tile with a nonzero delta
    if (EVAL && r0 < r1) {   // (chunks past the tail are empty: r0 may exceed n)
        const size_t lo = r0 > a.span ? r0 - a.span : 0;
        int mine = -1;
        for (size_t l = lo + lane; l < r0; l += 64)
            if (dl[l] != 0.0f) mine = (int)l;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mine = max(mine, __shfl_xor(mine, o, 64));
        lastnz = mine;
    }
    double bsse = 0.0, bub = 0.0;
    int bidx = -1;
    auto grid_mask = [&](int tt) {
        const size_t tb = r0 + (size_t)tt * kTR, te = tb + tile_rows(tt);
        unsigned km = 0;
        for (; next < te; next += a.span) km |= 1u << (unsigned)(next - tb);
        return km;
    };

    // ---- prologue: tile 0 into buffer 0 (gather: with the row indices of tiles 0 and 1)
    if constexpr (!SORTED) {
        if (tid < kTR) {
            if (T > 0 && tid < tile_rows(0)) srow[0][tid] = ord[r0 + tid];
            if (T > 1 && tid < tile_rows(1)) srow[1][tid] = ord[r0 + kTR + tid];
        }
        __syncthreads();
    }
    unsigned kmask = 0, nzmask = 0;
    if (T > 0) {
        const int nr = tile_rows(0);
        if constexpr (SORTED) {
            R.template load<0>(0);
            R.template store<0>(sq[0], sbx[0]);
        } else {
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int rr = i * RPS + ro;
                const bool in = rr < nr && bp;
                sq[0][rr][pos] = in ? (double)bp[(size_t)srow[0][rr] * bld] : 0.0;
            }
#pragma unroll
            for (int i = 0; i < NPC; ++i) {
                if (PPART && !pok[i]) continue;
                const bool in = prr[i] < nr && pp[i];
                const double v = in ? (double)pp[i][(size_t)srow[0][prr[i]] * pld[i]] : 0.0;
                if (pisb[i]) sbx[0][prr[i]][pj[i]].x = v;
                else sbx[0][prr[i]][pj[i]].y = v;
            }
        }
        const float dv = lane < nr ? dl[r0 + lane] : 0.0f;
        if (tid < kTR) sd[0][tid] = (double)dv;
        if (EVAL) {
            kmask = grid_mask(0);
            nzmask = (unsigned)__ballot(dv != 0.0f);
        }
    }
    // deltas of the tiles in flight (sorted: with the tile loads, two tiles ahead)
    float ndv[2] = {0.0f, 0.0f};
    if (T > 1) {
        if constexpr (NSR == 2) R.template load<1>(1);
        ndv[1] = lane < tile_rows(1) ? dl[r0 + kTR + lane] : 0.0f;
    }
    __syncthreads();

    // tile tt: LDS buffer tt & 1; tile tt + 1's loads in register set S
    auto body = [&](int tt, auto sc) {
        constexpr int S = decltype(sc)::value, S2 = S ^ 1;
        const int cur = tt & 1, nxt = cur ^ 1;
        const size_t tb0 = r0 + (size_t)tt * kTR;
        const int nr1 = tt + 1 < T ? tile_rows(tt + 1) : 0;
        const int nr2 = tt + 2 < T ? tile_rows(tt + 2) : 0;
        // ---- prefetch tile t+1 (gather: its row indices are in srow[nxt]; and tile t+2's indices)
        float bov[NB], pcv[NPC];
        int nrow = 0;
        if constexpr (NSR == 2) {
            if (nr2 > 0) R.template load<S2>(tt + 2);
        } else if constexpr (SORTED) {
            if (nr1 > 0) R.template load<0>(tt + 1);
        } else {
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int rr = i * RPS + ro;
                const bool in = rr < nr1 && bp;
                bov[i] = in ? bp[(size_t)srow[nxt][rr] * bld] : 0.0f;
            }
#pragma unroll
            for (int i = 0; i < NPC; ++i) {
                const bool in = prr[i] < nr1 && pp[i];
                pcv[i] = in ? pp[i][(size_t)srow[nxt][prr[i]] * pld[i]] : 0.0f;
            }
            if (tid < nr2) nrow = ord[tb0 + 2 * kTR + tid];
        }
        const float nd = ndv[S];
        ndv[S2] = lane < nr2 ? dl[tb0 + 2 * kTR + lane] : 0.0f;

        // ---- the recurrence over tile t (marsalgo.cc:576-649)
        if (wave_active) {
#pragma unroll
            for (int rr = 0; rr < kTR; ++rr) {
                const double d = sd[cur][rr];
                const double2 bx = sbx[cur][rr][slot];   // (b, Bx)
                double q[J];
#pragma unroll
                for (int s = 0; s < J; s += 2) {
                    const double2 v2 = *(const double2 *)&sq[cur][rr][t * J + s];
                    q[s] = v2.x;
                    q[s + 1] = v2.y;
                }
                q[0] = fma(isb, bx.x, q[0]);   // b and Bx on the group's first lane (its tile slots 0, 2 are 0)
                q[2] = fma(isb, bx.y, q[2]);
                const double fb = f[0];
#pragma unroll
                for (int s = 0; s < J; ++s) {
                    f[s] = fma(d, g[s], f[s]);
                    g[s] = fma(bx.x, q[s], g[s]);
                }
                // k += d^2 g_b + 2 d f_b with the b slot's (f, g) before the row = d (f_b + f_b')
                k2 = fma(d, fb + f[0], k2);
                if (EVAL && ((kmask >> rr) & 1u)) {
                    const size_t i = tb0 + rr;
                    const unsigned nzb = nzmask & (0xffffffffu >> (31 - rr));
                    const long long last = nzb ? (long long)(tb0 + 31 - __clz(nzb)) : (long long)lastnz;
                    const bool dup = (i >= g0 + a.span) && (last <= (long long)(i - a.span));
                    if (!dup) {
                        // fy - w over the y, bx and basis slots; k - |f|^2 over the bx and basis slots
                        const double f0 = f[0] * (1.0 - isb), f1 = f[1] * (1.0 - isb);
                        double uw = 0.0, ff = f0 * f0;
                        ff = fma(f1, f1, ff);
#pragma unroll
                        for (int s = 2; s < J; ++s) ff = fma(f[s], f[s], ff);
#pragma unroll
                        for (int s = 0; s < J; ++s) uw = fma(f[s], Y[s], uw);
                        double den = fma(isb, k2, -ff);
                        uw = group_sum<G>(uw);
                        den = group_sum<G>(den);
                        // the quotient only where it can beat the best so far (the product test has
                        // a margin far beyond its rounding); the winner's value is the quotient
                        const double u2 = uw * uw, dt = den + a.tol;
                        if (a.eps > 0) {
                            if (den > a.tol) {
                                // the screen: relative errors eps of the covariates move den = k - |f|^2
                                // by at most eps (k + |f|^2) and fy - w by eps sum |f_s Y_s|, which
                                // bounds the score.  Its quotients are reciprocals with one Newton
                                // step (relative error far below 2^-38 from any seed better than
                                // 2^-20), and the bound carries a relative margin of 2^-36 for them;
                                // the screen's own best is approximate (exact rescoring follows).
                                auto rcp = [](double x) {
                                    const double r = __builtin_amdgcn_rcp(x);
                                    return fma(r, fma(-x, r, 1.0), r);
                                };
                                const double sse = u2 * rcp(dt);
                                if (sse > bsse) {
                                    bsse = sse;
                                    bidx = (int)i;
                                }
                                double ua = 0.0;
#pragma unroll
                                for (int s = 0; s < J; ++s) ua = fma(fabs(f[s]), fabs(Y[s]), ua);
                                const double ffs = group_sum<G>(ff);
                                ua = group_sum<G>(ua);
                                const double rel = a.eps * ((den + 2.0 * ffs) * rcp(fmax(den, a.tol)) +
                                                            2.0 * ua * rcp(fmax(fabs(uw), 1e-300)));
                                bub = fmax(bub, sse * (1.0 + rel) * (1.0 + 0x1p-36));
                            }
                        } else if (den > a.tol && u2 >= bsse * dt * (1.0 - 1e-12)) {
                            const double sse = u2 / dt;
                            if (sse > bsse) {
                                bsse = sse;
                                bidx = (int)i;
                            }
                        }
                    }
                }
            }
        }
        if (EVAL) {
            if (nzmask) lastnz = (int)(tb0 + 31 - __clz(nzmask));
            if (tt + 1 < T) {
                kmask = grid_mask(tt + 1);
                nzmask = (unsigned)__ballot(nd != 0.0f);
            }
        }
        // buffer nxt last held tile t-1, which every wave finished before the
        // previous iteration's barrier: store tile t+1 into it
        if constexpr (SORTED) {
            if (nr1 > 0) R.template store<NSR == 2 ? S : 0>(sq[nxt], sbx[nxt]);
        } else {
#pragma unroll
            for (int i = 0; i < NB; ++i) sq[nxt][i * RPS + ro][pos] = (double)bov[i];
#pragma unroll
            for (int i = 0; i < NPC; ++i) {
                if (PPART && !pok[i]) continue;
                if (pisb[i]) sbx[nxt][prr[i]][pj[i]].x = (double)pcv[i];
                else sbx[nxt][prr[i]][pj[i]].y = (double)pcv[i];
            }
        }
        if (tid < kTR) {
            sd[nxt][tid] = (double)nd;
            if (!SORTED && tid < nr2) srow[cur][tid] = nrow;
        }
        __syncthreads();
    };
    for (int tt = 0; tt < T;) {
        body(tt, std::integral_constant<int, 1>{});
        if (++tt >= T) break;
        body(tt, std::integral_constant<int, 0>{});
        ++tt;
    }
    if (!EVAL && pg == 0 && wv == 0) {
        // the chunk's delta sum (carry shift), lane-strided then a fixed butterfly
        double dacc = 0.0;
        for (size_t l = r0 + lane; l < r1; l += 64) dacc += (double)dl[l];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) dacc += __shfl_xor(dacc, o, 64);
        if (lane == 0) a.dsum[(size_t)gi * a.C + chunk] = dacc;
    }
    if (!active) return;
    const size_t cs = (size_t)c * a.C + chunk;
    if (EVAL) {
        if (t == 0) {
            a.best[cs] = bsse;
            a.besti[cs] = bidx;
            if (a.eps > 0) a.bestub[cs] = bub;
        }
    } else {
        double *dst = a.sum + cs * a.L2 * 2;
#pragma unroll
        for (int s = 0; s < J; ++s) {
            dst[2 * (t * J + s)] = f[s];
            dst[2 * (t * J + s) + 1] = g[s];
        }
        if (t == 0) a.ksum[cs] = k2;
    }
}

/*
 * Chunk summaries of the sorted sweep (its first pass) on FP64 MFMA.  Run from the zero state,
 * the recurrence f += d g, g += b q over the chunk's rows l ends at
 *     g = G = sum_l b_l q_l,      f = sum_l b_l q_l (D - W_l) = D G - H,   H = sum_l b_l W_l q_l,
 * W_l = the chunk's deltas up to and including row l, D = all of them.  For the slots shared by
 * the X column's parents (y, the Bo columns) G and H are two contractions [b; b W] (candidates x
 * rows) x q (rows x slots): v_mfma_f64_16x16x4f64 with candidates as M, slots as N and 4 sorted
 * rows as K, wave w taking slots 16 w .. 16 w + 15.  The candidate's own slots have the same
 * closed forms with q = b (slot 0) and q = Bx (slot 2), and k = sum_l b_l^2 (D - W_l)^2; their
 * five sums (b^2, b^2 W, b Bx, b Bx W, b^2 (D - W)^2, every product of two f32 values exact)
 * are spread over the block, one candidate per thread and a fixed set of rows each, then
 * reduced in a fixed order.  D, needed from the first tile on, is summed over the chunk's
 * deltas before the tiles.  Tiles as in k_sweep (16 rows, LDS in FP64, double-buffered); the
 * prefix W of each tile is scanned by wave 0.  (Measured slower: a separate wave for the
 * candidates' own slots, and tile loads issued three tiles ahead.)
 */
template <int PT, int NC>
__global__ __launch_bounds__(64 * (PT / 16)) void k_sweep_sum(SweepArgs a)
{
    constexpr int NW = PT / 16, NT = 64 * NW, SQ = PT + 2, MT = (NC + 15) / 16;
    constexpr int NP = (NC * kTR + NT - 1) / NT;   // own-slot (candidate, row) pairs per thread and tile
    static_assert(NT % NC == 0, "a thread's pairs share one candidate");
    static_assert(5 * NW * NC <= 2 * kTR * SQ, "the own-slot reduction fits the tile buffers");
    __shared__ __attribute__((aligned(16))) double sq[2][kTR][SQ];
    __shared__ double2 sbx[2][kTR][NC];
    __shared__ double sw[2][kTR];
    __shared__ double sdp[NW];

    const int m = a.m;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, kr = lane >> 4, kc = lane & 15;
    const int chunk = blockIdx.x, pg = blockIdx.y, gi = blockIdx.z;
    const int xc = a.grp[3 * gi], c0 = a.grp[3 * gi + 1], np = a.grp[3 * gi + 2];
    const int sbase = pg * NC;
    if (sbase >= np) return;   // whole block idle (uniform)
    const int nact = min(NC, np - sbase);
    const size_t r0 = (size_t)chunk * a.clen;
    const size_t r1 = min(a.tail, r0 + a.clen);
    const float *dl = a.delta + (size_t)xc * a.n;
    const int T = r1 > r0 ? (int)((r1 - r0 + kTR - 1) / kTR) : 0;
    auto tile_rows = [&](int tt) { return (int)min((size_t)kTR, r1 - (r0 + (size_t)tt * kTR)); };

    // D: the chunk's delta sum (the carry's shift and the own slots' k), lane-strided per
    // thread, a fixed butterfly per wave, the waves added in order
    double Dc;
    {
        double dacc = 0.0;
        for (size_t l = r0 + tid; l < r1; l += NT) dacc += (double)dl[l];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) dacc += __shfl_xor(dacc, o, 64);
        if (lane == 0) sdp[wv] = dacc;
        __syncthreads();
        Dc = sdp[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) Dc += sdp[w];
    }
    // own slots: this thread's candidate and its rows of each tile
    const int oj = tid % NC;
    double oGb = 0.0, oHb = 0.0, oGx = 0.0, oHx = 0.0, oK = 0.0;

    // tiles and their deltas are fetched two tiles ahead (register sets 0 and 1)
    SortedRuns<PT, NC, NT, kTR, 2> R;
    R.init(a, tid, xc, m, c0, sbase, nact, r0, r1);
    v4d accG[MT], accH[MT];
#pragma unroll
    for (int u = 0; u < MT; ++u) accG[u] = accH[u] = v4d{0, 0, 0, 0};
    double wbase = 0.0;   // W at the end of the previous tile (wave 0)
    float dv[2];          // wave 0, lanes < 16: deltas of the tiles in flight
    auto load_d = [&](int tt) {
        const int nr = tt < T ? tile_rows(tt) : 0;
        return wv == 0 && lane < nr ? dl[r0 + (size_t)tt * kTR + lane] : 0.0f;
    };
    // the running prefix W of a tile's deltas d into sw[b] (wave 0)
    // (the 16 rows are one DPP row of lanes: row_shr moves, which shift in zeros, and a readlane,
    // instead of cross-lane permutes through the LDS unit while the other waves wait)
    static_assert(kTR == 16, "one tile is one DPP row");
    auto deltas = [&](float df, int b) {
        if (wv != 0) return;
        double w = (double)df;
        w += dpp_perm<0x111>(w);   // row_shr:1
        w += dpp_perm<0x112>(w);   // row_shr:2
        w += dpp_perm<0x114>(w);   // row_shr:4
        w += dpp_perm<0x118>(w);   // row_shr:8
        if (lane < kTR) sw[b][lane] = wbase + w;
        const unsigned long long x = __builtin_bit_cast(unsigned long long, w);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, kTR - 1);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), kTR - 1);
        wbase += __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
    };
    if (T > 0) {
        R.template load<0>(0);
        const float d0 = load_d(0);
        R.template store<0>(sq[0], sbx[0]);
        deltas(d0, 0);
        if (T > 1) {
            R.template load<1>(1);
            dv[1] = load_d(1);
        }
    }
    __syncthreads();
    // tile tt: in LDS buffer tt & 1; tile tt + 1 in register set S (loaded a tile earlier)
    auto body = [&](int tt, auto sc) {
        constexpr int S = decltype(sc)::value, S2 = S ^ 1;
        const int cur = tt & 1, nxt = cur ^ 1;
        const bool more = tt + 1 < T;
        if (tt + 2 < T) {
            R.template load<S2>(tt + 2);
            dv[S2] = load_d(tt + 2);
        }
        {
#pragma unroll
            for (int ks = 0; ks < kTR / 4; ++ks) {
                const int row = 4 * ks + kr;   // K step ks takes rows 4 ks .. 4 ks + 3
                const double q = sq[cur][row][16 * wv + kc];
                const double w = sw[cur][row];
#pragma unroll
                for (int u = 0; u < MT; ++u) {
                    const int j = 16 * u + kc;
                    const double b = j < NC ? sbx[cur][row][j < NC ? j : 0].x : 0.0;
                    accG[u] = mfma_f64(b, q, accG[u]);
                    accH[u] = mfma_f64(b * w, q, accH[u]);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const int e = tid + NT * i;
            if (NP * NT > NC * kTR && e >= NC * kTR) break;
            const int rr = e / NC;
            const double2 bx = sbx[cur][rr][oj];   // (b, Bx)
            const double w = sw[cur][rr];
            const double b2 = bx.x * bx.x, bb = bx.x * bx.y, dw = Dc - w;
            oGb += b2;
            oHb = fma(b2, w, oHb);
            oGx += bb;
            oHx = fma(bb, w, oHx);
            oK = fma(b2 * dw, dw, oK);
        }
        if (more) {
            R.template store<S>(sq[nxt], sbx[nxt]);
            deltas(dv[S], nxt);
        }
        __syncthreads();
    };
    for (int tt = 0; tt < T;) {
        body(tt, std::integral_constant<int, 1>{});
        if (++tt >= T) break;
        body(tt, std::integral_constant<int, 0>{});
        ++tt;
    }
    if (pg == 0 && tid == 0) a.dsum[(size_t)gi * a.C + chunk] = Dc;
    {
        // D layout of the FP64 MFMA: lane l holds rows (candidates) l/16 + 4 v, column (slot) l % 16
#pragma unroll
        for (int u = 0; u < MT; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = 16 * u + kr + 4 * v, sl = 16 * wv + kc;
                if (j >= nact || sl == 0 || sl == 2) continue;
                double *dst = a.sum + (((size_t)(c0 + sbase + j) * a.C + chunk) * a.L2 + sl) * 2;
                dst[0] = fma(Dc, accG[u][v], -accH[u][v]);
                dst[1] = accG[u][v];
            }
    }
    {
        // own slots: the threads of one candidate (lanes l = j mod NC of each wave) by a fixed
        // butterfly, then the waves in order through the (now free) tile buffers
        double v[5] = {oGb, oHb, oGx, oHx, oK};
#pragma unroll
        for (int q = 0; q < 5; ++q)
#pragma unroll
            for (int o = 32; o >= NC; o >>= 1) v[q] += __shfl_xor(v[q], o, 64);
        double *red = &sq[0][0][0];   // [5][NW][NC]
        if (lane < NC)
#pragma unroll
            for (int q = 0; q < 5; ++q) red[(q * NW + wv) * NC + lane] = v[q];
        __syncthreads();
        if (tid < nact) {
            double u[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                u[q] = red[(q * NW) * NC + tid];
                for (int w = 1; w < NW; ++w) u[q] += red[(q * NW + w) * NC + tid];
            }
            const size_t cs = (size_t)(c0 + sbase + tid) * a.C + chunk;
            double *dst = a.sum + cs * a.L2 * 2;
            dst[0] = fma(Dc, u[0], -u[1]);
            dst[1] = u[0];
            dst[4] = fma(Dc, u[2], -u[3]);
            dst[5] = u[2];
            a.ksum[cs] = u[4];
        }
    }
}

/* Turn chunk summaries into carry-ins, in chunk order.  One thread per (candidate, lane). */
__global__ void k_sweep_carry(SweepArgs a, const int *__restrict__ cgrp, int P)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = (int)(t / a.L2), k = (int)(t % a.L2);
    if (c >= P) return;
    const int gi = cgrp[c];
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

/* Best knot over chunks (first maximum in row order), hinge = linear + best; with eps > 0 also
 * hub = linear + the largest upper bound over every knot. */
__global__ void k_sweep_final(SweepArgs a, int P, const double *__restrict__ lin, double *__restrict__ hin,
                              int *__restrict__ knot, int *__restrict__ krow, double *__restrict__ hub)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P) return;
    double b = 0.0, ub = 0.0;
    int bi = -1;
    for (int ch = 0; ch < a.C; ++ch) {
        const double v = a.best[(size_t)c * a.C + ch];
        const int vi = a.besti[(size_t)c * a.C + ch];
        if (vi >= 0 && v > b) {
            b = v;
            bi = vi;
        }
        if (hub) ub = fmax(ub, a.bestub[(size_t)c * a.C + ch]);
    }
    if (hub) hub[c] = lin[c] + ub;
    knot[c] = bi;
    if (bi >= 0) {
        hin[c] = lin[c] + b;
        krow[c] = a.order[(size_t)a.grp[3 * a.cgrp[c]] * a.n + bi];
    } else {
        hin[c] = 0.0;
        krow[c] = -1;
    }
}

template <int PT, int NC, bool SORTED>
void launch_sweep_nc(const SweepArgs &a, dim3 grid, hipStream_t st, bool eval)
{
    const dim3 block(64 * SweepShape<PT, NC>::NW);
    if (eval)
        hipLaunchKernelGGL((k_sweep<PT, NC, true, SORTED>), grid, block, 0, st, a);
    else
        hipLaunchKernelGGL((k_sweep<PT, NC, false, SORTED>), grid, block, 0, st, a);
    MH_LAUNCH_CHECK();
}

template <int PT, bool SORTED>
void launch_sweep(const SweepArgs &a, dim3 grid, int nc, hipStream_t st, bool eval)
{
    switch (nc) {
    case 8: launch_sweep_nc<PT, 8, SORTED>(a, grid, st, eval); break;
    case 16: launch_sweep_nc<PT, 16, SORTED>(a, grid, st, eval); break;
    default: launch_sweep_nc<PT, 32, SORTED>(a, grid, st, eval); break;
    }
}

/* Slots per candidate (b, y, Bx and the m basis columns, padded): the sorted sweep has the finer
 * steps (its loads need no row-to-lane fit of the tile width). */
int sweep_slots(int m, bool sorted)
{
    static const int fine[] = {32, 48, 64, 96, 128, 192, 256, 512}, coarse[] = {32, 64, 128, 256, 512};
    for (int pt : sorted ? std::vector<int>(fine, fine + 8) : std::vector<int>(coarse, coarse + 5))
        if (m + 3 <= pt) return pt;
    return 512;
}

/* 512 slots: a wave per candidate, blocks of 8 candidates (16 would spill the scoring pass's
 * registers); both passes run the recurrence (the MFMA summary pass keeps one wave per 16 slots,
 * at most 256 slots per block). */
template <bool SORTED>
void launch_sweep_wide(const SweepArgs &a, dim3 grid, int nc, hipStream_t st, bool eval)
{
    require(nc == 8, "512-slot sweeps take blocks of 8 candidates");
    launch_sweep_nc<512, 8, SORTED>(a, grid, st, eval);
}

template <int PT>
void launch_sweep_sum(const SweepArgs &a, dim3 grid, int nc, hipStream_t st)
{
    const dim3 block(64 * (PT / 16));
    switch (nc) {
    case 8: hipLaunchKernelGGL((k_sweep_sum<PT, 8>), grid, block, 0, st, a); break;
    case 16: hipLaunchKernelGGL((k_sweep_sum<PT, 16>), grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL((k_sweep_sum<PT, 32>), grid, block, 0, st, a); break;
    }
    MH_LAUNCH_CHECK();
}

void launch_sweep_sorted(const SweepArgs &a, dim3 grid, int pt, int nc, hipStream_t st, bool eval)
{
    if (!eval) {  // the summary pass on MFMA
        switch (pt) {
        case 32: launch_sweep_sum<32>(a, grid, nc, st); break;
        case 48: launch_sweep_sum<48>(a, grid, nc, st); break;
        case 64: launch_sweep_sum<64>(a, grid, nc, st); break;
        case 96: launch_sweep_sum<96>(a, grid, nc, st); break;
        case 128: launch_sweep_sum<128>(a, grid, nc, st); break;
        case 192: launch_sweep_sum<192>(a, grid, nc, st); break;
        case 256: launch_sweep_sum<256>(a, grid, nc, st); break;
        default: launch_sweep_wide<true>(a, grid, nc, st, false); break;
        }
        return;
    }
    switch (pt) {
    case 32: launch_sweep<32, true>(a, grid, nc, st, eval); break;
    case 48: launch_sweep<48, true>(a, grid, nc, st, eval); break;
    case 64: launch_sweep<64, true>(a, grid, nc, st, eval); break;
    case 96: launch_sweep<96, true>(a, grid, nc, st, eval); break;
    case 128: launch_sweep<128, true>(a, grid, nc, st, eval); break;
    case 192: launch_sweep<192, true>(a, grid, nc, st, eval); break;
    case 256: launch_sweep<256, true>(a, grid, nc, st, eval); break;
    default: launch_sweep_wide<true>(a, grid, nc, st, eval); break;
    }
}

void launch_sweep_gather(const SweepArgs &a, dim3 grid, int pt, int nc, hipStream_t st, bool eval)
{
    switch (pt) {
    case 32: launch_sweep<32, false>(a, grid, nc, st, eval); break;
    case 64: launch_sweep<64, false>(a, grid, nc, st, eval); break;
    case 128: launch_sweep<128, false>(a, grid, nc, st, eval); break;
    case 256: launch_sweep<256, false>(a, grid, nc, st, eval); break;
    default: launch_sweep_wide<false>(a, grid, nc, st, eval); break;
    }
}

/* Extend the sorted replicas: for each listed (xc, k0), rows i < n of column xc's block get
 * B[:, k] and Bo[:, k] at row order[i] for k0 <= k < m (and x, y when k0 == 0).  The sources are
 * row-major, element (o, k) at sB / sBo + o sld + k - kbase: the basis's row-major copies (ld =
 * ldb, kbase 0), or, for an append's few new columns, a compact staging copy of just those columns
 * (a few MB: a sorted row's random read then hits the Infinity Cache instead of HBM). */
__global__ __launch_bounds__(256) void k_rep_gather(const int *__restrict__ todo, const int *__restrict__ order,
                                                    const float *__restrict__ X, const float *__restrict__ y,
                                                    const float *__restrict__ sB, const float *__restrict__ sBo,
                                                    size_t sld, int kbase, size_t ldx, size_t n, int m, int mcap,
                                                    size_t rstride, float *__restrict__ rep)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int xc = todo[2 * blockIdx.y], k0 = todo[2 * blockIdx.y + 1];
    const size_t o = (size_t)order[(size_t)xc * n + i];
    float *blk = rep + (size_t)xc * rstride;
    if (k0 == 0) {
        blk[i] = X[(size_t)xc * ldx + o];
        blk[ldx + i] = y[o];
    }
    const float *br = sB + o * sld - kbase, *bor = sBo + o * sld - kbase;
    for (int k = k0; k < m; ++k) {
        blk[(size_t)(2 + k) * ldx + i] = br[k];
        blk[(size_t)(2 + mcap + k) * ldx + i] = bor[k];
    }
}

/* Staging copy of basis columns k0 .. k0 + nk - 1: S[o][j] = B[o][k0 + j], S[o][nk + j] = Bo[o][k0 + j]. */
__global__ __launch_bounds__(256) void k_rep_stage(const float *__restrict__ BC, const float *__restrict__ BoC,
                                                   size_t ldx, size_t n, int k0, int nk, float *__restrict__ S)
{
    const size_t o = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= n) return;
    for (int j = 0; j < nk; ++j) {
        S[o * 2 * nk + j] = BC[(size_t)(k0 + j) * ldx + o];
        S[o * 2 * nk + nk + j] = BoC[(size_t)(k0 + j) * ldx + o];
    }
}

} // namespace

/* Make the stable descending order and f32 deltas of fit 0's X columns `xcs` resident (the columns
 * not sorted yet, in batches of at most 64 columns and about 1.5 GB of key / value space). */
static void ensure_sorted(mars_hip_ctx *ctx, const std::vector<int> &xcs)
{
    FitState &fs = ctx->fits[0];
    if (fs.sorted.size() != fs.px) fs.sorted.assign(fs.px, 0);
    std::vector<int> todo;
    for (int xc : xcs)
        if (!fs.sorted[xc]) {
            fs.sorted[xc] = 1;
            todo.push_back(xc);
        }
    if (todo.empty()) return;
    const size_t n = fs.n;
    hipStream_t st = ctx->stream;
    fs.order.ensure(fs.px * n);
    fs.delta.ensure(fs.px * n);
    const int nblocks = (int)((n + kSortTile - 1) / kSortTile);
    const size_t per = std::max<size_t>(1, std::min<size_t>(64, (size_t(3) << 29) / (16 * n + 1)));
    const size_t batch = std::min(per, todo.size());
    DBuf<unsigned> keys, hist;
    DBuf<int> vals, cols;
    keys.ensure(2 * n * batch);
    vals.ensure(2 * n * batch);
    hist.ensure((size_t)256 * nblocks * batch);
    cols.ensure(todo.size());
    h2d(ctx, cols.p, todo.data(), todo.size() * sizeof(int));
    KTimer kt(ctx, K_SORT, 0.0, 16.0 * 4 * n * todo.size());
    for (size_t b0 = 0; b0 < todo.size(); b0 += batch) {
        const unsigned nb = (unsigned)std::min(batch, todo.size() - b0);
        const int *cb = cols.p + b0;
        hipLaunchKernelGGL(k_sort_keys, dim3((unsigned)((n + 255) / 256), nb), dim3(256), 0, st,
                           (const float *)fs.X.p, fs.ldx, cb, n, keys.p, vals.p);
        MH_LAUNCH_CHECK();
        for (int pass = 0; pass < 4; ++pass) {
            const int shift = 8 * pass, half = pass & 1;
            hipLaunchKernelGGL(k_sort_hist, dim3(nblocks, nb), dim3(256), 0, st, (const unsigned *)keys.p, n, half,
                               shift, nblocks, hist.p);
            MH_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_sort_scan, dim3(nb), dim3(1024), 0, st, hist.p, 256 * nblocks);
            MH_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_sort_scatter, dim3(nblocks, nb), dim3(256), 0, st, keys.p, vals.p, n, half, shift,
                               nblocks, (const unsigned *)hist.p);
            MH_LAUNCH_CHECK();
        }
        // after 4 passes the values are back in half 0
        hipLaunchKernelGGL(k_sort_out, dim3((unsigned)((n + 255) / 256), nb), dim3(256), 0, st, (const int *)vals.p,
                           (const float *)fs.X.p, fs.ldx, (const float *)fs.s.p, cb, n, fs.order.p, fs.delta.p);
        MH_LAUNCH_CHECK();
    }
    // the temporaries are released once the stream has drained (queued kernels still use them)
    retire_device(keys.p, keys.cap * sizeof(unsigned));
    retire_device(vals.p, vals.cap * sizeof(int));
    retire_device(hist.p, hist.cap * sizeof(unsigned));
    retire_device(cols.p, cols.cap * sizeof(int));
    keys = DBuf<unsigned>{};
    vals = DBuf<int>{};
    hist = DBuf<unsigned>{};
    cols = DBuf<int>{};
}

void hinge_sweep_list(mars_hip_ctx *ctx, Cands &cs, const double *lin, const double *ybx, double *hin, int *knot,
                      int *krow, int min_span, int end_span, double eps, double *hub, const int *dpar, bool sorted)
{
    FitState &fs = ctx->fits[0];
    const int P = (int)cs.fit.size();
    const size_t n = fs.n;
    const int m = (int)fs.m;
    hipStream_t st = ctx->stream;

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
        c = e;
    }
    const int ng = (int)(grp.size() / 3);
    {
        std::vector<int> gx(ng);
        for (int g = 0; g < ng; ++g) gx[g] = grp[3 * g];
        ensure_sorted(ctx, gx);
    }
    const size_t tail = n - (size_t)end_span;
    require(m <= 509, "hinge sweep supports up to 509 basis functions");
    const int PT = sweep_slots(m, sorted);
    const int L2 = PT;
    // candidates per block: one block takes all parents of an X column, up to 32 (8 with 512 slots)
    const int NC = maxnp <= 8 || PT > 256 ? 8 : (maxnp <= 16 ? 16 : 32);
    const int pgroups = (maxnp + NC - 1) / NC;
    const int wpb = NC / (64 / (PT / sweep_j(PT, NC)));   // waves per block
    // about sweep_waves waves (65536), chunks of at least sweep_min_rows (512) sorted rows; a multiple of 8
    // chunks puts the blocks of one (column, chunk) on one XCD (shared L2).  Chunks start at multiples of
    // the 16-row tile (aligned float4 runs of the sorted sources).
    const int waves = (int)ctx->tune.sweep_waves;
    const int want = std::max(1, (waves + ng * pgroups * wpb - 1) / (ng * pgroups * wpb));
    const int maxc = (int)std::max<size_t>(1, tail / (size_t)ctx->tune.sweep_min_rows);
    int C = std::min(want, maxc);
    if (C >= 8) C -= C % 8;
    const size_t clen = ((tail + C - 1) / C + kTR - 1) / kTR * kTR;

    DBuf<int> dgrp, dcgrp;
    dgrp.ensure(grp.size());
    dcgrp.ensure(P);
    h2d(ctx, dgrp.p, grp.data(), grp.size() * sizeof(int));
    h2d(ctx, dcgrp.p, cgrp.data(), P * sizeof(int));
    ctx->hstate.ensure((size_t)P * C * L2 * 2 + (size_t)P * C + (size_t)ng * C);
    ctx->hbest.ensure((size_t)P * C);
    ctx->hbest_i.ensure((size_t)P * C);
    if (hub) ctx->hbest_ub.ensure((size_t)P * C);

    SweepArgs a{};
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
    a.ybx = ybx;
    a.grp = dgrp.p;
    a.cgrp = dcgrp.p;
    a.cpar = dpar ? dpar : cs.d_par.p;
    a.order = fs.order.p;
    a.delta = fs.delta.p;
    if (sorted) {
        a.rep = ctx->rep.p;
        a.hr = ctx->hr.p;
        a.rstride = ctx->rep_stride;
        a.zero = ctx->rep.p + ctx->rep_stride * fs.px;
        a.ldx = fs.ldx;
        a.mcap = (int)ctx->rep_mcap;
    }
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
    a.eps = hub ? eps : 0.0;
    a.bestub = hub ? ctx->hbest_ub.p : nullptr;

    const dim3 grid(C, pgroups, ng);
    // reference algorithmic work: n (m+1) (4 + 4/min_span) flops per candidate (marsalgo.cc:576-649);
    // bytes, each of the two passes: y, the m basis columns and the deltas per X column, b and Bx
    // per candidate
    {
        KTimer kt(ctx, K_SWEEP, (double)P * (double)tail * (m + 1) * (4.0 + 4.0 / min_span),
                  2.0 * ((double)ng * tail * (4.0 * m + 8.0) + (double)P * tail * 8.0));
        for (int pass = 0; pass < 2; ++pass) {
            const bool eval = pass == 1;
            if (eval) {
                const size_t threads = (size_t)P * L2;
                hipLaunchKernelGGL(k_sweep_carry, dim3((threads + 255) / 256), dim3(256), 0, st, a,
                                   dcgrp.p, P);
                MH_LAUNCH_CHECK();
            }
            if (sorted)
                launch_sweep_sorted(a, grid, PT, NC, st, eval);
            else
                launch_sweep_gather(a, grid, PT, NC, st, eval);
        }
        hipLaunchKernelGGL(k_sweep_final, dim3((P + 255) / 256), dim3(256), 0, st, a, P, lin, hin, knot, krow, hub);
        MH_LAUNCH_CHECK();
    }
    ctx_sync(ctx);
    dgrp.release();
    dcgrp.release();
}

/*
 * The sorted replicas: decided once per context (they take (2 + 2 max_terms) columns of n floats
 * per X column, and the candidates' columns up to max_terms per X column; they are used when both
 * fit in half the free device memory, and never with MARS_HIP_SORTED=0), then brought up to the
 * current basis for the listed X columns.  Returns false when the gather sweep serves instead.
 */
static bool replicas_ready(mars_hip_ctx *ctx, const std::vector<int> &xcs)
{
    FitState &fs = ctx->fits[0];
    hipStream_t st = ctx->stream;
    if (ctx->rep_state == 0) {
        ctx->rep_state = -1;
        const char *env = std::getenv("MARS_HIP_SORTED");
        if (!(env && env[0] == '0')) {
            const size_t mcap = std::max<size_t>(ctx->max_terms, 1);
            const size_t stride = (2 + 2 * mcap) * fs.ldx;
            const double need = 4.0 * (double)stride * fs.px + 4.0 * (double)fs.ldx * fs.px * mcap;
            size_t fr = 0, tot = 0;
            MH_CHECK(hipMemGetInfo(&fr, &tot));
            if (need <= 0.5 * (double)fr) {
                // (one more column of zeros after the blocks: the sweep's runs of absent columns)
                ctx->rep.ensure(stride * fs.px + fs.ldx);
                MH_CHECK(hipMemsetAsync(ctx->rep.p, 0, (stride * fs.px + fs.ldx) * sizeof(float), st));
                ctx->rep_mcap = mcap;
                ctx->rep_stride = stride;
                ctx->rep_m.assign(fs.px, 0);
                ctx->rep_state = 1;
            }
        }
    }
    if (ctx->rep_state < 0) return false;
    require(fs.m <= ctx->rep_mcap, "basis exceeds the sorted replicas' capacity");
    require(ctx->row_major, "the sorted replicas are gathered from the row-major basis copies");
    std::vector<int> todo;
    for (int xc : xcs)
        if (ctx->rep_m[xc] < (int)fs.m) {
            todo.push_back(xc);
            todo.push_back(ctx->rep_m[xc]);
            ctx->rep_m[xc] = (int)fs.m;
        }
    if (todo.empty()) return true;
    const int cnt = (int)(todo.size() / 2);
    double cols = 0;
    for (int i = 0; i < cnt; ++i) cols += 2.0 * ((double)fs.m - todo[2 * i + 1]) + (todo[2 * i + 1] == 0 ? 2 : 0);
    ctx->rep_todo.ensure(todo.size());
    h2d(ctx, ctx->rep_todo.p, todo.data(), todo.size() * sizeof(int));
    KTimer kt(ctx, K_SORT, 0.0, (8.0 * cols + 4.0 * cnt) * (double)fs.n);
    // an append's new columns, the same for every listed column: gathered from a staging copy
    int k0 = todo[1];
    for (int i = 0; i < cnt; ++i)
        if (todo[2 * i + 1] != k0) k0 = -1;
    const int nk = (int)fs.m - k0;
    const float *sB = fs.B.p, *sBo = fs.Bo.p;
    size_t sld = ctx->ldb;
    int kbase = 0;
    if (k0 > 0 && nk <= 8) {
        ctx->rep_stage.ensure(2 * (size_t)nk * fs.n);
        hipLaunchKernelGGL(k_rep_stage, dim3((unsigned)((fs.n + 255) / 256)), dim3(256), 0, st, (const float *)fs.BC.p,
                           (const float *)fs.BoC.p, fs.ldx, fs.n, k0, nk, ctx->rep_stage.p);
        MH_LAUNCH_CHECK();
        sB = ctx->rep_stage.p;
        sBo = ctx->rep_stage.p + nk;
        sld = 2 * (size_t)nk;
        kbase = k0;
    }
    for (int i0 = 0; i0 < cnt; i0 += 65535) {
        const int c = std::min(cnt - i0, 65535);
        hipLaunchKernelGGL(k_rep_gather, dim3((unsigned)((fs.n + 255) / 256), c), dim3(256), 0, st,
                           (const int *)ctx->rep_todo.p + 2 * i0, (const int *)fs.order.p, (const float *)fs.X.p,
                           (const float *)fs.y.p, sB, sBo, sld, kbase, fs.ldx, fs.n, (int)fs.m, (int)ctx->rep_mcap,
                           ctx->rep_stride, ctx->rep.p);
        MH_LAUNCH_CHECK();
    }
    return true;
}

/*
 * ctx->fits replaced, for the scope, by one view per X-column group of the list: a fit of one X
 * column (x, y, B, Bo from that column's sorted replica, its scale, Bo^T y) whose kept candidate
 * columns land at hr + c * ldx, column-major.  Scoring through the views runs the projection on
 * the sorted rows, so the normalised columns come out in the order the sweep walks.
 */
struct SortedViews {
    mars_hip_ctx *ctx;
    std::vector<FitState> saved;
    SortedViews(mars_hip_ctx *c, const std::vector<int> &gxc, const std::vector<int> &gc0) : ctx(c)
    {
        const FitState &fs = ctx->fits[0];
        std::vector<FitState> v(gxc.size());
        const size_t big = size_t(1) << 50;  // the views never grow their buffers
        for (size_t g = 0; g < gxc.size(); ++g) {
            FitState &f = v[g];
            f.view = true;
            f.n = fs.n;
            f.px = 1;
            f.ldx = fs.ldx;
            f.tol = fs.tol;
            f.yvar = fs.yvar;
            f.m = fs.m;
            float *blk = ctx->rep.p + (size_t)gxc[g] * ctx->rep_stride;
            f.X.p = blk;
            f.y.p = blk + fs.ldx;
            f.BC.p = blk + 2 * fs.ldx;
            f.BoC.p = blk + (2 + ctx->rep_mcap) * fs.ldx;
            f.s.p = fs.s.p + gxc[g];
            f.ybo.p = fs.ybo.p;
            f.rbuf.p = ctx->hr.p + (size_t)gc0[g] * fs.ldx;
            for (DBuf<float> *b : {&f.X, &f.y, &f.BC, &f.BoC, &f.s, &f.rbuf}) b->cap = big;
            f.ybo.cap = big;
            f.ldr = big;
        }
        saved.swap(ctx->fits);
        ctx->fits.swap(v);
        ctx->keep_cm = true;
    }
    ~SortedViews()
    {
        for (FitState &f : ctx->fits) f.detach();
        ctx->fits.swap(saved);
        ctx->keep_cm = false;
    }
};

void hinge_score(mars_hip_ctx *ctx, Cands &cs, double *lin, double *ybx, double *hin, int *knot, int *krow,
                 int min_span, int end_span, bool f32, double eps, double *hub)
{
    const int P = (int)cs.fit.size();
    if (P == 0) return;
    // the sweep's limit, checked before the scoring passes run
    require(ctx->fits[0].m <= 509, "hinge sweep supports up to 509 basis functions");
    std::vector<int> gxc, gc0;
    for (int c = 0; c < P; ++c)
        if (c == 0 || cs.xc[c] != cs.xc[c - 1]) {
            gxc.push_back(cs.xc[c]);
            gc0.push_back(c);
        }
    ensure_sorted(ctx, gxc);
    if (!replicas_ready(ctx, gxc)) {
        plain_score(ctx, cs, true, lin, ybx, nullptr, nullptr, false, nullptr, f32);
        hinge_sweep_list(ctx, cs, lin, ybx, hin, knot, krow, min_span, end_span, eps, hub);
        return;
    }
    const FitState &fs = ctx->fits[0];
    ctx->hr.ensure(fs.ldx * (size_t)P);
    Cands &vc = ctx->vcand;
    vc.fit.resize(P);
    vc.xc.assign(P, 0);
    vc.par = cs.par;
    for (int g = 0, c = 0; c < P; ++c) {
        if (g + 1 < (int)gc0.size() && c == gc0[g + 1]) ++g;
        vc.fit[c] = g;
    }
    {
        SortedViews views(ctx, gxc, gc0);
        plain_score(ctx, vc, true, lin, ybx, nullptr, nullptr, false, nullptr, f32);
    }
    hinge_sweep_list(ctx, cs, lin, ybx, hin, knot, krow, min_span, end_span, eps, hub, vc.d_par.p, true);
}

} // namespace mh
