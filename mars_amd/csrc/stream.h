// Wave-streaming FP64 contractions of the plain candidate path (gfx950).
//
// Each workgroup is a single wave64 that owns one item (fit, tile of at most
// 16*CG candidates, row range starting at a multiple of 32) and walks its rows
// in 32-row chunks, feeding v_mfma_f64_16x16x4f64 straight from registers.
// Which memory row an MFMA K (T-GEMM) or M (projection) index stands for is
// free, so it is chosen to make every lane's operands contiguous runs of the
// column-major copies of Bo and B (BoC, BC: ld = the 32-padded row count) and
// of X: 16-byte loads, no barriers inside the row walk, workgroups fully
// independent.
// 16-candidate groups past the tile's end are skipped, so small candidate
// sets (the mixed form's rescoring) cost in proportion.
//
//   k_tgemm_w:   T = Bo^T Bx over the item's rows                  (kernels.cc:332-345)
//                (F32: FP32 MFMA over each 64-row span, spans added in FP64 -- the precision
//                form's approximate first pass; its DGKS retry and norms stay FP64)
//   k_project_w: r = f32(Bx - Bo T) and sum r^2, r stored row-major or not,
//                or the normalised column's projection on y    (kernels.cc:347-416)
//                (T tile staged in LDS once per item)
// with Bx = f32(B[:,par] * f32(X[:,xc] * s[xc])) (kernels.cc:300-318), or, in
// the DGKS retry, the stored residual columns (kernels.cc:387-399).
//
// Rows past the fit's end inside the last chunk read as zero: X, BC and BoC
// are zero-padded to a multiple of 32 rows.
#pragma once

#include "common.h"
#include "kernels.h"

namespace mh {

constexpr int kWR = 32;      // rows per chunk (item row ranges start at multiples of 32)
constexpr int kSlots = 64;   // candidate slots per item in the partial-sum layouts

/* Candidate groups of 16 per item: bounds registers (T-GEMM accumulators) and
 * LDS (projection T tile: 32 candidates x at most 129 doubles for m <= 128; wider tiles halve the
 * basis reads of the projection, which outweighs their lower occupancy: config 5's plain fit
 * 32.5 -> 29.9 s with 32-candidate tiles up to 128 columns). */
constexpr int cg_tgemm(int mt) { return mt <= 2 ? 4 : (mt <= 8 ? 2 : 1); }
constexpr int cg_project(int mt) { return mt <= 8 ? 2 : 1; }
constexpr int waves_for(int mt) { return mt == 1 ? 2 : (mt <= 4 ? 3 : 2); }

/* Work items are listed fit by fit and tile by tile in row order, so neighbouring items read
 * the same columns and share their address translations.  Workgroups are dispatched round-robin
 * over the 8 XCDs; block b of a round of `cnt` blocks takes item xcd_item(b, cnt), which gives
 * each XCD (and so each CU) a contiguous run of the list instead of every 8th item. */
__device__ inline int xcd_item(int b, int cnt)
{
    const int x = b & 7, q = cnt >> 3, rm = cnt & 7;
    return x * q + (x < rm ? x : rm) + (b >> 3);
}

/* Per-lane candidate of an item slot. */
struct LaneCand {
    bool ok;
    float sx;
    const float *bc;   // B column of the parent (column-major)
    const float *xp;   // X column
    long long col;     // rbuf column
    int cand;          // candidate index (position in the scored list)
};

/* Lane setup for item slot `slot`: position I.c0 + slot, candidate list[pos] (or pos itself).
 * RB: the candidate's values come from its stored column (rbuf), not from B and X. */
template <bool RB>
__device__ inline LaneCand lane_cand(const FitDev &F, const Item &I, int slot, const int *cxc, const int *cpar,
                                     const int *list)
{
    LaneCand c{slot < I.nc, 0.0f, F.BC, F.X, 0, 0};
    if (!c.ok) return c;
    const int pos = I.c0 + slot;
    const int cand = list ? list[pos] : pos;
    c.cand = cand;
    // (an unstored candidate's offset is -1: clamped so an address formed for it stays valid)
    c.col = F.rbyl == 2 ? max(F.roff[cand], 0ll) : (F.rbyl ? pos - I.pad : cand - F.c0);
    if (!RB) {
        const int xc = cxc[cand];
        c.sx = F.s[xc];
        c.bc = F.BC + (size_t)cpar[cand] * F.ldx;
        c.xp = F.X + (size_t)xc * F.ldx;
    }
    return c;
}

/* Bx of rows row..row+3 (row a multiple of 4).  The loads are unconditional (addresses
 * clamped to valid memory) and masked afterwards, so a wave's loads stay in flight together;
 * rows past the fit's end read the zero padding of BC and X. */
template <bool RB>
__device__ inline float4 bx4(const FitDev &F, const LaneCand &c, long long row, long long r1)
{
    float4 v;
    if (RB) {
        const float *p = F.rbuf + c.col * F.rsc;
        float v0, v1, v2, v3;
        if (F.rsr == 1) {
            // column-major: rows row..row+3 lie inside the column's ldx rows (row < r1 <= n)
            const float4 q = gld4(p + row);
            v0 = q.x;
            v1 = q.y;
            v2 = q.z;
            v3 = q.w;
        } else {
            const long long last = r1 - 1;
            v0 = gld(p + (row < r1 ? row : last) * F.rsr);
            v1 = gld(p + (row + 1 < r1 ? row + 1 : last) * F.rsr);
            v2 = gld(p + (row + 2 < r1 ? row + 2 : last) * F.rsr);
            v3 = gld(p + (row + 3 < r1 ? row + 3 : last) * F.rsr);
        }
        v.x = row < r1 ? v0 : 0.f;
        v.y = row + 1 < r1 ? v1 : 0.f;
        v.z = row + 2 < r1 ? v2 : 0.f;
        v.w = row + 3 < r1 ? v3 : 0.f;
    } else {
        const float4 b = gld4(c.bc + row);
        const float4 x = gld4(c.xp + row);
        v.x = __fmul_rn(b.x, __fmul_rn(x.x, c.sx));
        v.y = __fmul_rn(b.y, __fmul_rn(x.y, c.sx));
        v.z = __fmul_rn(b.z, __fmul_rn(x.z, c.sx));
        v.w = __fmul_rn(b.w, __fmul_rn(x.w, c.sx));
    }
    if (!c.ok) v = float4{0.f, 0.f, 0.f, 0.f};
    return v;
}

/* Raw loads behind one float4 of Bx (bx_load), turned into Bx once they have arrived (bx_make):
 * the loads of later row steps can then be in flight while the MFMAs of this one run. */
struct BxRaw {
    float4 b, x;  // B and X runs, or (RB) the stored run in b
};

template <bool RB>
__device__ inline BxRaw bx_load(const FitDev &F, const LaneCand &c, long long row, long long r1)
{
    BxRaw w;
    if (RB) {
        const float *p = F.rbuf + c.col * F.rsc;
        if (F.rsr == 1) {
            w.b = gld4(p + row);  // column-major: rows row..row+3 lie inside the column's ldx rows
        } else {
            const long long last = r1 - 1;
            w.b.x = gld(p + (row < r1 ? row : last) * F.rsr);
            w.b.y = gld(p + (row + 1 < r1 ? row + 1 : last) * F.rsr);
            w.b.z = gld(p + (row + 2 < r1 ? row + 2 : last) * F.rsr);
            w.b.w = gld(p + (row + 3 < r1 ? row + 3 : last) * F.rsr);
        }
        w.x = w.b;
    } else {
        w.b = gld4(c.bc + row);
        w.x = gld4(c.xp + row);
    }
    return w;
}

template <bool RB>
__device__ inline float4 bx_make(const BxRaw &w, const LaneCand &c, long long row, long long r1)
{
    float4 v;
    if (RB) {
        v.x = row < r1 ? w.b.x : 0.f;
        v.y = row + 1 < r1 ? w.b.y : 0.f;
        v.z = row + 2 < r1 ? w.b.z : 0.f;
        v.w = row + 3 < r1 ? w.b.w : 0.f;
    } else {
        v.x = __fmul_rn(w.b.x, __fmul_rn(w.x.x, c.sx));
        v.y = __fmul_rn(w.b.y, __fmul_rn(w.x.y, c.sx));
        v.z = __fmul_rn(w.b.z, __fmul_rn(w.x.z, c.sx));
        v.w = __fmul_rn(w.b.w, __fmul_rn(w.x.w, c.sx));
    }
    if (!c.ok) v = float4{0.f, 0.f, 0.f, 0.f};
    return v;
}

/* Row steps of the T-GEMM whose loads are in flight at once: narrow tiles (little MFMA work per
 * step) keep four, wide ones two; tiles of more than 64 basis columns none ahead (registers). */
constexpr int tg_depth(int mt, int cg)
{
    return (mt >= 8 || (mt >= 5 && cg >= 2)) ? 1 : (cg == 1 && mt <= 4 ? 4 : 2);
}
constexpr int tg_waves(int mt, int cg) { return tg_depth(mt, cg) > 1 ? 2 : waves_for(mt); }
/* The projection's blocks hold more per step (a scalar Bo load per k step, Bx and y): two in
 * flight for 32-candidate tiles of at most 32 basis columns (the plain form's wide passes); the
 * narrow tiles of the rescoring measured faster at three waves per SIMD without a pipeline. */
constexpr int pr_depth(int mt, int cg) { return cg == 2 && mt <= 2 ? 2 : 1; }
constexpr int pr_waves(int mt, int cg) { return pr_depth(mt, cg) > 1 ? 2 : waves_for(mt); }

/*
 * T-GEMM partials: tpart[(item * 64 + slot) * ldT + k] = sum over the item's
 * rows of Bo[i][k] Bx[i][slot].  MFMA step q of a 4-row half-step at row r
 * takes rows r + 8 kr + q (kr = lane / 16), so each lane reads 4 consecutive
 * rows of its Bo columns 16u + lane%16 and of its candidates' Bx.  The item's
 * half-steps form a software pipeline: the loads of step s + PD - 1 are issued
 * before the MFMAs of step s (PD register sets, indexed at compile time).  A
 * launch covers basis columns kb .. kb + 16 MT - 1: fits of more than 256 basis
 * functions take one launch per 256 columns.
 */
template <int MT, int CG, bool RETRY, bool F32 = false>
__global__ __launch_bounds__(64, tg_waves(MT, CG)) void k_tgemm_w(const FitDev *__restrict__ fits,
                                                                  const Item *__restrict__ items, int nitems,
                                                                  const int *__restrict__ cxc,
                                                                  const int *__restrict__ cpar,
                                                                  const int *__restrict__ list, int ldT,
                                                                  double *__restrict__ tpart,
                                                                  const int *__restrict__ dn, int kb)
{
    if (dn) nitems = *dn;  // item count built on the device
    constexpr int PD = tg_depth(MT, CG);
    const int lane = threadIdx.x, kr = lane >> 4, kc = lane & 15;
    for (int round = 0; round < nitems; round += gridDim.x) {
        const int cnt = min((int)gridDim.x, nitems - round);
        if ((int)blockIdx.x >= cnt) break;
        const int it = round + xcd_item(blockIdx.x, cnt);
        const Item I = items[it];
        const FitDev F = fits[I.fit];
        LaneCand c[CG];
#pragma unroll
        for (int g = 0; g < CG; ++g) c[g] = lane_cand<RETRY>(F, I, 16 * g + kc, cxc, cpar, list);
        const float *bo[MT];
        bool kok[MT];
#pragma unroll
        for (int u = 0; u < MT; ++u) {
            const int k = kb + 16 * u + kc;
            kok[u] = k < F.m;
            bo[u] = F.BoC + (size_t)(kok[u] ? k : 0) * F.ldx;
        }
        v4d acc[CG][MT];
        v4f acc32[CG][MT];  // F32: one 32-row chunk (two steps) in FP32, then folded into FP64
#pragma unroll
        for (int g = 0; g < CG; ++g)
#pragma unroll
            for (int u = 0; u < MT; ++u) {
                acc[g][u] = v4d{0, 0, 0, 0};
                acc32[g][u] = v4f{0, 0, 0, 0};
            }
        // step s: rows r0 + 32 (s / 2) + 8 kr + 4 (s % 2) .. + 3 of this lane (a multiple of PD apart
        // from the loop's step, so its parity is the unrolled index's)
        const int nsteps = 2 * (int)((I.r1 - I.r0 + kWR - 1) / kWR);
        auto srow = [&](int st) { return I.r0 + (long long)kWR * (st >> 1) + 8 * kr + 4 * (st & 1); };
        float4 av[PD][MT];
        BxRaw bw[PD][CG];
        auto issue = [&](int j, int st) {
            if (st >= nsteps) return;  // (uniform)
            const long long row = srow(st);
#pragma unroll
            for (int u = 0; u < MT; ++u) av[j][u] = gld4(bo[u] + row);
#pragma unroll
            for (int g = 0; g < CG; ++g) bw[j][g] = bx_load<RETRY>(F, c[g], row, I.r1);
        };
        auto compute = [&](int j, int st) {
            if (st >= nsteps) return;  // (uniform)
            const long long row = srow(st);
            double a[MT][4], b[CG][4];
#pragma unroll
            for (int u = 0; u < MT; ++u) {
                a[u][0] = kok[u] ? av[j][u].x : 0.f;
                a[u][1] = kok[u] ? av[j][u].y : 0.f;
                a[u][2] = kok[u] ? av[j][u].z : 0.f;
                a[u][3] = kok[u] ? av[j][u].w : 0.f;
            }
#pragma unroll
            for (int g = 0; g < CG; ++g) {
                const float4 v = bx_make<RETRY>(bw[j][g], c[g], row, I.r1);
                b[g][0] = v.x;
                b[g][1] = v.y;
                b[g][2] = v.z;
                b[g][3] = v.w;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int g = 0; g < CG; ++g)
                    if (16 * g < I.nc)
#pragma unroll
                        for (int u = 0; u < MT; ++u) {
                            if (F32)
                                acc32[g][u] = mfma_f32((float)a[u][q], (float)b[g][q], acc32[g][u]);
                            else
                                acc[g][u] = mfma_f64(a[u][q], b[g][q], acc[g][u]);
                        }
            if (F32 && (st & 1)) {
#pragma unroll
                for (int g = 0; g < CG; ++g)
#pragma unroll
                    for (int u = 0; u < MT; ++u) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[g][u][v] += (double)acc32[g][u][v];
                        acc32[g][u] = v4f{0, 0, 0, 0};
                    }
            }
        };
#pragma unroll
        for (int j = 0; j + 1 < PD; ++j) issue(j, j);
#pragma unroll 1
        for (int s0 = 0; s0 < nsteps; s0 += PD) {
#pragma unroll
            for (int j = 0; j < PD; ++j) {
                issue((j + PD - 1) % PD, s0 + j + PD - 1);
                compute(j, s0 + j);
            }
        }
#pragma unroll
        for (int g = 0; g < CG; ++g) {
            const int slot = 16 * g + kc;
            if (slot >= I.nc) continue;
            double *dst = tpart + ((size_t)it * kSlots + slot) * ldT;
#pragma unroll
            for (int u = 0; u < MT; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v)  // D row l/16 + 4v (FP64 MFMA), 4 (l/16) + v (FP32 MFMA)
                    dst[kb + 16 * u + (F32 ? 4 * kr + v : kr + 4 * v)] = acc[g][u][v];
        }
    }
}

/*
 * T-GEMM partials by BF16-split products: each f32 operand x = hi + lo with hi = bf16(x) and
 * lo = bf16(x - hi), and x y ~ hi_x hi_y + hi_x lo_y + lo_x hi_y (the lo_x lo_y term and the
 * rounding of lo are below 2^-17 of |x y|), three v_mfma_f32_16x16x32_bf16 per 32-row chunk with
 * FP32 accumulation, chunks added in FP64.  Operand layout of the 16x16x32 BF16 MFMA: lane l holds
 * A[l%16][8 (l/16) .. 8 (l/16) + 7] and B[8 (l/16) .. + 7][l%16]; D as the FP32 16x16x4 MFMA.
 * So lane (kr, kc) supplies rows r + 8 kr .. r + 8 kr + 7 of its Bo column 16u + kc and of its
 * candidates' Bx: two float4 loads each, as k_tgemm_w.  Same partial layout as k_tgemm_w<F32>.
 */
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ inline void split_bf16(const float4 &x0, const float4 &x1, v8bf &hi, v8bf &lo)
{
    const float v[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const __bf16 h = (__bf16)v[j];
        hi[j] = h;
        lo[j] = (__bf16)(v[j] - (float)h);
    }
}

template <int MT, int CG>
__global__ __launch_bounds__(64, waves_for(MT)) void k_tgemm_bf16(const FitDev *__restrict__ fits,
                                                                   const Item *__restrict__ items, int nitems,
                                                                   const int *__restrict__ cxc,
                                                                   const int *__restrict__ cpar, int ldT,
                                                                   double *__restrict__ tpart)
{
    const int lane = threadIdx.x, kr = lane >> 4, kc = lane & 15;
    for (int round = 0; round < nitems; round += gridDim.x) {
        const int cnt = min((int)gridDim.x, nitems - round);
        if ((int)blockIdx.x >= cnt) break;
        const int it = round + xcd_item(blockIdx.x, cnt);
        const Item I = items[it];
        const FitDev F = fits[I.fit];
        LaneCand c[CG];
#pragma unroll
        for (int g = 0; g < CG; ++g) c[g] = lane_cand<false>(F, I, 16 * g + kc, cxc, cpar, nullptr);
        const float *bo[MT];
        bool kok[MT];
#pragma unroll
        for (int u = 0; u < MT; ++u) {
            const int k = 16 * u + kc;
            kok[u] = k < F.m;
            bo[u] = F.BoC + (size_t)(kok[u] ? k : 0) * F.ldx;
        }
        v4d acc[CG][MT];
#pragma unroll
        for (int g = 0; g < CG; ++g)
#pragma unroll
            for (int u = 0; u < MT; ++u) acc[g][u] = v4d{0, 0, 0, 0};
#pragma unroll 1
        for (long long r = I.r0; r < I.r1; r += kWR) {
            const long long row = r + 8 * kr;
            float4 av[MT][2], bv[CG][2];
#pragma unroll
            for (int u = 0; u < MT; ++u) {
                av[u][0] = gld4(bo[u] + row);
                av[u][1] = gld4(bo[u] + row + 4);
            }
#pragma unroll
            for (int g = 0; g < CG; ++g) {
                bv[g][0] = bx4<false>(F, c[g], row, I.r1);
                bv[g][1] = bx4<false>(F, c[g], row + 4, I.r1);
            }
            v8bf ah[MT], al[MT], bh[CG], bl[CG];
#pragma unroll
            for (int u = 0; u < MT; ++u) {
                if (!kok[u]) av[u][0] = av[u][1] = float4{0.f, 0.f, 0.f, 0.f};
                split_bf16(av[u][0], av[u][1], ah[u], al[u]);
            }
#pragma unroll
            for (int g = 0; g < CG; ++g) split_bf16(bv[g][0], bv[g][1], bh[g], bl[g]);
#pragma unroll
            for (int g = 0; g < CG; ++g) {
                if (16 * g >= I.nc) continue;
#pragma unroll
                for (int u = 0; u < MT; ++u) {
                    v4f d = v4f{0, 0, 0, 0};
                    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[u], bh[g], d, 0, 0, 0);
                    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[u], bl[g], d, 0, 0, 0);
                    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[u], bh[g], d, 0, 0, 0);
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[g][u][v] += (double)d[v];
                }
            }
        }
#pragma unroll
        for (int g = 0; g < CG; ++g) {
            const int slot = 16 * g + kc;
            if (slot >= I.nc) continue;
            double *dst = tpart + ((size_t)it * kSlots + slot) * ldT;
#pragma unroll
            for (int u = 0; u < MT; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) dst[16 * u + 4 * kr + v] = acc[g][u][v];  // D row 4 (l/16) + v
        }
    }
}

/* Projection modes. */
enum { kStoreR = 0, kSumSq = 1, kFinalY = 2 };

/*
 * Projection r = f32(Bx - Bo T) over the item's rows, then per MODE:
 *   kStoreR  r -> rbuf (row * rsr + column * rsc), spart[item * 64 + slot] = sum r^2;
 *            with `keep`, only candidates whose keep entry is nonzero store r;
 *   kSumSq   spart = sum r^2 only;
 *   kFinalY  Bxn = f32(r * scale), scale = f32(1 / sqrt(s + tol)) (0 when s <= tol,
 *            kernels.cc:401-416), spart = sum Bxn y (marsalgo.cc:500-505); candidates
 *            whose DGKS gate fired take r from their stored retry residual instead.
 * RB: Bx is the candidate's stored column (the DGKS retry).  TCAND: the item's
 * T rows are indexed by candidate instead of by position.
 * The item's T tile is staged once in LDS (MFMA B operand).  The A operand's
 * M index M stands for row 4 (M % 4) + M / 4 of the 16-row block, so the D
 * elements a lane holds (M = l/16 + 4v) are the 4 consecutive rows
 * 4 (l/16) + v, whose Bx it loads as one float4 each of B and X.  The 16-row
 * blocks are software-pipelined like the T-GEMM's half-steps.  WIDE (MT = 16
 * only) serves fits of more than 256 basis functions: the columns past the
 * first 256 are added to Bo T in the same FP64 sum, their Bo and T values read
 * straight from memory (a tile's T, 16 x ldT doubles, stays in the L2), before
 * the residual is rounded; a separate instantiation, so the kernels of smaller
 * bases keep their registers.
 */
template <int MT, int CG, int MODE, bool RB, bool TCAND, bool F32 = false, bool WIDE = false>
__global__ __launch_bounds__(64, pr_waves(MT, CG)) void k_project_w(const FitDev *__restrict__ fits,
                                                                    const Item *__restrict__ items, int nitems,
                                                                    const int *__restrict__ cxc,
                                                                    const int *__restrict__ cpar,
                                                                    const int *__restrict__ list,
                                                                    const double *__restrict__ T, int ldT,
                                                                    const double *__restrict__ ssum,
                                                                    const int *__restrict__ fired,
                                                                    const int *__restrict__ keep,
                                                                    double *__restrict__ spart,
                                                                    const int *__restrict__ dn)
{
    if (dn) nitems = *dn;  // item count built on the device
    constexpr int KQ = 4 * MT;           // k steps of 4
    constexpr int NC = 16 * CG, ST = 16 * MT + 1;
    constexpr int PD = pr_depth(MT, CG);  // 16-row blocks whose loads are in flight at once
    __shared__ double sT[NC][ST];
    const int lane = threadIdx.x, kr = lane >> 4, kc = lane & 15;
    // memory row of this lane's A-operand M index: the lane's D elements are then rows 4 (l/16) + v
    const int prow = F32 ? kc : 4 * (kc & 3) + (kc >> 2);
    for (int round = 0; round < nitems; round += gridDim.x) {
        const int cnt = min((int)gridDim.x, nitems - round);
        if ((int)blockIdx.x >= cnt) break;
        const int it = round + xcd_item(blockIdx.x, cnt);
        const Item I = items[it];
        const FitDev F = fits[I.fit];
        __syncthreads();
        {
            // only the live slots are loaded; eight loads in flight per lane
            const int live = ((I.nc + 15) & ~15) * 16 * MT;
#pragma unroll 8
            for (int e = lane; e < NC * 16 * MT; e += 64) {
                const int sl = e / (16 * MT), k = e - sl * (16 * MT);
                const bool ok = e < live && sl < I.nc && k < F.m;
                const int pos = I.c0 + (ok ? sl : 0);
                const int row = TCAND ? (list ? gld(list + pos) : pos) : pos;
                const double v = gld(T + (size_t)row * ldT + (ok ? k : 0));
                sT[sl][k] = ok ? v : 0.0;
            }
        }
        __syncthreads();
        LaneCand c[CG];
        double acc[CG];
        float scale[CG];
        bool fromr[CG], store[CG];
        static_assert(!WIDE || MT == 16, "columns past 256 follow a full 256-column tile");
        const double *trow[CG];  // WIDE: this lane's candidate's T row (columns past 256)
        const float *bo[KQ];
#pragma unroll
        for (int q = 0; q < KQ; ++q) {
            const int k = 4 * q + kr;
            bo[q] = F.BoC + (size_t)(k < F.m ? k : 0) * F.ldx;
        }
#pragma unroll
        for (int g = 0; g < CG; ++g) {
            c[g] = lane_cand<RB>(F, I, 16 * g + kc, cxc, cpar, list);
            if constexpr (WIDE) {
                const int sl = 16 * g + kc < I.nc ? 16 * g + kc : 0;
                const int pos = I.c0 + sl;
                const int row = TCAND ? (list ? gld(list + pos) : pos) : pos;
                trow[g] = T + (size_t)row * ldT;
            }
            acc[g] = 0.0;
            scale[g] = 0.0f;
            fromr[g] = false;
            store[g] = MODE == kStoreR && c[g].ok && (!keep || keep[c[g].cand] != 0);
            if (MODE == kFinalY && c[g].ok) {
                const int cand = c[g].cand;
                const double sv = ssum[cand];
                scale[g] = sv > F.tol ? (float)(1.0 / sqrt(sv + F.tol)) : 0.0f;
                fromr[g] = (fired[cand] & 1) != 0;
            }
        }
        // step s: the 16-row block at r0 + 16 s (item ranges start at multiples of 32, and the rows
        // of a block past the item's end lie past the fit's end, in the zero padding)
        const int nsteps = 2 * (int)((I.r1 - I.r0 + kWR - 1) / kWR);
        // k steps of 4 basis columns that hold live columns: the MFMAs of the tile's padding past m
        // are skipped (uniform over the item; its loads read column 0, a cache hit)
        const int kq = MT <= 8 ? (F.m + 3) >> 2 : KQ;  // (wide tiles: padding at most 15 of 16 MT columns)
        float avh[PD][KQ];
        BxRaw bw[PD][CG];
        float4 yh[PD];
        auto issue = [&](int j, int st) {
            if (st >= nsteps) return;  // (uniform)
            const long long r16 = I.r0 + 16LL * st;
            const long long row = r16 + 4 * kr;
#pragma unroll
            for (int q = 0; q < KQ; ++q) avh[j][q] = gld(bo[q] + r16 + prow);
#pragma unroll
            for (int g = 0; g < CG; ++g) bw[j][g] = bx_load<RB>(F, c[g], row, I.r1);
            yh[j] = float4{0.f, 0.f, 0.f, 0.f};
            if (MODE == kFinalY) yh[j] = gld4(F.y + row);
        };
        auto compute = [&](int j, int st) {
            if (st >= nsteps) return;  // (uniform)
            const long long row = I.r0 + 16LL * st + 4 * kr;
            float av[KQ];
#pragma unroll
            for (int q = 0; q < KQ; ++q) av[q] = q >= kq || 4 * q + kr >= F.m ? 0.0f : avh[j][q];
            const float yy[4] = {yh[j].x, yh[j].y, yh[j].z, yh[j].w};
#pragma unroll
            for (int g = 0; g < CG; ++g) {
                if (16 * g >= I.nc) continue;
                v4d d = v4d{0, 0, 0, 0};
                if (F32) {  // Bo T in FP32 (the approximate first pass of the precision form)
                    v4f d32 = v4f{0, 0, 0, 0};
#pragma unroll
                    for (int q = 0; q < KQ; ++q)
                        if (q < kq) d32 = mfma_f32(av[q], (float)sT[16 * g + kc][4 * q + kr], d32);
                    d = v4d{(double)d32[0], (double)d32[1], (double)d32[2], (double)d32[3]};
                } else {
#pragma unroll
                    for (int q = 0; q < KQ; ++q)
                        if (q < kq) d = mfma_f64((double)av[q], sT[16 * g + kc][4 * q + kr], d);
                }
                if constexpr (WIDE) {
                    // basis columns 256 .. m - 1 (uniform over the item)
                    const long long r16 = I.r0 + 16LL * st;
                    for (int k0 = 256; k0 < F.m; k0 += 256) {
#pragma unroll 8
                        for (int q = 0; q < KQ; ++q) {
                            const int k = k0 + 4 * q + kr;
                            const bool ok = k < F.m;
                            const int kk = ok ? k : 0;  // (clamped: every address formed is valid)
                            const float av1 = gld(F.BoC + (size_t)kk * F.ldx + r16 + prow);
                            const double bv1 = gld(trow[g] + kk);
                            const float a = ok ? av1 : 0.0f;
                            const double b = ok ? bv1 : 0.0;
                            if (F32) {
                                const v4f e = mfma_f32(a, (float)b, v4f{0, 0, 0, 0});
                                d = v4d{d[0] + e[0], d[1] + e[1], d[2] + e[2], d[3] + e[3]};
                            } else {
                                d = mfma_f64((double)a, b, d);
                            }
                        }
                    }
                }
                if (!c[g].ok) continue;
                const float4 bx = bx_make<RB>(bw[j][g], c[g], row, I.r1);
                const float bv[4] = {bx.x, bx.y, bx.z, bx.w};
                float rf[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) rf[v] = (float)((double)bv[v] - d[v]);
                float *rp = F.rbuf + c[g].col * F.rsc;
                if (MODE == kFinalY && fromr[g]) {
                    if (F.rsr == 1) {
                        const float4 q = gld4(rp + row);
                        rf[0] = q.x;
                        rf[1] = q.y;
                        rf[2] = q.z;
                        rf[3] = q.w;
                    } else {
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            if (row + v < I.r1) rf[v] = rp[(row + v) * F.rsr];
                    }
                }
                if (MODE == kStoreR && store[g]) {
                    if (F.rsr == 1) {
                        *reinterpret_cast<float4 *>(rp + row) = float4{rf[0], rf[1], rf[2], rf[3]};
                    } else {
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            if (row + v < I.r1) rp[(row + v) * F.rsr] = rf[v];
                    }
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    if (row + v >= I.r1) continue;
                    if (MODE == kFinalY) {
                        const float bxn = __fmul_rn(rf[v], scale[g]);
                        acc[g] = fma((double)bxn, (double)yy[v], acc[g]);
                    } else {
                        acc[g] = fma((double)rf[v], (double)rf[v], acc[g]);
                    }
                }
            }
        };
#pragma unroll
        for (int j = 0; j + 1 < PD; ++j) issue(j, j);
#pragma unroll 1
        for (int s0 = 0; s0 < nsteps; s0 += PD) {
#pragma unroll
            for (int j = 0; j < PD; ++j) {
                issue((j + PD - 1) % PD, s0 + j + PD - 1);
                compute(j, s0 + j);
            }
        }
#pragma unroll
        for (int g = 0; g < CG; ++g) {
            double s = acc[g];
            s = s + __shfl_xor(s, 16, 64);
            s = s + __shfl_xor(s, 32, 64);
            const int slot = 16 * g + kc;
            if (kr == 0 && slot < I.nc) spart[(size_t)it * kSlots + slot] = s;
        }
    }
}

} // namespace mh
