// T = Bo^T Bx for a tile of 64 candidates over a row range, staged through LDS.
//
// Shared by the FP64 path (v_mfma_f64_16x16x4f64, optionally on the DGKS
// residual columns) and the mixed path (v_mfma_f32_16x16x4f32 with FP32 sums
// over each 32-row chunk folded into FP64, plus exact FP64 |Bx|^2 and Bx.y).
//
// Per 32-row chunk the block stages
//   sBo[32][MT*16]  rows of Bo (one contiguous block of the row-major matrix;
//                   columns >= m read as 0),
//   sBx[32][64]     the candidates' column values Bx = f32(B[:,par]*f32(X[:,xc]*s[xc]))
//                   (or the residual columns of the DGKS retry),
// then wave w runs MT MFMAs per 4 rows for candidates 16w..16w+15.
#pragma once

#include "common.h"
#include "kernels.h"

namespace mh {

constexpr int kRC = 32;    // rows per LDS chunk
constexpr int kTileC = 64; // candidates per item

__device__ inline v4f mfma_f32(float a, float b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <int MT, bool F32, bool RETRY, bool NORMS>
__global__ __launch_bounds__(256) void k_tgemm_lds(const FitDev *__restrict__ fits,
                                                   const Item *__restrict__ items, int nitems,
                                                   const int *__restrict__ cxc,
                                                   const int *__restrict__ cpar,
                                                   const int *__restrict__ list, int ldT,
                                                   double *__restrict__ tpart,
                                                   double *__restrict__ npart)
{
    constexpr int KW = MT * 16;
    __shared__ float sBo[kRC][KW + 1];
    __shared__ float sBx[kRC][kTileC + 1];
    __shared__ double sN[2][4][kTileC];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int kr = lane >> 4, kc = lane & 15;
    const int sc = t & (kTileC - 1), g = t >> 6;  // staging role: candidate sc, rows g*8..g*8+7
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item I = items[it];
        const FitDev F = fits[I.fit];
        const bool sok = sc < I.nc;
        long long scol = 0;
        int spar = 0;
        float ssx = 0.f;
        const float *sx = F.X;
        if (sok) {
            if (RETRY) {
                scol = list[I.c0 + sc] - F.c0;
            } else {
                const int c = I.c0 + sc;
                spar = cpar[c];
                ssx = F.s[cxc[c]];
                sx = F.X + (size_t)cxc[c] * F.ldx;
            }
        }
        v4d acc[MT];
#pragma unroll
        for (int u = 0; u < MT; ++u) acc[u] = v4d{0, 0, 0, 0};
        double bxx = 0.0, bxy = 0.0;
        const bool wok = 16 * w < I.nc;
        for (long long r = I.r0; r < I.r1; r += kRC) {
            const int rc = (int)(I.r1 - r < kRC ? I.r1 - r : kRC);
            __syncthreads();
            for (int e = t; e < kRC * KW; e += 256) {
                const int i = e / KW, k = e - i * KW;
                sBo[i][k] = (i < rc && k < F.m) ? F.Bo[(r + i) * F.ldb + k] : 0.0f;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int i = g * 8 + j;
                float v = 0.f;
                if (sok && i < rc) {
                    const long long row = r + i;
                    if (RETRY)
                        v = F.rbuf[row * F.ldr + scol];
                    else
                        v = __fmul_rn(F.B[row * F.ldb + spar], __fmul_rn(sx[row], ssx));
                    if (NORMS) {
                        bxx = fma((double)v, (double)v, bxx);
                        bxy = fma((double)v, (double)F.y[row], bxy);
                    }
                }
                sBx[i][sc] = v;
            }
            __syncthreads();
            if (wok) {
                if (F32) {
                    v4f part[MT];
#pragma unroll
                    for (int u = 0; u < MT; ++u) part[u] = v4f{0, 0, 0, 0};
#pragma unroll
                    for (int q = 0; q < kRC; q += 4) {
                        const float b = sBx[q + kr][16 * w + kc];
#pragma unroll
                        for (int u = 0; u < MT; ++u) part[u] = mfma_f32(sBo[q + kr][u * 16 + kc], b, part[u]);
                    }
#pragma unroll
                    for (int u = 0; u < MT; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[u][v] += (double)part[u][v];
                } else {
#pragma unroll
                    for (int q = 0; q < kRC; q += 4) {
                        const double b = (double)sBx[q + kr][16 * w + kc];
#pragma unroll
                        for (int u = 0; u < MT; ++u)
                            acc[u] = mfma_f64((double)sBo[q + kr][u * 16 + kc], b, acc[u]);
                    }
                }
            }
        }
        const int slot = 16 * w + kc;
        if (wok && slot < I.nc) {
            double *dst = tpart + ((size_t)it * kTileC + slot) * ldT;
#pragma unroll
            for (int u = 0; u < MT; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    // D layouts: f64 -> row l/16 + 4v, f32 -> row 4(l/16) + v
                    const int k = u * 16 + (F32 ? 4 * kr + v : kr + 4 * v);
                    dst[k] = acc[u][v];
                }
        }
        if (NORMS) {
            sN[0][g][sc] = bxx;
            sN[1][g][sc] = bxy;
            __syncthreads();
            if (g == 0 && sok) {
                npart[((size_t)it * kTileC + sc) * 2] = ((sN[0][0][sc] + sN[0][1][sc]) + sN[0][2][sc]) + sN[0][3][sc];
                npart[((size_t)it * kTileC + sc) * 2 + 1] =
                    ((sN[1][0][sc] + sN[1][1][sc]) + sN[1][2][sc]) + sN[1][3][sc];
            }
        }
    }
}

} // namespace mh
