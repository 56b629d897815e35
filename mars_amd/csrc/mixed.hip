// Mixed-precision candidate scoring on gfx950 (linear delta-SSE).
//
// The basis only grows by appending columns (marsalgo.cc:760-875), so the
// projection of a candidate Bx = B[:,par] * x_xc on Bo's column k,
//     T[xc][par][k] = sum_i Bo[i,k] B[i,par] xs[i,xc],     xs = f32(X[:,xc] * s[xc]),
// never changes once both column k and parent par exist.  Each fit keeps a
// cache of T over every (X column, parent, basis column); when column j
// joins the basis only the entries with max(k, par) == j are new: 2j+1
// columns A[:,a] = f32(Bo[:,k] * B[:,par]) against all of X, one FP32 MFMA
// contraction (v_mfma_f32_16x16x4f32, FP32 sums over 32-row chunks folded
// into FP64).  The same pass forms the exact FP64 sums |Bx|^2 and Bx.y of the
// new parent (f32 products are exact in f64).  With r = Bx - Bo T and Bo
// orthonormal,
//     |r|^2 = |Bx|^2 - |T|^2,      r.y = Bx.y - T.ybo,
//     score = (r.y)^2 / (|r|^2 + tol)   (0 when |r|^2 <= tol)      (kernels.cc:401-416,
//                                                                   marsalgo.cc:500-505)
// which is accurate while |r|^2 is not a small difference of large numbers --
// exactly the DGKS gate's condition (9 |r|^2 >= |T|^2, kernels.h:17).
// Candidates that trip the gate, and every candidate whose approximate score
// lies within a relative window of the best ungated score of its fit, are
// rescored by the exact FP64 path (ortho.hip), so each step's winner and its
// delta-SSE are FP64 values computed with the reference's rounding points.
// Per step the cache costs O(n p m) per fit instead of O(n p m^2).
#include "internal.h"
#include "kernels.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

namespace mh {

namespace {

constexpr int kChunk = 32;         // rows per FP32 accumulation chunk (folded into FP64)
constexpr int kXT = 64;            // X columns per update item (4 groups of 16)
constexpr int kGroupInts = 7;      // fit, xc0, nc, j, a-tile, first list entry, item count

/*
 * Cache update for new basis column j of one fit.  Block item = (fit, X
 * columns [c0, c0+nc), rows [r0, r1), A tiles u0 .. u0+W-1), item.pad =
 * j << 8 | u0; wave w computes A tile u = u0 + w:
 *   U[a][xc] = sum_i A[i][a] xs[i][xc],   a = 16u + (0..15) < 2j+1,
 *   A[i][a] = f32(Bo[i][a] B[i][j])          (a <= j:  k = a,  parent j)
 *           = f32(Bo[i][j] B[i][a-j-1])      (a >  j:  k = j,  parent a-j-1)
 * and the u == 0 wave also |Bx|^2 and Bx.y for Bx = f32(B[:,j] xs[:,xc]).
 * Every 32-row chunk of xs = f32(X s) is staged once per block in LDS
 * (column-major, 36-float pitch: each 16-lane phase of a ds_read_b128 hits
 * distinct banks), double-buffered: the next chunk's X loads and each wave's
 * next A columns are in flight during the current chunk's MFMAs.  MFMA step q
 * of a 4-row half-step at row r takes row r + 8 (lane/16) + q, so every
 * operand is a float4 run of one column.  v_mfma_f32_16x16x4f32 accumulates
 * each chunk in FP32; the chunk sums are folded into FP64.
 */
constexpr int kTW = 8;         // A tiles (waves) per block at most
constexpr int kXP = 36;        // LDS floats per staged X column (32 rows + pad)

// (8 waves: a block holds a CU's 2 waves per SIMD at any register count up to 256, so no bound)
template <int W, bool F64 = false>
__global__ __launch_bounds__(64 * W, W == 4 ? 3 : 1) void k_tc_update(const FitDev *__restrict__ fits,
                                                      const Item *__restrict__ items, int nitems,
                                                      double *__restrict__ tpart, double *__restrict__ npart)
{
    constexpr int NXL = (kXT * kChunk / 4 + 64 * W - 1) / (64 * W);  // X float4 loads per thread and chunk
    __shared__ __attribute__((aligned(16))) float sx[2][kXT][kXP];
    __shared__ __attribute__((aligned(16))) float sjy[2][2][kChunk];  // B[:, j] and y of the chunk
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, kr = lane >> 4, kc = lane & 15;
    for (int bi = blockIdx.x; bi < nitems; bi += gridDim.x) {
        const Item I = items[bi];
        const FitDev F = fits[I.fit];
        const int j = I.pad >> 8, u = (I.pad & 255) + wv, na = 2 * j + 1;
        const int a = 16 * u + kc;
        const bool aok = a < na;
        const bool wact = 16 * u < na;   // this wave's tile holds A columns (uniform)
        // |Bx|^2 and Bx.y of the new parent (blocks with the first A tile): X-column group g goes to
        // active wave g mod (active waves), so no single wave carries them between barriers
        const bool norms = u == 0;
        const bool nblk = (I.pad & 255) == 0;
        const int nwa = max(1, min(W, (na + 15) / 16 - (I.pad & 255)));   // active waves of the block
        auto mine = [&](int g) {
            if constexpr (W >= 2) return nblk && g % nwa == wv;
            else return norms;
        };
        const int ka = a <= j ? a : j, pa = a <= j ? j : a - j - 1;
        const float *pbo = F.BoC + (size_t)(aok ? ka : 0) * F.ldx;
        const float *pb = F.BC + (size_t)(aok ? pa : 0) * F.ldx;
        const float *pbj = F.BC + (size_t)j * F.ldx;
        const int G = (I.nc + 15) >> 4;
        // this thread's share of the X chunk: float4 e -> column e/8, rows 4 (e%8) ..
        const float *xp[NXL];
        float xsc[NXL];
        int xcol[NXL], xrow[NXL];
#pragma unroll
        for (int q = 0; q < NXL; ++q) {
            const int e = tid + 64 * W * q;
            xcol[q] = e < kXT * kChunk / 4 ? e >> 3 : -1;
            xrow[q] = 4 * (e & 7);
            const bool ok = xcol[q] >= 0 && xcol[q] < I.nc;
            const int xc = I.c0 + (ok ? xcol[q] : 0);
            xp[q] = F.X + (size_t)xc * F.ldx + xrow[q];
            xsc[q] = ok ? F.s[xc] : 0.0f;
        }
        // threads 0..15 also stage B[:, j] (0..7) and y (8..15) of the chunk, 4 rows each
        const bool jyl = tid < 16;
        const float *jyp = (tid < 8 ? pbj : F.y) + 4 * (tid & 7);
        auto load_x = [&](float4 (&v)[NXL], float4 &jy, long long r) {
#pragma unroll
            for (int q = 0; q < NXL; ++q) v[q] = gld4(xp[q] + r);
            jy = gld4(jyp + r);
        };
        auto store_x = [&](const float4 (&v)[NXL], const float4 &jy, int buf) {
#pragma unroll
            for (int q = 0; q < NXL; ++q) {
                if (xcol[q] < 0) continue;
                const float s = xsc[q];
                const float4 w4 = make_float4(__fmul_rn(v[q].x, s), __fmul_rn(v[q].y, s), __fmul_rn(v[q].z, s),
                                              __fmul_rn(v[q].w, s));
                *(float4 *)&sx[buf][xcol[q]][xrow[q]] = w4;
            }
            if (jyl) *(float4 *)&sjy[buf][tid >> 3][4 * (tid & 7)] = jy;
        };
        // [register set][half]: A columns are fetched AS - 1 chunks ahead (4 waves: one, which leaves
        // the registers for spreading the norms over the waves)
        constexpr int AS = W == 4 ? 2 : 3;
        float4 ao[AS][2], ab[AS][2];
        auto load_a = [&](int set, long long r) {
#pragma unroll
            for (int hf = 0; hf < 2; ++hf) {
                const long long rb = r + 8 * kr + 4 * hf;
                ao[set][hf] = gld4(pbo + rb);
                ab[set][hf] = gld4(pb + rb);
            }
        };
        v4d acc[4];
        double bxx[4], bxy[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            acc[g] = v4d{0, 0, 0, 0};
            bxx[g] = 0.0;
            bxy[g] = 0.0;
        }
        auto compute = [&](int set, int buf) {
            v4f part[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) part[g] = v4f{0, 0, 0, 0};
#pragma unroll
            for (int hf = 0; hf < 2; ++hf) {
                const float ov[4] = {ao[set][hf].x, ao[set][hf].y, ao[set][hf].z, ao[set][hf].w};
                const float bv[4] = {ab[set][hf].x, ab[set][hf].y, ab[set][hf].z, ab[set][hf].w};
                float av[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) av[q] = aok ? __fmul_rn(ov[q], bv[q]) : 0.0f;
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    if (g >= G) continue;
                    const float4 x4 = *(const float4 *)&sx[buf][16 * g + kc][8 * kr + 4 * hf];
                    const float xs[4] = {x4.x, x4.y, x4.z, x4.w};
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (F64)
                            acc[g] = mfma_f64((double)av[q], (double)xs[q], acc[g]);
                        else
                            part[g] = mfma_f32(av[q], xs[q], part[g]);
                    }
                    if (mine(g)) {
                        const float4 j4 = *(const float4 *)&sjy[buf][0][8 * kr + 4 * hf];
                        const float4 y4 = *(const float4 *)&sjy[buf][1][8 * kr + 4 * hf];
                        const float jv[4] = {j4.x, j4.y, j4.z, j4.w};
                        const float yy[4] = {y4.x, y4.y, y4.z, y4.w};
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const double bx = (double)__fmul_rn(jv[q], xs[q]);
                            bxx[g] = fma(bx, bx, bxx[g]);
                            bxy[g] = fma(bx, (double)yy[q], bxy[g]);
                        }
                    }
                }
            }
            if (!F64) {
#pragma unroll
                for (int g = 0; g < 4; ++g)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[g][v] += (double)part[g][v];
            }
        };
        // prologue: chunk 0 into buffer 0, A of chunks 0 (and 1) into register sets 0 (and 1)
        float4 xv[NXL], jy;
        const long long r0 = I.r0, r1 = I.r1;
        if (r0 < r1) {
            load_x(xv, jy, r0);
            if (wact) {
                load_a(0, r0);
                if constexpr (AS == 3) load_a(1, r0 + kChunk < r1 ? r0 + kChunk : r0);
            }
            store_x(xv, jy, 0);
        }
        __syncthreads();
        // chunk c at row r: A in register set c % AS (a constant in each unrolled copy)
        auto step = [&](long long r, int c, auto setc) {
            constexpr int set = decltype(setc)::value, ahead = (set + AS - 1) % AS;
            const int cur = c & 1, nxt = cur ^ 1;
            // a missing chunk ahead reloads the current one (in bounds, unused)
            const long long rn = r + kChunk < r1 ? r + kChunk : r;
            const long long ra = r + (AS - 1) * kChunk < r1 ? r + (AS - 1) * kChunk : r;
            load_x(xv, jy, rn);
            if (wact) {
                load_a(ahead, ra);
                compute(set, cur);
            }
            // buffer nxt last held chunk c-1, finished by every wave before the previous barrier
            store_x(xv, jy, nxt);
            __syncthreads();
        };
        {
            long long r = r0;
            int c = 0;
#pragma unroll 1
            while (r < r1) {
                step(r, c, std::integral_constant<int, 0>{});
                r += kChunk, ++c;
                if (r >= r1) break;
                step(r, c, std::integral_constant<int, 1>{});
                r += kChunk, ++c;
                if constexpr (AS == 3) {
                    if (r >= r1) break;
                    step(r, c, std::integral_constant<int, 2 % AS>{});
                    r += kChunk, ++c;
                }
            }
        }
        const int it = bi * W + wv;   // sub-item of this wave
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            if (g >= G) continue;
            const int slot = 16 * g + kc;
            // D row 4(l/16)+v (FP32 MFMA) or l/16 + 4v (FP64 MFMA) is A column 16u + that row
            double *dst = tpart + ((size_t)it * kXT + slot) * 16;
            if (slot < I.nc)
#pragma unroll
                for (int v = 0; v < 4; ++v) dst[F64 ? kr + 4 * v : 4 * kr + v] = acc[g][v];
            if (mine(g)) {
                double s0 = bxx[g], s1 = bxy[g];
                s0 = s0 + __shfl_xor(s0, 16, 64);
                s0 = s0 + __shfl_xor(s0, 32, 64);
                s1 = s1 + __shfl_xor(s1, 16, 64);
                s1 = s1 + __shfl_xor(s1, 32, 64);
                if (kr == 0 && slot < I.nc) {   // at the first A tile's sub-item, whichever wave summed them
                    const size_t nit = W >= 2 ? (size_t)bi * W : (size_t)it;
                    npart[(nit * kXT + slot) * 2] = s0;
                    npart[(nit * kXT + slot) * 2 + 1] = s1;
                }
            }
        }
        __syncthreads();   // the next item restages sx
    }
}

/* Sum each (fit, A tile, X-column block) group's row partials and write the new cache entries.
 * Block (group, 64-entry slice): lane l takes entry 64 y + l of the group (X column e / 16, A
 * column e % 16), wave w the w-th of four equal segments of the group's row partials (each summed
 * in row order), and the segment sums are added in wave order; slice 0 of the u == 0 groups also
 * sums |Bx|^2 and Bx.y of the group's X columns the same way. */
__global__ __launch_bounds__(256) void k_tc_reduce(const FitDev *__restrict__ fits, const int *__restrict__ groups,
                                                   const int *__restrict__ glist, const double *__restrict__ tpart,
                                                   const double *__restrict__ npart)
{
    __shared__ double sh[4][64][3];
    const int *G = groups + (size_t)blockIdx.x * kGroupInts;
    const int fit = G[0], xc0 = G[1], nc = G[2], j = G[3], u = G[4], first = G[5], count = G[6];
    const FitDev F = fits[fit];
    const int na = 2 * j + 1;
    const long long ldb = F.ldb;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int seg = (count + 3) / 4, lo = min(count, w * seg), hi = min(count, lo + seg);
    const int e = (int)blockIdx.y * 64 + lane, xl = e >> 4, al = e & 15, a = 16 * u + al;
    const bool ok = xl < nc && a < na;
    const bool nok = u == 0 && blockIdx.y == 0 && lane < nc;
    double s = 0.0, s0 = 0.0, s1 = 0.0;
    if (ok) s = range_sum(tpart, lo, hi, [&](int i) { return ((size_t)glist[first + i] * kXT + xl) * 16 + al; });
    if (nok) {
        s0 = range_sum(npart, lo, hi, [&](int i) { return ((size_t)glist[first + i] * kXT + lane) * 2; });
        s1 = range_sum(npart, lo, hi, [&](int i) { return ((size_t)glist[first + i] * kXT + lane) * 2 + 1; });
    }
    sh[w][lane][0] = s;
    sh[w][lane][1] = s0;
    sh[w][lane][2] = s1;
    __syncthreads();
    if (w != 0) return;
    for (int v = 1; v < 4; ++v) {
        s += sh[v][lane][0];
        s0 += sh[v][lane][1];
        s1 += sh[v][lane][2];
    }
    if (ok) {
        const long long xc = xc0 + xl;
        const int k = a <= j ? a : j, par = a <= j ? j : a - j - 1;
        F.tc[(xc * ldb + par) * ldb + k] = s;
    }
    if (nok) {
        F.tcn[((xc0 + lane) * ldb + j) * 2] = s0;
        F.tcn[((xc0 + lane) * ldb + j) * 2 + 1] = s1;
    }
}

/*
 * Approximate score, gate flag and an interval [lo, hi] that contains the plain form's score,
 * per candidate, from the cache; 16 lanes per candidate.
 *
 * Interval (u = 2^-24, the f32 unit roundoff).  The cached T differs from the plain form's FP64
 * T = Bo^T Bx by at most gamma sum_i |Bo_ik| |Bx_i| <= gamma |Bx| per entry (Cauchy-Schwarz,
 * unit Bo columns), gamma = 40 u: the rounded products f32(Bo B) against Bo f32(B x) (2u) and
 * at most 32 FP32 accumulations per row chunk (32u), with margin.  So with |dT| <= sqrt(m) gamma |Bx|:
 *     |d(r.y)|  <= |dT| |ybo|,
 *     |d|r|^2|  <= 2 |T| |dT| + |dT|^2 + m 8u |T|^2
 * (the last term bounds T^T (Bo^T Bo - I) T, the orthogonality loss of the stored f32 columns,
 * which the closed form |r|^2 = |Bx|^2 - |T|^2 omits).  That bounds the real
 * |r.y| / sqrt(|r|^2 + tol); the plain form rounds r, the scale and each normalised entry to
 * f32, which moves its ybx by at most 6u (|r.y| / |r| <= |y| = 1), taken as 8u.  The plain
 * score ybx^2 then lies in [lo, hi].  A candidate whose hi is below the largest lo of its fit's
 * ungated candidates cannot be the step's winner and keeps its approximate score.
 */
__global__ __launch_bounds__(256) void k_tc_score(const FitDev *__restrict__ fits, const int *__restrict__ cfit,
                                                  const int *__restrict__ cxc, const int *__restrict__ cpar,
                                                  int count, double *__restrict__ score, int *__restrict__ flag,
                                                  double *__restrict__ lo, double *__restrict__ hi,
                                                  const int *__restrict__ dcount)
{
    if (dcount) count = *dcount;
    const int c = blockIdx.x * 16 + (threadIdx.x >> 4), l = threadIdx.x & 15;
    const bool ok = c < count;
    double tt = 0.0, ty = 0.0, yy = 0.0;
    const FitDev *F = ok ? fits + cfit[c] : fits;
    size_t row = 0;
    if (ok) {
        row = (size_t)cxc[c] * F->ldb + cpar[c];
        const double *T = F->tc + row * F->ldb;
        for (int k = l; k < F->m; k += 16) {
            const double v = T[k], yb = F->ybo[k];
            tt = fma(v, v, tt);
            ty = fma(v, yb, ty);
            yy = fma(yb, yb, yy);
        }
    }
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) {
        tt += __shfl_xor(tt, o, 16);
        ty += __shfl_xor(ty, o, 16);
        yy += __shfl_xor(yy, o, 16);
    }
    if (!ok || l != 0) return;
    const double bxx = F->tcn[row * 2], bxy = F->tcn[row * 2 + 1];
    const double s = bxx - tt;
    const double ry = bxy - ty;
    const bool gate = !(s > F->tol) || s * kGate < tt;
    flag[c] = gate ? 1 : 0;
    score[c] = gate ? 0.0 : (ry * ry) / (s + F->tol);
    constexpr double u = 0x1p-24;
    const double m = (double)F->m;
    const double dT = 40.0 * u * sqrt(m) * sqrt(fmax(bxx, 0.0));
    const double dry = dT * sqrt(yy);
    const double ds = 2.0 * sqrt(tt) * dT + dT * dT + 8.0 * u * m * tt;
    const double slo = s - ds, shi = s + ds, ary = fabs(ry);
    const double yhi = slo > 0.0 ? (ary + dry) / sqrt(slo + F->tol) : INFINITY;
    const double ylo = slo > F->tol ? fmax(ary - dry, 0.0) / sqrt(shi + F->tol) : 0.0;
    const double h = yhi + 8.0 * u, lw = fmax(ylo - 8.0 * u, 0.0);
    hi[c] = gate ? INFINITY : h * h;
    lo[c] = gate ? 0.0 : lw * lw;
}

/* mx[f] = largest lower score bound of fit f's ungated candidates (0 if none); one block per fit. */
__global__ __launch_bounds__(256) void k_fit_max(const int *__restrict__ c0, const int *__restrict__ cnt,
                                                 const double *__restrict__ score, const int *__restrict__ flag,
                                                 double *__restrict__ mx)
{
    __shared__ double sh[256];
    const int f = blockIdx.x, t = threadIdx.x;
    double v = 0.0;
    for (int c = c0[f] + t; c < c0[f] + cnt[f]; c += 256)
        if (!flag[c]) v = fmax(v, score[c]);
    sh[t] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) sh[t] = fmax(sh[t], sh[t + s]);
        __syncthreads();
    }
    if (t == 0) mx[f] = sh[0];
}

/*
 * Rescoring list, in candidate order: every gated candidate and every ungated one whose
 * score interval reaches its fit's largest lower bound (k_tc_score).  Two passes over
 * tiles of kSelTile candidates (256 threads, 4 coalesced candidates each): k_select_count
 * writes each tile's list and gated counts, k_select_write places the tile's entries after
 * those of the tiles before it; out_n[0] = list length, out_n[1] = gated count.  keep[i] = 1
 * when the rescoring stores list entry i's residual for the final pass: with store policy 0,
 * when the entry is ungated (an exact s > tol is then almost certain); policy 1 all, 2 none.
 */
constexpr int kSelTile = 1024;

__device__ inline bool sel_take(int c, const int *cfit, const double *hi, const int *flag, const double *mx)
{
    return flag[c] != 0 || hi[c] >= mx[cfit[c]];
}

__device__ inline int wave_sum_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_select_count(int P, const int *__restrict__ cfit,
                                                      const double *__restrict__ hi,
                                                      const int *__restrict__ flag, const double *__restrict__ mx,
                                                      int *__restrict__ blk, const int *__restrict__ dcount)
{
    if (dcount) P = *dcount;
    __shared__ int sn[4], sg[4];
    const int t = threadIdx.x, base = blockIdx.x * kSelTile;
    int n = 0, g = 0;
#pragma unroll
    for (int i = 0; i < kSelTile / 256; ++i) {
        const int c = base + 256 * i + t;
        if (c < P) {
            n += sel_take(c, cfit, hi, flag, mx) ? 1 : 0;
            g += flag[c] != 0 ? 1 : 0;
        }
    }
    n = wave_sum_i(n);
    g = wave_sum_i(g);
    if ((t & 63) == 0) {
        sn[t >> 6] = n;
        sg[t >> 6] = g;
    }
    __syncthreads();
    if (t == 0) {
        blk[2 * blockIdx.x] = sn[0] + sn[1] + sn[2] + sn[3];
        blk[2 * blockIdx.x + 1] = sg[0] + sg[1] + sg[2] + sg[3];
    }
}

__global__ __launch_bounds__(256) void k_select_write(int P, const int *__restrict__ cfit,
                                                      const double *__restrict__ hi,
                                                      const int *__restrict__ flag, const double *__restrict__ mx,
                                                      const int *__restrict__ blk,
                                                      int *__restrict__ list, int *__restrict__ keep,
                                                      int policy, int *__restrict__ out_n,
                                                      const int *__restrict__ dcount)
{
    if (dcount) P = *dcount;
    __shared__ int sw[4], sg[4];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63, b = blockIdx.x, base = b * kSelTile;
    // this tile's offset: the list counts of the tiles before it (the last tile also sums the gated counts)
    int off = 0, gat = 0;
    const bool last = b == (int)gridDim.x - 1;
    for (int j = t; j < b; j += 256) {
        off += blk[2 * j];
        if (last) gat += blk[2 * j + 1];
    }
    off = wave_sum_i(off);
    gat = wave_sum_i(gat);
    if (lane == 0) {
        sw[w] = off;
        sg[w] = gat;
    }
    __syncthreads();
    int run = sw[0] + sw[1] + sw[2] + sw[3];
    if (last && t == 0) {
        out_n[0] = run + blk[2 * b];
        out_n[1] = sg[0] + sg[1] + sg[2] + sg[3] + blk[2 * b + 1];
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll 1
    for (int i = 0; i < kSelTile / 256; ++i) {
        const int c = base + 256 * i + t;
        const bool tk = c < P && sel_take(c, cfit, hi, flag, mx);
        const unsigned long long m = __ballot(tk);
        if (lane == 0) sw[w] = __popcll(m);
        __syncthreads();
        int woff = run;
        for (int k = 0; k < w; ++k) woff += sw[k];
        if (tk) {
            const int pos = woff + __popcll(m & below);
            list[pos] = c;
            keep[pos] = policy == 1 ? 1 : policy == 2 ? 0 : (flag[c] == 0 ? 1 : 0);
        }
        run += sw[0] + sw[1] + sw[2] + sw[3];
        __syncthreads();
    }
}

/* Which rescored candidates keep their first residual in rbuf, from MARS_HIP_RESCORE_STORE:
 * "ungated" (default, 0), "all" (1) or "none" (2).  The scores do not depend on it: a candidate
 * without a stored residual that ends with s > tol has r recomputed from Bo and T. */
int store_policy()
{
    const char *e = std::getenv("MARS_HIP_RESCORE_STORE");
    if (!e || !*e || !std::strcmp(e, "ungated")) return 0;
    if (!std::strcmp(e, "all")) return 1;
    if (!std::strcmp(e, "none")) return 2;
    throw Error(MARS_HIP_EINVAL, std::string("MARS_HIP_RESCORE_STORE must be ungated, all or none, not '") + e + "'");
}

/* Check of the rescoring's premise, one block per fit: with E the largest exact score of the
 * fit's rescored candidates, no candidate left unrescored may have an upper bound at or above
 * E.  Violations (none while the interval bounds hold) are counted in stat[3]. */
__global__ __launch_bounds__(256) void k_window_check(const int *__restrict__ c0, const int *__restrict__ cnt,
                                                      const double *__restrict__ lin, const double *__restrict__ hi,
                                                      const int *__restrict__ done,
                                                      unsigned long long *__restrict__ stat)
{
    __shared__ double sh[256];
    __shared__ int sn[256];
    const int f = blockIdx.x, t = threadIdx.x;
    double e = 0.0;
    for (int c = c0[f] + t; c < c0[f] + cnt[f]; c += 256)
        if (done[c]) e = fmax(e, lin[c]);
    sh[t] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) sh[t] = fmax(sh[t], sh[t + s]);
        __syncthreads();
    }
    const double E = sh[0];
    int v = 0;
    for (int c = c0[f] + t; c < c0[f] + cnt[f]; c += 256)
        if (!done[c] && hi[c] >= E && E > 0.0) ++v;
    sn[t] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) sn[t] += sn[t + s];
        __syncthreads();
    }
    if (t == 0 && sn[0]) atomicAdd(stat + 3, (unsigned long long)sn[0]);
}

/* stat[1] += rescored count, stat[2] += gated count (nsel: [count, gated]). */
__global__ void k_stat_add(const int *__restrict__ nsel, unsigned long long *__restrict__ stat)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        stat[1] += (unsigned long long)nsel[0];
        stat[2] += (unsigned long long)nsel[1];
    }
}

__global__ void k_zero_i(int *__restrict__ p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

/* done[idx[i]] = 1 */
__global__ void k_mark(const int *__restrict__ idx, int n, int *__restrict__ done)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) done[idx[i]] = 1;
}

/* lin[idx[i]] = val[i] */
__global__ void k_put(const int *__restrict__ idx, const double *__restrict__ val, int n, double *__restrict__ lin)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) lin[idx[i]] = val[i];
}

constexpr size_t kMaxMixedTerms = 512;  // cache: p_x * ldb^2 doubles per fit (the append's limit too)

} // namespace

/* Allocate the caches of fits that have none yet. */
void ensure_tcache(mars_hip_ctx *ctx)
{
    for (FitState &fs : ctx->fits) {
        require(fs.m <= kMaxMixedTerms, "the mixed-precision form supports up to 512 basis functions");
        if (fs.tc.p) continue;
        fs.tc.ensure(fs.px * ctx->ldb * ctx->ldb);
        fs.tcn.ensure(fs.px * ctx->ldb * 2);
        fs.mc = 0;
    }
}

namespace {

/* Bring every fit's cache up to its current basis, one new column per fit per round. */
void tc_update(mars_hip_ctx *ctx)
{
    const size_t F = ctx->fits.size();
    hipStream_t st = ctx->stream;
    while (true) {
        std::vector<int> jf(F, -1);
        double work = 0, flops = 0, bytes = 0;
        for (size_t f = 0; f < F; ++f) {
            const FitState &fs = ctx->fits[f];
            if (fs.mc >= fs.m) continue;
            jf[f] = (int)fs.mc;
            const size_t tiles = (2 * fs.mc + 1 + 15) / 16;
            work += (double)fs.n * tiles * ((fs.px + kXT - 1) / kXT);
            flops += 2.0 * fs.n * (2.0 * fs.mc + 1) * fs.px;
            bytes += 4.0 * fs.n * (2.0 * (fs.mc + 1) + fs.px + 1);
        }
        if (work == 0) break;
        // Block items (fit, X-column block, group of W A tiles, row range), one wave per tile;
        // every X chunk is staged once per block.  W: the largest tile count of the round,
        // rounded up to a power of two (at most kTW; larger counts take several groups).
        int umax = 1;
        for (size_t f = 0; f < F; ++f)
            if (jf[f] >= 0) umax = std::max(umax, (2 * jf[f] + 1 + 15) / 16);
        const int W = umax <= 1 ? 1 : (umax <= 2 ? 2 : (umax <= 4 ? 4 : kTW));
        double bwork = 0;
        for (size_t f = 0; f < F; ++f) {
            if (jf[f] < 0) continue;
            const int tiles = (2 * jf[f] + 1 + 15) / 16;
            bwork += (double)ctx->fits[f].n * ((tiles + W - 1) / W) * ((ctx->fits[f].px + kXT - 1) / kXT);
        }
        const double target = ctx->tune.update_waves / W;   // waves per launch (default 16384)
        std::vector<Item> items;
        std::vector<int> groups, glist;
        for (size_t f = 0; f < F; ++f) {
            if (jf[f] < 0) continue;
            const FitState &fs = ctx->fits[f];
            const long long n = (long long)fs.n;
            const int tiles = (2 * jf[f] + 1 + 15) / 16;
            long long splits = std::max<long long>(1, std::llround(target * n / bwork));
            splits = std::min<long long>(splits, std::max<long long>(1, n / 256));
            const long long rows = ((n + splits - 1) / splits + kChunk - 1) / kChunk * kChunk;
            for (size_t xc0 = 0; xc0 < fs.px; xc0 += kXT) {
                const int nc = (int)std::min<size_t>(kXT, fs.px - xc0);
                for (int u0 = 0; u0 < tiles; u0 += W) {
                    const int first_block = (int)items.size();
                    for (long long r0 = 0; r0 < n; r0 += rows)
                        items.push_back(Item{(int)f, (int)xc0, nc, (jf[f] << 8) | u0, r0, std::min(n, r0 + rows)});
                    const int nb = (int)items.size() - first_block;
                    for (int u = u0; u < std::min(tiles, u0 + W); ++u) {
                        const int h[kGroupInts] = {(int)f, (int)xc0, nc, jf[f], u, (int)glist.size(), nb};
                        groups.insert(groups.end(), h, h + kGroupInts);
                        for (int bb = 0; bb < nb; ++bb) glist.push_back((first_block + bb) * W + (u - u0));
                    }
                }
            }
        }
        const size_t nh = groups.size();
        groups.insert(groups.end(), glist.begin(), glist.end());
        const size_t ni = items.size(), ng = nh / kGroupInts, nsub = ni * W;
        ctx->d_items.ensure(ni);
        ctx->d_cmap.ensure(groups.size());
        h2d_defer(ctx, ctx->d_items.p, items.data(), ni * sizeof(Item));
        h2d_defer(ctx, ctx->d_cmap.p, groups.data(), groups.size() * sizeof(int));
        h2d_flush(ctx);  // with the fits and candidates deferred before the update
        ctx->tpart.ensure(nsub * kXT * 16);
        ctx->aux.ensure(nsub * kXT * 2);
        {
            KTimer kt(ctx, K_MIXED_A, flops, bytes);  // the update kernel alone
            const dim3 grid((unsigned)std::max<size_t>(1, std::min<size_t>(ni, 65536)));
            const Item *di = (const Item *)ctx->d_items.p;
            const FitDev *dfit = (const FitDev *)ctx->d_fits.p;
            // MARS_HIP_CACHE=f64: the same update on FP64 MFMA (a measurement variant: it shows how much
            // of the mixed form's gain is the FP32 arithmetic rather than the cache)
            const char *ce = std::getenv("MARS_HIP_CACHE");
            const bool f64 = ce && std::strcmp(ce, "f64") == 0;
            auto go = [&](auto w, auto d) {
                constexpr int WW = decltype(w)::value;
                constexpr bool DD = decltype(d)::value;
                hipLaunchKernelGGL((k_tc_update<WW, DD>), grid, dim3(64 * WW), 0, st, dfit, di, (int)ni, ctx->tpart.p,
                                   ctx->aux.p);
            };
            using t_ = std::true_type;
            using f_ = std::false_type;
            switch (W) {
            case 1: f64 ? go(std::integral_constant<int, 1>{}, t_{}) : go(std::integral_constant<int, 1>{}, f_{}); break;
            case 2: f64 ? go(std::integral_constant<int, 2>{}, t_{}) : go(std::integral_constant<int, 2>{}, f_{}); break;
            case 4: f64 ? go(std::integral_constant<int, 4>{}, t_{}) : go(std::integral_constant<int, 4>{}, f_{}); break;
            default:
                f64 ? go(std::integral_constant<int, kTW>{}, t_{}) : go(std::integral_constant<int, kTW>{}, f_{});
                break;
            }
            MH_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_tc_reduce, dim3((unsigned)ng, kXT * 16 / 64), dim3(256), 0, st,
                           (const FitDev *)ctx->d_fits.p, (const int *)ctx->d_cmap.p, (const int *)ctx->d_cmap.p + nh,
                           (const double *)ctx->tpart.p, (const double *)ctx->aux.p);
        MH_LAUNCH_CHECK();
        for (size_t f = 0; f < F; ++f)
            if (jf[f] >= 0) ctx->fits[f].mc += 1;
    }
}

} // namespace

void mixed_prepare(mars_hip_ctx *ctx)
{
    PendingScope pending_scope(ctx);
    ensure_tcache(ctx);
    Cands none;  // the cache update reads no candidate list
    upload_fits(ctx, none);
    h2d_flush(ctx);
    tc_update(ctx);
}

void mixed_score(mars_hip_ctx *ctx, bool host_lists, const int *dP, int Pb)
{
    PendingScope pending_scope(ctx);
    Cands &cs = ctx->cand;
    const bool dev_cands = dP != nullptr;  // candidates generated on the device (count *dP <= Pb)
    const int P = dev_cands ? Pb : (int)cs.fit.size();
    if (P == 0) return;
    hipStream_t st = ctx->stream;
    const size_t F = ctx->fits.size();
    ensure_tcache(ctx);
    if (!dev_cands) {
        index_cands(ctx, cs);
        upload_fits(ctx, cs);
        upload_cands(ctx, cs);
        h2d_flush(ctx);  // k_tc_score reads them whether or not the update below has work
    }
    tc_update(ctx);
    ctx->cgate.ensure(P);
    ctx->lin.ensure(P);
    ctx->clo.ensure(P);
    ctx->chi.ensure(P);
    const int nblk = (P + kSelTile - 1) / kSelTile;
    ctx->ival.ensure(2 * F + 2 + 2 * (size_t)nblk);  // (with device candidates: already so sized)
    ctx->red.ensure(F);
    ctx->rsel.ensure(P);
    ctx->rkeep.ensure(P);
    ctx->cdone.ensure(P);
    int *d_nsel = ctx->ival.p + 2 * F, *d_blk = d_nsel + 2;
    if (!dev_cands) {
        std::vector<int> ranges(2 * F);
        for (size_t f = 0; f < F; ++f) {
            ranges[f] = cs.fit_c0[f];
            ranges[F + f] = cs.fit_P[f];
        }
        h2d_defer(ctx, ctx->ival.p, ranges.data(), 2 * F * sizeof(int));
        h2d_flush(ctx);
    }
    const int policy = store_policy();
    const bool device = !host_lists;
    if (device)  // buffers and fit views of the exact rescoring, before anything is queued below
        plain_score_device(ctx, cs, ctx->rsel.p, d_nsel, P, ctx->rkeep.p, ctx->lin.p, nullptr, ctx->cdone.p,
                           ctx->dredo.p, true, 1);
    // approximate scores, the rescoring list, and (device lists) the exact rescoring itself
    auto body = [&] {
        {
            double bytes = 0;
            for (size_t f = 0; f < F; ++f) bytes += 8.0 * cs.fit_P[f] * (ctx->fits[f].m + 2);
            KTimer kt(ctx, K_MIXED_B, 0.0, bytes);
            hipLaunchKernelGGL(k_tc_score, dim3((unsigned)((P + 15) / 16)), dim3(256), 0, st,
                               (const FitDev *)ctx->d_fits.p, (const int *)cs.d_fit.p, (const int *)cs.d_xc.p,
                               (const int *)cs.d_par.p, P, ctx->lin.p, ctx->cgate.p, ctx->clo.p, ctx->chi.p, dP);
            MH_LAUNCH_CHECK();
        }
        KTimer kt(ctx, K_RECHECK, 0.0, 0.0);
        hipLaunchKernelGGL(k_fit_max, dim3((unsigned)F), dim3(256), 0, st, (const int *)ctx->ival.p,
                           (const int *)ctx->ival.p + F, (const double *)ctx->clo.p, (const int *)ctx->cgate.p,
                           ctx->red.p);
        MH_LAUNCH_CHECK();
        // An exact gated score above a fit's largest lower bound only raises the maximum, so every
        // candidate that can win is on the list.
        hipLaunchKernelGGL(k_select_count, dim3((unsigned)nblk), dim3(256), 0, st, P, (const int *)cs.d_fit.p,
                           (const double *)ctx->chi.p, (const int *)ctx->cgate.p, (const double *)ctx->red.p, d_blk,
                           dP);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_select_write, dim3((unsigned)nblk), dim3(256), 0, st, P, (const int *)cs.d_fit.p,
                           (const double *)ctx->chi.p, (const int *)ctx->cgate.p, (const double *)ctx->red.p,
                           (const int *)d_blk, ctx->rsel.p, ctx->rkeep.p, policy, d_nsel, dP);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_zero_i, dim3((unsigned)std::min(1024, (P + 255) / 256)), dim3(256), 0, st, ctx->cdone.p,
                           P);
        MH_LAUNCH_CHECK();
        if (!device) return;
        hipLaunchKernelGGL(k_stat_add, dim3(1), dim3(64), 0, st, (const int *)d_nsel, ctx->dstat.p);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_zero_i, dim3(1), dim3(64), 0, st, ctx->dredo.p, 1);
        MH_LAUNCH_CHECK();
        plain_score_device(ctx, cs, ctx->rsel.p, d_nsel, P, ctx->rkeep.p, ctx->lin.p, nullptr, ctx->cdone.p,
                           ctx->dredo.p, true, 2);
        hipLaunchKernelGGL(k_window_check, dim3((unsigned)F), dim3(256), 0, st, (const int *)ctx->ival.p,
                           (const int *)ctx->ival.p + F, (const double *)ctx->lin.p, (const double *)ctx->chi.p,
                           (const int *)ctx->cdone.p, ctx->dstat.p);
        MH_LAUNCH_CHECK();
    };
    if (device && dev_cands && ctx->use_graphs) {
        // one graph per basis-size class: every argument below is fixed for the fit
        size_t mmax = 1;
        for (const FitState &fs : ctx->fits) mmax = std::max(mmax, fs.m);
        std::vector<unsigned long long> key = {1ull, (unsigned long long)((mmax + 15) / 16),
                                               (unsigned long long)P, (unsigned long long)F,
                                               (unsigned long long)policy};
        graph_key_buffers(ctx, key);
        tuning_key(ctx, key);
        KTimer kt(ctx, K_RECHECK, 0.0, 0.0);
        run_graph(ctx, key, body);
        return;
    }
    body();
    if (device) return;
    {
        require(!dev_cands, "internal: host-built rescoring lists need host-built candidates");
        // the list comes back to the host, which builds the rescoring's work lists
        int *pin = ctx->sel_pin.get<int>(2 + (size_t)P);  // [count, gated count, list...]
        MH_CHECK(hipMemcpyAsync(pin, d_nsel, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        ctx_sync(ctx);
        const int R = pin[0];
        require(R >= 0 && R <= P, "rescoring list length out of range");
        ctx->gated += pin[1];
        ctx->rescored += R;
        if (R > 0) {
            MH_CHECK(hipMemcpyAsync(pin + 2, ctx->rsel.p, R * sizeof(int), hipMemcpyDeviceToHost, st));
            ctx_sync(ctx);
            const int *sub = pin + 2;
            Cands &rc = ctx->rcand;
            rc.fit.resize(R);
            rc.xc.resize(R);
            rc.par.resize(R);
            for (int i = 0; i < R; ++i) {
                rc.fit[i] = cs.fit[sub[i]];
                rc.xc[i] = cs.xc[sub[i]];
                rc.par[i] = cs.par[sub[i]];
            }
            ctx->rlin.ensure(R);
            ctx->rybx.ensure(R);
            plain_score(ctx, rc, false, ctx->rlin.p, ctx->rybx.p, nullptr, nullptr, true, ctx->rkeep.p);
            hipLaunchKernelGGL(k_put, dim3((R + 255) / 256), dim3(256), 0, st, (const int *)ctx->rsel.p,
                               (const double *)ctx->rlin.p, R, ctx->lin.p);
            MH_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_mark, dim3((R + 255) / 256), dim3(256), 0, st, (const int *)ctx->rsel.p, R,
                               ctx->cdone.p);
            MH_LAUNCH_CHECK();
        }
    }
    hipLaunchKernelGGL(k_window_check, dim3((unsigned)F), dim3(256), 0, st, (const int *)ctx->ival.p,
                       (const int *)ctx->ival.p + F, (const double *)ctx->lin.p, (const double *)ctx->chi.p,
                       (const int *)ctx->cdone.p, ctx->dstat.p);
    MH_LAUNCH_CHECK();
}

} // namespace mh
