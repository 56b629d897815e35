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

namespace mh {

namespace {

constexpr int kChunk = 32;         // rows per FP32 accumulation chunk (folded into FP64)
constexpr int kXT = 64;            // X columns per update item (4 waves x 16)
constexpr double kWindow = 1e-4;   // relative near-tie window rescored in FP64
constexpr int kGroupInts = 6;      // fit, xc0, nc, j, first item, item count

/*
 * Cache update for new basis column j = item.pad of one fit: for the item's
 * X columns [c0, c0+nc) and rows [r0, r1), partial sums of
 *   U[a][xc] = sum_i A[i][a] xs[i][xc],   a < 2j+1,
 *   A[i][a] = f32(Bo[i][a] B[i][j])          (a <= j:  k = a,  parent j)
 *           = f32(Bo[i][j] B[i][a-j-1])      (a >  j:  k = j,  parent a-j-1)
 * and of |Bx|^2, Bx.y for Bx = f32(B[:,j] xs[:,xc]).
 * Per 32-row chunk the block stages the Bo and B rows (columns <= j) and the
 * 64 scaled X columns with 16-byte loads, forms the chunk's A tile once in
 * LDS, and wave w runs the MFMAs of X columns 16w..16w+15.  Rows r0 are
 * multiples of 32; X and y are padded to a multiple of 32 rows, so the vector
 * loads stay inside their buffers.
 */
template <int MT>
__global__ __launch_bounds__(256) void k_tc_update(const FitDev *__restrict__ fits,
                                                   const Item *__restrict__ items, int nitems, int ldA,
                                                   double *__restrict__ tpart, double *__restrict__ npart)
{
    constexpr int SW = 8 * MT;  // staged basis columns: j + 1 <= SW since 2j + 1 <= 16 MT
    constexpr int KW = 16 * MT;
    __shared__ float sBo[kChunk][SW + 4];
    __shared__ float sB[kChunk][SW + 4];
    __shared__ float sA[kChunk][KW + 4];
    __shared__ float sX[kChunk][kXT + 4];
    __shared__ double sN[2][4][kXT];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int kr = lane >> 4, kc = lane & 15;
    const int sc = t & (kXT - 1), g = t >> 6;  // staging role: X column sc, rows 8g..8g+7
    const float4 zero4 = {0.f, 0.f, 0.f, 0.f};
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item I = items[it];
        const FitDev F = fits[I.fit];
        const int j = I.pad, na = 2 * j + 1, mtf = (na + 15) / 16, q4 = (j + 4) / 4;
        const bool sok = sc < I.nc;
        const int xc = I.c0 + (sok ? sc : 0);
        const float sx = F.s[xc];
        const float *xp = F.X + (size_t)xc * F.ldx;
        v4d acc[MT];
#pragma unroll
        for (int u = 0; u < MT; ++u) acc[u] = v4d{0, 0, 0, 0};
        double bxx = 0.0, bxy = 0.0;
        const bool wok = 16 * w < I.nc;
        for (long long r = I.r0; r < I.r1; r += kChunk) {
            const int rc = (int)(I.r1 - r < kChunk ? I.r1 - r : kChunk);
            __syncthreads();
            for (int e = t; e < kChunk * q4; e += 256) {
                const int i = e / q4, q = e - i * q4;
                float4 a = zero4, b = zero4;
                if (i < rc) {
                    a = *(const float4 *)(F.Bo + (r + i) * F.ldb + 4 * q);
                    b = *(const float4 *)(F.B + (r + i) * F.ldb + 4 * q);
                }
                *(float4 *)&sBo[i][4 * q] = a;
                *(float4 *)&sB[i][4 * q] = b;
            }
            float xv[8];
            {
                float4 x0 = zero4, x1 = zero4;
                if (sok) {
                    const float4 *src = (const float4 *)(xp + r + 8 * g);
                    x0 = src[0];
                    x1 = src[1];
                }
                xv[0] = x0.x; xv[1] = x0.y; xv[2] = x0.z; xv[3] = x0.w;
                xv[4] = x1.x; xv[5] = x1.y; xv[6] = x1.z; xv[7] = x1.w;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    xv[u] = (8 * g + u < rc) ? __fmul_rn(xv[u], sx) : 0.0f;
                    sX[8 * g + u][sc] = xv[u];
                }
            }
            __syncthreads();
            for (int e = t; e < kChunk * na; e += 256) {
                const int i = e / na, a = e - i * na;
                const int ka = a <= j ? a : j, pa = a <= j ? j : a - j - 1;
                sA[i][a] = __fmul_rn(sBo[i][ka], sB[i][pa]);
            }
            for (int e = t; e < kChunk * (16 * mtf - na); e += 256) {
                const int i = e / (16 * mtf - na), a = na + e - i * (16 * mtf - na);
                sA[i][a] = 0.0f;
            }
            if (sok) {
                const float4 *yp = (const float4 *)(F.y + r + 8 * g);
                const float4 y0 = yp[0], y1 = yp[1];
                const float yv[8] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z, y1.w};
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const double bx = (double)__fmul_rn(sB[8 * g + u][j], xv[u]);
                    bxx = fma(bx, bx, bxx);
                    bxy = fma(bx, (double)yv[u], bxy);
                }
            }
            __syncthreads();
            if (wok) {
                v4f part[MT];
#pragma unroll
                for (int u = 0; u < MT; ++u) part[u] = v4f{0, 0, 0, 0};
#pragma unroll
                for (int q = 0; q < kChunk; q += 4) {
                    const float b = sX[q + kr][16 * w + kc];
#pragma unroll
                    for (int u = 0; u < MT; ++u)
                        if (u < mtf) part[u] = mfma_f32(sA[q + kr][16 * u + kc], b, part[u]);
                }
#pragma unroll
                for (int u = 0; u < MT; ++u)
                    if (u < mtf)
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[u][v] += (double)part[u][v];
            }
        }
        const int slot = 16 * w + kc;
        if (wok && slot < I.nc) {
            double *dst = tpart + ((size_t)it * kXT + slot) * ldA;
#pragma unroll
            for (int u = 0; u < MT; ++u)
                if (u < mtf)
#pragma unroll
                    for (int v = 0; v < 4; ++v) dst[16 * u + 4 * kr + v] = acc[u][v];  // D row 4(l/16)+v
        }
        sN[0][g][sc] = bxx;
        sN[1][g][sc] = bxy;
        __syncthreads();
        if (g == 0 && sok) {
            npart[((size_t)it * kXT + sc) * 2] = ((sN[0][0][sc] + sN[0][1][sc]) + sN[0][2][sc]) + sN[0][3][sc];
            npart[((size_t)it * kXT + sc) * 2 + 1] = ((sN[1][0][sc] + sN[1][1][sc]) + sN[1][2][sc]) + sN[1][3][sc];
        }
    }
}

/* Sum each (fit, X-column tile) group's row partials in item order and write the new cache entries. */
__global__ __launch_bounds__(256) void k_tc_reduce(const FitDev *__restrict__ fits, const int *__restrict__ groups,
                                                   int ldA, const double *__restrict__ tpart,
                                                   const double *__restrict__ npart)
{
    const int *G = groups + (size_t)blockIdx.x * kGroupInts;
    const int fit = G[0], xc0 = G[1], nc = G[2], j = G[3], first = G[4], count = G[5];
    const FitDev F = fits[fit];
    const int na = 2 * j + 1;
    const long long ldb = F.ldb;
    for (int e = threadIdx.x; e < nc * na; e += 256) {
        const int xl = e / na, a = e - xl * na;
        double s = 0.0;
        for (int i = 0; i < count; ++i) s += tpart[((size_t)(first + i) * kXT + xl) * ldA + a];
        const long long xc = xc0 + xl;
        const int k = a <= j ? a : j, par = a <= j ? j : a - j - 1;
        F.tc[(xc * ldb + par) * ldb + k] = s;
    }
    for (int xl = threadIdx.x; xl < nc; xl += 256) {
        double s0 = 0.0, s1 = 0.0;
        for (int i = 0; i < count; ++i) {
            s0 += npart[((size_t)(first + i) * kXT + xl) * 2];
            s1 += npart[((size_t)(first + i) * kXT + xl) * 2 + 1];
        }
        F.tcn[((xc0 + xl) * ldb + j) * 2] = s0;
        F.tcn[((xc0 + xl) * ldb + j) * 2 + 1] = s1;
    }
}

/* Approximate score and gate flag per candidate from the cache; 16 lanes per candidate. */
__global__ __launch_bounds__(256) void k_tc_score(const FitDev *__restrict__ fits, const int *__restrict__ cfit,
                                                  const int *__restrict__ cxc, const int *__restrict__ cpar,
                                                  int count, double *__restrict__ score, int *__restrict__ flag)
{
    const int c = blockIdx.x * 16 + (threadIdx.x >> 4), l = threadIdx.x & 15;
    const bool ok = c < count;
    double tt = 0.0, ty = 0.0;
    const FitDev *F = ok ? fits + cfit[c] : fits;
    size_t row = 0;
    if (ok) {
        row = (size_t)cxc[c] * F->ldb + cpar[c];
        const double *T = F->tc + row * F->ldb;
        for (int k = l; k < F->m; k += 16) {
            const double v = T[k];
            tt = fma(v, v, tt);
            ty = fma(v, F->ybo[k], ty);
        }
    }
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) {
        tt += __shfl_xor(tt, o, 16);
        ty += __shfl_xor(ty, o, 16);
    }
    if (!ok || l != 0) return;
    const double bxx = F->tcn[row * 2], bxy = F->tcn[row * 2 + 1];
    const double s = bxx - tt;
    const double ry = bxy - ty;
    const bool gate = !(s > F->tol) || s * kGate < tt;
    flag[c] = gate ? 1 : 0;
    score[c] = gate ? 0.0 : (ry * ry) / (s + F->tol);
}

template <typename... A>
void launch_update(int mt, dim3 g, hipStream_t st, A... args)
{
    switch (mt) {
    case 1: hipLaunchKernelGGL(k_tc_update<1>, g, dim3(256), 0, st, args...); break;
    case 2: hipLaunchKernelGGL(k_tc_update<2>, g, dim3(256), 0, st, args...); break;
    case 3: hipLaunchKernelGGL(k_tc_update<3>, g, dim3(256), 0, st, args...); break;
    case 4: hipLaunchKernelGGL(k_tc_update<4>, g, dim3(256), 0, st, args...); break;
    case 5: hipLaunchKernelGGL(k_tc_update<5>, g, dim3(256), 0, st, args...); break;
    case 6: hipLaunchKernelGGL(k_tc_update<6>, g, dim3(256), 0, st, args...); break;
    case 7: hipLaunchKernelGGL(k_tc_update<7>, g, dim3(256), 0, st, args...); break;
    case 8: hipLaunchKernelGGL(k_tc_update<8>, g, dim3(256), 0, st, args...); break;
    case 9: hipLaunchKernelGGL(k_tc_update<9>, g, dim3(256), 0, st, args...); break;
    case 10: hipLaunchKernelGGL(k_tc_update<10>, g, dim3(256), 0, st, args...); break;
    case 11: hipLaunchKernelGGL(k_tc_update<11>, g, dim3(256), 0, st, args...); break;
    case 12: hipLaunchKernelGGL(k_tc_update<12>, g, dim3(256), 0, st, args...); break;
    case 13: hipLaunchKernelGGL(k_tc_update<13>, g, dim3(256), 0, st, args...); break;
    case 14: hipLaunchKernelGGL(k_tc_update<14>, g, dim3(256), 0, st, args...); break;
    case 15: hipLaunchKernelGGL(k_tc_update<15>, g, dim3(256), 0, st, args...); break;
    default: hipLaunchKernelGGL(k_tc_update<16>, g, dim3(256), 0, st, args...); break;
    }
    MH_LAUNCH_CHECK();
}

constexpr size_t kMaxMixedTerms = 128;  // 2j+1 <= 255 cache columns per update: MT <= 16

/* Allocate the caches of fits that have none yet. */
void ensure_tcache(mars_hip_ctx *ctx)
{
    for (FitState &fs : ctx->fits) {
        require(fs.m <= kMaxMixedTerms, "the mixed-precision form supports up to 128 basis functions");
        if (fs.tc.p) continue;
        fs.tc.ensure(fs.px * ctx->ldb * ctx->ldb);
        fs.tcn.ensure(fs.px * ctx->ldb * 2);
        fs.mc = 0;
    }
}

/* Bring every fit's cache up to its current basis, one new column per fit per round. */
void tc_update(mars_hip_ctx *ctx)
{
    const size_t F = ctx->fits.size();
    hipStream_t st = ctx->stream;
    while (true) {
        std::vector<int> jf(F, -1);
        size_t jmax = 0;
        double work = 0, flops = 0, bytes = 0;
        for (size_t f = 0; f < F; ++f) {
            const FitState &fs = ctx->fits[f];
            if (fs.mc >= fs.m) continue;
            jf[f] = (int)fs.mc;
            jmax = std::max(jmax, fs.mc);
            work += (double)fs.n * ((fs.px + kXT - 1) / kXT);
            flops += 2.0 * fs.n * (2.0 * fs.mc + 1) * fs.px;
            bytes += 4.0 * fs.n * (2.0 * (fs.mc + 1) + fs.px + 1);
        }
        if (work == 0) break;
        const int mt = (int)((2 * jmax + 1 + 15) / 16);
        const int ldA = 16 * mt;
        std::vector<Item> items;
        std::vector<int> groups;
        for (size_t f = 0; f < F; ++f) {
            if (jf[f] < 0) continue;
            const FitState &fs = ctx->fits[f];
            const long long n = (long long)fs.n;
            long long splits = std::max<long long>(1, std::llround(2048.0 * n / work));
            splits = std::min<long long>(splits, std::max<long long>(1, n / 256));
            const long long rows = ((n + splits - 1) / splits + kChunk - 1) / kChunk * kChunk;
            for (size_t xc0 = 0; xc0 < fs.px; xc0 += kXT) {
                const int nc = (int)std::min<size_t>(kXT, fs.px - xc0);
                const int first = (int)items.size();
                for (long long r0 = 0; r0 < n; r0 += rows)
                    items.push_back(Item{(int)f, (int)xc0, nc, jf[f], r0, std::min(n, r0 + rows)});
                const int g[kGroupInts] = {(int)f, (int)xc0, nc, jf[f], first, (int)items.size() - first};
                groups.insert(groups.end(), g, g + kGroupInts);
            }
        }
        const size_t ni = items.size(), ng = groups.size() / kGroupInts;
        ctx->d_items.ensure(ni);
        ctx->d_cmap.ensure(groups.size());
        MH_CHECK(hipMemcpyAsync(ctx->d_items.p, items.data(), ni * sizeof(Item), hipMemcpyHostToDevice, st));
        MH_CHECK(hipMemcpyAsync(ctx->d_cmap.p, groups.data(), groups.size() * sizeof(int), hipMemcpyHostToDevice,
                                st));
        ctx->tpart.ensure(ni * kXT * ldA);
        ctx->aux.ensure(ni * kXT * 2);
        {
            KTimer kt(ctx, K_MIXED_A, flops, bytes);
            const dim3 grid((unsigned)std::max<size_t>(1, std::min<size_t>(ni, 4096)));
            launch_update(mt, grid, st, (const FitDev *)ctx->d_fits.p, (const Item *)ctx->d_items.p, (int)ni, ldA,
                          ctx->tpart.p, ctx->aux.p);
            hipLaunchKernelGGL(k_tc_reduce, dim3((unsigned)ng), dim3(256), 0, st, (const FitDev *)ctx->d_fits.p,
                               (const int *)ctx->d_cmap.p, ldA, (const double *)ctx->tpart.p,
                               (const double *)ctx->aux.p);
            MH_LAUNCH_CHECK();
        }
        for (size_t f = 0; f < F; ++f)
            if (jf[f] >= 0) ctx->fits[f].mc += 1;
    }
}

} // namespace

void mixed_score(mars_hip_ctx *ctx)
{
    Cands &cs = ctx->cand;
    const int P = (int)cs.fit.size();
    if (P == 0) return;
    hipStream_t st = ctx->stream;
    const size_t F = ctx->fits.size();
    index_cands(ctx, cs);
    ensure_tcache(ctx);
    upload_fits(ctx, cs);
    upload_cands(ctx, cs);
    tc_update(ctx);
    ctx->flag.ensure(P);
    ctx->lin.ensure(P);
    {
        double bytes = 0;
        for (size_t f = 0; f < F; ++f) bytes += 8.0 * cs.fit_P[f] * (ctx->fits[f].m + 2);
        KTimer kt(ctx, K_MIXED_B, 0.0, bytes);
        hipLaunchKernelGGL(k_tc_score, dim3((unsigned)((P + 15) / 16)), dim3(256), 0, st,
                           (const FitDev *)ctx->d_fits.p, (const int *)cs.d_fit.p, (const int *)cs.d_xc.p,
                           (const int *)cs.d_par.p, P, ctx->lin.p, ctx->flag.p);
        MH_LAUNCH_CHECK();
    }
    std::vector<double> approx(P);
    std::vector<int> flags(P);
    MH_CHECK(hipMemcpyAsync(approx.data(), ctx->lin.p, P * sizeof(double), hipMemcpyDeviceToHost, st));
    MH_CHECK(hipMemcpyAsync(flags.data(), ctx->flag.p, P * sizeof(int), hipMemcpyDeviceToHost, st));
    MH_CHECK(hipStreamSynchronize(st));

    KTimer kt(ctx, K_RECHECK, 0.0, 0.0);
    // Rescore exactly: the gated candidates, and every ungated one within the window of
    // its fit's best ungated approximation.  An exact gated score above that best only
    // raises the maximum, so the window taken here contains the final one.
    std::vector<int> sub;
    for (size_t f = 0; f < F; ++f) {
        const int b = cs.fit_c0[f], e = b + cs.fit_P[f];
        double mx = 0.0;
        for (int c = b; c < e; ++c)
            if (!flags[c]) mx = std::max(mx, approx[c]);
        for (int c = b; c < e; ++c)
            if (flags[c] || (mx > 0.0 && approx[c] >= mx * (1.0 - kWindow))) sub.push_back(c);
    }
    if (!sub.empty()) {
        const size_t R = sub.size();
        Cands &rc = ctx->rcand;
        rc.fit.resize(R);
        rc.xc.resize(R);
        rc.par.resize(R);
        for (size_t i = 0; i < R; ++i) {
            rc.fit[i] = cs.fit[sub[i]];
            rc.xc[i] = cs.xc[sub[i]];
            rc.par[i] = cs.par[sub[i]];
        }
        ctx->rlin.ensure(R);
        ctx->rybx.ensure(R);
        plain_score(ctx, rc, false, ctx->rlin.p, ctx->rybx.p);
        std::vector<double> exact(R);
        MH_CHECK(hipMemcpyAsync(exact.data(), ctx->rlin.p, R * sizeof(double), hipMemcpyDeviceToHost, st));
        MH_CHECK(hipStreamSynchronize(st));
        for (size_t i = 0; i < R; ++i) approx[sub[i]] = exact[i];
    }
    MH_CHECK(hipMemcpyAsync(ctx->lin.p, approx.data(), P * sizeof(double), hipMemcpyHostToDevice, st));
    ctx->rescored += (long)sub.size();
}

} // namespace mh
