// Mixed-precision candidate scoring on gfx950 (linear delta-SSE).
//
// Bulk pass (FP32 MFMA, v_mfma_f32_16x16x4f32): T = Bo^T Bx accumulated in
// FP32 within 256-row chunks and in FP64 across chunks; in the same pass the
// exact FP64 sums |Bx|^2 and Bx.y (f32 products are exact in f64).  The
// residual is never formed: with r = Bx - Bo T and Bo orthonormal,
//     |r|^2 = |Bx|^2 - |T|^2,      r.y = Bx.y - T.ybo,
//     score = (r.y)^2 / (|r|^2 + tol)   (0 when |r|^2 <= tol)      (kernels.cc:401-416,
//                                                                   marsalgo.cc:500-505)
// which is accurate while |r|^2 is not a small difference of large numbers --
// exactly the DGKS gate's condition (9 |r|^2 >= |T|^2, kernels.h:17).
// Candidates that trip the gate, and every candidate whose approximate score
// lies within a relative window of its fit's best, are rescored by the exact
// FP64 path (ortho.hip), so each step's winner and its delta-SSE are FP64
// values computed with the reference's rounding points.
#include "internal.h"
#include "kernels.h"
#include "tgemm.h"

#include <algorithm>

namespace mh {

namespace {

constexpr int kTile = kTileC;      // candidates per item (4 waves x 16)
constexpr int kChunk = kRC;        // rows per FP32 accumulation chunk (folded into FP64)
constexpr double kWindow = 1e-4;   // relative near-tie window rescored in FP64

/* Per candidate: T, |Bx|^2, Bx.y from the item partials; approximate score and gate flag. */
__global__ __launch_bounds__(64) void k_mixed_reduce(const FitDev *__restrict__ fits,
                                                     const int *__restrict__ cfit, int count,
                                                     const int *__restrict__ cmap,
                                                     const double *__restrict__ tpart,
                                                     const double *__restrict__ npart, int ldT,
                                                     double *__restrict__ score, int *__restrict__ flag)
{
    const int c = blockIdx.x, lane = threadIdx.x;
    if (c >= count) return;
    const FitDev F = fits[cfit[c]];
    const int first = cmap[3 * c], cnt = cmap[3 * c + 1], slot = cmap[3 * c + 2];
    double tt = 0.0, ty = 0.0;
    for (int k = lane; k < ldT; k += 64) {
        double s = 0.0;
        for (int i = 0; i < cnt; ++i) s += tpart[((size_t)(first + i) * kTile + slot) * ldT + k];
        if (k < F.m) {
            tt = fma(s, s, tt);
            ty = fma(s, F.ybo[k], ty);
        }
    }
    tt = wave_sum(tt);
    ty = wave_sum(ty);
    if (lane != 0) return;
    double bxx = 0.0, bxy = 0.0;
    for (int i = 0; i < cnt; ++i) {
        bxx += npart[((size_t)(first + i) * kTile + slot) * 2];
        bxy += npart[((size_t)(first + i) * kTile + slot) * 2 + 1];
    }
    const double s = bxx - tt;
    const double ry = bxy - ty;
    const bool gate = !(s > F.tol) || s * kGate < tt;
    flag[c] = gate ? 1 : 0;
    score[c] = gate ? 0.0 : (ry * ry) / (s + F.tol);
}

/* lin[idx[i]] = val[i] */
__global__ void k_put(const int *__restrict__ idx, const double *__restrict__ val, int n,
                      double *__restrict__ lin)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) lin[idx[i]] = val[i];
}

template <typename... A>
void launch_mixed_t(int mt, dim3 g, hipStream_t st, A... args)
{
    switch (mt) {
    case 1: hipLaunchKernelGGL((k_tgemm_lds<1, true, false, true>), g, dim3(256), 0, st, args...); break;
    case 2: hipLaunchKernelGGL((k_tgemm_lds<2, true, false, true>), g, dim3(256), 0, st, args...); break;
    case 3: hipLaunchKernelGGL((k_tgemm_lds<3, true, false, true>), g, dim3(256), 0, st, args...); break;
    case 4: hipLaunchKernelGGL((k_tgemm_lds<4, true, false, true>), g, dim3(256), 0, st, args...); break;
    case 5: hipLaunchKernelGGL((k_tgemm_lds<5, true, false, true>), g, dim3(256), 0, st, args...); break;
    case 6: hipLaunchKernelGGL((k_tgemm_lds<6, true, false, true>), g, dim3(256), 0, st, args...); break;
    case 7: hipLaunchKernelGGL((k_tgemm_lds<7, true, false, true>), g, dim3(256), 0, st, args...); break;
    default: hipLaunchKernelGGL((k_tgemm_lds<8, true, false, true>), g, dim3(256), 0, st, args...); break;
    }
    MH_LAUNCH_CHECK();
}

} // namespace

/* Items of the bulk pass (shared with the FP64 path's work-list builder). */
Pass make_items(const Cands &cs, const mars_hip_ctx *ctx, int tile, double target, long long align);

void mixed_score(mars_hip_ctx *ctx)
{
    Cands &cs = ctx->cand;
    const int P = (int)cs.fit.size();
    if (P == 0) return;
    hipStream_t st = ctx->stream;
    const size_t F = ctx->fits.size();
    index_cands(ctx, cs);
    upload_fits(ctx, cs);
    upload_cands(ctx, cs);
    size_t mmax = 1;
    double flops = 0, bytes = 0;
    for (size_t f = 0; f < F; ++f) {
        mmax = std::max(mmax, ctx->fits[f].m);
        flops += 2.0 * ctx->fits[f].n * ctx->fits[f].m * cs.fit_P[f];
        bytes += 4.0 * ctx->fits[f].n * (2.0 * ctx->fits[f].m + 2);
    }
    require(mmax <= 128, "the mixed-precision form supports up to 128 basis functions");
    const int mt = (int)((mmax + 15) / 16);
    const int ldT = mt * 16;
    Pass p = make_items(cs, ctx, kTile, 2048, kChunk);
    ctx->d_items.ensure(p.items.size() + 1);
    ctx->d_cmap.ensure(p.cmap.size() + 1);
    MH_CHECK(hipMemcpyAsync(ctx->d_items.p, p.items.data(), p.items.size() * sizeof(Item),
                            hipMemcpyHostToDevice, st));
    MH_CHECK(hipMemcpyAsync(ctx->d_cmap.p, p.cmap.data(), p.cmap.size() * sizeof(int), hipMemcpyHostToDevice, st));
    ctx->tpart.ensure(p.items.size() * kTile * ldT + 1);
    ctx->aux.ensure(p.items.size() * kTile * 2 + 1);
    ctx->flag.ensure(P);
    ctx->lin.ensure(P);
    {
        KTimer kt(ctx, K_MIXED_A, flops, bytes);
        const dim3 g((unsigned)std::max<size_t>(1, std::min<size_t>(p.items.size(), 4096)));
        launch_mixed_t(mt, g, st, (const FitDev *)ctx->d_fits.p, (const Item *)ctx->d_items.p,
                       (int)p.items.size(), (const int *)cs.d_xc.p, (const int *)cs.d_par.p, (const int *)nullptr, ldT,
                       ctx->tpart.p, ctx->aux.p);
        hipLaunchKernelGGL(k_mixed_reduce, dim3(P), dim3(64), 0, st, ctx->d_fits.p, cs.d_fit.p, P, ctx->d_cmap.p,
                           ctx->tpart.p, ctx->aux.p, ldT, ctx->lin.p, ctx->flag.p);
        MH_LAUNCH_CHECK();
    }
    std::vector<double> approx(P);
    std::vector<int> flags(P);
    MH_CHECK(hipMemcpyAsync(approx.data(), ctx->lin.p, P * sizeof(double), hipMemcpyDeviceToHost, st));
    MH_CHECK(hipMemcpyAsync(flags.data(), ctx->flag.p, P * sizeof(int), hipMemcpyDeviceToHost, st));
    MH_CHECK(hipStreamSynchronize(st));

    KTimer kt(ctx, K_RECHECK, 0.0, 0.0);
    // 1) rescore the gated candidates exactly; 2) rescore everything within the window of
    //    the (partly exact) per-fit maximum.  Candidates stay grouped by fit.
    std::vector<int> sub;
    auto rescore = [&](const std::vector<int> &ids) {
        if (ids.empty()) return;
        Cands rc;
        rc.fit.reserve(ids.size());
        for (int c : ids) {
            rc.fit.push_back(cs.fit[c]);
            rc.xc.push_back(cs.xc[c]);
            rc.par.push_back(cs.par[c]);
        }
        DBuf<double> lin2, ybx2;
        lin2.ensure(ids.size());
        ybx2.ensure(ids.size());
        plain_score(ctx, rc, false, lin2.p, ybx2.p);
        std::vector<double> exact(ids.size());
        MH_CHECK(hipMemcpyAsync(exact.data(), lin2.p, ids.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        MH_CHECK(hipStreamSynchronize(st));
        for (size_t i = 0; i < ids.size(); ++i) approx[ids[i]] = exact[i];
        rc.d_fit.release();
        rc.d_xc.release();
        rc.d_par.release();
        lin2.release();
        ybx2.release();
    };
    std::vector<char> done(P, 0);
    for (int c = 0; c < P; ++c)
        if (flags[c]) {
            sub.push_back(c);
            done[c] = 1;
        }
    rescore(sub);
    sub.clear();
    for (size_t f = 0; f < F; ++f) {
        const int b = cs.fit_c0[f], e = b + cs.fit_P[f];
        double mx = 0.0;
        for (int c = b; c < e; ++c) mx = std::max(mx, approx[c]);
        if (mx <= 0.0) continue;
        for (int c = b; c < e; ++c)
            if (!done[c] && approx[c] >= mx * (1.0 - kWindow)) {
                sub.push_back(c);
                done[c] = 1;
            }
    }
    rescore(sub);
    MH_CHECK(hipMemcpyAsync(ctx->lin.p, approx.data(), P * sizeof(double), hipMemcpyHostToDevice, st));
    long nre = 0;
    for (char d : done) nre += d;
    ctx->rescored += nre;
    ctx->scored += P;
}

} // namespace mh
