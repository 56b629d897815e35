// Plain (FP64) candidate orthonormalization on gfx950.
//
// Restates mars::orthonormalize (kernels.cc:283-417) for a whole block of
// candidates (X column, parent basis) at once, keeping the reference's f32
// rounding points (SURVEY Appendix A R5-R10 in the design notes):
//   Bx  = f32(B[:,par] * f32(X[:,xc] * s[xc]))          -- rebuilt on the fly, never stored
//   T   = Bo^T Bx                     (FP64 MFMA, split over rows, fixed-order reduce)
//   r   = f32(Bx - Bo T),  s = sum r^2  (FP64 MFMA; r kept row-major in HBM)
//   DGKS: s > tol and 9 s < |T|^2  ->  T2 = Bo^T r, r = f32(r - Bo T2), s = sum r^2
//         (flagged columns compacted on the device; no host round trip)
//   Bxn = f32(r * f32(1/sqrt(s + tol))) (0 if s <= tol),  ybx = sum Bxn y
// All reductions run in a fixed order, so results are run-to-run identical.
#include "internal.h"
#include "kernels.h"

#include <algorithm>

namespace mh {

namespace {

/* ---------------------------------------------------------------------- */
/* T = Bo^T Bx (or Bo^T r for the DGKS retry): per-(split, wave) partials  */
/* grid (cand tiles of 32, row splits, basis-tile groups), 4 waves          */
/* ---------------------------------------------------------------------- */
template <int MT, bool RETRY>
__global__ __launch_bounds__(256) void k_tgemm(
    const float *__restrict__ Bo, const float *__restrict__ B, size_t ldb, int m, size_t n,
    const float *__restrict__ X, size_t ldx, const float *__restrict__ sc,
    const int *__restrict__ cxc, const int *__restrict__ cpar, int P,
    const float *__restrict__ rbuf, size_t ldr, const int *__restrict__ list,
    const int *__restrict__ count, size_t rps, int ldT, double *__restrict__ tpart)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int kr = lane >> 4, kc = lane & 15;
    const int ct = blockIdx.x, split = blockIdx.y, mt0 = blockIdx.z * 8;
    const int Pc = RETRY ? *count : P;
    if (ct * 32 >= Pc) return;
    const size_t r0 = (size_t)split * rps;
    const size_t r1 = r0 + rps < n ? r0 + rps : n;

    bool cok[2];
    int cid[2] = {0, 0}, par[2] = {0, 0};
    float sx[2] = {0.f, 0.f};
    const float *xcol[2] = {X, X};
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int c = ct * 32 + u * 16 + kc;
        cok[u] = c < Pc;
        if (cok[u]) {
            if (RETRY) {
                cid[u] = list[c];
            } else {
                const int xc = cxc[c];
                par[u] = cpar[c];
                sx[u] = sc[xc];
                xcol[u] = X + (size_t)xc * ldx;
            }
        }
    }

    v4d acc[MT][2];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[t][u] = v4d{0, 0, 0, 0};

    for (size_t q = r0 + 4 * w; q < r1; q += 16) {
        const size_t row = q + kr;
        const bool rok = row < r1;
        double a[MT];
#pragma unroll
        for (int t = 0; t < MT; ++t) {
            const int k = (mt0 + t) * 16 + kc;
            a[t] = (rok && k < m) ? (double)Bo[row * ldb + k] : 0.0;
        }
        double b[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            float v = 0.f;
            if (rok && cok[u]) {
                if (RETRY)
                    v = rbuf[row * ldr + cid[u]];
                else
                    v = __fmul_rn(B[row * ldb + par[u]], __fmul_rn(xcol[u][row], sx[u]));
            }
            b[u] = (double)v;
        }
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int u = 0; u < 2; ++u) acc[t][u] = mfma_f64(a[t], b[u], acc[t][u]);
    }

    double *dst = tpart + (size_t)(split * 4 + w) * P * ldT;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int c = ct * 32 + u * 16 + kc;
        if (!cok[u]) continue;
#pragma unroll
        for (int t = 0; t < MT; ++t) {
            if ((mt0 + t) * 16 >= ldT) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int k = (mt0 + t) * 16 + kr + 4 * v;
                dst[(size_t)c * ldT + k] = acc[t][u][v];
            }
        }
    }
}

/* T[c][k] = sum over partials in order; tnorm[c] = |T[c][:m]|^2.  One wave per candidate. */
__global__ __launch_bounds__(64) void k_treduce(const double *__restrict__ tpart, int nparts, int P,
                                                int ldT, int m, const int *__restrict__ count,
                                                double *__restrict__ T, double *__restrict__ tnorm)
{
    const int c = blockIdx.x, lane = threadIdx.x;
    if (count && c >= *count) return;
    double tt = 0.0;
    for (int k = lane; k < ldT; k += 64) {
        double s = 0.0;
        for (int p = 0; p < nparts; ++p) s += tpart[((size_t)p * P + c) * ldT + k];
        T[(size_t)c * ldT + k] = s;
        if (k < m) tt = fma(s, s, tt);
    }
    tt = wave_sum(tt);
    if (lane == 0 && tnorm) tnorm[c] = tt;
}

/* ---------------------------------------------------------------------- */
/* r = f32(src - Bo T) with s partials per row group.                      */
/* grid (row groups, cand tiles of 64), 4 waves x 16-row tiles             */
/* ---------------------------------------------------------------------- */
template <bool RETRY>
__global__ __launch_bounds__(256) void k_project(
    const float *__restrict__ Bo, const float *__restrict__ B, size_t ldb, int m, size_t n,
    const float *__restrict__ X, size_t ldx, const float *__restrict__ sc,
    const int *__restrict__ cxc, const int *__restrict__ cpar, int P,
    const double *__restrict__ T, int ldT, float *__restrict__ rbuf, size_t ldr,
    const int *__restrict__ list, const int *__restrict__ count, size_t rpg,
    double *__restrict__ spart)
{
    __shared__ double sh[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int kr = lane >> 4, kc = lane & 15;
    const int g = blockIdx.x, ct = blockIdx.y;
    const int Pc = RETRY ? *count : P;
    if (ct * 64 >= Pc) return;
    const size_t g0 = (size_t)g * rpg;
    const size_t g1 = g0 + rpg < n ? g0 + rpg : n;

    bool cok[4];
    int slot[4], col[4], par[4] = {0, 0, 0, 0};
    float sx[4] = {0.f, 0.f, 0.f, 0.f};
    const float *xcol[4] = {X, X, X, X};
    double sacc[4] = {0, 0, 0, 0};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        slot[u] = ct * 64 + u * 16 + kc;
        cok[u] = slot[u] < Pc;
        col[u] = slot[u];
        if (cok[u]) {
            if (RETRY) {
                col[u] = list[slot[u]];
            } else {
                const int xc = cxc[slot[u]];
                par[u] = cpar[slot[u]];
                sx[u] = sc[xc];
                xcol[u] = X + (size_t)xc * ldx;
            }
        }
    }

    for (size_t t0 = g0 + 16 * w; t0 < g1; t0 += 64) {
        v4d acc[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] = v4d{0, 0, 0, 0};
        const size_t arow = t0 + kc;
        const bool aok = arow < g1;
        for (int k0 = 0; k0 < m; k0 += 4) {
            const int k = k0 + kr;
            const bool kok = k < m;
            const double a = (aok && kok) ? (double)Bo[arow * ldb + k] : 0.0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double b = (cok[u] && kok) ? T[(size_t)slot[u] * ldT + k] : 0.0;
                acc[u] = mfma_f64(a, b, acc[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (!cok[u]) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const size_t row = t0 + kr + 4 * v;
                if (row >= g1) continue;
                float src;
                if (RETRY)
                    src = rbuf[row * ldr + col[u]];
                else
                    src = __fmul_rn(B[row * ldb + par[u]], __fmul_rn(xcol[u][row], sx[u]));
                const float rf = (float)((double)src - acc[u][v]);
                rbuf[row * ldr + col[u]] = rf;
                const double rd = (double)rf;
                sacc[u] = fma(rd, rd, sacc[u]);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        double s = sacc[u];
        s = s + __shfl_xor(s, 16, 64);
        s = s + __shfl_xor(s, 32, 64);
        if (kr == 0) sh[w][u * 16 + kc] = s;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int t = threadIdx.x;
        const double tot = ((sh[0][t] + sh[1][t]) + sh[2][t]) + sh[3][t];
        const int c = ct * 64 + t;
        if (c < Pc) spart[(size_t)g * P + c] = tot;
    }
}

/* s[c] = sum of group partials; first pass also evaluates the DGKS gate. */
__global__ void k_sreduce(const double *__restrict__ spart, int groups, int P,
                          const int *__restrict__ list, const int *__restrict__ count,
                          const double *__restrict__ tnorm, double tol, double gate,
                          double *__restrict__ ssum, int *__restrict__ flag)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int Pc = list ? *count : P;
    if (c >= Pc) return;
    double s = 0.0;
    for (int g = 0; g < groups; ++g) s += spart[(size_t)g * P + c];
    if (list) {
        ssum[list[c]] = s;
    } else {
        ssum[c] = s;
        flag[c] = (s > tol && s * gate < tnorm[c]) ? 1 : 0;
    }
}

/* Ordered compaction of the DGKS flags (single block of 1024 threads). */
__global__ __launch_bounds__(1024) void k_compact(const int *__restrict__ flag, int P,
                                                  int *__restrict__ list, int *__restrict__ count,
                                                  long long *__restrict__ dgks)
{
    __shared__ int sh[1024];
    const int t = threadIdx.x;
    const int chunk = (P + 1023) / 1024;
    const int b = t * chunk, e = min(P, b + chunk);
    int cnt = 0;
    for (int i = b; i < e; ++i) cnt += flag[i];
    sh[t] = cnt;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = t >= o ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    int pos = sh[t] - cnt;
    for (int i = b; i < e; ++i)
        if (flag[i]) list[pos++] = i;
    if (t == 1023) {
        *count = sh[1023];
        *dgks += sh[1023];
    }
}

/* Normalize and project on y: Bxn = f32(r * scale), ybx partial = sum Bxn y. */
__global__ __launch_bounds__(256) void k_finalize(float *__restrict__ rbuf, size_t ldr, int P,
                                                  const double *__restrict__ ssum, double tol,
                                                  const float *__restrict__ y, size_t n,
                                                  size_t rpg, int store,
                                                  double *__restrict__ ypart)
{
    const int c = blockIdx.y * 256 + threadIdx.x;
    if (c >= P) return;
    const size_t g0 = (size_t)blockIdx.x * rpg;
    const size_t g1 = g0 + rpg < n ? g0 + rpg : n;
    const double s = ssum[c];
    const float scale = (s > tol) ? (float)(1.0 / sqrt(s + tol)) : 0.0f;
    double acc = 0.0;
    for (size_t row = g0; row < g1; ++row) {
        const float bxn = __fmul_rn(rbuf[row * ldr + c], scale);
        if (store) rbuf[row * ldr + c] = bxn;
        acc = fma((double)bxn, (double)y[row], acc);
    }
    ypart[(size_t)blockIdx.x * P + c] = acc;
}

__global__ void k_yreduce(const double *__restrict__ ypart, int groups, int P,
                          double *__restrict__ ybx, double *__restrict__ lin)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P) return;
    double s = 0.0;
    for (int g = 0; g < groups; ++g) s += ypart[(size_t)g * P + c];
    ybx[c] = s;
    lin[c] = s * s;
}

template <bool RETRY>
void launch_tgemm(mars_hip_ctx *ctx, const XSrc &xs, const PlainPlan &pl, const int *list,
                  const int *count)
{
    const int zt = (pl.mt + 7) / 8;
    dim3 grid((pl.P + 31) / 32, pl.ks, zt);
    const int mt = std::min(pl.mt, 8);
    auto args = [&](auto kern) {
        hipLaunchKernelGGL(kern, grid, dim3(256), 0, ctx->stream, ctx->Bo.p, ctx->B.p, ctx->ldb,
                           (int)ctx->m, ctx->n, xs.X, xs.ldx, xs.s, ctx->cand.xc.p,
                           ctx->cand.par.p, pl.P, ctx->rbuf.p, pl.ldr, list, count, pl.rps,
                           pl.ldT, ctx->tpart.p);
    };
    switch (mt) {
    case 1: args(k_tgemm<1, RETRY>); break;
    case 2: args(k_tgemm<2, RETRY>); break;
    case 3: args(k_tgemm<3, RETRY>); break;
    case 4: args(k_tgemm<4, RETRY>); break;
    case 5: args(k_tgemm<5, RETRY>); break;
    case 6: args(k_tgemm<6, RETRY>); break;
    case 7: args(k_tgemm<7, RETRY>); break;
    default: args(k_tgemm<8, RETRY>); break;
    }
    MH_LAUNCH_CHECK();
}

} // namespace

size_t block_capacity(const mars_hip_ctx *ctx)
{
    /* residual buffer budget: 8 GiB, at least one full X column of parents */
    const size_t budget = (size_t)8 << 30;
    size_t cap = budget / (sizeof(float) * std::max<size_t>(ctx->n, 1));
    cap = cap / 64 * 64;
    return std::max<size_t>(cap, ((ctx->max_terms + 63) / 64) * 64);
}

void upload_cands(mars_hip_ctx *ctx)
{
    const size_t P = ctx->cand.h_xc.size();
    ctx->cand.xc.ensure(P);
    ctx->cand.par.ensure(P);
    if (!P) return;
    MH_CHECK(hipMemcpyAsync(ctx->cand.xc.p, ctx->cand.h_xc.data(), P * sizeof(int),
                            hipMemcpyHostToDevice, ctx->stream));
    MH_CHECK(hipMemcpyAsync(ctx->cand.par.p, ctx->cand.h_par.data(), P * sizeof(int),
                            hipMemcpyHostToDevice, ctx->stream));
}

PlainPlan plain_orthonormalize(mars_hip_ctx *ctx, const XSrc &xs, int P, bool keep_bxn)
{
    PlainPlan pl;
    const size_t n = ctx->n;
    pl.P = P;
    pl.mt = std::max<int>(1, (int)((ctx->m + 15) / 16));
    pl.ldT = pl.mt * 16;
    const int ctiles = (P + 31) / 32;
    pl.ks = std::max(1, std::min<int>((1024 + ctiles - 1) / ctiles, (int)std::max<size_t>(1, n / 256)));
    pl.rps = ((n + pl.ks - 1) / pl.ks + 15) / 16 * 16;
    pl.ks = (int)((n + pl.rps - 1) / pl.rps);
    pl.rpg = std::max<size_t>(256, ((n + 1023) / 1024 + 63) / 64 * 64);
    pl.groups = (int)((n + pl.rpg - 1) / pl.rpg);
    pl.ldr = ((size_t)P + 63) / 64 * 64;

    const int nparts = pl.ks * 4;
    ctx->tpart.ensure((size_t)nparts * P * pl.ldT);
    ctx->T.ensure((size_t)P * pl.ldT);
    ctx->tnorm.ensure(P);
    ctx->spart.ensure((size_t)pl.groups * P);
    ctx->ssum.ensure(P);
    ctx->ybx.ensure(P);
    ctx->lin.ensure(P);
    ctx->rbuf.ensure(n * pl.ldr);
    ctx->fired_list.ensure(P + 1);
    ctx->fired_count.ensure(2);
    ctx->knot.ensure(P);   // reused as the DGKS flag array here
    ctx->dgks.ensure(1);

    hipStream_t st = ctx->stream;
    int *flag = ctx->knot.p;

    const double m_live = (double)ctx->m, nd = (double)n;
    // Pass 1: T = Bo^T Bx
    {
        KTimer kt(ctx, K_TGEMM, 2.0 * nd * m_live * P, 4.0 * nd * (2 * m_live + 1));
        launch_tgemm<false>(ctx, xs, pl, nullptr, nullptr);
        hipLaunchKernelGGL(k_treduce, dim3(P), dim3(64), 0, st, ctx->tpart.p, nparts, P, pl.ldT,
                           (int)ctx->m, nullptr, ctx->T.p, ctx->tnorm.p);
        MH_LAUNCH_CHECK();
    }
    // Pass 2: r = f32(Bx - Bo T), s
    {
    KTimer kt(ctx, K_PROJECT, 2.0 * nd * m_live * P, 4.0 * nd * (2 * m_live + 1) + 4.0 * nd * P);
    hipLaunchKernelGGL(k_project<false>, dim3(pl.groups, (P + 63) / 64), dim3(256), 0, st,
                       ctx->Bo.p, ctx->B.p, ctx->ldb, (int)ctx->m, n, xs.X, xs.ldx, xs.s,
                       ctx->cand.xc.p, ctx->cand.par.p, P, ctx->T.p, pl.ldT, ctx->rbuf.p, pl.ldr,
                       nullptr, nullptr, pl.rpg, ctx->spart.p);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sreduce, dim3((P + 255) / 256), dim3(256), 0, st, ctx->spart.p,
                       pl.groups, P, nullptr, nullptr, ctx->tnorm.p, ctx->tol, kGate, ctx->ssum.p,
                       flag);
    MH_LAUNCH_CHECK();
    }
    // DGKS: compact flagged columns, retry them (grids sized for P, idle beyond the count)
    {
    KTimer kt(ctx, K_DGKS, 0.0, 0.0);
    hipLaunchKernelGGL(k_compact, dim3(1), dim3(1024), 0, st, flag, P, ctx->fired_list.p,
                       ctx->fired_count.p, ctx->dgks.p);
    MH_LAUNCH_CHECK();
    launch_tgemm<true>(ctx, xs, pl, ctx->fired_list.p, ctx->fired_count.p);
    hipLaunchKernelGGL(k_treduce, dim3(P), dim3(64), 0, st, ctx->tpart.p, nparts, P, pl.ldT,
                       (int)ctx->m, ctx->fired_count.p, ctx->T.p, nullptr);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_project<true>, dim3(pl.groups, (P + 63) / 64), dim3(256), 0, st,
                       ctx->Bo.p, ctx->B.p, ctx->ldb, (int)ctx->m, n, xs.X, xs.ldx, xs.s,
                       ctx->cand.xc.p, ctx->cand.par.p, P, ctx->T.p, pl.ldT, ctx->rbuf.p, pl.ldr,
                       ctx->fired_list.p, ctx->fired_count.p, pl.rpg, ctx->spart.p);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sreduce, dim3((P + 255) / 256), dim3(256), 0, st, ctx->spart.p,
                       pl.groups, P, ctx->fired_list.p, ctx->fired_count.p, nullptr, ctx->tol,
                       kGate, ctx->ssum.p, nullptr);
    MH_LAUNCH_CHECK();
    }
    // Pass 3: normalize, ybx
    {
    KTimer kt(ctx, K_FINALIZE, 2.0 * nd * P, (keep_bxn ? 8.0 : 4.0) * nd * P + 4.0 * nd);
    hipLaunchKernelGGL(k_finalize, dim3(pl.groups, (P + 255) / 256), dim3(256), 0, st,
                       ctx->rbuf.p, pl.ldr, P, ctx->ssum.p, ctx->tol, ctx->y.p, n, pl.rpg,
                       keep_bxn ? 1 : 0, ctx->spart.p);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_yreduce, dim3((P + 255) / 256), dim3(256), 0, st, ctx->spart.p,
                       pl.groups, P, ctx->ybx.p, ctx->lin.p);
    MH_LAUNCH_CHECK();
    }
    return pl;
}

} // namespace mh
