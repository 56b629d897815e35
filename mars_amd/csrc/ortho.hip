// Plain (FP64) candidate orthonormalization on gfx950, batched over fits.
//
// Restates mars::orthonormalize (kernels.cc:283-417) for every candidate
// (fit, X column, parent basis) of one forward step at once, keeping the
// reference's f32 rounding points:
//   Bx  = f32(B[:,par] * f32(X[:,xc] * s[xc]))          rebuilt on the fly, never stored
//   T   = Bo^T Bx                     FP64 MFMA, split over rows, fixed-order reduce
//   r   = f32(Bx - Bo T),  s = sum r^2  FP64 MFMA; r, when kept, goes to rbuf
//         (column-major for linear scoring, row-major for the hinge sweep)
//   DGKS: s > tol and 9 s < |T|^2  ->  T2 = Bo^T r, r = f32(r - Bo T2), s = sum r^2
//   Bxn = f32(r * f32(1/sqrt(s + tol))) (0 if s <= tol),  ybx = sum Bxn y
// Work is described by items (fit, candidate tile, row range) consumed by
// persistent grids; every reduction runs in a fixed order, so results are
// run-to-run identical for a given batch (row ranges are sized from the whole
// batch's work, so a fit's summation order depends on the batch it is in).
#include "internal.h"
#include "kernels.h"
#include "stream.h"

#include <algorithm>
#include <numeric>

namespace mh {

namespace {

constexpr int kTileF = 64;   // candidates per finalize item (x 4 row phases)
// basis functions the FP64 scoring passes take: 256 per T-GEMM launch and in the projection's
// LDS tile, the columns beyond in further launches and straight from memory; 512 is also the
// device append's limit
constexpr size_t kMaxScoreTerms = 512;
// work items per projection / T-GEMM pass (ctx->tune.pass_items, default 32768): about 32 one-wave
// items per SIMD, so a pass keeps every CU busy however few candidates each fit has
constexpr long long kMaxSplits = 2048;  // row ranges per candidate tile at most
// residual bytes that the plain form's DGKS retry keeps at once (one arena per context)
constexpr size_t kRetryArena = size_t(2) << 30;
// residual bytes of the device-driven rescoring: half for stored entries, half for fired ones
constexpr size_t kRescoreArena = size_t(2) << 30;

/* T[c] = sum of the candidate's item partials; tnorm[c] = |T[c][:m]|^2.  One block of S waves per
 * candidate (or list position): lane l takes basis columns l, l + 64, ..., wave w the w-th of S
 * equal segments of the item partials (each summed left to right), and the segment sums are added
 * in wave order: a fixed order for a given S (S = 1: the partials left to right). */
__global__ __launch_bounds__(1024) void k_treduce(const FitDev *__restrict__ fits,
                                                  const int *__restrict__ cfit,
                                                  const int *__restrict__ list, int count,
                                                  const int *__restrict__ cmap,
                                                  const double *__restrict__ tpart, int ldT,
                                                  double *__restrict__ T, double *__restrict__ tnorm,
                                                  const int *__restrict__ dcount)
{
    __shared__ double sh[16][64];
    if (dcount) count = *dcount;  // list length built on the device
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, S = blockDim.x >> 6;
    for (int c = blockIdx.x; c < count; c += gridDim.x) {
        const int cid = list ? list[c] : c;
        const int m = fits[cfit[cid]].m;
        const int first = cmap[3 * c], cnt = cmap[3 * c + 1], slot = cmap[3 * c + 2];
        const int seg = (cnt + S - 1) / S, lo = min(cnt, w * seg), hi = min(cnt, lo + seg);
        double tt = 0.0;
        for (int k0 = 0; k0 < ldT; k0 += 64) {
            const int k = k0 + lane;
            double s = 0.0;
            if (k < ldT) s = range_sum(tpart, lo, hi, [&](int i) { return ((size_t)(first + i) * 64 + slot) * ldT + k; });
            if (S > 1) {
                sh[w][lane] = s;
                __syncthreads();
                if (w == 0)
                    for (int v = 1; v < S; ++v) s += sh[v][lane];
                __syncthreads();
            }
            if (w == 0 && k < ldT) {
                T[(size_t)c * ldT + k] = s;
                if (k < m) tt = fma(s, s, tt);
            }
        }
        if (w == 0) {
            tt = wave_sum(tt);
            if (lane == 0 && tnorm) tnorm[c] = tt;
        }
    }
}

/* s[c] = sum of item partials; the first pass also sets flag[c]: bit 0 the DGKS gate
 * fired (kernels.cc:387-388), bit 1 s > tol (the candidate can score above 0), bit 2 its
 * residual was not stored (keep[c] == 0).  L lanes per candidate (a power of two dividing 64):
 * lane j of the group sums the j-th of L equal segments of the partials left to right and the
 * group adds the segment sums by a fixed butterfly (L = 1: the partials left to right). */
__global__ void k_sreduce(const int *__restrict__ list, int count, const int *__restrict__ cmap,
                          const double *__restrict__ spart, int tile,
                          const double *__restrict__ tnorm, const FitDev *__restrict__ fits,
                          const int *__restrict__ cfit, const int *__restrict__ keep,
                          double *__restrict__ ssum, int *__restrict__ flag, const int *__restrict__ dcount, int L)
{
    if (dcount) count = *dcount;
    const int t = blockIdx.x * blockDim.x + threadIdx.x, sub = t % L, ng = gridDim.x * blockDim.x / L;
    for (int c = t / L; c < count; c += ng) {
        const int first = cmap[3 * c], cnt = cmap[3 * c + 1], slot = cmap[3 * c + 2];
        const int seg = (cnt + L - 1) / L, lo = min(cnt, sub * seg), hi = min(cnt, lo + seg);
        double s = range_sum(spart, lo, hi, [&](int i) { return (size_t)(first + i) * tile + slot; });
        s = group_sum(s, L);
        if (sub != 0) continue;
        if (list) {
            ssum[list[c]] = s;
        } else {
            const double tol = fits[cfit[c]].tol;
            ssum[c] = s;
            flag[c] = ((s > tol && s * kGate < tnorm[c]) ? 1 : 0) | (s > tol ? 2 : 0) |
                      ((keep && keep[c] == 0) ? 4 : 0);
        }
    }
}

/*
 * Normalize in place and project on y (row-major columns kept for the hinge sweep or a
 * download): Bxn = f32(r * scale) replaces r, ybx partial = sum Bxn y.
 * 256 threads = 256/TF row phases x TF candidates: thread (g, slot) takes rows
 * r0 + g, r0 + g + 256/TF, ... (eight loads in flight), and the phase sums are
 * added in a fixed order.  TF = 16 serves small candidate sets, TF = kTileF the others.
 */
template <int TF>
__global__ __launch_bounds__(256) void k_finalize(const FitDev *__restrict__ fits,
                                                  const Item *__restrict__ items, int nitems,
                                                  const double *__restrict__ ssum,
                                                  double *__restrict__ ypart)
{
    constexpr int NPH = 256 / TF;
    __shared__ double sh[NPH][TF];
    const int slot = threadIdx.x % TF, g = threadIdx.x / TF;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item I = items[it];
        double acc = 0.0;
        if (slot < I.nc) {
            const FitDev F = fits[I.fit];
            const int c = I.c0 + slot;
            const long long col = c - F.c0;
            const double s = ssum[c];
            const float scale = (s > F.tol) ? (float)(1.0 / sqrt(s + F.tol)) : 0.0f;
            float *base = F.rbuf + col * F.rsc;
            long long row = I.r0 + g;
            for (; row + 7 * NPH < I.r1; row += 8 * NPH) {
                float r[8], yv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    r[u] = base[(row + NPH * u) * F.rsr];
                    yv[u] = F.y[row + NPH * u];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const float bxn = __fmul_rn(r[u], scale);
                    base[(row + NPH * u) * F.rsr] = bxn;
                    acc = fma((double)bxn, (double)yv[u], acc);
                }
            }
            for (; row < I.r1; row += NPH) {
                const float bxn = __fmul_rn(base[row * F.rsr], scale);
                base[row * F.rsr] = bxn;
                acc = fma((double)bxn, (double)F.y[row], acc);
            }
        }
        sh[g][slot] = acc;
        __syncthreads();
        if (g == 0 && slot < I.nc) {
            double t = sh[0][slot];
#pragma unroll
            for (int q = 1; q < NPH; ++q) t += sh[q][slot];
            ypart[(size_t)it * kTileF + slot] = t;
        }
        __syncthreads();
    }
}

/*
 * The same over column-major residuals (linear scoring): wave w of 4 takes the item's slots
 * w, w + 4, ...; its lanes read the slot's column as float4 (rows r0 + 4 lane + 256 u, four
 * in flight), and the wave's sum is formed in a fixed order.  With `list`, item positions
 * index the list of candidates.  A candidate whose residual was not stored (keep entry 0)
 * gets 0 here; if its s > tol, a later pass recomputes it.
 * wb: the normalised column Bxn = f32(r scale) also replaces r (the sorted hinge path keeps it;
 * a candidate with s <= tol gets a zero column), except for candidates whose `skip` flag has
 * bit 0 set (DGKS fired: the retry still needs r; their ybx partial is 0 until it runs).
 */
__global__ __launch_bounds__(256) void k_finalize_cm(const FitDev *__restrict__ fits,
                                                     const Item *__restrict__ items, int nitems,
                                                     const int *__restrict__ list,
                                                     const int *__restrict__ keep,
                                                     const double *__restrict__ ssum,
                                                     double *__restrict__ ypart, const int *__restrict__ dn,
                                                     const int *__restrict__ skip, int wb)
{
    if (dn) nitems = *dn;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item I = items[it];
        const FitDev F = fits[I.fit];
        for (int slot = w; slot < I.nc; slot += 4) {
            const int c = list ? list[I.c0 + slot] : I.c0 + slot;
            if (skip && (skip[c] & 1)) {
                if (lane == 0) ypart[(size_t)it * kTileF + slot] = 0.0;
                continue;
            }
            const double s = ssum[c];
            const float scale = (s > F.tol) ? (float)(1.0 / sqrt(s + F.tol)) : 0.0f;
            const long long cc = F.rbyl == 2 ? max(F.roff[c], 0ll) : (F.rbyl ? I.c0 + slot - I.pad : c - F.c0);
            float *col = F.rbuf + cc * F.rsc;
            double acc = 0.0;
            // s <= tol: scale 0, so ybx = 0 exactly (kernels.cc:401-416) and r is not read; most
            // rescored candidates (duplicates of basis functions) end here
            const bool live = s > F.tol && (!keep || keep[c] != 0);
            const long long rend = (live || wb) ? I.r1 : I.r0;
            for (long long r = I.r0 + 4 * lane; r < rend; r += 1024) {
                float4 rv[4], yv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const long long rr = r + 256 * u;
                    const long long ra = rr < rend ? rr : r;  // clamped: stays inside the column
                    rv[u] = live ? gld4(col + ra) : float4{0.f, 0.f, 0.f, 0.f};
                    yv[u] = live ? gld4(F.y + ra) : float4{0.f, 0.f, 0.f, 0.f};
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const long long rr = r + 256 * u;
                    if (rr >= rend) continue;
                    const float a[4] = {rv[u].x, rv[u].y, rv[u].z, rv[u].w};
                    const float b[4] = {yv[u].x, yv[u].y, yv[u].z, yv[u].w};
                    float o[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        o[v] = __fmul_rn(a[v], scale);
                        if (rr + v < rend) acc = fma((double)o[v], (double)b[v], acc);
                    }
                    if (wb) {
                        if (rr + 3 < rend) {
                            *(float4 *)(col + rr) = float4{o[0], o[1], o[2], o[3]};
                        } else {
#pragma unroll
                            for (int v = 0; v < 4; ++v)
                                if (rr + v < rend) col[rr + v] = o[v];
                        }
                    }
                }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
            if (lane == 0) ypart[(size_t)it * kTileF + slot] = acc;
        }
    }
}

/* ybx = sum of item partials, lin = ybx^2, for candidate c (or list[c]). */
__global__ void k_yreduce(int count, const int *__restrict__ cmap, const double *__restrict__ ypart,
                          const int *__restrict__ list, double *__restrict__ ybx, double *__restrict__ lin,
                          const int *__restrict__ dcount, int L)
{
    if (dcount) count = *dcount;
    const int t = blockIdx.x * blockDim.x + threadIdx.x, sub = t % L, ng = gridDim.x * blockDim.x / L;
    for (int c = t / L; c < count; c += ng) {
        const int first = cmap[3 * c], cnt = cmap[3 * c + 1], slot = cmap[3 * c + 2];
        const int seg = (cnt + L - 1) / L, lo = min(cnt, sub * seg), hi = min(cnt, lo + seg);
        double s = range_sum(ypart, lo, hi, [&](int i) { return (size_t)(first + i) * kTileF + slot; });
        s = group_sum(s, L);
        if (sub != 0) continue;
        const int cand = list ? list[c] : c;
        if (ybx) ybx[cand] = s;
        lin[cand] = s * s;
    }
}

/* lin = ybx = 0 (candidates whose s <= tol: scale 0, kernels.cc:401-416). */
__global__ void k_zero2(int count, double *__restrict__ a, double *__restrict__ b)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= count) return;
    if (a) a[c] = 0.0;
    b[c] = 0.0;
}

/*
 * Work list: for each fit, its (list) positions [b_f, e_f) cut into tiles of
 * `tile`, each tile cut into row ranges of about `rows_f` rows (multiple of
 * `align`).  cmap gets (first item, item count, slot) per position.
 */
Pass make_pass(const std::vector<int> &fb, const std::vector<int> &fe, const std::vector<long long> &n,
               const std::vector<long long> &rows, int tile, long long align)
{
    Pass p;
    const int total = fe.empty() ? 0 : *std::max_element(fe.begin(), fe.end());
    p.cmap.assign(3 * (size_t)std::max(total, 0), 0);
    size_t count = 0;
    for (size_t f = 0; f < fb.size(); ++f) {
        if (fe[f] <= fb[f]) continue;
        const long long r = std::max<long long>(align, (rows[f] + align - 1) / align * align);
        count += (size_t)((fe[f] - fb[f] + tile - 1) / tile) * (size_t)((n[f] + r - 1) / r);
    }
    p.items.reserve(count);
    for (size_t f = 0; f < fb.size(); ++f) {
        if (fe[f] <= fb[f]) continue;
        long long r = std::max<long long>(align, (rows[f] + align - 1) / align * align);
        for (int c0 = fb[f]; c0 < fe[f]; c0 += tile) {
            const int nc = std::min(tile, fe[f] - c0);
            const int first = (int)p.items.size();
            for (long long r0 = 0; r0 < n[f]; r0 += r) {
                Item it;
                it.fit = (int)f;
                it.c0 = c0;
                it.nc = nc;
                it.pad = fb[f];
                it.r0 = r0;
                it.r1 = std::min(n[f], r0 + r);
                p.items.push_back(it);
            }
            const int cnt = (int)p.items.size() - first;
            for (int c = c0; c < c0 + nc; ++c) {
                p.cmap[3 * (size_t)c] = first;
                p.cmap[3 * (size_t)c + 1] = cnt;
                p.cmap[3 * (size_t)c + 2] = c - c0;
            }
        }
    }
    return p;
}

/* Rows per item so that the whole pass has roughly `target` items. */
std::vector<long long> split_rows(const std::vector<long long> &n, const std::vector<int> &cnt,
                                  int tile, double target, long long minrows)
{
    double work = 0;
    for (size_t f = 0; f < n.size(); ++f) work += (double)n[f] * ((cnt[f] + tile - 1) / tile);
    std::vector<long long> r(n.size());
    for (size_t f = 0; f < n.size(); ++f) {
        const double tiles = (double)((cnt[f] + tile - 1) / tile);
        const double want = work > 0 ? target * n[f] * tiles / work / std::max(tiles, 1.0) : 1;
        long long splits = std::max<long long>(1, (long long)want);
        splits = std::min<long long>(splits, std::max<long long>(1, n[f] / minrows));
        // a tile's partial sums are added by one wave per candidate: bound how many there are
        splits = std::min<long long>(splits, kMaxSplits);
        r[f] = (n[f] + splits - 1) / splits;
    }
    return r;
}

struct Uploaded {
    Item *items = nullptr;
    int *cmap = nullptr;
};

/* Upload items and cmap at offsets of `items` / `cmap`, which the caller has sized already (no
 * reallocation here: pointers into earlier passes of the same buffers stay valid). */
void upload_pass(mars_hip_ctx *ctx, const Pass &p, DBuf<Item> &items, DBuf<int> &cmap, size_t item_off,
                 size_t cmap_off, Uploaded &u)
{
    require(item_off + p.items.size() <= items.cap && cmap_off + p.cmap.size() <= cmap.cap,
            "internal: work list buffer too small");
    u.items = items.p + item_off;
    u.cmap = cmap.p + cmap_off;
    if (!p.items.empty())
        h2d_defer(ctx, u.items, p.items.data(), p.items.size() * sizeof(Item));
    if (!p.cmap.empty())
        h2d_defer(ctx, u.cmap, p.cmap.data(), p.cmap.size() * sizeof(int));
}

int grid_for(size_t items)
{
    return (int)std::max<size_t>(1, std::min<size_t>(items, 4096));
}

#define MH_MT_SWITCH(mt, GO)                                                                 \
    switch (mt) {                                                                            \
    case 1: GO(1); break;   case 2: GO(2); break;   case 3: GO(3); break;   case 4: GO(4); break;   \
    case 5: GO(5); break;   case 6: GO(6); break;   case 7: GO(7); break;   case 8: GO(8); break;   \
    case 9: GO(9); break;   case 10: GO(10); break; case 11: GO(11); break; case 12: GO(12); break; \
    case 13: GO(13); break; case 14: GO(14); break; case 15: GO(15); break; default: GO(16); break;  \
    }

int grid_w(size_t items)
{
    return (int)std::max<size_t>(1, std::min<size_t>(items, 65536));
}

template <bool RETRY, bool F32 = false>
void launch_tgemm(mars_hip_ctx *ctx, int mt, const Uploaded &u, int nitems, const Cands &cs,
                  const int *list, int ldT, bool narrow, const int *dn = nullptr, const FitDev *fits = nullptr)
{
    if (nitems == 0) return;
    const dim3 grid(grid_w(nitems)), block(64);
    const FitDev *fv = fits ? fits : (const FitDev *)ctx->d_fits.p;
    // basis columns [kb, kb + 256) per launch: fits of more than 256 basis functions take several
    for (int kb = 0; kb < 16 * mt; kb += 256) {
#define MH_GO1(M, CG)                                                                                 \
    hipLaunchKernelGGL((k_tgemm_w<M, CG, RETRY, F32>), grid, block, 0, ctx->stream, fv, \
                       (const Item *)u.items, nitems, (const int *)cs.d_xc.p, (const int *)cs.d_par.p, list, ldT, \
                       ctx->tpart.p, dn, kb)
#define MH_GO(M)                      \
    if (narrow) MH_GO1(M, 1);         \
    else MH_GO1(M, cg_tgemm(M))
        MH_MT_SWITCH(std::min(mt - kb / 16, 16), MH_GO)
#undef MH_GO
#undef MH_GO1
        MH_LAUNCH_CHECK();
    }
}

void launch_tgemm_bf16(mars_hip_ctx *ctx, int mt, const Uploaded &u, int nitems, const Cands &cs, int ldT,
                       bool narrow)
{
    if (nitems == 0) return;
    require(mt <= 16, "the BF16-split T-GEMM covers up to 256 basis columns");
    const dim3 grid(grid_w(nitems)), block(64);
#define MH_GO1(M, CG)                                                                                           \
    hipLaunchKernelGGL((k_tgemm_bf16<M, CG>), grid, block, 0, ctx->stream, (const FitDev *)ctx->d_fits.p,       \
                       (const Item *)u.items, nitems, (const int *)cs.d_xc.p, (const int *)cs.d_par.p, ldT,     \
                       ctx->tpart.p)
#define MH_GO(M)                      \
    if (narrow) MH_GO1(M, 1);         \
    else MH_GO1(M, cg_tgemm(M))
    MH_MT_SWITCH(mt, MH_GO)
#undef MH_GO
#undef MH_GO1
    MH_LAUNCH_CHECK();
}

template <int MODE, bool RB, bool TCAND, bool F32 = false>
void launch_project(mars_hip_ctx *ctx, int mt, const Uploaded &u, int nitems, const Cands &cs,
                    const int *list, const double *T, int ldT, bool narrow, const int *keep = nullptr,
                    const int *dn = nullptr, const FitDev *fits = nullptr)
{
    if (nitems == 0) return;
    const dim3 grid(grid_w(nitems)), block(64);
    const FitDev *fv = fits ? fits : (const FitDev *)ctx->d_fits.p;
#define MH_GO2(M, CG, W)                                                                                  \
    hipLaunchKernelGGL((k_project_w<M, CG, MODE, RB, TCAND, F32, W>), grid, block, 0, ctx->stream,        \
                       fv, (const Item *)u.items, nitems, (const int *)cs.d_xc.p, \
                       (const int *)cs.d_par.p, list, T, ldT, (const double *)ctx->ssum.p,                \
                       (const int *)ctx->flag.p, keep, ctx->spart.p, dn)
#define MH_GO1(M, CG) MH_GO2(M, CG, false)
#define MH_GO(M)                      \
    if (narrow) MH_GO1(M, 1);         \
    else MH_GO1(M, cg_project(M))
    // more than 256 basis functions: the 256-column tile plus the columns past it from memory
    if (mt > 16) MH_GO2(16, 1, true);
    else MH_MT_SWITCH(mt, MH_GO)
#undef MH_GO
#undef MH_GO1
#undef MH_GO2
    MH_LAUNCH_CHECK();
}


constexpr int kRedGrid = 1024;       // blocks of the grid-stride reductions over device-sized lists

/* Threads per candidate of the ordered reductions, from the list's length bound: the passes split
 * each candidate tile's rows to fill the GPU, so few candidates have many partials each and get
 * more threads; many candidates get one lane (L) or one wave (S) each. */
int red_lanes(size_t bound) { return bound >= 8192 ? 1 : bound >= 1024 ? 8 : bound >= 128 ? 32 : 64; }
int red_waves(size_t bound) { return bound >= 1024 ? 1 : bound >= 128 ? 4 : 16; }

void launch_treduce(hipStream_t st, int S, size_t bound, const FitDev *fits, const int *cfit, const int *list,
                    int count, const int *cmap, const double *tpart, int ldT, double *T, double *tnorm,
                    const int *dcount)
{
    const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>(bound, 4 * kRedGrid));
    hipLaunchKernelGGL(k_treduce, dim3(g), dim3(64 * S), 0, st, fits, cfit, list, count, cmap, tpart, ldT, T, tnorm,
                       dcount);
    MH_LAUNCH_CHECK();
}

void launch_sreduce(hipStream_t st, int L, size_t bound, const int *list, int count, const int *cmap,
                    const double *spart, int tile, const double *tnorm, const FitDev *fits, const int *cfit,
                    const int *keep, double *ssum, int *flag, const int *dcount)
{
    const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>((bound * L + 255) / 256, 4 * kRedGrid));
    hipLaunchKernelGGL(k_sreduce, dim3(g), dim3(256), 0, st, list, count, cmap, spart, tile, tnorm, fits, cfit, keep,
                       ssum, flag, dcount, L);
    MH_LAUNCH_CHECK();
}

void launch_yreduce(hipStream_t st, int L, size_t bound, int count, const int *cmap, const double *ypart,
                    const int *list, double *ybx, double *lin, const int *dcount)
{
    const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>((bound * L + 255) / 256, 4 * kRedGrid));
    hipLaunchKernelGGL(k_yreduce, dim3(g), dim3(256), 0, st, count, cmap, ypart, list, ybx, lin, dcount, L);
    MH_LAUNCH_CHECK();
}

} // namespace

void index_cands(const mars_hip_ctx *ctx, Cands &cs)
{
    const size_t F = ctx->fits.size();
    cs.fit_c0.assign(F, 0);
    cs.fit_P.assign(F, 0);
    for (size_t c = 0; c < cs.fit.size(); ++c) {
        const int f = cs.fit[c];
        if (cs.fit_P[f] == 0) cs.fit_c0[f] = (int)c;
        cs.fit_P[f] += 1;
    }
    for (size_t f = 0; f < F; ++f)
        if (cs.fit_P[f] == 0) cs.fit_c0[f] = (int)cs.fit.size();
}

void upload_cands(mars_hip_ctx *ctx, Cands &cs)
{
    const size_t P = cs.fit.size();
    cs.d_fit.ensure(P);
    cs.d_xc.ensure(P);
    cs.d_par.ensure(P);
    if (!P) return;
    h2d_defer(ctx, cs.d_fit.p, cs.fit.data(), P * sizeof(int));
    h2d_defer(ctx, cs.d_xc.p, cs.xc.data(), P * sizeof(int));
    h2d_defer(ctx, cs.d_par.p, cs.par.data(), P * sizeof(int));
}

void ensure_rbuf(mars_hip_ctx *ctx, const Cands &cs)
{
    for (size_t f = 0; f < ctx->fits.size(); ++f) {
        FitState &fs = ctx->fits[f];
        const size_t need = ((size_t)std::max(cs.fit_P[f], 1) + 63) / 64 * 64;
        if (need > fs.ldr) {
            // grow geometrically in columns so the buffer is reallocated O(log) times per fit; the
            // old block is retired (queued kernels may still read it), not freed mid-step
            const size_t ldr = std::max(need, std::min<size_t>(2 * fs.ldr, ctx->max_terms * fs.px + 64));
            fs.rbuf.ensure(fs.ldx * ldr);  // room for either layout (column-major uses ldx rows)
            fs.ldr = ldr;
        }
    }
}

void upload_fits(mars_hip_ctx *ctx, const Cands &cs, const float *xover, const float *sover,
                 const std::vector<size_t> *arena_off)
{
    std::vector<FitDev> h(ctx->fits.size());
    for (size_t f = 0; f < ctx->fits.size(); ++f) {
        const FitState &fs = ctx->fits[f];
        FitDev &d = h[f];
        d.X = fs.X.p;
        d.s = fs.s.p;
        d.ldx = (long long)fs.ldx;
        if (f == 0 && xover) {
            d.X = xover;
            d.s = sover;
        }
        d.B = fs.B.p;
        d.Bo = fs.Bo.p;
        d.BoC = fs.BoC.p;
        d.BC = fs.BC.p;
        d.y = fs.y.p;
        d.ybo = fs.ybo.p;
        d.rbuf = fs.rbuf.p;
        d.v = fs.v.p;
        d.tc = fs.tc.p;
        d.tcn = fs.tcn.p;
        d.n = (long long)fs.n;
        d.ldb = (long long)ctx->ldb;
        d.rsr = ctx->rbuf_cm ? 1 : (long long)fs.ldr;
        d.rsc = ctx->rbuf_cm ? (long long)fs.ldx : 1;
        d.rbyl = 0;
        d.px = (int)fs.px;
        d.act = f < ctx->act_h.size() ? ctx->act_h[f] : 1;
        d.pad_ = 0;
        if (arena_off) {  // compact column-major residuals in the retry arena
            d.rbuf = ctx->rarena.p + (*arena_off)[f];
            d.rsr = 1;
            d.rsc = (long long)fs.ldx;
            d.rbyl = 1;
        }
        d.m = (int)fs.m;
        d.c0 = cs.fit_c0.empty() ? 0 : cs.fit_c0[f];
        d.tol = fs.tol;
    }
    ctx->d_fits.ensure(h.size());
    h2d_defer(ctx, ctx->d_fits.p, h.data(), h.size() * sizeof(FitDev));
}

void reserve_step_buffers(mars_hip_ctx *ctx)
{
    const size_t parents = ctx->max_terms > 1 ? ctx->max_terms - 1 : 1;
    size_t P = 0;
    for (const FitState &fs : ctx->fits) P += fs.px * parents;
    P = std::min<size_t>(P, size_t(1) << 22);  // beyond this the buffers grow when needed
    const size_t F = ctx->fits.size();
    const size_t ldT = 16 * ((std::min<size_t>(ctx->max_terms, kMaxScoreTerms) + 15) / 16);
    const size_t items = 3 * ((size_t)ctx->tune.pass_items + P / 16 + F) + 1;  // the three passes' work lists
    for (Cands *c : {&ctx->cand, &ctx->rcand}) {
        c->d_fit.ensure(P);
        c->d_xc.ensure(P);
        c->d_par.ensure(P);
    }
    for (DBuf<double> *b : {&ctx->lin, &ctx->ybx, &ctx->rlin, &ctx->rybx, &ctx->tnorm, &ctx->ssum}) b->ensure(P);
    for (DBuf<int> *b : {&ctx->flag, &ctx->rsel, &ctx->rkeep, &ctx->fired_list}) b->ensure(P);
    // T rows: every candidate in the plain form, the rescored few in the mixed form
    ctx->T.ensure((ctx->form == MARS_HIP_FORM_MIXED ? std::min<size_t>(P, 65536) : P) * ldT);
    ctx->d_items.ensure(items);
    ctx->d_cmap.ensure(9 * P + 1);
    ctx->spart.ensure(items * kSlots);
}

namespace {

/* A few candidates per fit (the mixed form's rescoring, DGKS retries) run on 16-wide candidate
 * tiles: the wide tiles would be mostly zero padding in the FP64 MFMA work that bounds these
 * kernels. */
bool narrow_for(const std::vector<int> &c, int avg)
{
    long long tot = 0, act = 0;
    for (int v : c) {
        tot += v;
        act += v > 0;
    }
    return tot <= (long long)avg * act;
}

} // namespace

bool narrow_tiles(const mars_hip_ctx *ctx, const std::vector<int> &cnt) { return narrow_for(cnt, ctx->tune.narrow_avg); }

namespace {

/*
 * DGKS retry of the plain form's linear scoring (kernels.cc:387-399), for the fired candidates
 * fired[rb[f], re[f]) of each fit: store the first residual r = f32(Bx - Bo T), T2 = Bo^T r,
 * r = f32(r - Bo T2) in place, s = sum r^2, then ybx = sum f32(r scale) y.  The residuals go to a
 * shared arena, column-major and numbered by position in the fit's part of the fired list, in
 * blocks of at most kRetryArena bytes: the memory is bounded whatever the number of fired
 * candidates or rows (the reference bounds its candidate block the same way,
 * orthonormalize.cu:476-489).
 */
void retry_compact(mars_hip_ctx *ctx, Cands &cs, const std::vector<int> &fired, const std::vector<int> &rb,
                   const std::vector<int> &re, const std::vector<long long> &n, int mt, int ldT, double *ybx,
                   double *lin)
{
    hipStream_t st = ctx->stream;
    const size_t F = ctx->fits.size();
    ctx->dgks_host += (long)fired.size();
    // blocks of the fired list (grouped by fit), each within the arena budget
    size_t k0 = 0;
    std::vector<int> fit_of(fired.size());
    for (size_t f = 0; f < F; ++f)
        for (int i = rb[f]; i < re[f]; ++i) fit_of[i] = (int)f;
    while (k0 < fired.size()) {
        size_t k1 = k0, bytes = 0;
        while (k1 < fired.size()) {
            const size_t col = ctx->fits[fit_of[k1]].ldx * sizeof(float);
            if (k1 > k0 && bytes + col > kRetryArena) break;
            bytes += col;
            ++k1;
        }
        const int nf = (int)(k1 - k0);
        std::vector<int> bb(F, 0), be(F, 0), fcnt(F, 0);
        std::vector<size_t> off(F, 0);
        size_t used = 0;
        for (size_t f = 0; f < F; ++f) {
            const int b = std::max<int>(rb[f], (int)k0), e = std::min<int>(re[f], (int)k1);
            bb[f] = std::min<int>(b, (int)k1) - (int)k0;
            be[f] = std::max(b, e) - (int)k0;
            fcnt[f] = be[f] - bb[f];
            off[f] = used;
            used += (size_t)fcnt[f] * ctx->fits[f].ldx;
        }
        ctx->rarena.ensure(std::max<size_t>(used, 1));
        upload_fits(ctx, cs, nullptr, nullptr, &off);
        h2d_defer(ctx, ctx->fired_list.p, fired.data() + k0, nf * sizeof(int));
        const bool rnarrow = narrow_for(fcnt, ctx->tune.narrow_avg);
        const int rtT = 16 * (rnarrow ? 1 : cg_tgemm(mt)), rtP = 16 * (rnarrow ? 1 : cg_project(mt));
        Pass rT = make_pass(bb, be, n, split_rows(n, fcnt, rtT, ctx->tune.retry_items, 256), rtT, kWR);
        Pass rP = make_pass(bb, be, n, split_rows(n, fcnt, rtP, ctx->tune.retry_items, 256), rtP, kWR);
        Pass fF = make_pass(bb, be, n, split_rows(n, fcnt, 16, 2048, 512), 16, 32);
        Uploaded vT, vP, wF;
        ctx->d_ritems.ensure(rT.items.size() + rP.items.size() + fF.items.size() + 1);
        ctx->d_rcmap.ensure(rT.cmap.size() + rP.cmap.size() + fF.cmap.size() + 1);
        upload_pass(ctx, rT, ctx->d_ritems, ctx->d_rcmap, 0, 0, vT);
        upload_pass(ctx, rP, ctx->d_ritems, ctx->d_rcmap, rT.items.size(), rT.cmap.size(), vP);
        upload_pass(ctx, fF, ctx->d_ritems, ctx->d_rcmap, rT.items.size() + rP.items.size(),
                    rT.cmap.size() + rP.cmap.size(), wF);
        h2d_flush(ctx);
        ctx->tpart.ensure(std::max<size_t>(rT.items.size(), 1) * kSlots * ldT);
        ctx->spart.ensure(std::max(rP.items.size() * kSlots, fF.items.size() * kTileF) + 1);
        ctx->T2.ensure((size_t)nf * ldT);
        const int *fl = ctx->fired_list.p;
        launch_project<kStoreR, false, true>(ctx, mt, vP, (int)rP.items.size(), cs, fl, ctx->T.p, ldT, rnarrow);
        launch_tgemm<true>(ctx, mt, vT, (int)rT.items.size(), cs, fl, ldT, rnarrow);
        launch_treduce(st, red_waves(nf), nf, ctx->d_fits.p, cs.d_fit.p, fl, (int)nf, vT.cmap, ctx->tpart.p, ldT,
                       ctx->T2.p, nullptr, nullptr);
        launch_project<kStoreR, true, false>(ctx, mt, vP, (int)rP.items.size(), cs, fl, ctx->T2.p, ldT, rnarrow);
        launch_sreduce(st, red_lanes(nf), nf, fl, (int)nf, vP.cmap, ctx->spart.p, kSlots, nullptr, ctx->d_fits.p,
                       cs.d_fit.p, nullptr, ctx->ssum.p, nullptr, nullptr);
        hipLaunchKernelGGL(k_finalize_cm, dim3(grid_for(fF.items.size())), dim3(256), 0, st, ctx->d_fits.p,
                           wF.items, (int)fF.items.size(), fl, nullptr, ctx->ssum.p, ctx->spart.p, nullptr, nullptr,
                           0);
        MH_LAUNCH_CHECK();
        launch_yreduce(st, red_lanes(nf), nf, (int)nf, wF.cmap, ctx->spart.p, fl, ybx, lin, nullptr);
        k0 = k1;
        if (k0 < fired.size()) ctx_sync(ctx);  // the next block reuses the arena and the work lists
    }
}

} // namespace

namespace {

/* ---------------------------------------------------------------------- */
/* Device-built lists and work items (the device-driven rescoring)         */
/* ---------------------------------------------------------------------- */

constexpr int kCompactTile = 1024;   // list entries per compaction block (256 threads x 4)

/* rfit/rxc/rpar[i] = cfit/cxc/cpar[list[i]] for i < *dcount. */
__global__ void k_rc_gather(const int *__restrict__ list, const int *__restrict__ dcount,
                            const int *__restrict__ cfit, const int *__restrict__ cxc, const int *__restrict__ cpar,
                            int *__restrict__ rfit, int *__restrict__ rxc, int *__restrict__ rpar)
{
    const int n = *dcount;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int c = list[i];
        rfit[i] = cfit[c];
        rxc[i] = cxc[c];
        rpar[i] = cpar[c];
    }
}

/* first[f], cnt[f] of a list grouped by fit in increasing fit order; the fit of entry i is
 * lfit[sub ? sub[i] : i].  One thread per fit, binary search. */
__global__ void k_list_ranges(const int *__restrict__ lfit, const int *__restrict__ sub,
                              const int *__restrict__ dcount, int F, int *__restrict__ first, int *__restrict__ cnt)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int n = *dcount;
    auto fit_at = [&](int i) { return lfit[sub ? sub[i] : i]; };
    auto lower = [&](int key) {
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (fit_at(mid) < key) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    const int b = lower(f), e = lower(f + 1);
    first[f] = b;
    cnt[f] = e - b;
}

/* Residual columns of the entries marked keep: ldx floats each from the arena (bump allocator
 * shared by the call; the order of the offsets does not matter).  An entry that does not fit
 * in `cap` floats is not stored (keep 0): its residual is recomputed by the final pass. */
__global__ void k_alloc_kept(const FitDev *__restrict__ fits, const int *__restrict__ rfit,
                             const int *__restrict__ dcount, int *__restrict__ keep, long long *__restrict__ roff,
                             unsigned long long *__restrict__ used, unsigned long long cap)
{
    const int n = *dcount;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        long long off = -1;
        if (keep && keep[i]) {
            const unsigned long long need = (unsigned long long)fits[rfit[i]].ldx;
            const unsigned long long at = atomicAdd(used, need);
            if (at + need <= cap) off = (long long)at;
            else keep[i] = 0;
        }
        roff[i] = off;
    }
}

/* Fired entries (list `fl`) without a stored residual take arena columns for the DGKS retry;
 * when the arena is full, redo[0] = 1 (the caller scores the step again on the host-driven
 * path: the retry cannot run without the residual). */
__global__ void k_alloc_fired(const FitDev *__restrict__ fits, const int *__restrict__ rfit,
                              const int *__restrict__ fl, const int *__restrict__ dnf, long long *__restrict__ roff,
                              unsigned long long *__restrict__ used, unsigned long long cap, int *__restrict__ redo)
{
    const int n = *dnf;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const int i = fl[j];
        if (roff[i] >= 0) continue;
        const unsigned long long need = (unsigned long long)fits[rfit[i]].ldx;
        const unsigned long long at = atomicAdd(used, need);
        if (at + need <= cap) roff[i] = (long long)at;
        else redo[0] = 1;
    }
}

__device__ inline int wave_isum(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

/* Order-preserving compaction of the positions i < *dcount with (flag[i] & mask) == want:
 * k_compact_count writes each tile's count, k_compact_write places the tile's entries after
 * those of the tiles before it and writes the length to *out_n. */
__global__ __launch_bounds__(256) void k_compact_count(const int *__restrict__ dcount, const int *__restrict__ flag,
                                                       int mask, int want, int *__restrict__ blk)
{
    __shared__ int sn[4];
    const int n = *dcount, t = threadIdx.x, base = blockIdx.x * kCompactTile;
    int c = 0;
#pragma unroll
    for (int u = 0; u < kCompactTile / 256; ++u) {
        const int i = base + 256 * u + t;
        c += (i < n && (flag[i] & mask) == want) ? 1 : 0;
    }
    c = wave_isum(c);
    if ((t & 63) == 0) sn[t >> 6] = c;
    __syncthreads();
    if (t == 0) blk[blockIdx.x] = sn[0] + sn[1] + sn[2] + sn[3];
}

__global__ __launch_bounds__(256) void k_compact_write(const int *__restrict__ dcount, const int *__restrict__ flag,
                                                       int mask, int want, const int *__restrict__ blk,
                                                       int *__restrict__ out, int *__restrict__ out_n)
{
    __shared__ int sw[4];
    const int n = *dcount, t = threadIdx.x, w = t >> 6, lane = t & 63, b = blockIdx.x;
    // every block of the grid has a count (0 past the list's end): the last one writes the total
    int off = 0;
    for (int j = t; j < b; j += 256) off += blk[j];
    off = wave_isum(off);
    if (lane == 0) sw[w] = off;
    __syncthreads();
    int run = sw[0] + sw[1] + sw[2] + sw[3];
    __syncthreads();
    if (b == (int)gridDim.x - 1 && t == 0) out_n[0] = run + blk[b];
    if (b * kCompactTile >= n) return;
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll 1
    for (int u = 0; u < kCompactTile / 256; ++u) {
        const int i = b * kCompactTile + 256 * u + t;
        const bool take = i < n && (flag[i] & mask) == want;
        const unsigned long long m = __ballot(take);
        if (lane == 0) sw[w] = __popcll(m);
        __syncthreads();
        int woff = run;
        for (int k = 0; k < w; ++k) woff += sw[k];
        if (take) out[woff + __popcll(m & below)] = i;
        run += sw[0] + sw[1] + sw[2] + sw[3];
        __syncthreads();
    }
}

/* One work list for k_pass_build. */
struct PassSpec {
    const int *first, *cnt;  // each fit's positions [first, first + cnt) in the device list
    Item *items;
    int *cmap, *nitems;
    double target;           // work items the pass aims at
    long long minrows, maxsplits, align;
    int tile, cap;
};

/*
 * Work lists of passes over a device list, built on the device the way make_pass and
 * split_rows build them on the host: each fit's positions [first, first + cnt) cut into tiles of
 * `tile`, each tile into row ranges of about n * target / work rows (at least minrows, at most
 * maxsplits ranges, multiples of align); cmap gets (first item, item count, slot) per position.
 * One block of 1024 threads per list (blockIdx.x selects s0, s1 or s2): the fits are split into
 * contiguous ranges per thread for the item counts, and the items themselves are written by the
 * whole block in index order.
 */
__global__ __launch_bounds__(1024) void k_pass_build(const FitDev *__restrict__ fits, int F, PassSpec s0,
                                                     PassSpec s1, PassSpec s2)
{
    const PassSpec S = blockIdx.x == 0 ? s0 : (blockIdx.x == 1 ? s1 : s2);
    const int *__restrict__ first = S.first;
    const int *__restrict__ cnt = S.cnt;
    const int tile = S.tile, cap = S.cap;
    __shared__ double sw[1024];
    __shared__ long long ss[1024];  // inclusive prefix of the item counts of the threads' fit ranges
    const int t = threadIdx.x;
    const int chunk = (F + 1023) / 1024, f0 = min(F, t * chunk), f1 = min(F, f0 + chunk);
    double w = 0.0;
    for (int f = f0; f < f1; ++f) w += (double)fits[f].n * (double)((cnt[f] + tile - 1) / tile);
    sw[t] = w;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) sw[t] = sw[t] + sw[t + s];
        __syncthreads();
    }
    const double work = sw[0];
    auto row_len = [&](int f) {
        const long long n = fits[f].n;
        long long splits = work > 0 ? (long long)(S.target * (double)n / work) : 1;
        splits = max(splits, 1ll);
        splits = min(splits, max(1ll, n / S.minrows));
        splits = min(splits, S.maxsplits);
        const long long rows = (n + splits - 1) / splits;
        return max(S.align, (rows + S.align - 1) / S.align * S.align);
    };
    long long mine = 0;
    for (int f = f0; f < f1; ++f) {
        if (cnt[f] <= 0) continue;
        const long long n = fits[f].n, r = row_len(f);
        mine += (long long)((cnt[f] + tile - 1) / tile) * ((n + r - 1) / r);
    }
    ss[t] = mine;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const long long v = t >= o ? ss[t - o] : 0;
        __syncthreads();
        ss[t] += v;
        __syncthreads();
    }
    const long long total = ss[1023];
    if (t == 0) S.nitems[0] = (int)min(total, (long long)cap);
    // Items are numbered fit by fit, tile by tile, row range by row range.  Item i belongs to the
    // first thread range whose prefix exceeds i, and to the fit of that range whose items cover i;
    // the first row range of each tile also maps the tile's candidates.
    for (long long i = t; i < total; i += 1024) {
        int lo = 0, hi = 1023;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (ss[mid] > i) hi = mid;
            else lo = mid + 1;
        }
        long long at = lo ? ss[lo - 1] : 0, r = 1, n = 0;
        int f = min(F, lo * chunk), nr = 1;
        const int fe = min(F, f + chunk);
        for (; f < fe; ++f) {
            if (cnt[f] <= 0) continue;
            n = fits[f].n;
            r = row_len(f);
            nr = (int)((n + r - 1) / r);
            const long long k = (long long)((cnt[f] + tile - 1) / tile) * nr;
            if (i < at + k) break;
            at += k;
        }
        if (f >= fe) continue;  // (unreachable: the range's counts cover i)
        const int fst = first[f], fc = cnt[f];
        const long long local = i - at;
        const int tl = (int)(local / nr), q = (int)(local - (long long)tl * nr);
        const int c0 = fst + tl * tile, nc = min(tile, fst + fc - c0);
        if (i < cap) {
            Item it;
            it.fit = f;
            it.c0 = c0;
            it.nc = nc;
            it.pad = fst;
            it.r0 = (long long)q * r;
            it.r1 = min(n, (long long)(q + 1) * r);
            S.items[i] = it;
        }
        if (q == 0)
            for (int c = c0; c < c0 + nc; ++c) {
                S.cmap[3 * (size_t)c] = (int)min(i, (long long)cap);
                S.cmap[3 * (size_t)c + 1] = i + nr <= cap ? nr : 0;
                S.cmap[3 * (size_t)c + 2] = c - c0;
            }
    }
}

/* Algorithmic work of one exact-scoring stage over a device list, added to stat[0] (flops) and
 * stat[1] (HBM bytes): 4 n m flops per entry (one T-GEMM and one projection pass,
 * the orthonormalization's 4 n m per candidate), `row_bytes` n bytes per entry (its own columns
 * over the stage's passes), and 8 n m bytes per fit with entries (its Bo columns, once per pass).
 * Entry i's fit is cfit[list ? list[i] : i]; fcnt (F entries) counts each fit's entries.
 * Integer sums: the totals do not depend on the order of the atomic adds. */
__global__ __launch_bounds__(256) void k_work_count(const FitDev *__restrict__ fits, const int *__restrict__ cfit,
                                                    const int *__restrict__ list, const int *__restrict__ dcount,
                                                    const int *__restrict__ fcnt, int F, int row_bytes,
                                                    unsigned long long *__restrict__ stat)
{
    __shared__ unsigned long long sh[2][256];
    const int n = *dcount, stride = gridDim.x * blockDim.x;
    unsigned long long fl = 0, by = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const FitDev &Fv = fits[cfit[list ? list[i] : i]];
        fl += 4ull * (unsigned long long)Fv.n * (unsigned long long)Fv.m;
        by += (unsigned long long)row_bytes * (unsigned long long)Fv.n;
    }
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < F; f += stride)
        if (fcnt[f] > 0) by += 8ull * (unsigned long long)fits[f].n * (unsigned long long)fits[f].m;
    sh[0][threadIdx.x] = fl;
    sh[1][threadIdx.x] = by;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (sh[0][0]) atomicAdd(stat, sh[0][0]);
        if (sh[1][0]) atomicAdd(stat + 1, sh[1][0]);
    }
}

/* lin[list[i]] = val[i], done[list[i]] = 1 for i < *dcount. */
__global__ void k_put_dev(const int *__restrict__ list, const int *__restrict__ dcount,
                          const double *__restrict__ val, double *__restrict__ lin, int *__restrict__ done)
{
    const int n = *dcount;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        lin[list[i]] = val[i];
        if (done) done[list[i]] = 1;
    }
}

/* *acc += *n */
__global__ void k_count_add(const int *__restrict__ n, unsigned long long *__restrict__ acc)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *acc += (unsigned long long)*n;
}

__global__ void k_fill_i(int *__restrict__ p, int n, int v)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *__restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

/* Device pass: items, cmap and the device item count. */
struct DevPass {
    Item *items;
    int *cmap;
    int *n;
    int cap;
};

} // namespace

/*
 * FP64 scoring (the plain form's arithmetic and rounding points) with every list, work item
 * and count built on the device: no host synchronisation.
 *   list != null: the candidates cs[list[i]], i < *dcount (increasing, so grouped by fit);
 *                 bound: the host's upper bound on *dcount; lin[list[i]] = score, done[list[i]] = 1;
 *                 keep[i]: store entry i's first residual for the final pass.
 *   list == null: every candidate of cs (bound = its size); lin[c], ybx[c] = score, ybx; no
 *                 residual is stored except the DGKS-fired ones'.
 * narrow: 16-candidate tiles (few candidates per fit) instead of the T-GEMM / projection widths.
 * Residuals go to the context's arena (kRescoreArena bytes: at most half for stored entries, the
 * rest for the DGKS retry of fired ones); an entry that does not fit is recomputed, and a fired
 * one that does not fit sets *redo (the caller scores again with host-built lists).
 */
void plain_score_device(mars_hip_ctx *ctx, const Cands &cs, const int *list, const int *dcount, int bound,
                        int *keep, double *lin, double *ybx, int *done, int *redo, bool narrow, int phase)
{
    PendingScope pending_scope(ctx);
    if (bound <= 0) return;
    hipStream_t st = ctx->stream;
    const int F = (int)ctx->fits.size();
    size_t mmax = 1;
    for (const FitState &fs : ctx->fits) mmax = std::max(mmax, fs.m);
    require(mmax <= kMaxScoreTerms, "the FP64 scoring path supports up to 512 basis functions");
    const int mt = (int)((mmax + 15) / 16), ldT = 16 * mt;
    const size_t B = (size_t)bound;
    const bool ident = list == nullptr;
    const bool none_stored = keep == nullptr;
    const int tileT = narrow ? 16 : 16 * cg_tgemm(mt), tileP = narrow ? 16 : 16 * cg_project(mt);

    // buffers, sized for the bound (no reallocation once the pipeline is queued)
    Cands &rc = ctx->rcand;
    if (!ident) {
        rc.d_fit.ensure(B);
        rc.d_xc.ensure(B);
        rc.d_par.ensure(B);
    }
    const Cands &sc = ident ? cs : rc;  // the candidates as the passes see them
    ctx->roff.ensure(B);
    ctx->T.ensure(B * ldT);
    ctx->T2.ensure(B * ldT);
    ctx->tnorm.ensure(B);
    ctx->ssum.ensure(B);
    ctx->flag.ensure(B);
    ctx->rlin.ensure(B);
    ctx->rybx.ensure(B);
    ctx->fired_list.ensure(B);
    ctx->live_list.ensure(B);
    ctx->dl_int.ensure(6 * (size_t)F + 16 + 2 * ((B + kCompactTile - 1) / kCompactTile + 1));
    ctx->dl_u64.ensure(2);
    const size_t capT = (size_t)ctx->tune.pass_items + B / 16 + F + 64;   // T-GEMM
    const size_t capP = (size_t)ctx->tune.pass_items + B / 16 + F + 64;   // projection
    const size_t capF = 2048 + B / 16 + F + 64;                 // finalize
    const size_t capR = (size_t)ctx->tune.retry_items + B / 16 + F + 64;  // DGKS retry passes
    const size_t capL = (size_t)ctx->tune.pass_items + B / 16 + F + 64;   // final pass of unstored entries
    ctx->d_items.ensure(capT + capP + capF);
    ctx->d_ritems.ensure(2 * capR + capF);
    ctx->d_litems.ensure(capL);
    ctx->d_cmap.ensure(9 * B + 9);
    ctx->d_rcmap.ensure(9 * B + 9);
    ctx->d_lcmap.ensure(3 * B + 3);
    ctx->tpart.ensure(std::max(capT, capR) * kSlots * ldT);
    ctx->spart.ensure(std::max(std::max(capP, capR) * kSlots, std::max(capF, capL) * kTileF) + 1);
    const size_t arena = kRescoreArena / sizeof(float);
    ctx->rarena.ensure(arena);

    int *ip = ctx->dl_int.p;
    int *rfirst = ip, *rcnt = ip + F, *ffirst = ip + 2 * F, *fcnt = ip + 3 * F, *lfirst = ip + 4 * F,
        *lcnt = ip + 5 * F;
    int *nT = ip + 6 * F, *nF = nT + 1, *nRT = nT + 2, *nRP = nT + 3, *nFF = nT + 4, *nL = nT + 5, *dnf = nT + 6,
        *dnl = nT + 7, *nP = nT + 8, *dall = nT + 9;
    int *blk = nT + 16;
    unsigned long long *used = ctx->dl_u64.p;
    DevPass pT{ctx->d_items.p, ctx->d_cmap.p, nT, (int)capT};
    DevPass pP{ctx->d_items.p + capT, ctx->d_cmap.p + 3 * B + 3, nP, (int)capP};
    DevPass pF{ctx->d_items.p + capT + capP, ctx->d_cmap.p + 6 * B + 6, nF, (int)capF};
    DevPass rT{ctx->d_ritems.p, ctx->d_rcmap.p, nRT, (int)capR};
    DevPass rP{ctx->d_ritems.p + capR, ctx->d_rcmap.p + 3 * B + 3, nRP, (int)capR};
    DevPass fF{ctx->d_ritems.p + 2 * capR, ctx->d_rcmap.p + 6 * B + 6, nFF, (int)capF};
    DevPass pL{ctx->d_litems.p, ctx->d_lcmap.p, nL, (int)capL};

    // fits as seen by these passes: residual columns at per-entry arena offsets
    {
        std::vector<FitDev> h((size_t)F);
        for (int f = 0; f < F; ++f) {
            const FitState &fs = ctx->fits[f];
            FitDev &d = h[f];
            d.X = fs.X.p;
            d.s = fs.s.p;
            d.B = fs.B.p;
            d.Bo = fs.Bo.p;
            d.BoC = fs.BoC.p;
            d.BC = fs.BC.p;
            d.y = fs.y.p;
            d.ybo = fs.ybo.p;
            d.rbuf = ctx->rarena.p;
            d.v = fs.v.p;
            d.tc = fs.tc.p;
            d.tcn = fs.tcn.p;
            d.n = (long long)fs.n;
            d.ldx = (long long)fs.ldx;
            d.ldb = (long long)ctx->ldb;
            d.rsr = 1;
            d.rsc = 1;
            d.m = (int)fs.m;
            d.c0 = 0;
            d.rbyl = 2;
            d.px = (int)fs.px;
            d.act = f < (int)ctx->act_h.size() ? ctx->act_h[f] : 1;
            d.roff = ctx->roff.p;
            d.tol = fs.tol;
        }
        ctx->d_fits2.ensure((size_t)F);
        if (phase != 2) {  // (phase 2 may be captured into a graph: nothing but kernel launches)
            h2d_defer(ctx, ctx->d_fits2.p, h.data(), h.size() * sizeof(FitDev));
            h2d_flush(ctx);
        }
    }
    if (phase == 1) return;
    const FitDev *fits = ctx->d_fits2.p;
    const dim3 gB((unsigned)std::min<size_t>((B + 255) / 256, kRedGrid)), b256(256);
    const dim3 gF((unsigned)((F + 255) / 256));
    const unsigned nblk = (unsigned)((B + kCompactTile - 1) / kCompactTile);
    auto spec = [&](const DevPass &p, const int *first, const int *cnt, int tile, double target, long long minrows,
                    long long align) {
        PassSpec q;
        q.first = first;
        q.cnt = cnt;
        q.items = p.items;
        q.cmap = p.cmap;
        q.nitems = p.n;
        q.target = target;
        q.minrows = minrows;
        q.maxsplits = (long long)kMaxSplits;
        q.align = align;
        q.tile = tile;
        q.cap = p.cap;
        return q;
    };
    // independent lists are built by one launch, a block each
    auto build = [&](const std::vector<PassSpec> &v) {
        require(!v.empty() && v.size() <= 3, "internal: 1 to 3 work lists per build");
        const PassSpec &a = v[0], &b = v.size() > 1 ? v[1] : v[0], &c = v.size() > 2 ? v[2] : v[0];
        hipLaunchKernelGGL(k_pass_build, dim3((unsigned)v.size()), dim3(1024), 0, st, fits, F, a, b, c);
        MH_LAUNCH_CHECK();
    };
    Uploaded uT{pT.items, pT.cmap}, uP{pP.items, pP.cmap}, vT{rT.items, rT.cmap}, vP{rP.items, rP.cmap},
        wF{fF.items, fF.cmap}, uL{pL.items, pL.cmap};
    const int gT = (int)std::min<size_t>(capT, 65536), gP = (int)std::min<size_t>(capP, 65536),
              gR = (int)std::min<size_t>(capR, 65536), gL = (int)std::min<size_t>(capL, 65536);
    double *rl = ident ? lin : ctx->rlin.p;   // scores by entry
    double *ry = ident && ybx ? ybx : ctx->rybx.p;
    const int *cnt_dev = dcount;
    if (ident && !dcount) {  // the count is the host's: written on the device so every kernel reads it alike
        hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, st, dall, 1, bound);
        MH_LAUNCH_CHECK();
        cnt_dev = dall;
    }
    double flops = 0, bytes = 0;
    if (ident) {
        for (int f = 0; f < F; ++f) {
            const FitState &fs = ctx->fits[f];
            const double nf = (double)fs.n;
            flops += 2.0 * nf * fs.m * (f < (int)cs.fit_P.size() ? cs.fit_P[f] : 0);
            bytes += 4.0 * nf * (2.0 * fs.m + 1);
        }
    }

    // candidates of the list, their ranges, residual columns
    if (!ident) {
        hipLaunchKernelGGL(k_rc_gather, gB, b256, 0, st, list, dcount, (const int *)cs.d_fit.p,
                           (const int *)cs.d_xc.p, (const int *)cs.d_par.p, rc.d_fit.p, rc.d_xc.p, rc.d_par.p);
        MH_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_list_ranges, gF, b256, 0, st, (const int *)sc.d_fit.p, nullptr, cnt_dev, F, rfirst, rcnt);
    MH_LAUNCH_CHECK();
    if (!ident) {
        // the rescoring's algorithmic work, counted on the device (its list is the device's): per entry
        // B and X columns in the T-GEMM and the projection (16 n bytes) and its stored residual
        // written and read back by the final pass (8 n bytes)
        hipLaunchKernelGGL(k_work_count, gB, b256, 0, st, fits, (const int *)sc.d_fit.p, nullptr, cnt_dev,
                           (const int *)rcnt, F, 24, ctx->dstat.p + 5);
        MH_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, used, 1);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_alloc_kept, gB, b256, 0, st, fits, (const int *)sc.d_fit.p, cnt_dev, keep, ctx->roff.p,
                       used, (unsigned long long)(arena / 2));
    MH_LAUNCH_CHECK();
    {
        std::vector<PassSpec> v{spec(pT, rfirst, rcnt, tileT, ctx->tune.pass_items, 256, kWR)};
        if (tileP != tileT) v.push_back(spec(pP, rfirst, rcnt, tileP, ctx->tune.pass_items, 256, kWR));
        if (!none_stored) v.push_back(spec(pF, rfirst, rcnt, 16, 2048, 512, 32));
        build(v);
    }
    const DevPass &pPr = tileP != tileT ? pP : pT;
    const Uploaded &uPr = tileP != tileT ? uP : uT;
    const int gPr = tileP != tileT ? gP : gT;

    // T = Bo^T Bx, s = sum r^2 (residuals stored for the entries kept), DGKS flags
    {
        KTimer kt(ctx, K_TGEMM, flops, bytes);
        launch_tgemm<false>(ctx, mt, uT, gT, sc, nullptr, ldT, narrow, nT, fits);
        launch_treduce(st, red_waves(B), B, fits, (const int *)sc.d_fit.p, nullptr, bound, (const int *)pT.cmap,
                       (const double *)ctx->tpart.p, ldT, ctx->T.p, ctx->tnorm.p, cnt_dev);
    }
    {
        KTimer kt(ctx, K_PROJECT, flops, bytes);
        if (none_stored)
            launch_project<kSumSq, false, false>(ctx, mt, uPr, gPr, sc, nullptr, ctx->T.p, ldT, narrow, nullptr,
                                                 pPr.n, fits);
        else
            launch_project<kStoreR, false, false>(ctx, mt, uPr, gPr, sc, nullptr, ctx->T.p, ldT, narrow, keep,
                                                  pPr.n, fits);
        launch_sreduce(st, red_lanes(B), B, nullptr, bound, (const int *)pPr.cmap, (const double *)ctx->spart.p, kSlots,
                       (const double *)ctx->tnorm.p, fits, (const int *)sc.d_fit.p, (const int *)keep, ctx->ssum.p,
                       ctx->flag.p, cnt_dev);
    }
    // final pass over the stored residuals (unstored entries with s > tol get 0 here)
    {
        KTimer kt(ctx, K_FINALIZE, 0.0, 0.0);
        if (none_stored) {
            hipLaunchKernelGGL(k_zero2, dim3((unsigned)((B + 255) / 256)), b256, 0, st, bound, ry, rl);
            MH_LAUNCH_CHECK();
        } else {
            hipLaunchKernelGGL(k_finalize_cm, dim3((unsigned)std::min<size_t>(capF, 4096)), b256, 0, st, fits,
                               (const Item *)pF.items, (int)capF, nullptr, (const int *)keep,
                               (const double *)ctx->ssum.p, ctx->spart.p, (const int *)nF, nullptr, 0);
            MH_LAUNCH_CHECK();
            launch_yreduce(st, red_lanes(B), B, bound, (const int *)pF.cmap, (const double *)ctx->spart.p, nullptr, ry,
                           rl, cnt_dev);
        }
    }
    // DGKS retry of the fired entries (kernels.cc:387-399)
    {
        KTimer kt(ctx, K_DGKS, 0.0, 0.0);
        hipLaunchKernelGGL(k_compact_count, dim3(nblk), b256, 0, st, cnt_dev, (const int *)ctx->flag.p, 1, 1, blk);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_compact_write, dim3(nblk), b256, 0, st, cnt_dev, (const int *)ctx->flag.p, 1, 1,
                           (const int *)blk, ctx->fired_list.p, dnf);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_list_ranges, gF, b256, 0, st, (const int *)sc.d_fit.p, (const int *)ctx->fired_list.p,
                           (const int *)dnf, F, ffirst, fcnt);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_count_add, dim3(1), dim3(64), 0, st, (const int *)dnf, ctx->dstat.p + 4);
        MH_LAUNCH_CHECK();
        if (!ident) {  // the retry of each fired entry: its first residual read twice, the second written
            hipLaunchKernelGGL(k_work_count, gB, b256, 0, st, fits, (const int *)sc.d_fit.p,
                               (const int *)ctx->fired_list.p, (const int *)dnf, (const int *)fcnt, F, 12,
                               ctx->dstat.p + 5);
            MH_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_alloc_fired, dim3(4), b256, 0, st, fits, (const int *)sc.d_fit.p,
                           (const int *)ctx->fired_list.p, (const int *)dnf, ctx->roff.p, used,
                           (unsigned long long)arena, redo);
        MH_LAUNCH_CHECK();
        build({spec(rT, ffirst, fcnt, 16, ctx->tune.retry_items, 256, kWR),
               spec(rP, ffirst, fcnt, 16, ctx->tune.retry_items, 256, kWR), spec(fF, ffirst, fcnt, 16, 2048, 512, 32)});
        const int *fl = ctx->fired_list.p;
        // the first residual of every fired entry (recomputed where it was stored already)
        launch_project<kStoreR, false, true>(ctx, mt, vP, gR, sc, fl, ctx->T.p, ldT, true, nullptr, nRP, fits);
        launch_tgemm<true>(ctx, mt, vT, gR, sc, fl, ldT, true, nRT, fits);
        // (the fired count is the device's and usually small: a few waves or lanes per entry)
        launch_treduce(st, 4, 256, fits, (const int *)sc.d_fit.p, fl, bound, (const int *)rT.cmap,
                       (const double *)ctx->tpart.p, ldT, ctx->T2.p, nullptr, (const int *)dnf);
        launch_project<kStoreR, true, false>(ctx, mt, vP, gR, sc, fl, ctx->T2.p, ldT, true, nullptr, nRP, fits);
        launch_sreduce(st, 16, 1024, fl, bound, (const int *)rP.cmap, (const double *)ctx->spart.p, kSlots, nullptr,
                       fits, (const int *)sc.d_fit.p, nullptr, ctx->ssum.p, nullptr, (const int *)dnf);
        hipLaunchKernelGGL(k_finalize_cm, dim3(256), b256, 0, st, fits, (const Item *)fF.items, (int)capF, fl, nullptr,
                           (const double *)ctx->ssum.p, ctx->spart.p, (const int *)nFF, nullptr, 0);
        MH_LAUNCH_CHECK();
        launch_yreduce(st, 16, 1024, bound, (const int *)fF.cmap, (const double *)ctx->spart.p, fl, ry, rl,
                       (const int *)dnf);
    }
    // entries with s > tol whose residual was not stored and that did not fire: recompute r
    {
        KTimer kt(ctx, K_FINALIZE, 0.0, 0.0);
        const int mask = none_stored ? 3 : 7, want = none_stored ? 2 : 6;
        hipLaunchKernelGGL(k_compact_count, dim3(nblk), b256, 0, st, cnt_dev, (const int *)ctx->flag.p, mask, want,
                           blk);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_compact_write, dim3(nblk), b256, 0, st, cnt_dev, (const int *)ctx->flag.p, mask, want,
                           (const int *)blk, ctx->live_list.p, dnl);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_list_ranges, gF, b256, 0, st, (const int *)sc.d_fit.p, (const int *)ctx->live_list.p,
                           (const int *)dnl, F, lfirst, lcnt);
        MH_LAUNCH_CHECK();
        // the live entries of a full scoring are most of the list: they take the wide tiles
        const int tileL = none_stored ? tileP : 16;
        build({spec(pL, lfirst, lcnt, tileL, ctx->tune.pass_items, 256, kWR)});
        launch_project<kFinalY, false, true>(ctx, mt, uL, gL, sc, ctx->live_list.p, ctx->T.p, ldT, tileL == 16,
                                             nullptr, nL, fits);
        launch_yreduce(st, red_lanes(B), B, bound, (const int *)pL.cmap, (const double *)ctx->spart.p,
                       (const int *)ctx->live_list.p, ry, rl, (const int *)dnl);
    }
    if (!ident) {
        hipLaunchKernelGGL(k_put_dev, gB, b256, 0, st, list, dcount, (const double *)rl, lin, done);
        MH_LAUNCH_CHECK();
    }
}

void plain_score(mars_hip_ctx *ctx, Cands &cs, bool keep_bxn, double *lin, double *ybx,
                 const float *xover, const float *sover, bool rescore, const int *keep, bool f32)
{
    require(!f32 || (keep_bxn && !rescore), "internal: the FP32 first pass serves the hinge path");
    PendingScope pending_scope(ctx);
    const bool store_r = keep_bxn || rescore;  // residuals stay in rbuf (all, or those `keep` marks)
    if (!rescore) keep = nullptr;
    const int P = (int)cs.fit.size();
    if (P == 0) return;
    hipStream_t st = ctx->stream;
    const size_t F = ctx->fits.size();
    index_cands(ctx, cs);
    // the plain form's linear scoring stores no residual in the fits' buffers: its DGKS-fired
    // candidates keep theirs in the bounded retry arena (retry_compact)
    if (store_r) ensure_rbuf(ctx, cs);
    // normalised columns kept for the gather sweep or a download stay row-major; residuals that
    // only linear scoring reads, and the sorted hinge path's columns (keep_cm), are column-major
    // (contiguous per candidate)
    ctx->rbuf_cm = !keep_bxn || ctx->keep_cm;
    upload_fits(ctx, cs, xover, sover);
    upload_cands(ctx, cs);

    size_t mmax = 1;
    std::vector<long long> n(F);
    std::vector<int> fb(F), fe(F), cnt(F);
    double flops = 0, bytes = 0;
    for (size_t f = 0; f < F; ++f) {
        mmax = std::max(mmax, ctx->fits[f].m);
        n[f] = (long long)ctx->fits[f].n;
        fb[f] = cs.fit_c0[f];
        cnt[f] = cs.fit_P[f];
        fe[f] = fb[f] + cnt[f];
        flops += 2.0 * n[f] * ctx->fits[f].m * cnt[f];
        bytes += 4.0 * n[f] * (2.0 * ctx->fits[f].m + 1);
    }
    const int mt = (int)((mmax + 15) / 16);
    require(mmax <= kMaxScoreTerms, "the FP64 scoring path supports up to 512 basis functions");
    const int ldT = mt * 16;
    const bool narrow = narrow_for(cnt, ctx->tune.narrow_avg);
    const int tileT = 16 * (narrow ? 1 : cg_tgemm(mt)), tileP = 16 * (narrow ? 1 : cg_project(mt));

    // equal tile widths give the two passes the same work list: it is built and sent once
    const bool shared = tileT == tileP;
    Pass pT = make_pass(fb, fe, n, split_rows(n, cnt, tileT, ctx->tune.pass_items, 256), tileT, kWR);
    Pass pPown;
    if (!shared) pPown = make_pass(fb, fe, n, split_rows(n, cnt, tileP, ctx->tune.pass_items, 256), tileP, kWR);
    const Pass &pP = shared ? pT : pPown;
    Pass pF;  // finalize work list, over stored residuals (hinge path, rescoring)
    const int tileF = narrow ? 16 : kTileF;
    if (store_r) pF = make_pass(fb, fe, n, split_rows(n, cnt, tileF, 2048, 512), tileF, 32);
    Uploaded uT, uP, uF;
    size_t io = 0, co = 0;
    const size_t total_items = pT.items.size() + pPown.items.size() + pF.items.size();
    const size_t total_cmap = pT.cmap.size() + pPown.cmap.size() + pF.cmap.size();
    ctx->d_items.ensure(total_items + 1);
    ctx->d_cmap.ensure(total_cmap + 1);
    upload_pass(ctx, pT, ctx->d_items, ctx->d_cmap, io, co, uT);
    io += pT.items.size(); co += pT.cmap.size();
    if (shared) {
        uP = uT;
    } else {
        upload_pass(ctx, pP, ctx->d_items, ctx->d_cmap, io, co, uP);
        io += pP.items.size(); co += pP.cmap.size();
    }
    upload_pass(ctx, pF, ctx->d_items, ctx->d_cmap, io, co, uF);
    h2d_flush(ctx);  // fits, candidates and the three work lists in one copy

    ctx->tpart.ensure(std::max<size_t>(pT.items.size(), 1) * kSlots * ldT);
    ctx->T.ensure((size_t)P * ldT);
    ctx->tnorm.ensure(P);
    ctx->spart.ensure(std::max(pP.items.size() * kSlots, pF.items.size() * kTileF) + 1);
    ctx->ssum.ensure(P);
    ctx->flag.ensure(P);
    ctx->fired_list.ensure(P);

    // Pass 1: T = Bo^T Bx (FP32 MFMA in the precision form's approximate pass)
    {
        KTimer kt(ctx, f32 ? (ctx->split_bf16 && mt <= 16 ? K_TGEMM_BF16 : K_TGEMM32) : K_TGEMM, flops, bytes);
        // (past 256 basis columns the FP32 T-GEMM serves the BF16-split form: its error is smaller
        // than the split's, which the screening bound assumes)
        if (f32 && ctx->split_bf16 && mt <= 16)
            launch_tgemm_bf16(ctx, mt, uT, (int)pT.items.size(), cs, ldT, narrow);
        else if (f32)
            launch_tgemm<false, true>(ctx, mt, uT, (int)pT.items.size(), cs, nullptr, ldT, narrow);
        else
            launch_tgemm<false>(ctx, mt, uT, (int)pT.items.size(), cs, nullptr, ldT, narrow);
        launch_treduce(st, red_waves(P), P, ctx->d_fits.p, cs.d_fit.p, nullptr, P, uT.cmap, ctx->tpart.p, ldT, ctx->T.p,
                       ctx->tnorm.p, nullptr);
    }
    // Pass 2: s = sum f32(Bx - Bo T)^2 and the DGKS gate; r itself is kept only for the
    // hinge sweep (which needs the normalized columns)
    {
        KTimer kt(ctx, f32 ? K_PROJECT32 : K_PROJECT, flops, bytes + 0.0);
        if (f32)
            launch_project<kStoreR, false, false, true>(ctx, mt, uP, (int)pP.items.size(), cs, nullptr, ctx->T.p,
                                                        ldT, narrow);
        else if (store_r)
            launch_project<kStoreR, false, false>(ctx, mt, uP, (int)pP.items.size(), cs, nullptr, ctx->T.p, ldT,
                                                  narrow, keep);
        else
            launch_project<kSumSq, false, false>(ctx, mt, uP, (int)pP.items.size(), cs, nullptr, ctx->T.p, ldT,
                                                 narrow);
        launch_sreduce(st, red_lanes(P), P, nullptr, P, uP.cmap, ctx->spart.p, kSlots, ctx->tnorm.p, ctx->d_fits.p,
                       cs.d_fit.p, keep, ctx->ssum.p, ctx->flag.p, nullptr);
    }
    // Stored column-major residuals (the mixed form's rescoring, the sorted hinge path) are
    // finalized before the host has read the DGKS flags: the device runs pass 3 over every
    // candidate while the flags come back, and the few fired candidates are finalized again after
    // their retry (a kept column is normalised in place only once its residual is final).
    const bool spec = store_r && ctx->rbuf_cm;
    const int nF = (int)pF.items.size();
    std::vector<int> flag_v;
    const int *flags = nullptr;
    if (spec) {
        if (ctx->rb_cap < (size_t)P) {
            // an earlier readback may still be in flight into the old block: retire it
            if (ctx->rb_pin) retire_host(ctx->rb_pin);
            ctx->rb_pin = nullptr;
            const size_t cap = std::max<size_t>((size_t)P, std::max<size_t>(2 * ctx->rb_cap, 65536));
            ctx->rb_cap = 0;
            MH_CHECK(hipHostMalloc((void **)&ctx->rb_pin, cap * sizeof(int), hipHostMallocDefault));
            ctx->rb_cap = cap;
        }
        MH_CHECK(hipMemcpyAsync(ctx->rb_pin, ctx->flag.p, P * sizeof(int), hipMemcpyDeviceToHost, st));
        MH_CHECK(hipEventRecord(ctx->ev_rb, st));
        {
            double fb2 = 0;
            for (size_t f = 0; f < F; ++f) fb2 += 4.0 * n[f] * cnt[f] + 4.0 * n[f];
            KTimer kt(ctx, K_FINALIZE, 0.0, fb2);
            hipLaunchKernelGGL(k_finalize_cm, dim3(grid_for(pF.items.size())), dim3(256), 0, st, ctx->d_fits.p,
                               uF.items, nF, nullptr, keep, ctx->ssum.p, ctx->spart.p, nullptr,
                               keep_bxn ? (const int *)ctx->flag.p : nullptr, keep_bxn ? 1 : 0);
            MH_LAUNCH_CHECK();
            launch_yreduce(st, red_lanes(P), P, P, uF.cmap, ctx->spart.p, nullptr, ybx, lin, nullptr);
        }
        MH_CHECK(hipEventSynchronize(ctx->ev_rb));
        flags = ctx->rb_pin;
    } else {
        flag_v.resize(P);
        MH_CHECK(hipMemcpyAsync(flag_v.data(), ctx->flag.p, P * sizeof(int), hipMemcpyDeviceToHost, st));
        ctx_sync(ctx);
        flags = flag_v.data();
    }
    // DGKS retry of the flagged columns, grouped by fit (kernels.cc:387-399)
    // late: candidates with s > tol whose residual was not stored and no retry will store
    std::vector<int> fired, live, late;
    std::vector<int> rb(F, 0), re(F, 0), lb(F, 0), le(F, 0), tb(F, 0), te(F, 0);
    bool fired_unstored = false;
    for (size_t f = 0; f < F; ++f) {
        rb[f] = (int)fired.size();
        lb[f] = (int)live.size();
        tb[f] = (int)late.size();
        for (int c = fb[f]; c < fe[f]; ++c) {
            if (flags[c] & 1) fired.push_back(c);
            if (flags[c] & 2) live.push_back(c);
            if ((flags[c] & 7) == 6) late.push_back(c);
            if ((flags[c] & 5) == 5) fired_unstored = true;
        }
        re[f] = (int)fired.size();
        le[f] = (int)live.size();
        te[f] = (int)late.size();
    }
    if (!fired.empty() && store_r) {
        const int nf = (int)fired.size();
        ctx->dgks_host += nf;
        KTimer kt(ctx, K_DGKS, 0.0, 0.0);
        h2d_defer(ctx, ctx->fired_list.p, fired.data(), nf * sizeof(int));
        std::vector<int> fcnt(F);
        for (size_t f = 0; f < F; ++f) fcnt[f] = re[f] - rb[f];
        const bool rnarrow = narrow_for(fcnt, ctx->tune.narrow_avg);
        const int rtT = 16 * (rnarrow ? 1 : cg_tgemm(mt)), rtP = 16 * (rnarrow ? 1 : cg_project(mt));
        Pass rT = make_pass(rb, re, n, split_rows(n, fcnt, rtT, ctx->tune.retry_items, 256), rtT, kWR);
        Pass rP = make_pass(rb, re, n, split_rows(n, fcnt, rtP, ctx->tune.retry_items, 256), rtP, kWR);
        Uploaded vT, vP;
        ctx->d_ritems.ensure(rT.items.size() + rP.items.size() + 1);
        ctx->d_rcmap.ensure(rT.cmap.size() + rP.cmap.size() + 1);
        upload_pass(ctx, rT, ctx->d_ritems, ctx->d_rcmap, 0, 0, vT);
        upload_pass(ctx, rP, ctx->d_ritems, ctx->d_rcmap, rT.items.size(), rT.cmap.size(), vP);
        h2d_flush(ctx);
        ctx->tpart.ensure(std::max<size_t>(rT.items.size(), 1) * kSlots * ldT);
        ctx->spart.ensure(rP.items.size() * kSlots + 1);
        ctx->T2.ensure((size_t)nf * ldT);
        if (fired_unstored)  // the first residual of the fired columns, for the retry
            launch_project<kStoreR, false, true>(ctx, mt, vP, (int)rP.items.size(), cs, ctx->fired_list.p,
                                                 ctx->T.p, ldT, rnarrow);
        launch_tgemm<true>(ctx, mt, vT, (int)rT.items.size(), cs, ctx->fired_list.p, ldT, rnarrow);
        launch_treduce(st, red_waves(nf), nf, ctx->d_fits.p, cs.d_fit.p, ctx->fired_list.p, nf, vT.cmap, ctx->tpart.p,
                       ldT, ctx->T2.p, nullptr, nullptr);
        launch_project<kStoreR, true, false>(ctx, mt, vP, (int)rP.items.size(), cs, ctx->fired_list.p, ctx->T2.p,
                                             ldT, rnarrow);
        launch_sreduce(st, red_lanes(nf), nf, ctx->fired_list.p, nf, vP.cmap, ctx->spart.p, kSlots, nullptr,
                       ctx->d_fits.p, cs.d_fit.p, nullptr, ctx->ssum.p, nullptr, nullptr);
        if (spec) {  // pass 3 again for the fired candidates, over their retried residuals
            Pass fF = make_pass(rb, re, n, split_rows(n, fcnt, 16, 2048, 512), 16, 32);
            Uploaded wF;
            ctx->d_litems.ensure(fF.items.size() + 1);
            ctx->d_lcmap.ensure(fF.cmap.size() + 1);
            upload_pass(ctx, fF, ctx->d_litems, ctx->d_lcmap, 0, 0, wF);
            h2d_flush(ctx);
            ctx->spart.ensure(fF.items.size() * kTileF + 1);
            hipLaunchKernelGGL(k_finalize_cm, dim3(grid_for(fF.items.size())), dim3(256), 0, st, ctx->d_fits.p,
                               wF.items, (int)fF.items.size(), (const int *)ctx->fired_list.p, nullptr,
                               ctx->ssum.p, ctx->spart.p, nullptr, nullptr, keep_bxn ? 1 : 0);
            MH_LAUNCH_CHECK();
            launch_yreduce(st, red_lanes(nf), nf, nf, wF.cmap, ctx->spart.p, ctx->fired_list.p, ybx, lin, nullptr);
        }
    }
    if (spec) {
        // candidates with s > tol whose residual was not stored: recompute r and finish them
        // (the final pass above gave them 0)
        const int nl = (int)late.size();
        if (nl > 0) {
            KTimer kt(ctx, K_FINALIZE, 0.0, 0.0);
            std::vector<int> lcnt(F);
            for (size_t f = 0; f < F; ++f) lcnt[f] = te[f] - tb[f];
            const bool lnarrow = narrow_for(lcnt, ctx->tune.narrow_avg);
            const int ltP = 16 * (lnarrow ? 1 : cg_project(mt));
            Pass pL = make_pass(tb, te, n, split_rows(n, lcnt, ltP, ctx->tune.pass_items, 256), ltP, kWR);
            Uploaded uL;
            ctx->live_list.ensure(nl);
            ctx->d_litems.ensure(pL.items.size() + 1);
            ctx->d_lcmap.ensure(pL.cmap.size() + 1);
            h2d_defer(ctx, ctx->live_list.p, late.data(), nl * sizeof(int));
            upload_pass(ctx, pL, ctx->d_litems, ctx->d_lcmap, 0, 0, uL);
            h2d_flush(ctx);
            ctx->spart.ensure(pL.items.size() * kSlots + 1);
            launch_project<kFinalY, false, true>(ctx, mt, uL, (int)pL.items.size(), cs, ctx->live_list.p,
                                                 ctx->T.p, ldT, lnarrow);
            launch_yreduce(st, red_lanes(nl), nl, nl, uL.cmap, ctx->spart.p, ctx->live_list.p, ybx, lin, nullptr);
        }
        return;
    }
    // Pass 3: Bxn = f32(r * scale), ybx = sum Bxn y
    {
        double fb2 = 0;
        for (size_t f = 0; f < F; ++f) fb2 += (keep_bxn ? 8.0 : 4.0) * n[f] * cnt[f] + 4.0 * n[f];
        KTimer kt(ctx, K_FINALIZE, store_r ? 0.0 : flops, fb2);
        if (store_r) {
            // Bxn = f32(r scale), ybx = sum Bxn y; a candidate with s <= tol has scale 0, so ybx = 0
            // (row-major normalised columns, kept for the hinge sweep or a download)
            const dim3 gF(grid_for(pF.items.size()));
            if (narrow)
                hipLaunchKernelGGL((k_finalize<16>), gF, dim3(256), 0, st, ctx->d_fits.p, uF.items, nF,
                                   ctx->ssum.p, ctx->spart.p);
            else
                hipLaunchKernelGGL((k_finalize<kTileF>), gF, dim3(256), 0, st, ctx->d_fits.p, uF.items, nF,
                                   ctx->ssum.p, ctx->spart.p);
            MH_LAUNCH_CHECK();
            launch_yreduce(st, red_lanes(P), P, P, uF.cmap, ctx->spart.p, nullptr, ybx, lin, nullptr);
        } else {
            // only candidates with s > tol can score above 0; the others are 0 exactly.  The
            // fired ones are finished by their retry (retry_compact); the rest recompute r.
            hipLaunchKernelGGL(k_zero2, dim3((P + 255) / 256), dim3(256), 0, st, P, ybx, lin);
            MH_LAUNCH_CHECK();
            {
                std::vector<int> keep_live;
                keep_live.reserve(live.size());
                for (size_t f = 0; f < F; ++f) {
                    const int b = (int)keep_live.size();
                    for (int i = lb[f]; i < le[f]; ++i)
                        if (!(flags[live[i]] & 1)) keep_live.push_back(live[i]);
                    lb[f] = b;
                    le[f] = (int)keep_live.size();
                }
                live.swap(keep_live);
            }
            const int nl = (int)live.size();
            if (nl > 0) {
                std::vector<int> lcnt(F);
                for (size_t f = 0; f < F; ++f) lcnt[f] = le[f] - lb[f];
                const bool lnarrow = narrow_for(lcnt, ctx->tune.narrow_avg);
                const int ltP = 16 * (lnarrow ? 1 : cg_project(mt));
                Pass pL = make_pass(lb, le, n, split_rows(n, lcnt, ltP, ctx->tune.pass_items, 256), ltP, kWR);
                Uploaded uL;
                ctx->live_list.ensure(nl);
                ctx->d_litems.ensure(pL.items.size() + 1);
                ctx->d_lcmap.ensure(pL.cmap.size() + 1);
                h2d_defer(ctx, ctx->live_list.p, live.data(), nl * sizeof(int));
                upload_pass(ctx, pL, ctx->d_litems, ctx->d_lcmap, 0, 0, uL);
                h2d_flush(ctx);
                ctx->spart.ensure(pL.items.size() * kSlots + 1);
                launch_project<kFinalY, false, true>(ctx, mt, uL, (int)pL.items.size(), cs, ctx->live_list.p,
                                                     ctx->T.p, ldT, lnarrow);
                launch_yreduce(st, red_lanes(nl), nl, nl, uL.cmap, ctx->spart.p, ctx->live_list.p, ybx, lin, nullptr);
            }
            if (!fired.empty()) {
                KTimer kt2(ctx, K_DGKS, 0.0, 0.0);
                retry_compact(ctx, cs, fired, rb, re, n, mt, ldT, ybx, lin);
            }
        }
        MH_LAUNCH_CHECK();
    }
}

} // namespace mh
