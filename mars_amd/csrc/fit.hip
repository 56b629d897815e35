// Resident fit state on the device: constructor, batched append, step
// evaluation, mirror mode and the C ABI of libmars_hip (include/mars_hip.h).
#include "internal.h"
#include "kernels.h"

#include <atomic>
#include <algorithm>
#include <mutex>
#include <functional>
#include <cmath>
#include <cstring>

namespace mh {

thread_local std::string g_last_error;

void set_last_error(const char *msg) { g_last_error = msg; }

namespace {

/* per-fit scalar slots */
enum { S_YN = 0, S_WN, S_YBO0, S_MEAN, S_YVAR, S_BADX, S_PROJ, S_PROJ2, S_VV, S_COUNT };

constexpr int kInitBlocks = 256;

/* ---------------------------------------------------------------------- */
/* Constructor kernels (marsalgo.cc:266-370), one fit at a time            */
/* ---------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_init1(const float *__restrict__ y, const float *__restrict__ w,
                                               size_t n, double *__restrict__ yd,
                                               double *__restrict__ sw, double *__restrict__ part)
{
    __shared__ double sh[256];
    double ay = 0.0, aw = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)kInitBlocks * 256) {
        double yi = (double)y[i];
        double si = sqrt((double)w[i]);
        if (!isfinite(yi)) yi = si = 0.0;
        const double d = yi * si;
        yd[i] = d;
        sw[i] = si;
        ay = fma(d, d, ay);
        aw = fma(si, si, aw);
    }
    ay = block_sum_256(ay, sh);
    aw = block_sum_256(aw, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = ay;
        part[kInitBlocks + blockIdx.x] = aw;
    }
}

__global__ __launch_bounds__(256) void k_init2(double *__restrict__ yd, const double *__restrict__ sw,
                                               size_t n, const double *__restrict__ part,
                                               float *__restrict__ y32, float *__restrict__ B,
                                               float *__restrict__ Bo, size_t ldb, float *__restrict__ BoC,
                                               float *__restrict__ BC,
                                               double *__restrict__ part2, double *__restrict__ scal)
{
    __shared__ double sh[256];
    const double yn = sqrt(sum_partials(part, kInitBlocks, sh));
    const double wn = sqrt(sum_partials(part + kInitBlocks, kInitBlocks, sh));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[S_YN] = yn;
        scal[S_WN] = wn;
    }
    double ab = 0.0, am = 0.0;
    if (yn > 0.0 && wn > 0.0) {
        for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)kInitBlocks * 256) {
            const double d = yd[i] / yn;
            const double s = sw[i] / wn;
            yd[i] = d;
            const float yf = (float)d;
            const float b = (float)s;
            y32[i] = yf;
            if (B) B[i * ldb] = b;
            if (Bo) Bo[i * ldb] = b;
            BoC[i] = b;
            BC[i] = b;
            ab = fma((double)b, (double)yf, ab);
            am += d;
        }
    }
    ab = block_sum_256(ab, sh);
    am = block_sum_256(am, sh);
    if (threadIdx.x == 0) {
        part2[blockIdx.x] = ab;
        part2[kInitBlocks + blockIdx.x] = am;
    }
}

__global__ __launch_bounds__(256) void k_init3(const double *__restrict__ yd, size_t n,
                                               const double *__restrict__ part2,
                                               double *__restrict__ part3, double *__restrict__ scal)
{
    __shared__ double sh[256];
    const double ybo0 = sum_partials(part2, kInitBlocks, sh);
    const double mean = sum_partials(part2 + kInitBlocks, kInitBlocks, sh) / (double)n;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[S_YBO0] = ybo0;
        scal[S_MEAN] = mean;
    }
    double av = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)kInitBlocks * 256) {
        const double dv = yd[i] - mean;
        av = fma(dv, dv, av);
    }
    av = block_sum_256(av, sh);
    if (threadIdx.x == 0) part3[blockIdx.x] = av;
}

__global__ __launch_bounds__(256) void k_init4(const double *__restrict__ part3, size_t n,
                                               double *__restrict__ scal, double *__restrict__ ybo)
{
    __shared__ double sh[256];
    const double v = sum_partials(part3, kInitBlocks, sh);
    if (threadIdx.x == 0) {
        scal[S_YVAR] = v / (double)n;
        ybo[0] = scal[S_YBO0];
    }
}

/* Column sums of squares, 64 blocks per column. */
__global__ __launch_bounds__(256) void k_colsq(const float *__restrict__ X, size_t ldx, size_t n,
                                               double *__restrict__ part)
{
    __shared__ double sh[256];
    const float *x = X + (size_t)blockIdx.y * ldx;
    double a = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += 64 * 256) {
        const double v = (double)x[i];
        a = fma(v, v, a);
    }
    a = block_sum_256(a, sh);
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * 64 + blockIdx.x] = a;
}

__global__ void k_colscale(const double *__restrict__ part, int px, size_t n, float *__restrict__ s,
                           double *__restrict__ scal)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= px) return;
    double sq = 0.0;
    for (int b = 0; b < 64; ++b) sq += part[(size_t)j * 64 + b];
    const double rms = sqrt(sq / (double)n);
    if (!isfinite(rms)) scal[S_BADX] = 1.0;
    s[j] = rms > 0.0 ? (float)(1.0 / rms) : 1.0f;
}

/* ---------------------------------------------------------------------- */
/* Append kernels (marsalgo.cc:760-875, kernels.cc:239-281), batched: grid */
/* (nb, fits); per-fit partials at part[f * nb + b], scalars at scal[f*S]   */
/* ---------------------------------------------------------------------- */
__device__ inline float relu0(float a) { return a < 0.0f ? 0.0f : a; }

__global__ __launch_bounds__(256) void k_app_build(const FitDev *__restrict__ fits,
                                                   const AppendArg *__restrict__ args, int nb,
                                                   double *__restrict__ part, double *__restrict__ scal)
{
    __shared__ double sh[256];
    const int f = blockIdx.y;
    const AppendArg A = args[f];
    if (!A.active) return;
    const FitDev F = fits[f];
    const float s = F.s[A.xcol];
    const float *xp = F.X + (size_t)A.xcol * F.ldx;
    double a = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < F.n; i += (long long)nb * 256) {
        const float sb = __fmul_rn(s, F.BC[(size_t)A.bcol * F.ldx + i]);
        float v;
        if (A.type == 0)
            v = __fmul_rn(sb, xp[i]);
        else if (A.type == 1)
            v = __fmul_rn(sb, relu0(__fsub_rn(xp[i], A.h)));
        else
            v = __fmul_rn(sb, relu0(__fsub_rn(A.h, xp[i])));
        const double d = (double)v;
        F.v[i] = d;
        a = fma(d, d, a);
    }
    a = block_sum_256(a, sh);
    if (threadIdx.x == 0) part[(size_t)f * nb + blockIdx.x] = a;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[f * S_COUNT + S_PROJ] = 0.0;
        scal[f * S_COUNT + S_PROJ2] = 0.0;
    }
}

/*
 * Modified Gram-Schmidt of the residual v against Bo's m columns (kernels.cc:239-281), in
 * two passes over the rows.  MGS's coefficient of column j is c_j = q_j . v_j with
 * v_j = v - sum_{l<j} c_l q_l, which equals d_j - sum_{l<j} c_l G_jl for d = Q^T v and the
 * Gram matrix G = Q^T Q of the stored f32 columns (all products exact in FP64).  So:
 *   k_app_dots   partials of d (kAppG columns per block), and on the first sweep of an
 *                append also the Gram row of column m-1 (the newest one) and the stored
 *                raw-norm-normalized B column (marsalgo.cc:829-836);
 *   k_app_solve  per fit: d and the new Gram row from their partials, then the recurrence
 *                for c in column order, and sum c^2 for the DGKS gate;
 *   k_app_sub    v <- v - c_0 q_0 - c_1 q_1 - ... row by row (MGS's order of updates), and the
 *                partials of |v|^2.
 * `gate` (retry sweep) makes a launch a no-op for fits whose DGKS gate did not fire.
 */
constexpr int kAppG = 8;              // Bo columns per k_app_dots block
constexpr int kMaxAppendCols = 512;   // basis columns the device append orthogonalizes against

__global__ __launch_bounds__(256) void k_app_dots(const FitDev *__restrict__ fits,
                                                  const AppendArg *__restrict__ args, int nb,
                                                  const double *__restrict__ bpart, int first,
                                                  const int *__restrict__ gate, double *__restrict__ dpart,
                                                  double *__restrict__ gpart, const FitW *__restrict__ W)
{
    __shared__ double sh[4][2 * kAppG];
    __shared__ double shr[256];
    const int f = blockIdx.y, g0 = kAppG * (int)blockIdx.z;
    if (!args[f].active) return;
    const FitDev F = fits[f];
    const int m = F.m;
    if (gate && !gate[f]) return;
    // the first group also writes the B column; the others only serve columns they hold
    if (g0 >= m && !(first && blockIdx.z == 0)) return;
    const int kn = min(kAppG, m - g0);                       // live columns of this group
    const bool grow = first && m >= 2;                      // Gram row of column m - 1
    float rn = 1.0f;
    if (first && blockIdx.z == 0) rn = (float)sqrt(sum_partials(bpart + (size_t)f * nb, nb, shr));
    float *B = W[f].B, *BC = W[f].BC + (size_t)m * F.ldx;
    const float *qn = F.BoC + (size_t)(m - 1) * F.ldx;
    // every column load is issued unconditionally (columns past the group's end read column
    // g0 again and are masked), so a row's loads are in flight together
    const float *qc[kAppG];
#pragma unroll
    for (int k = 0; k < kAppG; ++k) qc[k] = F.BoC + (size_t)(g0 + (k < kn ? k : 0)) * F.ldx;
    double d[kAppG], gr[kAppG];
#pragma unroll
    for (int k = 0; k < kAppG; ++k) d[k] = gr[k] = 0.0;
    // a contiguous row range per block: its 16 columns span few pages
    const long long span = (F.n + (long long)nb * 256 - 1) / ((long long)nb * 256) * 256;
    const long long r1 = min(F.n, span * (blockIdx.x + 1));
    for (long long i = span * blockIdx.x + threadIdx.x; i < r1; i += 256) {
        const double vi = gld(F.v + i);
        const double qm = (double)gld(qn + i);  // column m - 1 (only used when grow)
        float q[kAppG];
#pragma unroll
        for (int k = 0; k < kAppG; ++k) q[k] = gld(qc[k] + i);
        if (first && blockIdx.z == 0) {
            const float b = __fdiv_rn((float)vi, rn);
            if (B) B[i * F.ldb + m] = b;
            BC[i] = b;
        }
#pragma unroll
        for (int k = 0; k < kAppG; ++k) {
            const double dk = fma((double)q[k], vi, d[k]);
            const double gk = fma(qm, (double)q[k], gr[k]);
            d[k] = k < kn ? dk : d[k];
            gr[k] = (grow && k < kn) ? gk : gr[k];
        }
    }
    // fixed-order block sums: wave butterflies, then the four waves in order
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < kAppG; ++k) {
        const double a = wave_sum(d[k]), b = wave_sum(gr[k]);
        if (lane == 0) {
            sh[w][k] = a;
            sh[w][kAppG + k] = b;
        }
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < kn) {
        const double a = ((sh[0][t] + sh[1][t]) + sh[2][t]) + sh[3][t];
        dpart[((size_t)f * F.ldb + g0 + t) * nb + blockIdx.x] = a;
    } else if (grow && t >= kAppG && t - kAppG < kn) {
        const int k = t - kAppG;
        const double b = ((sh[0][t] + sh[1][t]) + sh[2][t]) + sh[3][t];
        gpart[((size_t)f * F.ldb + g0 + k) * nb + blockIdx.x] = b;
    }
}

/* One block of 256 per fit: reduce the partials, extend G, run the MGS recurrence. */
__global__ __launch_bounds__(256) void k_app_solve(const FitDev *__restrict__ fits,
                                                   const AppendArg *__restrict__ args, int nb, int first,
                                                   const int *__restrict__ gate,
                                                   const double *__restrict__ dpart,
                                                   const double *__restrict__ gpart, const FitW *__restrict__ W,
                                                   double *__restrict__ coef, double *__restrict__ scal)
{
    __shared__ double sd[kMaxAppendCols], sg[kMaxAppendCols];
    const int f = blockIdx.x;
    if (!args[f].active) return;
    if (gate && !gate[f]) return;
    const FitDev F = fits[f];
    const int m = F.m;
    double *G = W[f].G;  // G[j * ldb + l], l < j
    // thread k sums the partials of d_k (and of the Gram row entry) in block order
    for (int k = threadIdx.x; k < m; k += 256) {
        const double *pd = dpart + ((size_t)f * F.ldb + k) * nb;
        const double a = ordered_sum(pd, nb, [](int b) { return (size_t)b; });
        sd[k] = a;
        if (first && k < m - 1) {
            const double *pg = gpart + ((size_t)f * F.ldb + k) * nb;
            const double g = ordered_sum(pg, nb, [](int b) { return (size_t)b; });
            G[(size_t)(m - 1) * F.ldb + k] = g;
            sg[k] = g;
        }
    }
    __syncthreads();
    if (threadIdx.x >= 64) return;
    // c_j = d_j - sum_{l<j} c_l G_jl on one wave, without barriers: lane l % 64 keeps c_l in
    // slot l / 64; every lane forms the same c_j and the same sum of c^2
    const int lane = threadIdx.x;
    double cown[kMaxAppendCols / 64];
#pragma unroll
    for (int u = 0; u < kMaxAppendCols / 64; ++u) cown[u] = 0.0;
    double proj = 0.0;
    for (int j = 0; j < m; ++j) {
        double part = 0.0;
#pragma unroll
        for (int u = 0; u < kMaxAppendCols / 64; ++u) {
            const int l = lane + 64 * u;
            if (l < j) part = fma(cown[u], (first && j == m - 1) ? sg[l] : G[(size_t)j * F.ldb + l], part);
        }
        const double c = sd[j] - wave_sum(part);
#pragma unroll
        for (int u = 0; u < kMaxAppendCols / 64; ++u)
            if (j == lane + 64 * u) cown[u] = c;
        proj = fma(c, c, proj);
    }
#pragma unroll
    for (int u = 0; u < kMaxAppendCols / 64; ++u) {
        const int j = lane + 64 * u;
        if (j < m) coef[(size_t)f * F.ldb + j] = cown[u];
    }
    if (lane == 0) scal[f * S_COUNT + (first ? S_PROJ : S_PROJ2)] = proj;
}

__global__ __launch_bounds__(256) void k_app_sub(const FitDev *__restrict__ fits,
                                                 const AppendArg *__restrict__ args, int nb,
                                                 const int *__restrict__ gate,
                                                 const double *__restrict__ coef, double *__restrict__ part)
{
    __shared__ double sh[256];
    __shared__ double sc[kMaxAppendCols];
    const int f = blockIdx.y;
    if (!args[f].active) return;
    if (gate && !gate[f]) return;
    const FitDev F = fits[f];
    const int m = F.m;
    for (int j = threadIdx.x; j < m; j += 256) sc[j] = coef[(size_t)f * F.ldb + j];
    __syncthreads();
    double a = 0.0;
    const long long span = (F.n + (long long)nb * 256 - 1) / ((long long)nb * 256) * 256;
    const long long r1 = min(F.n, span * (blockIdx.x + 1));
    for (long long i = span * blockIdx.x + threadIdx.x; i < r1; i += 256) {
        double vi = F.v[i];
        for (int j = 0; j < m; ++j) vi = fma(-sc[j], (double)F.BoC[(size_t)j * F.ldx + i], vi);
        F.v[i] = vi;
        a = fma(vi, vi, a);
    }
    a = block_sum_256(a, sh);
    if (threadIdx.x == 0) part[(size_t)f * nb + blockIdx.x] = a;
}

/* DGKS gate of the append (kernels.cc:376-385); the second call records |v|^2 after the retry. */
__global__ __launch_bounds__(256) void k_app_gate(const FitDev *__restrict__ fits,
                                                  const AppendArg *__restrict__ args, int nb,
                                                  const double *__restrict__ pa,
                                                  const double *__restrict__ pb,
                                                  int *__restrict__ gate, double *__restrict__ scal,
                                                  long long *__restrict__ dgks, int retry_done)
{
    __shared__ double sh[256];
    const int f = blockIdx.x;
    if (!args[f].active) return;
    const FitDev F = fits[f];
    const double *part = ((F.m + 1) % 2 == 0 ? pa : pb) + (size_t)f * nb;
    const double vv = sum_partials(part, nb, sh);
    if (threadIdx.x != 0) return;
    double *sc = scal + f * S_COUNT;
    if (!retry_done) {
        const int fire = (vv > F.tol && vv * kGate < sc[S_PROJ]) ? 1 : 0;
        gate[f] = fire;
        if (fire) atomicAdd((unsigned long long *)dgks, 1ull);
        sc[S_VV] = vv;
    } else if (gate[f]) {
        sc[S_VV] = vv;
    }
}

__global__ __launch_bounds__(256) void k_app_store(const FitDev *__restrict__ fits,
                                                   const AppendArg *__restrict__ args, int nb,
                                                   const FitW *__restrict__ W,
                                                   const double *__restrict__ scal,
                                                   double *__restrict__ part)
{
    __shared__ double sh[256];
    const int f = blockIdx.y;
    if (!args[f].active) return;
    const FitDev F = fits[f];
    const double w = sqrt(scal[f * S_COUNT + S_VV]);
    const bool ok = w * w > F.tol;
    float *Bo = W[f].Bo, *BoC = W[f].BoC + (size_t)F.m * F.ldx;
    double a = 0.0;
    if (ok) {
        for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < F.n; i += (long long)nb * 256) {
            const float b = (float)(F.v[i] / w);
            if (Bo) Bo[i * F.ldb + F.m] = b;
            BoC[i] = b;
            a = fma((double)b, (double)F.y[i], a);
        }
    }
    a = block_sum_256(a, sh);
    if (threadIdx.x == 0) part[(size_t)f * nb + blockIdx.x] = a;
}

__global__ __launch_bounds__(256) void k_app_final(const FitDev *__restrict__ fits,
                                                   const AppendArg *__restrict__ args, int nb,
                                                   const double *__restrict__ part,
                                                   const double *__restrict__ scal,
                                                   const FitW *__restrict__ W,
                                                   const int *__restrict__ gate,
                                                   AppendOut *__restrict__ out)
{
    __shared__ double sh[256];
    const int f = blockIdx.x;
    if (!args[f].active) return;
    const FitDev F = fits[f];
    const double yb = sum_partials(part + (size_t)f * nb, nb, sh);
    if (threadIdx.x != 0) return;
    const double w = sqrt(scal[f * S_COUNT + S_VV]);
    double *ybo = W[f].ybo;
    AppendOut o;
    o.fired = gate[f];
    o.ybo_m = yb;
    o.status = 0;
    o.mse = -1.0;
    if (w * w > F.tol) {
        double sq = yb * yb;
        for (int k = 0; k < F.m; ++k) sq = fma(ybo[k], ybo[k], sq);
        const double mse = (1.0 - sq) / (double)F.n;
        if (mse >= -F.tol) {
            o.status = 1;
            o.mse = mse > 0.0 ? mse : 0.0;
            ybo[F.m] = yb;
        }
    }
    out[f] = o;
}

/* ---------------------------------------------------------------------- */
/* Dense (p_x, m) outputs and the argmax (mars.py:342-343)                */
/* ---------------------------------------------------------------------- */
__global__ void k_dense_init(double *__restrict__ lin, double *__restrict__ hin,
                             double *__restrict__ cut, int *__restrict__ knot, size_t total)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    lin[i] = 0.0;
    hin[i] = 0.0;
    cut[i] = __builtin_nan("");
    knot[i] = -1;
}

__global__ void k_scatter(const int *__restrict__ cxc, const int *__restrict__ cpar, int P, int m,
                          const double *__restrict__ clin, const double *__restrict__ chin,
                          const int *__restrict__ cknot, const int *__restrict__ crow,
                          const float *__restrict__ X, size_t ldx, int hinge,
                          double *__restrict__ lin, double *__restrict__ hin,
                          double *__restrict__ cut, int *__restrict__ knot)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P) return;
    const size_t d = (size_t)cxc[c] * m + cpar[c];
    lin[d] = clin[c];
    if (hinge && cknot[c] >= 0) {
        hin[d] = chin[c];
        cut[d] = (double)X[(size_t)cxc[c] * ldx + crow[c]];
        knot[d] = cknot[c];
    }
}

/* First maximum in row-major order, like numpy.argmax over values >= 0.
 * One block per fit over its contiguous range [off[f], off[f] + len[f]);
 * `hin` may be null.  Index 0 is returned when every value is 0. */
__global__ __launch_bounds__(1024) void k_argmax(const double *__restrict__ lin,
                                                 const double *__restrict__ hin,
                                                 const int *__restrict__ knot,
                                                 const int *__restrict__ off,
                                                 const int *__restrict__ len,
                                                 BestOut *__restrict__ out)
{
    __shared__ double sv[2][1024];
    __shared__ int si[2][1024];
    const int f = blockIdx.x, t = threadIdx.x;
    const int base = off[f], total = len[f];
    const int chunk = (total + 1023) / 1024;
    const int b = t * chunk, e = min(total, b + chunk);
    double v1 = -1.0, v2 = -1.0;
    int i1 = -1, i2 = -1;
    for (int i = b; i < e; ++i) {
        if (lin[base + i] > v1) { v1 = lin[base + i]; i1 = i; }
        if (hin && hin[base + i] > v2) { v2 = hin[base + i]; i2 = i; }
    }
    sv[0][t] = v1; si[0][t] = i1;
    sv[1][t] = v2; si[1][t] = i2;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) {
            for (int q = 0; q < 2; ++q) {
                const double a = sv[q][t], bb = sv[q][t + s];
                const int ia = si[q][t], ib = si[q][t + s];
                if (bb > a || (bb == a && ib >= 0 && (ia < 0 || ib < ia))) {
                    sv[q][t] = bb;
                    si[q][t] = ib;
                }
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        BestOut o;
        o.v1 = sv[0][0] > 0.0 ? sv[0][0] : 0.0;
        o.c1 = sv[0][0] > 0.0 ? si[0][0] : 0;
        o.v2 = sv[1][0] > 0.0 ? sv[1][0] : 0.0;
        o.c2 = sv[1][0] > 0.0 ? si[1][0] : 0;
        o.knot2 = (knot && total) ? knot[base + o.c2] : -1;
        o.pad1 = 0;
        out[f] = o;
    }
}

__global__ void k_put_bcol(float *__restrict__ dst, size_t ldb, int col, const float *__restrict__ src,
                           size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i * ldb + col] = src[i];
}

__global__ void k_gather_cols(const float *__restrict__ rbuf, size_t ldr, size_t n,
                              float *__restrict__ out, size_t ldo)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < n) out[(size_t)j * ldo + i] = rbuf[i * ldr + j];
}

} // namespace

/* ---------------------------------------------------------------------- */
/* Host side                                                              */
/* ---------------------------------------------------------------------- */
void ctx_base_init(mars_hip_ctx *ctx, size_t max_terms)
{
    require(max_terms >= 1, "max_terms must be >= 1");
    MH_CHECK(hipGetDevice(&ctx->device));
    MH_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    MH_CHECK(hipEventCreate(&ctx->ev0));
    MH_CHECK(hipEventCreate(&ctx->ev1));
    MH_CHECK(hipEventCreateWithFlags(&ctx->ev_rb, hipEventDisableTiming));
    MH_CHECK(hipEventCreateWithFlags(&ctx->ev_app, hipEventDisableTiming));
    ctx->max_terms = max_terms;
    ctx->ldb = (max_terms + 3) / 4 * 4;
    ctx->dgks.ensure(1);
    ctx->fired_count.ensure(2);
    MH_CHECK(hipMemsetAsync(ctx->dgks.p, 0, sizeof(long long), ctx->stream));
    ctx->dstat.ensure(kDStat);
    MH_CHECK(hipMemsetAsync(ctx->dstat.p, 0, kDStat * sizeof(unsigned long long), ctx->stream));
    ctx->dredo.ensure(1);
    MH_CHECK(hipMemsetAsync(ctx->dredo.p, 0, sizeof(int), ctx->stream));
    ctx->scratch = scratch_ptr();  // this device's no-alloc scratch (resolved once per device)
}

void read_dstat(mars_hip_ctx *ctx, unsigned long long *out)
{
    std::fill(out, out + kDStat, 0ull);
    if (!ctx->dstat.p) return;
    MH_CHECK(hipMemcpyAsync(out, ctx->dstat.p, kDStat * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                            ctx->stream));
    MH_CHECK(hipStreamSynchronize(ctx->stream));
}

void fit_state_alloc(mars_hip_ctx *ctx, FitState &fs, size_t n, double tol)
{
    require(n > 0, "empty dataset");
    fs.n = n;
    fs.tol = tol;
    if (ctx->row_major) {
        fs.B.ensure(n * ctx->ldb);
        fs.Bo.ensure(n * ctx->ldb);
        MH_CHECK(hipMemsetAsync(fs.B.p, 0, n * ctx->ldb * sizeof(float), ctx->stream));
        MH_CHECK(hipMemsetAsync(fs.Bo.p, 0, n * ctx->ldb * sizeof(float), ctx->stream));
    }
    fs.ldx = (n + 31) / 32 * 32;
    fs.BoC.ensure(fs.ldx * ctx->ldb);
    MH_CHECK(hipMemsetAsync(fs.BoC.p, 0, fs.ldx * ctx->ldb * sizeof(float), ctx->stream));
    fs.BC.ensure(fs.ldx * ctx->ldb);
    MH_CHECK(hipMemsetAsync(fs.BC.p, 0, fs.ldx * ctx->ldb * sizeof(float), ctx->stream));
    fs.y.ensure(fs.ldx + 32);
    MH_CHECK(hipMemsetAsync(fs.y.p, 0, (fs.ldx + 32) * sizeof(float), ctx->stream));
    fs.ybo.ensure(ctx->max_terms);
    fs.v.ensure(n);
    // room for the residuals of 64 candidates (either layout), allocated with the fit so that a
    // forward pass does not begin with one large allocation per fit; grown when a step needs more.
    // A batch stores none there: its scoring keeps residuals in the context's bounded arena (the
    // host-built fallback grows the buffer when it runs).
    if (ctx->row_major) {
        fs.ldr = 64;
        fs.rbuf.ensure(fs.ldx * fs.ldr);
    }
    fs.G.ensure(ctx->ldb * ctx->ldb);
}

constexpr size_t kStageMax = 16u << 20;  // larger copies (one-time data uploads) go direct

/* Room for `bytes` in the pinned staging space; returns its offset.  A full space is recycled
 * after a synchronisation (every copy out of it has completed). */
size_t stage_take(mars_hip_ctx *ctx, size_t bytes)
{
    const size_t need = (bytes + 255) / 256 * 256;
    if (ctx->stage_off + need > ctx->stage_cap) {
        ctx_sync(ctx);
        if (need > ctx->stage_cap) {
            if (ctx->stage) MH_CHECK(hipHostFree(ctx->stage));
            ctx->stage = nullptr;
            ctx->stage_dev = nullptr;
            ctx->stage_cap = 0;
            const size_t cap = std::max<size_t>(need, 4u << 20);
            // coherent: the scatter kernel reads it directly, uncached
            MH_CHECK(hipHostMalloc((void **)&ctx->stage, cap, hipHostMallocCoherent));
            MH_CHECK(hipHostGetDevicePointer((void **)&ctx->stage_dev, ctx->stage, 0));
            ctx->stage_cap = cap;
        }
    }
    const size_t off = ctx->stage_off;
    ctx->stage_off += need;
    return off;
}

void h2d(mars_hip_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return;
    if (bytes > kStageMax) {
        MH_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        return;
    }
    const size_t off = stage_take(ctx, bytes);  // may (re)allocate ctx->stage: read it afterwards
    char *h = ctx->stage + off;
    std::memcpy(h, src, bytes);
    MH_CHECK(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, ctx->stream));
}

constexpr unsigned kScatterChunk = 4096;  // bytes per k_upload_scatter block (4 words per thread)

/* Block (x, y) copies bytes [x * kScatterChunk, ...) of upload y from the pinned staging space;
 * each thread issues its four host reads before any store, so their latencies overlap. */
__global__ __launch_bounds__(256) void k_upload_scatter(const char *__restrict__ stage, UpTable t)
{
    const UpDesc u = t.d[blockIdx.y];
    const unsigned long long b0 = (unsigned long long)blockIdx.x * kScatterChunk;
    if (b0 >= u.bytes) return;
    const unsigned long long w0 = b0 / 4 + threadIdx.x, w1 = min(u.bytes, b0 + kScatterChunk) / 4;
    const unsigned *src = (const unsigned *)(stage + u.off);
    unsigned *dst = (unsigned *)u.dst;
    unsigned v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (w0 + 256 * k < w1) v[k] = src[w0 + 256 * k];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (w0 + 256 * k < w1) dst[w0 + 256 * k] = v[k];
}

void h2d_defer(mars_hip_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return;
    if (bytes > kStageMax || bytes % 4 != 0) {
        h2d(ctx, dst, src, bytes);
        return;
    }
    // a second upload to a pending destination replaces the first (the scatter kernel's blocks
    // run in any order, so one flush never writes a destination twice)
    int same = -1;
    for (int i = 0; i < ctx->up_pend.n; ++i)
        if (ctx->up_pend.d[i].dst == dst) same = i;
    if (same >= 0 && ctx->up_pend.d[same].bytes != bytes) {
        h2d_flush(ctx);
        same = -1;
    }
    if (same < 0 && ctx->up_pend.n == kUpMax) h2d_flush(ctx);
    const size_t off = stage_take(ctx, bytes);  // may synchronise, which flushes what is pending
    std::memcpy(ctx->stage + off, src, bytes);
    if (same >= 0 && ctx->up_pend.n > 0)
        ctx->up_pend.d[same].off = off;
    else
        ctx->up_pend.d[ctx->up_pend.n++] = UpDesc{dst, off, bytes};
}

void h2d_flush(mars_hip_ctx *ctx)
{
    UpTable &t = ctx->up_pend;
    if (t.n == 0) return;
    unsigned long long most = 0;
    for (int i = 0; i < t.n; ++i) most = std::max(most, t.d[i].bytes);
    const UpTable send = t;
    t.n = 0;
    hipLaunchKernelGGL(k_upload_scatter, dim3((unsigned)((most + kScatterChunk - 1) / kScatterChunk), (unsigned)send.n),
                       dim3(256), 0, ctx->stream, (const char *)ctx->stage_dev, send);
    MH_LAUNCH_CHECK();
}

namespace {
thread_local std::vector<std::pair<void *, size_t>> g_retired_dev;
thread_local std::vector<void *> g_retired_host;
std::atomic<long long> g_dev_live{0}, g_dev_peak{0};
}

void dev_account(long long delta)
{
    const long long now = g_dev_live.fetch_add(delta) + delta;
    long long peak = g_dev_peak.load();
    while (now > peak && !g_dev_peak.compare_exchange_weak(peak, now)) {
    }
}

void retire_device(void *p, size_t bytes) { g_retired_dev.emplace_back(p, bytes); }
void retire_host(void *p) { g_retired_host.push_back(p); }

void free_retired()
{
    if (g_retired_dev.empty() && g_retired_host.empty()) return;
    // another context on this thread may still have work queued on a retired block
    MH_CHECK(hipDeviceSynchronize());
    for (auto &pb : g_retired_dev) {
        MH_CHECK(hipFree(pb.first));
        dev_account(-(long long)pb.second);
    }
    for (void *p : g_retired_host) MH_CHECK(hipHostFree(p));
    g_retired_dev.clear();
    g_retired_host.clear();
}

void run_graph(mars_hip_ctx *ctx, const std::vector<unsigned long long> &key, const std::function<void()> &launches)
{
    if (!ctx->use_graphs) {
        launches();
        return;
    }
    auto it = ctx->graphs.find(key);
    if (it == ctx->graphs.end()) {
        // capture: the launches' own timers are off (a graph replay runs none of their host code)
        const bool timing = ctx->timing;
        ctx->timing = false;
        hipGraph_t g = nullptr;
        MH_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        try {
            launches();
        } catch (...) {
            (void)hipStreamEndCapture(ctx->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            ctx->timing = timing;
            throw;
        }
        MH_CHECK(hipStreamEndCapture(ctx->stream, &g));
        ctx->timing = timing;
        hipGraphExec_t ex = nullptr;
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        MH_CHECK(e);
        it = ctx->graphs.emplace(key, ex).first;
        ctx->graph_captures += 1;
    }
    MH_CHECK(hipGraphLaunch(it->second, ctx->stream));
    ctx->graph_launches += 1;
}

void tuning_key(const mars_hip_ctx *ctx, std::vector<unsigned long long> &key)
{
    const mars_hip_tuning &t = ctx->tune;
    for (double v : {t.pass_items, t.retry_items, t.update_waves, t.sweep_waves})
        key.push_back(__builtin_bit_cast(unsigned long long, v));
    key.push_back((unsigned long long)t.sweep_min_rows);
    key.push_back((unsigned long long)t.append_rows);
    key.push_back((unsigned long long)t.narrow_avg);
}

void graph_key_buffers(mars_hip_ctx *ctx, std::vector<unsigned long long> &key)
{
    const void *p[] = {ctx->cand.d_fit.p, ctx->cand.d_xc.p, ctx->cand.d_par.p, ctx->rcand.d_fit.p, ctx->rcand.d_xc.p,
                       ctx->rcand.d_par.p, ctx->lin.p, ctx->ybx.p, ctx->cgate.p, ctx->clo.p, ctx->chi.p, ctx->cdone.p,
                       ctx->rsel.p, ctx->rkeep.p, ctx->ival.p, ctx->red.p, ctx->dstat.p, ctx->dredo.p, ctx->d_fits.p,
                       ctx->d_fits2.p, ctx->roff.p, ctx->rarena.p, ctx->T.p, ctx->T2.p, ctx->tnorm.p, ctx->ssum.p,
                       ctx->flag.p, ctx->rlin.p, ctx->rybx.p, ctx->fired_list.p, ctx->live_list.p, ctx->dl_int.p,
                       ctx->dl_u64.p, ctx->d_items.p, ctx->d_ritems.p, ctx->d_litems.p, ctx->d_cmap.p,
                       ctx->d_rcmap.p, ctx->d_lcmap.p, ctx->tpart.p, ctx->spart.p, ctx->dcand.p};
    for (const void *q : p) key.push_back((unsigned long long)(uintptr_t)q);
}

void ctx_sync(mars_hip_ctx *ctx)
{
    h2d_flush(ctx);
    MH_CHECK(hipStreamSynchronize(ctx->stream));
    free_retired();
    ctx->stage_off = 0;
}

void ctx_free(mars_hip_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    try {
        free_retired();
    } catch (...) {
    }
    if (ctx->stage) (void)hipHostFree(ctx->stage);
    ctx->stage = nullptr;
    for (auto &fs : ctx->fits) fs.release();
    DBuf<double> *db[] = {&ctx->tpart, &ctx->T, &ctx->T2, &ctx->tnorm, &ctx->spart, &ctx->ssum, &ctx->ybx,
                          &ctx->lin, &ctx->hin, &ctx->aux, &ctx->red, &ctx->vcol, &ctx->scal,
                          &ctx->dense, &ctx->hstate, &ctx->hbest, &ctx->rlin, &ctx->rybx,
                          &ctx->adot, &ctx->agram, &ctx->acoef};
    for (auto *b : db) b->release();
    DBuf<int> *ib[] = {&ctx->flag, &ctx->fired_list, &ctx->fired_count, &ctx->knot, &ctx->krow,
                       &ctx->ival, &ctx->d_cmap, &ctx->dknot, &ctx->hbest_i, &ctx->cand.d_fit,
                       &ctx->cand.d_xc, &ctx->cand.d_par, &ctx->rcand.d_fit, &ctx->rcand.d_xc,
                       &ctx->rcand.d_par};
    for (auto *b : ib) b->release();
    ctx->d_fits.release();
    ctx->d_items.release();
    ctx->d_ritems.release();
    ctx->d_rcmap.release();
    ctx->d_litems.release();
    ctx->d_lcmap.release();
    ctx->live_list.release();
    ctx->rsel.release();
    ctx->rkeep.release();
    ctx->dgks.release();
    ctx->aout.release();
    ctx->aarg.release();
    ctx->d_fitw.release();
    ctx->bout.release();
    ctx->rarena.release();
    ctx->roff.release();
    ctx->d_fits2.release();
    for (DBuf<int> *b : {&ctx->dl_int, &ctx->cgate, &ctx->cdone, &ctx->dredo, &ctx->bdeg, &ctx->dcand})
        b->release();
    for (DBuf<double> *b : {&ctx->clo, &ctx->chi, &ctx->hbest2, &ctx->hbest_ub, &ctx->hub}) b->release();
    ctx->rknot.release();
    ctx->rep.release();
    ctx->hr.release();
    ctx->rep_stage.release();
    for (DBuf<int> *b : {&ctx->vcand.d_fit, &ctx->vcand.d_xc, &ctx->vcand.d_par, &ctx->rep_todo}) b->release();
    ctx->dl_u64.release();
    ctx->dstat.release();
    ctx->bhas.release();
    for (auto &g : ctx->graphs) (void)hipGraphExecDestroy(g.second);
    ctx->graphs.clear();
    for (auto &r : ctx->pend) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : ctx->evpool) (void)hipEventDestroy(e);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->ev_rb) (void)hipEventDestroy(ctx->ev_rb);
    if (ctx->ev_app) (void)hipEventDestroy(ctx->ev_app);
    if (ctx->aout_pin) (void)hipHostFree(ctx->aout_pin);
    ctx->aout_pin = nullptr;
    ctx->sel_pin.release();
    ctx->best_pin.release();
    ctx->aarg_pin.release();
    if (ctx->rb_pin) (void)hipHostFree(ctx->rb_pin);
    ctx->rb_pin = nullptr;
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

/* MarsAlgo constructor on the device for one fit (marsalgo.cc:266-370). */
void fit_init(mars_hip_ctx *ctx, FitState &fs, const float *X, size_t ldX, const float *y, const float *w,
              size_t n, size_t px)
{
    require(X && y && w, "null input array");
    require(ldX >= n || px <= 1, "invalid leading dimension for X");
    require(px >= 1, "X has no columns");
    fit_state_alloc(ctx, fs, n, (n * 0.02) * 2.220446049250313e-16);
    fs.px = px;
    fs.X.ensure(fs.ldx * px + 32);
    fs.s.ensure(px);
    hipStream_t st = ctx->stream;
    MH_CHECK(hipMemsetAsync(fs.X.p, 0, (fs.ldx * px + 32) * sizeof(float), st));
    MH_CHECK(hipMemcpy2DAsync(fs.X.p, fs.ldx * sizeof(float), X, std::max(ldX, n) * sizeof(float), n * sizeof(float),
                              px, hipMemcpyHostToDevice, st));
    DBuf<float> yw;
    yw.ensure(2 * n);
    h2d(ctx, yw.p, y, n * sizeof(float));
    h2d(ctx, yw.p + n, w, n * sizeof(float));
    DBuf<double> tmp, part, scal, cpart;
    tmp.ensure(2 * n);
    part.ensure(4 * kInitBlocks);
    scal.ensure(S_COUNT);
    cpart.ensure(px * 64);
    MH_CHECK(hipMemsetAsync(scal.p, 0, S_COUNT * sizeof(double), st));
    double *yd = tmp.p, *sw = tmp.p + n, *p1 = part.p, *p2 = part.p + 2 * kInitBlocks;
    hipLaunchKernelGGL(k_init1, dim3(kInitBlocks), dim3(256), 0, st, yw.p, yw.p + n, n, yd, sw, p1);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_init2, dim3(kInitBlocks), dim3(256), 0, st, yd, sw, n, p1, fs.y.p, fs.B.p, fs.Bo.p,
                       ctx->ldb, fs.BoC.p, fs.BC.p, p2, scal.p);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_init3, dim3(kInitBlocks), dim3(256), 0, st, yd, n, p2, p1, scal.p);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_init4, dim3(1), dim3(256), 0, st, p1, n, scal.p, fs.ybo.p);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_colsq, dim3(64, px), dim3(256), 0, st, fs.X.p, fs.ldx, n, cpart.p);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_colscale, dim3((px + 255) / 256), dim3(256), 0, st, cpart.p, (int)px, n, fs.s.p,
                       scal.p);
    MH_LAUNCH_CHECK();
    double sc[S_COUNT];
    MH_CHECK(hipMemcpyAsync(sc, scal.p, sizeof(sc), hipMemcpyDeviceToHost, st));
    ctx_sync(ctx);
    yw.release();
    tmp.release();
    part.release();
    scal.release();
    cpart.release();
    require(sc[S_YN] > 0.0 && sc[S_WN] > 0.0, "target Y is all zero or NANs");
    require(sc[S_BADX] == 0.0, "not all columns in X are finite");
    fs.yvar = sc[S_YVAR];
    fs.ybo_h.assign(1, sc[S_YBO0]);
    fs.m = 1;
    fs.sorted.assign(px, 0);
}

/*
 * Append one basis function to every fit with args[f].active (device
 * counterpart of MarsAlgo::append, marsalgo.cc:760-875).  Validation happens
 * on the host with the reference's messages; out[f] receives the result.
 */
void fit_append_batch(mars_hip_ctx *ctx, std::vector<AppendArg> &args, std::vector<AppendOut> &out,
                      const std::function<void()> &before_readback, bool device_args, std::vector<BestOut> *best)
{
    PendingScope pending_scope(ctx);
    require(ctx->resident, "context was not created by mars_hip_fit_create");
    const size_t F = ctx->fits.size();
    require(args.size() == F, "one append argument per fit");
    size_t mmax = 0, nmax = 0;
    for (size_t f = 0; f < F; ++f) {
        if (!args[f].active) continue;
        const FitState &fs = ctx->fits[f];
        if (fs.m >= ctx->max_terms) throw Error(MARS_HIP_EFULL, "basis matrix is full");
        if (!device_args && (args[f].bcol < 0 || (size_t)args[f].bcol >= fs.m)) {
            char msg[80];
            snprintf(msg, sizeof(msg), "invalid basis column number: %d", args[f].bcol);
            throw Error(MARS_HIP_EINVAL, msg);
        }
        if (!device_args) {  // device arguments come from the fit's own candidates
            require(args[f].xcol >= 0 && (size_t)args[f].xcol < fs.px, "invalid X column index");
            require(args[f].type >= 0 && args[f].type <= 2, "invalid basis type");
        }
        require(fs.m < (size_t)kMaxAppendCols, "the device append supports up to 512 basis functions");
        mmax = std::max(mmax, fs.m);
        nmax = std::max(nmax, fs.n);
    }
    out.assign(F, AppendOut{-1.0, 0.0, 0, 0});
    if (mmax == 0) return;
    hipStream_t st = ctx->stream;
    // rows per block: 8192 for large batches (4096 and 16384 measured slower on 256 fits), 4096
    // when few fits must fill the device
    const size_t rpb = ctx->tune.append_rows > 0 ? (size_t)ctx->tune.append_rows : (F >= 64 ? 8192 : 4096);
    const int nb = (int)std::min<size_t>(256, std::max<size_t>(1, (nmax + rpb - 1) / rpb));
    Cands none;
    none.fit_c0.assign(F, 0);
    upload_fits(ctx, none);
    ctx->aarg.ensure(F);
    ctx->aout.ensure(F);
    ctx->red.ensure(3 * F * nb);
    ctx->scal.ensure(F * S_COUNT);
    ctx->flag.ensure(F);
    if (!device_args) h2d_defer(ctx, ctx->aarg.p, args.data(), F * sizeof(AppendArg));
    std::vector<FitW> fw(F);
    for (size_t f = 0; f < F; ++f) {
        FitState &fs = ctx->fits[f];
        fs.G.ensure(ctx->ldb * ctx->ldb);  // Gram rows of the stored columns, grown one per append
        fw[f] = FitW{fs.B.p, fs.Bo.p, fs.BoC.p, fs.BC.p, fs.ybo.p, fs.G.p};
    }
    ctx->d_fitw.ensure(F);
    h2d_defer(ctx, ctx->d_fitw.p, fw.data(), F * sizeof(FitW));
    double *pa = ctx->red.p, *pb = pa + F * nb, *pc = pb + F * nb;
    ctx->adot.ensure(F * ctx->ldb * nb);
    ctx->agram.ensure(F * ctx->ldb * nb);
    ctx->acoef.ensure(F * ctx->ldb);
    int *gate = ctx->flag.p;
    // algorithmic bytes of one sweep, each array once: the m f32 Bo columns read by the dots and
    // by the subtraction (8·n·m), the FP64 residual built, read by the dots, the subtraction and
    // the store and rewritten (40·n), the parent and X columns, the new B and Bo columns, the
    // newest Bo column and y (24·n).  The retry sweep, which runs only where the DGKS gate
    // fires, is not counted.
    double abytes = 0;
    for (size_t f = 0; f < F; ++f)
        if (args[f].active) abytes += 8.0 * (double)ctx->fits[f].n * ((double)ctx->fits[f].m + 8.0);
    {
        h2d_flush(ctx);  // fits, arguments and column pointers in one copy
        KTimer kt(ctx, K_APPEND, 0.0, abytes);
        const dim3 grid(nb, F);
        hipLaunchKernelGGL(k_app_build, grid, dim3(256), 0, st, ctx->d_fits.p, ctx->aarg.p, nb, pa, ctx->scal.p);
        MH_LAUNCH_CHECK();
        for (int pass = 0; pass < 2; ++pass) {
            // the retry sweep is launched without reading the gates back: its blocks return at
            // once for fits whose DGKS gate did not fire
            const int first = pass == 0 ? 1 : 0;
            const int *gp = pass == 0 ? nullptr : gate;
            const unsigned kg = (unsigned)((mmax + kAppG - 1) / kAppG);
            hipLaunchKernelGGL(k_app_dots, dim3(nb, F, kg), dim3(256), 0, st, ctx->d_fits.p, ctx->aarg.p, nb,
                               (const double *)pa, first, gp, ctx->adot.p, ctx->agram.p, ctx->d_fitw.p);
            MH_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_app_solve, dim3(F), dim3(256), 0, st, ctx->d_fits.p, ctx->aarg.p, nb, first, gp,
                               (const double *)ctx->adot.p, (const double *)ctx->agram.p, ctx->d_fitw.p,
                               ctx->acoef.p, ctx->scal.p);
            MH_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_app_sub, grid, dim3(256), 0, st, ctx->d_fits.p, ctx->aarg.p, nb, gp,
                               (const double *)ctx->acoef.p, pb);
            MH_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_app_gate, dim3(F), dim3(256), 0, st, ctx->d_fits.p, ctx->aarg.p, nb, pb, pb,
                               gate, ctx->scal.p, ctx->dgks.p, pass);
            MH_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_app_store, grid, dim3(256), 0, st, ctx->d_fits.p, ctx->aarg.p, nb, ctx->d_fitw.p,
                           ctx->scal.p, pc);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_app_final, dim3(F), dim3(256), 0, st, ctx->d_fits.p, ctx->aarg.p, nb, pc,
                           ctx->scal.p, ctx->d_fitw.p, gate, ctx->aout.p);
        MH_LAUNCH_CHECK();
    }
    // the results come back through pinned memory behind an event, so work the caller queues
    // after them (before_readback) runs on the device while the host reads them
    if (ctx->aout_cap < F) {
        if (ctx->aout_pin) retire_host(ctx->aout_pin);
        ctx->aout_pin = nullptr;
        ctx->aout_cap = 0;
        MH_CHECK(hipHostMalloc((void **)&ctx->aout_pin, F * sizeof(AppendOut), hipHostMallocDefault));
        ctx->aout_cap = F;
    }
    MH_CHECK(hipMemcpyAsync(ctx->aout_pin, ctx->aout.p, F * sizeof(AppendOut), hipMemcpyDeviceToHost, st));
    AppendArg *apin = nullptr;
    BestOut *bpin = nullptr;
    if (device_args) {
        apin = ctx->aarg_pin.get<AppendArg>(F);
        MH_CHECK(hipMemcpyAsync(apin, ctx->aarg.p, F * sizeof(AppendArg), hipMemcpyDeviceToHost, st));
        if (best) {
            bpin = ctx->best_pin.get<BestOut>(F);
            MH_CHECK(hipMemcpyAsync(bpin, ctx->bout.p, F * sizeof(BestOut), hipMemcpyDeviceToHost, st));
        }
    }
    MH_CHECK(hipEventRecord(ctx->ev_app, st));
    if (before_readback) before_readback();
    MH_CHECK(hipEventSynchronize(ctx->ev_app));
    std::copy(ctx->aout_pin, ctx->aout_pin + F, out.begin());
    if (device_args) {
        args.assign(apin, apin + F);
        if (best) best->assign(bpin, bpin + F);
    }
    resolve_timing(ctx);
    for (size_t f = 0; f < F; ++f) {
        if (!args[f].active) continue;
        if (out[f].fired) ctx->dgks_host += 1;
        if (out[f].status == 1) {
            ctx->fits[f].ybo_h.push_back(out[f].ybo_m);
            ctx->fits[f].m += 1;
        } else {
            out[f].mse = -1.0;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Append without allocation (mars_hip_fit_append_noalloc)                  */
/* ---------------------------------------------------------------------- */

/* The append's scratch for callers that cannot allocate: per-launch state (the fit's device
 * view, column pointers, argument, result, DGKS gate, scalars) and the partial sums of the
 * append's reductions.  Each GPU holds its own copy for the life of the process.  Host code
 * never uses this name as an address: it uses the per-device pointer that hipGetSymbolAddress
 * returns (scratch_ptr). */
constexpr size_t kNoAllocBytes = size_t(4) << 20;
static __device__ __align__(256) double scratch[kNoAllocBytes / sizeof(double)];

namespace {

std::mutex g_scratch_mu;
std::vector<char *> g_scratch_dev;  // per device: the device address of `scratch` (null: not resolved)

/* Byte offsets of the no-alloc append's pieces inside `scratch` (single fit). */
constexpr size_t kNaFit = 0, kNaFitW = 256, kNaArg = 512, kNaOut = 768, kNaGate = 1024, kNaScal = 1280,
                 kNaRed = 2048, kNaCoef = 8192, kNaDots = 12288;
static_assert(sizeof(FitDev) <= kNaFitW - kNaFit && sizeof(FitW) <= kNaArg - kNaFitW &&
                  sizeof(AppendArg) <= kNaOut - kNaArg && sizeof(AppendOut) <= kNaGate - kNaOut &&
                  S_COUNT * sizeof(double) <= kNaRed - kNaScal && 3 * 256 * sizeof(double) <= kNaCoef - kNaRed &&
                  kMaxAppendCols * sizeof(double) <= kNaDots - kNaCoef,
              "no-alloc scratch layout");

__global__ void k_noalloc_args(FitDev fd, FitW fw, AppendArg a, FitDev *__restrict__ dfd, FitW *__restrict__ dfw,
                               AppendArg *__restrict__ darg)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *dfd = fd;
        *dfw = fw;
        *darg = a;
    }
}

} // namespace

/* Resolve the current device's address of `scratch` (once per device; the device must be
 * current).  Called when a context is created. */
char *scratch_ptr()
{
    int dev = 0;
    MH_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_dev.size() <= dev) g_scratch_dev.resize(dev + 1, nullptr);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        MH_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = static_cast<char *>(p);
    }
    return g_scratch_dev[dev];
}

/* The append of fit_append_batch for a single-fit context, with every scratch buffer inside
 * `scratch` and the arguments passed by value: no device or pinned allocation, no staged upload.
 * The result comes back by a copy to the caller's memory and a stream synchronisation. */
void fit_append_noalloc(mars_hip_ctx *ctx, const AppendArg &a, double *mse)
{
    require(ctx->resident && ctx->fits.size() == 1, "mars_hip_fit_append_noalloc needs a single-fit resident context");
    FitState &fs = ctx->fits[0];
    if (fs.m >= ctx->max_terms) throw Error(MARS_HIP_EFULL, "basis matrix is full");
    if (a.bcol < 0 || (size_t)a.bcol >= fs.m) {
        char msg[80];
        snprintf(msg, sizeof(msg), "invalid basis column number: %d", a.bcol);
        throw Error(MARS_HIP_EINVAL, msg);
    }
    require(a.xcol >= 0 && (size_t)a.xcol < fs.px, "invalid X column index");
    require(fs.m < (size_t)kMaxAppendCols, "the device append supports up to 512 basis functions");
    const int nb = (int)std::min<size_t>(256, std::max<size_t>(1, (fs.n + 4095) / 4096));  // as fit_append_batch
    const size_t need = kNaDots + 2 * ctx->ldb * (size_t)nb * sizeof(double);
    if (need > kNoAllocBytes) {
        char msg[160];
        snprintf(msg, sizeof(msg),
                 "mars_hip_fit_append_noalloc: needs %zu scratch bytes for max_terms %zu and %zu rows, above its "
                 "%zu-byte maximum",
                 need, ctx->max_terms, fs.n, kNoAllocBytes);
        throw Error(MARS_HIP_EINVAL, msg);
    }
    char *base = ctx->scratch;
    require(base != nullptr, "no-alloc scratch was not resolved for this context's device");
    FitDev *dfit = reinterpret_cast<FitDev *>(base + kNaFit);
    FitW *dfw = reinterpret_cast<FitW *>(base + kNaFitW);
    AppendArg *darg = reinterpret_cast<AppendArg *>(base + kNaArg);
    AppendOut *dout = reinterpret_cast<AppendOut *>(base + kNaOut);
    int *gate = reinterpret_cast<int *>(base + kNaGate);
    double *scal = reinterpret_cast<double *>(base + kNaScal);
    double *pa = reinterpret_cast<double *>(base + kNaRed), *pb = pa + nb, *pc = pb + nb;
    double *coef = reinterpret_cast<double *>(base + kNaCoef);
    double *adot = reinterpret_cast<double *>(base + kNaDots);
    double *agram = adot + ctx->ldb * (size_t)nb;

    FitDev d{};
    d.X = fs.X.p;
    d.s = fs.s.p;
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
    d.ldx = (long long)fs.ldx;
    d.ldb = (long long)ctx->ldb;
    d.rsr = 1;
    d.rsc = (long long)fs.ldx;
    d.m = (int)fs.m;
    d.px = (int)fs.px;
    d.act = 1;
    d.tol = fs.tol;
    const FitW w{fs.B.p, fs.Bo.p, fs.BoC.p, fs.BC.p, fs.ybo.p, fs.G.p};
    hipStream_t st = ctx->stream;
    // every scratch entry a kernel reads is written earlier in this same sequence, so the contents
    // a previous call left behind are never read
    hipLaunchKernelGGL(k_noalloc_args, dim3(1), dim3(64), 0, st, d, w, a, dfit, dfw, darg);
    MH_LAUNCH_CHECK();
    const dim3 grid(nb, 1);
    hipLaunchKernelGGL(k_app_build, grid, dim3(256), 0, st, (const FitDev *)dfit, (const AppendArg *)darg, nb, pa,
                       scal);
    MH_LAUNCH_CHECK();
    for (int pass = 0; pass < 2; ++pass) {
        const int first = pass == 0 ? 1 : 0;
        const int *gp = pass == 0 ? nullptr : gate;
        const unsigned kg = (unsigned)((fs.m + kAppG - 1) / kAppG);
        hipLaunchKernelGGL(k_app_dots, dim3(nb, 1, kg), dim3(256), 0, st, (const FitDev *)dfit,
                           (const AppendArg *)darg, nb, (const double *)pa, first, gp, adot, agram,
                           (const FitW *)dfw);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_app_solve, dim3(1), dim3(256), 0, st, (const FitDev *)dfit, (const AppendArg *)darg, nb,
                           first, gp, (const double *)adot, (const double *)agram, (const FitW *)dfw, coef, scal);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_app_sub, grid, dim3(256), 0, st, (const FitDev *)dfit, (const AppendArg *)darg, nb, gp,
                           (const double *)coef, pb);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_app_gate, dim3(1), dim3(256), 0, st, (const FitDev *)dfit, (const AppendArg *)darg, nb,
                           (const double *)pb, (const double *)pb, gate, scal, ctx->dgks.p, pass);
        MH_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_app_store, grid, dim3(256), 0, st, (const FitDev *)dfit, (const AppendArg *)darg, nb,
                       (const FitW *)dfw, (const double *)scal, pc);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_app_final, dim3(1), dim3(256), 0, st, (const FitDev *)dfit, (const AppendArg *)darg, nb,
                       (const double *)pc, (const double *)scal, (const FitW *)dfw, (const int *)gate, dout);
    MH_LAUNCH_CHECK();
    AppendOut o{};
    MH_CHECK(hipMemcpyAsync(&o, dout, sizeof(AppendOut), hipMemcpyDeviceToHost, st));
    MH_CHECK(hipStreamSynchronize(st));
    if (o.fired) ctx->dgks_host += 1;
    if (o.status == 1) {
        fs.ybo_h.push_back(o.ybo_m);
        fs.m += 1;
        *mse = o.mse;
    } else {
        *mse = -1.0;
    }
}

double host_dsse(const FitState &fs)
{
    double s = 0.0;
    for (double v : fs.ybo_h) s = std::fma(v, v, s);
    return s;
}

/* Per-fit best linear candidate (first maximum in row-major (xcol, bcol) order). */
void argmax_launch(mars_hip_ctx *ctx, const Cands &cs, const double *lin)
{
    PendingScope pending_scope(ctx);
    const size_t F = ctx->fits.size();
    ctx->ival.ensure(2 * F);
    ctx->bout.ensure(F);
    std::vector<int> ol(2 * F);
    for (size_t f = 0; f < F; ++f) {
        ol[f] = cs.fit_c0[f];
        ol[F + f] = cs.fit_P[f];
    }
    h2d_defer(ctx, ctx->ival.p, ol.data(), 2 * F * sizeof(int));
    h2d_flush(ctx);
    {
        KTimer kt(ctx, K_ARGMAX, 0.0, 8.0 * cs.fit.size());
        hipLaunchKernelGGL(k_argmax, dim3(F), dim3(1024), 0, ctx->stream, lin, nullptr, nullptr, ctx->ival.p,
                           ctx->ival.p + F, ctx->bout.p);
        MH_LAUNCH_CHECK();
    }
}

/* ---------------------------------------------------------------------- */
/* Precision form with hinges: exact rescoring of the near winners         */
/* ---------------------------------------------------------------------- */
constexpr double kHingeWindow = 2e-3;  // relative window around the approximate best linear score
// relative error of the approximate normalised candidate columns taken by the screen: the FP32
// projection's T error, summed over 64-row spans whose rounding errors add like a random walk,
// is about u sqrt(64 m) |Bx| against |r| >= |Bx| / sqrt(10) for an ungated candidate: 1.5e-6 at
// m = 64 (taken as 3e-6); the BF16 split's products err by 2^-17 each: 3e-5.  The worst case
// (every rounding error in one direction) is 20x wider; the rescoring rounds that follow catch
// a candidate whose approximation erred beyond its bound.
constexpr double kEpsF32 = 3e-6, kEpsBf16 = 3e-5;

/* mx[0] = largest linear, mx[1] = largest hinge score (one block). */
__global__ __launch_bounds__(1024) void k_two_max(const double *__restrict__ lin, const double *__restrict__ hin,
                                                  int P, double *__restrict__ mx)
{
    __shared__ double s1[1024], s2[1024];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int c = t; c < P; c += 1024) {
        a = fmax(a, lin[c]);
        b = fmax(b, hin[c]);
    }
    s1[t] = a;
    s2[t] = b;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) {
            s1[t] = fmax(s1[t], s1[t + s]);
            s2[t] = fmax(s2[t], s2[t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        mx[0] = s1[0];
        mx[1] = s2[0];
    }
}

/* take[c] = 1 when the candidate's approximate linear score is within the window of the best one,
 * the upper bound of its hinge scores (hub) reaches the window of the best approximate hinge
 * score, its DGKS gate fired (flag bit 0), or its residual energy is within a factor 1000 of the
 * degeneracy threshold (the approximation may sit on either side of it). */
__global__ void k_hsel(const double *__restrict__ lin, const double *__restrict__ hub, const int *__restrict__ flag,
                       const double *__restrict__ ssum, double tol, int P, const double *__restrict__ mx, double w,
                       int *__restrict__ take)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P) return;
    take[c] = ((flag[c] & 1) || ssum[c] <= 1e3 * tol || (mx[0] > 0 && lin[c] >= mx[0] * (1.0 - w)) ||
               (mx[1] > 0 && hub[c] >= mx[1] * (1.0 - w)))
                  ? 1
                  : 0;
}

/* Exact values of the rescored candidates into the full arrays; they are marked done. */
__global__ void k_hput(const int *__restrict__ idx, int R, const double *__restrict__ rl, const double *__restrict__ rh,
                       const int *__restrict__ rk, const int *__restrict__ rr, double *__restrict__ lin,
                       double *__restrict__ hin, int *__restrict__ knot, int *__restrict__ krow, int *__restrict__ done)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    const int c = idx[i];
    lin[c] = rl[i];
    hin[c] = rh[i];
    knot[c] = rk[i];
    krow[c] = rr[i];
    done[c] = 1;
}

/* With approximate scores within a relative w / 2 of the exact ones (the window's premise), a
 * candidate left approximate, below (1 - w) times the best approximate score, stays below the
 * exact best by a factor 1 + w / 2.  take[c] = 1 for every approximate candidate that does not
 * (it may beat the exact best: the approximation erred by more than w / 2 somewhere, as a noise
 * column can); *more = their count.  One block. */
__global__ __launch_bounds__(1024) void k_hmore(const double *__restrict__ lin, const double *__restrict__ hin,
                                                const double *__restrict__ hub, const int *__restrict__ done, int P,
                                                double w, int *__restrict__ take, int *__restrict__ more)
{
    __shared__ double s1[1024], s2[1024];
    __shared__ int sv[1024];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int c = t; c < P; c += 1024)
        if (done[c]) {
            a = fmax(a, lin[c]);
            b = fmax(b, hin[c]);
        }
    s1[t] = a;
    s2[t] = b;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) {
            s1[t] = fmax(s1[t], s1[t + s]);
            s2[t] = fmax(s2[t], s2[t + s]);
        }
        __syncthreads();
    }
    const double e1 = s1[0], e2 = s2[0];
    int v = 0;
    for (int c = t; c < P; c += 1024) {
        const bool m = !done[c] && ((lin[c] > 0 && lin[c] * (1.0 + 0.5 * w) >= e1) ||
                                    (hub[c] > 0 && hub[c] >= e2));
        take[c] = m ? 1 : 0;
        v += m ? 1 : 0;
    }
    sv[t] = v;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) sv[t] += sv[t + s];
        __syncthreads();
    }
    if (t == 0) *more = sv[0];
}

/*
 * Precision form of a hinge eval: the candidates' first projection runs on FP32 MFMA (DGKS retry,
 * norms and the FP64 hinge sweep as in the plain form), then every candidate near the best
 * approximate linear or hinge score, every one whose DGKS gate fired and every near-degenerate
 * one is scored again by the plain form (FP64 projection and sweep), and its exact values replace
 * the approximate ones.  A hinge candidate's screen is the upper bound of its knots' scores under
 * a relative error eps of its covariates (the sweep computes it; an ill-conditioned knot, whose
 * denominator k - |f|^2 cancels, gets a wide bound).  Rounds of rescoring follow while an
 * approximate candidate could still beat the exact best (k_hmore), so the step's winner and its
 * scores are plain-form values.
 */
void hinge_eval_mixed(mars_hip_ctx *ctx, Cands &cs, int min_span, int end_span)
{
    hipStream_t st = ctx->stream;
    const int P = (int)cs.fit.size();
    const double tol = ctx->fits[0].tol;
    ctx->hub.ensure(std::max(P, 1));
    ctx->hin.ensure(std::max(P, 1));
    ctx->knot.ensure(std::max(P, 1));
    ctx->krow.ensure(std::max(P, 1));
    hinge_score(ctx, cs, ctx->lin.p, ctx->ybx.p, ctx->hin.p, ctx->knot.p, ctx->krow.p, min_span, end_span, true,
                ctx->split_bf16 ? kEpsBf16 : kEpsF32, ctx->hub.p);
    KTimer kt(ctx, K_HRECHECK, 0.0, 0.0);
    ctx->red.ensure(2);
    ctx->rsel.ensure(P);
    ctx->cdone.ensure(P);
    ctx->dcand.ensure(2);
    hipLaunchKernelGGL(k_two_max, dim3(1), dim3(1024), 0, st, (const double *)ctx->lin.p, (const double *)ctx->hin.p, P,
                       ctx->red.p);
    MH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_hsel, dim3((P + 255) / 256), dim3(256), 0, st, (const double *)ctx->lin.p,
                       (const double *)ctx->hub.p, (const int *)ctx->flag.p, (const double *)ctx->ssum.p, tol, P,
                       (const double *)ctx->red.p, kHingeWindow, ctx->rsel.p);
    MH_LAUNCH_CHECK();
    std::vector<int> zeros(P, 0);
    h2d(ctx, ctx->cdone.p, zeros.data(), P * sizeof(int));
    std::vector<int> take(P);
    int more = 0;
    for (int round = 0;; ++round) {
        MH_CHECK(hipMemcpyAsync(take.data(), ctx->rsel.p, P * sizeof(int), hipMemcpyDeviceToHost, st));
        ctx_sync(ctx);
        std::vector<int> idx;
        Cands &rc = ctx->rcand;
        rc.fit.clear();
        rc.xc.clear();
        rc.par.clear();
        for (int c = 0; c < P; ++c)
            if (take[c]) {
                idx.push_back(c);
                rc.fit.push_back(0);
                rc.xc.push_back(cs.xc[c]);
                rc.par.push_back(cs.par[c]);
            }
        const int R = (int)idx.size();
        if (R == 0) break;
        if (round == 4) {  // (has not happened: the approximation errs beyond w / 2 only for noise columns)
            ctx->window_misses += R;
            break;
        }
        ctx->rescored += R;
        ctx->rlin.ensure(R);
        ctx->rybx.ensure(R);
        ctx->hbest2.ensure(R);
        ctx->rknot.ensure(2 * (size_t)R + 1);
        ctx->live_list.ensure(R);
        hinge_score(ctx, rc, ctx->rlin.p, ctx->rybx.p, ctx->hbest2.p, ctx->rknot.p, ctx->rknot.p + R, min_span,
                    end_span);
        h2d(ctx, ctx->live_list.p, idx.data(), R * sizeof(int));
        hipLaunchKernelGGL(k_hput, dim3((R + 255) / 256), dim3(256), 0, st, (const int *)ctx->live_list.p, R,
                           (const double *)ctx->rlin.p, (const double *)ctx->hbest2.p, (const int *)ctx->rknot.p,
                           (const int *)ctx->rknot.p + R, ctx->lin.p, ctx->hin.p, ctx->knot.p, ctx->krow.p,
                           ctx->cdone.p);
        MH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_hmore, dim3(1), dim3(1024), 0, st, (const double *)ctx->lin.p, (const double *)ctx->hin.p,
                           (const double *)ctx->hub.p, (const int *)ctx->cdone.p, P, kHingeWindow, ctx->rsel.p,
                           ctx->dcand.p);
        MH_LAUNCH_CHECK();
        MH_CHECK(hipMemcpyAsync(&more, ctx->dcand.p, sizeof(int), hipMemcpyDeviceToHost, st));
        ctx_sync(ctx);
        if (more == 0) break;
        ctx->hinge_rounds += 1;
    }
}

bool valid_form(int form)
{
    return form == MARS_HIP_FORM_PLAIN || form == MARS_HIP_FORM_MIXED || form == MARS_HIP_FORM_MIXED_BF16;
}

void set_form(mars_hip_ctx *ctx, int form)
{
    ctx->split_bf16 = form == MARS_HIP_FORM_MIXED_BF16;
    ctx->form = form == MARS_HIP_FORM_PLAIN ? MARS_HIP_FORM_PLAIN : MARS_HIP_FORM_MIXED;
}

void argmax_launch_ranges(mars_hip_ctx *ctx, const double *lin, size_t bound)
{
    const size_t F = ctx->fits.size();
    ctx->bout.ensure(F);
    KTimer kt(ctx, K_ARGMAX, 0.0, 8.0 * (double)bound);
    hipLaunchKernelGGL(k_argmax, dim3(F), dim3(1024), 0, ctx->stream, lin, nullptr, nullptr, ctx->ival.p,
                       ctx->ival.p + F, ctx->bout.p);
    MH_LAUNCH_CHECK();
}

void fit_eval(mars_hip_ctx *ctx, const unsigned char *mask, size_t rows, size_t stride, int min_span,
              int end_span, int linear_only, double *lin_out, double *hin_out, double *cut_out,
              mars_hip_best *best)
{
    PendingScope pending_scope(ctx);
    require(ctx->resident, "context was not created by mars_hip_fit_create");
    require(ctx->fits.size() == 1, "use mars_hip_batch_fit on batch contexts");
    FitState &fs = ctx->fits[0];
    require(mask != nullptr, "null mask");
    require(rows == fs.px, "invalid dimension for mask");
    require(stride >= fs.m, "invalid mask row stride");
    require(min_span >= 1, "min_span must be >= 1");
    require(end_span >= 0 && (size_t)end_span < fs.n, "end_span must be in [0, n)");

    const size_t m = fs.m, total = rows * m;
    hipStream_t st = ctx->stream;

    Cands &cs = ctx->cand;
    cs.fit.clear();
    cs.xc.clear();
    cs.par.clear();
    for (size_t i = 0; i < rows; ++i)
        for (size_t j = 0; j < m; ++j)
            if (mask[i * stride + j]) {
                cs.fit.push_back(0);
                cs.xc.push_back((int)i);
                cs.par.push_back((int)j);
            }
    const int P = (int)cs.fit.size();
    index_cands(ctx, cs);

    ctx->dense.ensure(3 * std::max<size_t>(total, 1));
    ctx->dknot.ensure(std::max<size_t>(total, 1));
    double *dlin = ctx->dense.p, *dhin = dlin + total, *dcut = dhin + total;
    int *dknot = ctx->dknot.p;
    if (total) {
        hipLaunchKernelGGL(k_dense_init, dim3((total + 255) / 256), dim3(256), 0, st, dlin, dhin, dcut, dknot,
                           total);
        MH_LAUNCH_CHECK();
    }
    MH_CHECK(hipEventRecord(ctx->ev0, st));
    ctx->scored += P;
    if (P > 0) {
        ctx->lin.ensure(P);
        ctx->ybx.ensure(P);
        if (ctx->form == MARS_HIP_FORM_MIXED && linear_only) {
            mixed_score(ctx, true);  // one fit: the eval synchronises anyway
        } else if (ctx->form == MARS_HIP_FORM_MIXED) {
            require(ctx->row_major, "hinge sweeps need the row-major basis copies");
            hinge_eval_mixed(ctx, cs, min_span, end_span);
        } else {
            require(linear_only || ctx->row_major, "hinge sweeps need the row-major basis copies");
            if (linear_only) {
                plain_score(ctx, cs, false, ctx->lin.p, ctx->ybx.p);
            } else {
                ctx->hin.ensure(P);
                ctx->knot.ensure(P);
                ctx->krow.ensure(P);
                hinge_score(ctx, cs, ctx->lin.p, ctx->ybx.p, ctx->hin.p, ctx->knot.p, ctx->krow.p, min_span,
                            end_span);
            }
        }
        upload_cands(ctx, cs);
        h2d_flush(ctx);
        hipLaunchKernelGGL(k_scatter, dim3((P + 255) / 256), dim3(256), 0, st, cs.d_xc.p, cs.d_par.p, P, (int)m,
                           ctx->lin.p, ctx->hin.p, ctx->knot.p, ctx->krow.p, fs.X.p, fs.ldx, linear_only ? 0 : 1,
                           dlin, dhin, dcut, dknot);
        MH_LAUNCH_CHECK();
    }
    ctx->ival.ensure(2);
    ctx->bout.ensure(1);
    const int ol[2] = {0, (int)total};
    h2d(ctx, ctx->ival.p, ol, sizeof(ol));
    hipLaunchKernelGGL(k_argmax, dim3(1), dim3(1024), 0, st, dlin, dhin, dknot, ctx->ival.p, ctx->ival.p + 1,
                       ctx->bout.p);
    MH_LAUNCH_CHECK();
    MH_CHECK(hipEventRecord(ctx->ev1, st));
    BestOut bo;
    MH_CHECK(hipMemcpyAsync(&bo, ctx->bout.p, sizeof(bo), hipMemcpyDeviceToHost, st));
    if (lin_out && total) MH_CHECK(hipMemcpyAsync(lin_out, dlin, total * 8, hipMemcpyDeviceToHost, st));
    if (hin_out && total) MH_CHECK(hipMemcpyAsync(hin_out, dhin, total * 8, hipMemcpyDeviceToHost, st));
    if (cut_out && total) MH_CHECK(hipMemcpyAsync(cut_out, dcut, total * 8, hipMemcpyDeviceToHost, st));
    ctx_sync(ctx);
    double cut2 = NAN;
    if (total) {
        require(bo.c2 >= 0 && (size_t)bo.c2 < total, "internal: argmax index out of range");
        MH_CHECK(hipMemcpy(&cut2, dcut + bo.c2, 8, hipMemcpyDeviceToHost));
    }
    float ms = 0.f;
    MH_CHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->eval_ms += ms;
    resolve_timing(ctx);
    if (best) {
        best->sse0 = host_dsse(fs);
        best->sse1 = bo.v1;
        best->xcol1 = m ? bo.c1 / (int)m : 0;
        best->bcol1 = m ? bo.c1 % (int)m : 0;
        best->sse2 = bo.v2;
        best->xcol2 = m ? bo.c2 / (int)m : 0;
        best->bcol2 = m ? bo.c2 % (int)m : 0;
        best->knot2 = bo.knot2;
        best->cut2 = cut2;
    }
}

/* ---------------------------------------------------------------------- */
/* Mirror mode (the reference's mars::cuda boundary)                       */
/* ---------------------------------------------------------------------- */
void mirror_sync_basis(mars_hip_ctx *ctx, size_t m, const float *B, size_t ldB, const float *Bo, size_t ldBo)
{
    require(!ctx->resident, "resident contexts own their basis");
    require(m <= ctx->max_terms, "m exceeds max_terms");
    require(m == 0 || (B && Bo), "null basis pointer");
    FitState &fs = ctx->fits[0];
    const size_t n = fs.n;
    hipStream_t st = ctx->stream;
    if (m > fs.m_synced) {
        DBuf<float> col;
        col.ensure(n);
        std::vector<float> h(n);
        for (size_t j = fs.m_synced; j < m; ++j) {
            h2d(ctx, col.p, B + j * ldB, n * sizeof(float));
            hipLaunchKernelGGL(k_put_bcol, dim3((n + 255) / 256), dim3(256), 0, st, fs.B.p, ctx->ldb, (int)j, col.p,
                               n);
            MH_LAUNCH_CHECK();
            MH_CHECK(hipMemcpyAsync(fs.BC.p + j * fs.ldx, col.p, n * sizeof(float), hipMemcpyDeviceToDevice, st));
            ctx_sync(ctx);
            for (size_t i = 0; i < n; ++i) h[i] = Bo[i * ldBo + j];
            h2d(ctx, col.p, h.data(), n * sizeof(float));
            hipLaunchKernelGGL(k_put_bcol, dim3((n + 255) / 256), dim3(256), 0, st, fs.Bo.p, ctx->ldb, (int)j, col.p,
                               n);
            MH_LAUNCH_CHECK();
            MH_CHECK(hipMemcpyAsync(fs.BoC.p + j * fs.ldx, col.p, n * sizeof(float), hipMemcpyDeviceToDevice, st));
            ctx_sync(ctx);
        }
        col.release();
        fs.m_synced = m;
    }
    fs.m = m;
}

void mirror_set_target(mars_hip_ctx *ctx, const float *y)
{
    require(y != nullptr, "null target");
    FitState &fs = ctx->fits[0];
    if (fs.y_key == y) return;
    h2d(ctx, fs.y.p, y, fs.n * sizeof(float));
    ctx_sync(ctx);
    fs.y_key = y;
}

void mirror_sync_xcols(mars_hip_ctx *ctx, size_t px, const float *X, size_t ldX, const float *s, const int *cols,
                       size_t ncols)
{
    require(X && s, "null X or scale pointer");
    FitState &fs = ctx->fits[0];
    const size_t n = fs.n;
    if (fs.px == 0) {
        fs.px = px;
        fs.X.ensure(fs.ldx * px + 32);
        MH_CHECK(hipMemsetAsync(fs.X.p, 0, (fs.ldx * px + 32) * sizeof(float), ctx->stream));
        fs.s.ensure(px);
        fs.xres.assign(px, 0);
    }
    require(px == fs.px, "p_x changed between calls");
    require(ncols == 0 || cols != nullptr, "null column list");
    h2d(ctx, fs.s.p, s, px * sizeof(float));
    for (size_t k = 0; k < ncols; ++k) {
        const int c = cols[k];
        require(c >= 0 && (size_t)c < px, "X column out of range");
        if (fs.xres[c]) continue;
        h2d(ctx, fs.X.p + (size_t)c * fs.ldx, X + (size_t)c * ldX, n * sizeof(float));
        fs.xres[c] = 1;
    }
    ctx_sync(ctx);
}

void mirror_orthonormalize(mars_hip_ctx *ctx, size_t m, size_t p, const float *x, const int *mask, float *Bx,
                           size_t ldBx, double *ybx, long *fired)
{
    require(!ctx->resident, "use mars_hip_fit_eval on resident contexts");
    FitState &fs = ctx->fits[0];
    require(m == fs.m && m <= fs.m_synced && m >= 1, "basis not synced up to m");
    require(p <= ctx->max_terms, "p exceeds max_terms");
    require(x && mask, "null x or mask");
    require(Bx == nullptr || ldBx >= fs.n, "invalid leading dimension for Bx");
    if (p == 0) return;
    const size_t n = fs.n;
    hipStream_t st = ctx->stream;
    DBuf<float> xs1;
    xs1.ensure(fs.ldx + 32);
    MH_CHECK(hipMemsetAsync(xs1.p, 0, (fs.ldx + 32) * sizeof(float), st));
    const float one = 1.0f;
    h2d(ctx, xs1.p, x, n * sizeof(float));
    h2d(ctx, xs1.p + fs.ldx, &one, sizeof(float));
    Cands &cs = ctx->cand;
    cs.fit.assign(p, 0);
    cs.xc.assign(p, 0);
    cs.par.assign(mask, mask + p);
    for (int v : cs.par) require(v >= 0 && (size_t)v < m, "mask entry out of range");
    const long d0 = ctx->dgks_host;
    ctx->lin.ensure(p);
    ctx->ybx.ensure(p);
    plain_score(ctx, cs, Bx != nullptr, ctx->lin.p, ctx->ybx.p, xs1.p, xs1.p + fs.ldx);
    if (Bx) {
        DBuf<float> out;
        out.ensure(n * p);
        hipLaunchKernelGGL(k_gather_cols, dim3((n + 255) / 256, p), dim3(256), 0, st, fs.rbuf.p, fs.ldr, n, out.p, n);
        MH_LAUNCH_CHECK();
        MH_CHECK(hipMemcpy2DAsync(Bx, ldBx * sizeof(float), out.p, n * sizeof(float), n * sizeof(float), p,
                                  hipMemcpyDeviceToHost, st));
        ctx_sync(ctx);
        out.release();
    }
    if (ybx) MH_CHECK(hipMemcpyAsync(ybx, ctx->ybx.p, p * sizeof(double), hipMemcpyDeviceToHost, st));
    ctx_sync(ctx);
    xs1.release();
    resolve_timing(ctx);
    if (fired) *fired += ctx->dgks_host - d0;
}

void mirror_orthonormalize_batch(mars_hip_ctx *ctx, size_t m, size_t P, const int *src_xcol, const int *src_basis,
                                 double *ybx, long *fired)
{
    require(!ctx->resident, "use mars_hip_fit_eval on resident contexts");
    FitState &fs = ctx->fits[0];
    require(m == fs.m && m <= fs.m_synced && m >= 1, "basis not synced up to m");
    require(fs.px > 0, "no X columns resident (call mars_hip_context_sync_xcols)");
    if (P == 0) return;
    require(src_xcol && src_basis && ybx, "null argument");
    Cands &cs = ctx->cand;
    cs.fit.assign(P, 0);
    cs.xc.assign(src_xcol, src_xcol + P);
    cs.par.assign(src_basis, src_basis + P);
    for (size_t j = 0; j < P; ++j) {
        require(src_xcol[j] >= 0 && (size_t)src_xcol[j] < fs.px && fs.xres[src_xcol[j]], "src_xcol not resident");
        require(src_basis[j] >= 0 && (size_t)src_basis[j] < m, "src_basis out of range");
    }
    const long d0 = ctx->dgks_host;
    ctx->lin.ensure(P);
    ctx->ybx.ensure(P);
    plain_score(ctx, cs, false, ctx->lin.p, ctx->ybx.p);
    MH_CHECK(hipMemcpyAsync(ybx, ctx->ybx.p, P * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ctx_sync(ctx);
    resolve_timing(ctx);
    if (fired) *fired += ctx->dgks_host - d0;
}

} // namespace mh

/* ---------------------------------------------------------------------- */
/* C ABI                                                                  */
/* ---------------------------------------------------------------------- */
using namespace mh;

extern "C" {

const char *mars_hip_last_error(void) { return g_last_error.c_str(); }

int mars_hip_context_create(size_t n, size_t max_terms, double tol, mars_hip_ctx **out)
{
    return guarded([&] {
        require(out != nullptr, "null output pointer");
        *out = nullptr;
        auto *ctx = new mars_hip_ctx;
        try {
            ctx_base_init(ctx, max_terms);
            ctx->fits.resize(1);
            fit_state_alloc(ctx, ctx->fits[0], n, tol);
            ctx_sync(ctx);
        } catch (...) {
            ctx_free(ctx);
            throw;
        }
        *out = ctx;
    });
}

int mars_hip_context_destroy(mars_hip_ctx *ctx)
{
    return guarded([&] { ctx_free(ctx); });
}

int mars_hip_context_batch_capacity(mars_hip_ctx *ctx, size_t *capacity)
{
    return guarded([&] {
        require(ctx && capacity, "null argument");
        // the residual buffer is (n x P) f32 and grows on demand; report the
        // candidate count whose buffer fits in 8 GiB
        const size_t n = std::max<size_t>(ctx->fits[0].n, 1);
        *capacity = std::max<size_t>(ctx->max_terms, ((size_t)8 << 30) / (4 * n));
    });
}

int mars_hip_context_sync_basis(mars_hip_ctx *ctx, size_t m, const float *B, size_t ldB, const float *Bo,
                                size_t ldBo)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        mirror_sync_basis(ctx, m, B, ldB, Bo, ldBo);
    });
}

int mars_hip_context_set_target(mars_hip_ctx *ctx, const float *y)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        mirror_set_target(ctx, y);
    });
}

int mars_hip_context_sync_xcols(mars_hip_ctx *ctx, size_t p_x, const float *X, size_t ldX, const float *s,
                                const int *cols, size_t ncols)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        require(!ctx->resident, "resident contexts own their X");
        require(ldX >= ctx->fits[0].n, "invalid leading dimension for X");
        mirror_sync_xcols(ctx, p_x, X, ldX, s, cols, ncols);
    });
}

int mars_hip_orthonormalize(mars_hip_ctx *ctx, size_t m, size_t p, const float *x, const int *mask, float *Bx,
                            size_t ldBx, double *ybx, long *dgks_fired)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        mirror_orthonormalize(ctx, m, p, x, mask, Bx, ldBx, ybx, dgks_fired);
    });
}

int mars_hip_orthonormalize_batch(mars_hip_ctx *ctx, size_t m, size_t P, const int *src_xcol,
                                  const int *src_basis, double *ybx, long *dgks_fired)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        mirror_orthonormalize_batch(ctx, m, P, src_xcol, src_basis, ybx, dgks_fired);
    });
}

int mars_hip_fit_create(const float *X, size_t ldX, const float *y, const float *w, size_t n, size_t p_x,
                        size_t max_terms, int form, mars_hip_ctx **out)
{
    return guarded([&] {
        require(out != nullptr, "null output pointer");
        require(valid_form(form), "unknown precision form");
        *out = nullptr;
        auto *ctx = new mars_hip_ctx;
        try {
            ctx_base_init(ctx, max_terms);
            ctx->fits.resize(1);
            ctx->resident = true;
            set_form(ctx, form);
            form = ctx->form;
            fit_init(ctx, ctx->fits[0], X, ldX, y, w, n, p_x);
            if (form == MARS_HIP_FORM_MIXED) ensure_tcache(ctx);
            reserve_step_buffers(ctx);
        } catch (...) {
            ctx_free(ctx);
            throw;
        }
        *out = ctx;
    });
}

int mars_hip_fit_info(mars_hip_ctx *ctx, int *nbasis, double *yvar, double *dsse)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        require(!ctx->fits.empty(), "empty context");
        const FitState &fs = ctx->fits[0];
        if (nbasis) *nbasis = (int)fs.m;
        if (yvar) *yvar = fs.yvar;
        if (dsse) *dsse = host_dsse(fs);
    });
}

int mars_hip_fit_set_form(mars_hip_ctx *ctx, int form)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        require(valid_form(form), "unknown precision form");
        set_form(ctx, form);
    });
}

int mars_hip_fit_eval(mars_hip_ctx *ctx, const unsigned char *mask, size_t mask_rows, size_t mask_stride,
                      int min_span, int end_span, int linear_only, double *linear, double *hinge, double *cut,
                      mars_hip_best *best)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        fit_eval(ctx, mask, mask_rows, mask_stride, min_span, end_span, linear_only, linear, hinge, cut, best);
    });
}

int mars_hip_fit_append(mars_hip_ctx *ctx, char type, int xcol, int bcol, float h, double *mse)
{
    return guarded([&] {
        require(ctx != nullptr && mse != nullptr, "null argument");
        require(ctx->fits.size() == 1, "use mars_hip_batch_fit on batch contexts");
        if (ctx->fits[0].m >= ctx->max_terms) throw Error(MARS_HIP_EFULL, "basis matrix is full");
        AppendArg a{0, xcol, bcol, 1, h, 0};
        if (type == 'l') a.type = 0;
        else if (type == '+') a.type = 1;
        else if (type == '-') a.type = 2;
        else throw Error(MARS_HIP_EINVAL, "invalid basis type");
        std::vector<AppendOut> o;
        std::vector<AppendArg> args{a};
        fit_append_batch(ctx, args, o);
        *mse = o[0].mse;
    });
}

int mars_hip_fit_append_noalloc(mars_hip_ctx *ctx, char type, int xcol, int bcol, float h, double *mse)
{
    return guarded([&] {
        require(ctx != nullptr && mse != nullptr, "null argument");
        AppendArg a{0, xcol, bcol, 1, h, 0};
        if (type == 'l') a.type = 0;
        else if (type == '+') a.type = 1;
        else if (type == '-') a.type = 2;
        else throw Error(MARS_HIP_EINVAL, "invalid basis type");
        fit_append_noalloc(ctx, a, mse);
    });
}

int mars_hip_fit_basis(mars_hip_ctx *ctx, float *B, float *Bo, size_t *m_out)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        require(!ctx->fits.empty(), "empty context");
        const FitState &fs = ctx->fits[0];
        const size_t n = fs.n, m = fs.m, ld = fs.ldx;
        if (m_out) *m_out = m;
        std::vector<float> h(ld * std::max<size_t>(m, 1));
        float *dst[2] = {B, Bo};
        const float *src[2] = {fs.BC.p, fs.BoC.p};  // column-major copies (always present)
        for (int q = 0; q < 2; ++q) {
            if (!dst[q] || m == 0) continue;
            MH_CHECK(hipMemcpyAsync(h.data(), src[q], ld * m * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
            ctx_sync(ctx);
            for (size_t i = 0; i < n; ++i)
                for (size_t j = 0; j < m; ++j) dst[q][i * m + j] = h[j * ld + i];
        }
    });
}

int mars_hip_timing(mars_hip_ctx *ctx, int enable)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        ctx_sync(ctx);
        resolve_timing(ctx);
        ctx->timing = enable != 0;
    });
}

int mars_hip_kernel_stats(mars_hip_ctx *ctx, mars_hip_kstat *out, int capacity, int *count)
{
    return guarded([&] {
        require(ctx != nullptr && count != nullptr, "null argument");
        ctx_sync(ctx);
        resolve_timing(ctx);
        unsigned long long d[kDStat];
        read_dstat(ctx, d);
        int k = 0;
        for (int c = 0; c < K_NCLASS; ++c) {
            if (ctx->kacc[c].launches == 0) continue;
            if (out && k < capacity) {
                mars_hip_kstat &r = out[k];
                std::memset(r.name, 0, sizeof(r.name));
                std::strncpy(r.name, kClassName[c], sizeof(r.name) - 1);
                r.launches = ctx->kacc[c].launches;
                r.ms = ctx->kacc[c].ms;
                r.flops = ctx->kacc[c].flops;
                r.bytes = ctx->kacc[c].bytes;
                // the device-driven exact rescoring: its list lengths, and so its flops, are the device's
                if (c == K_RECHECK) {
                    r.flops += (double)d[5];
                    r.bytes += (double)d[6];
                }
            }
            ++k;
        }
        *count = k;
    });
}

int mars_hip_score_stats(mars_hip_ctx *ctx, long *scored, long *rescored, long *gated)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        unsigned long long d[kDStat];
        read_dstat(ctx, d);
        if (scored) *scored = ctx->scored + (long)d[0];
        if (rescored) *rescored = ctx->rescored + (long)d[1];
        if (gated) *gated = ctx->gated + (long)d[2];
    });
}

int mars_hip_check_stats(mars_hip_ctx *ctx, long *window_violations, long *host_redo_steps, long *graph_captures,
                         long *graph_launches)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        unsigned long long d[kDStat];
        read_dstat(ctx, d);
        if (window_violations) *window_violations = (long)d[3] + ctx->window_misses;
        if (host_redo_steps) *host_redo_steps = ctx->redo_steps;
        if (graph_captures) *graph_captures = ctx->graph_captures;
        if (graph_launches) *graph_launches = ctx->graph_launches;
    });
}

int mars_hip_stats(mars_hip_ctx *ctx, long *dgks_fired, double *eval_ms)
{
    return guarded([&] {
        require(ctx != nullptr, "null context");
        unsigned long long d[kDStat];
        read_dstat(ctx, d);
        if (dgks_fired) *dgks_fired = ctx->dgks_host + (long)d[4];
        if (eval_ms) *eval_ms = ctx->eval_ms;
    });
}

int mars_hip_default_tuning(mars_hip_tuning *out)
{
    return guarded([&] {
        require(out != nullptr, "null argument");
        *out = mars_hip_ctx().tune;
    });
}

int mars_hip_get_tuning(mars_hip_ctx *ctx, mars_hip_tuning *out)
{
    return guarded([&] {
        require(ctx != nullptr && out != nullptr, "null argument");
        *out = ctx->tune;
    });
}

int mars_hip_set_tuning(mars_hip_ctx *ctx, const mars_hip_tuning *t)
{
    return guarded([&] {
        require(ctx != nullptr && t != nullptr, "null argument");
        auto in = [](double v, double lo, double hi) { return v >= lo && v <= hi; };
        require(in(t->pass_items, 256, 1e6) && in(t->retry_items, 256, 1e6) && in(t->update_waves, 256, 1e6) &&
                    in(t->sweep_waves, 256, 1e6),
                "tuning: item and wave targets must be in [256, 1e6]");
        require(t->sweep_min_rows >= 32 && t->sweep_min_rows <= (1l << 20), "tuning: sweep_min_rows out of range");
        require(t->append_rows == 0 || (t->append_rows % 256 == 0 && t->append_rows >= 256 &&
                                        t->append_rows <= (1l << 20)),
                "tuning: append_rows must be 0 or a multiple of 256 in [256, 2^20]");
        require(t->narrow_avg >= 0 && t->narrow_avg <= 4096, "tuning: narrow_avg out of range");
        ctx->tune = *t;
        ctx->tune.pad = 0;
    });
}

int mars_hip_device_bytes(size_t *live, size_t *peak, int reset_peak)
{
    return guarded([&] {
        const long long now = g_dev_live.load();
        if (live) *live = (size_t)std::max<long long>(now, 0);
        if (peak) *peak = (size_t)std::max<long long>(g_dev_peak.load(), 0);
        if (reset_peak) g_dev_peak.store(now);
    });
}

} // extern "C"
