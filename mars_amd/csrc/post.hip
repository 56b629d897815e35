// After the forward pass, on gfx950: the basis matrix of a model
// (mars.py:423-450) and Gram products A^T B with float64 accumulation
// (mars.py:456-490), the O(n M^2) part of the backward pass (mars.py:607-698).
//
// Gram: f32 inputs are widened to f64 in registers and contracted with
// v_mfma_f64_16x16x4f64, so every product is exact (24 + 24 bits fit the
// 53-bit mantissa) and every sum is an f64 sum.  The reference forms each
// 128-row chunk's product in the input precision before widening it, so its
// f32 results differ from these by that chunk rounding (about 1e-7 relative).
#include "internal.h"
#include "kernels.h"

#include <algorithm>
#include <vector>

namespace mh {
namespace {

enum TermCode { kIntercept = 0, kLinear = 1, kPos = 2, kNeg = 3 };

/*
 * B columns in term order, one thread per row (a term's parent is an earlier
 * column of the same row).  'l': B_b * x in the input precision.  '+' / '-':
 * the knot is f64, so x - h (h - x) and its product with B_b are f64 values
 * rounded once to the output precision -- numpy's promotion of an f32 array
 * against the model's f64 scalar.  max(d, 0) keeps a NaN d, as numpy.maximum.
 */
template <typename T>
__global__ __launch_bounds__(256) void k_expand(const T *__restrict__ X, size_t ldx, size_t n,
                                                const int *__restrict__ code, const int *__restrict__ basis,
                                                const int *__restrict__ input, const double *__restrict__ hinge,
                                                int M, T *B, size_t ldb)
{
    const size_t r = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    for (int i = 0; i < M; ++i) {
        const int c = code[i];
        T v = T(1);
        if (c != kIntercept) {
            const T b = B[(size_t)basis[i] * ldb + r];
            const T x = X[(size_t)input[i] * ldx + r];
            if (c == kLinear) {
                v = b * x;
            } else {
                const double d = c == kPos ? (double)x - hinge[i] : hinge[i] - (double)x;
                const double h = d != d ? d : (d > 0.0 ? d : 0.0);
                v = (T)((double)b * h);
            }
        }
        B[(size_t)i * ldb + r] = v;
    }
}

constexpr int kGT = 8;  // 16-column B tiles per wave

__device__ inline void load4(const float *p, double (&v)[4])
{
    const float4 f = gld4(p);
    v[0] = f.x;
    v[1] = f.y;
    v[2] = f.z;
    v[3] = f.w;
}

__device__ inline void load4(const double *p, double (&v)[4])
{
    const v4d d = *(const v4d *)p;
    v[0] = d[0];
    v[1] = d[1];
    v[2] = d[2];
    v[3] = d[3];
}

/*
 * Partial A^T B of one row split for a 16 x (16 kGT) output strip: A tile ti,
 * B tiles kGT*tg ... .  Both inputs are zero-padded device copies (rows to a
 * multiple of 16, columns to a multiple of 16, leading dimension npad).  Lane
 * (kr = l/16, kc = l%16) holds rows r + 4kr .. r + 4kr + 3 of A column
 * 16 ti + kc and of each B tile's column kc; MFMA step q contracts row
 * r + 4kr + q as k index kr, so the four steps cover rows r .. r+15.
 */
template <typename T>
__global__ __launch_bounds__(64) void k_gram(const T *__restrict__ A, const T *__restrict__ Bm, size_t npad,
                                             int tbn, int groups, int S, size_t rows, double *__restrict__ part)
{
    const int lane = threadIdx.x, kr = lane >> 4, kc = lane & 15;
    const int w = blockIdx.x;
    const int split = w % S, strip = w / S;
    const int ti = strip / groups, tg = strip % groups;
    const size_t r0 = (size_t)split * rows, r1 = min(npad, r0 + rows);
    const T *pa = A + (size_t)(16 * ti + kc) * npad;
    const T *pb[kGT];
#pragma unroll
    for (int t = 0; t < kGT; ++t) {
        const int tj = min(tg * kGT + t, tbn - 1);  // tiles past the end repeat the last (discarded)
        pb[t] = Bm + (size_t)(16 * tj + kc) * npad;
    }
    v4d acc[kGT];
#pragma unroll
    for (int t = 0; t < kGT; ++t) acc[t] = v4d{0, 0, 0, 0};
#pragma unroll 1
    for (size_t r = r0; r < r1; r += 16) {
        const size_t rb = r + 4 * kr;
        double a[4], b[kGT][4];
        load4(pa + rb, a);
#pragma unroll
        for (int t = 0; t < kGT; ++t) load4(pb[t] + rb, b[t]);
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int t = 0; t < kGT; ++t) acc[t] = mfma_f64(a[q], b[t][q], acc[t]);
    }
    // D lane l element v: output row l/16 + 4v, column l%16 of the tile
    double *dst = part + (size_t)w * kGT * 256;
#pragma unroll
    for (int t = 0; t < kGT; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) dst[t * 256 + (kr + 4 * v) * 16 + kc] = acc[t][v];
}

/* out[i + j ldo] = sum of the splits' partials, in split order. */
__global__ __launch_bounds__(256) void k_gram_reduce(const double *__restrict__ part, int S, int groups,
                                                     int ma, int mb, double *__restrict__ out, size_t ldo)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)ma * mb) return;
    const int i = (int)(e % ma), j = (int)(e / ma);
    const int ti = i >> 4, tj = j >> 4;
    const int strip = ti * groups + tj / kGT, t = tj % kGT;
    double s = 0.0;
    for (int sp = 0; sp < S; ++sp)
        s += part[((size_t)(strip * S + sp) * kGT + t) * 256 + (i & 15) * 16 + (j & 15)];
    out[(size_t)j * ldo + i] = s;
}

/* A zero-padded (npad x mpad) device copy of the host column-major (n x m, ld) matrix. */
template <typename T>
void upload_padded(DBuf<T> &d, const T *h, size_t ld, size_t n, size_t m, size_t npad, size_t mpad)
{
    d.ensure(npad * mpad);
    MH_CHECK(hipMemset(d.p, 0, npad * mpad * sizeof(T)));
    if (n > 0 && m > 0)
        MH_CHECK(hipMemcpy2D(d.p, npad * sizeof(T), h, ld * sizeof(T), n * sizeof(T), m, hipMemcpyHostToDevice));
}

template <typename T>
void gram(const T *A, size_t ldA, size_t n, size_t ma, const T *B, size_t ldB, size_t mb, double *out,
          size_t ldo)
{
    require(A != nullptr && out != nullptr, "null pointer");
    require(ma >= 1, "A needs at least one column");
    require(ldA >= n, "ldA must be >= n");
    const bool sym = B == nullptr;
    if (sym) {
        B = A;
        ldB = ldA;
        mb = ma;
    }
    require(mb >= 1, "B needs at least one column");
    require(ldB >= n, "ldB must be >= n");
    require(ldo >= ma, "ldOut must be >= the column count of A");
    require(ma <= (1u << 20) && mb <= (1u << 20), "too many columns");
    const size_t npad = std::max<size_t>(16, (n + 15) / 16 * 16);
    const size_t mapad = (ma + 15) / 16 * 16, mbpad = (mb + 15) / 16 * 16;
    DBuf<T> dA, dB;
    upload_padded(dA, A, ldA, n, ma, npad, mapad);
    if (!sym) upload_padded(dB, B, ldB, n, mb, npad, mbpad);
    const T *pB = sym ? dA.p : dB.p;
    const int ta = (int)(mapad / 16), tbn = (int)(mbpad / 16);
    const int groups = (tbn + kGT - 1) / kGT;
    const int strips = ta * groups;
    const size_t maxs = std::max<size_t>(1, npad / 256);
    size_t S = std::min<size_t>(maxs, std::max<size_t>(1, 4096 / (size_t)strips));
    const size_t rows = ((npad + S - 1) / S + 15) / 16 * 16;
    S = (npad + rows - 1) / rows;
    DBuf<double> part, dout;
    part.ensure((size_t)strips * S * kGT * 256);
    dout.ensure(ma * mb);
    hipLaunchKernelGGL(k_gram<T>, dim3((unsigned)(strips * S)), dim3(64), 0, 0, (const T *)dA.p, pB, npad, tbn,
                       groups, (int)S, rows, part.p);
    MH_LAUNCH_CHECK();
    const size_t ne = ma * mb;
    hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, 0, (const double *)part.p,
                       (int)S, groups, (int)ma, (int)mb, dout.p, ma);
    MH_LAUNCH_CHECK();
    MH_CHECK(hipMemcpy2D(out, ldo * sizeof(double), dout.p, ma * sizeof(double), ma * sizeof(double), mb,
                         hipMemcpyDeviceToHost));
    if (sym)  // A^T A: the upper triangle, mirrored
        for (size_t j = 0; j < ma; ++j)
            for (size_t i = j + 1; i < ma; ++i) out[j * ldo + i] = out[i * ldo + j];
}

template <typename T>
void expand(const T *X, size_t ldX, size_t n, size_t px, size_t M, const char *type, const int *basis,
            const int *input, const double *hinge, T *B, size_t ldB)
{
    require(M == 0 || (type && basis && input && hinge), "null model pointer");
    require(M == 0 || B != nullptr, "null output pointer");
    require(ldB >= n, "ldB must be >= n");
    std::vector<int> code(M);
    for (size_t i = 0; i < M; ++i) {
        switch (type[i]) {
        case 'i': code[i] = kIntercept; break;
        case 'l': code[i] = kLinear; break;
        case '+': code[i] = kPos; break;
        case '-': code[i] = kNeg; break;
        default: throw Error(MARS_HIP_EINVAL, "unknown basis type");
        }
        if (code[i] != kIntercept) {
            require(input[i] >= 0 && (size_t)input[i] < px, "input column out of range");
            require(basis[i] >= 0 && (size_t)basis[i] < i, "parent must be an earlier term");
        }
    }
    if (M == 0 || n == 0) return;
    require(X != nullptr, "null X");
    require(ldX >= n, "ldX must be >= n");
    DBuf<T> dX, dB;
    DBuf<int> dc;
    DBuf<double> dh;
    dX.ensure(std::max<size_t>(1, n * px));
    if (px > 0)
        MH_CHECK(hipMemcpy2D(dX.p, n * sizeof(T), X, ldX * sizeof(T), n * sizeof(T), px, hipMemcpyHostToDevice));
    dB.ensure(n * M);
    dc.ensure(3 * M);
    dh.ensure(M);
    MH_CHECK(hipMemcpy(dc.p, code.data(), M * sizeof(int), hipMemcpyHostToDevice));
    MH_CHECK(hipMemcpy(dc.p + M, basis, M * sizeof(int), hipMemcpyHostToDevice));
    MH_CHECK(hipMemcpy(dc.p + 2 * M, input, M * sizeof(int), hipMemcpyHostToDevice));
    MH_CHECK(hipMemcpy(dh.p, hinge, M * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_expand<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, (const T *)dX.p, n, n,
                       (const int *)dc.p, (const int *)dc.p + M, (const int *)dc.p + 2 * M, (const double *)dh.p,
                       (int)M, dB.p, n);
    MH_LAUNCH_CHECK();
    MH_CHECK(hipMemcpy2D(B, ldB * sizeof(T), dB.p, n * sizeof(T), n * sizeof(T), M, hipMemcpyDeviceToHost));
}

} // namespace
} // namespace mh

using namespace mh;

extern "C" {

int mars_hip_expand(const float *X, size_t ldX, size_t n, size_t p_x, size_t M, const char *type, const int *basis,
                    const int *input, const double *hinge, float *B, size_t ldB)
{
    return guarded([&] { expand<float>(X, ldX, n, p_x, M, type, basis, input, hinge, B, ldB); });
}

int mars_hip_expand_f64(const double *X, size_t ldX, size_t n, size_t p_x, size_t M, const char *type,
                        const int *basis, const int *input, const double *hinge, double *B, size_t ldB)
{
    return guarded([&] { expand<double>(X, ldX, n, p_x, M, type, basis, input, hinge, B, ldB); });
}

int mars_hip_gram(const float *A, size_t ldA, size_t n, size_t ma, const float *B, size_t ldB, size_t mb,
                  double *out, size_t ldOut)
{
    return guarded([&] { gram<float>(A, ldA, n, ma, B, ldB, mb, out, ldOut); });
}

int mars_hip_gram_f64(const double *A, size_t ldA, size_t n, size_t ma, const double *B, size_t ldB, size_t mb,
                      double *out, size_t ldOut)
{
    return guarded([&] { gram<double>(A, ldA, n, ma, B, ldB, mb, out, ldOut); });
}

} // extern "C"
