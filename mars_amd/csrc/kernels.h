// Device helpers shared by the libmars_hip translation units (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

namespace mh {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;
constexpr int kRedThreads = 256;

/*
 * Load through the global address space.  Pointers read from device structs
 * are generic to the compiler, which then emits flat loads; every pointer the
 * kernels dereference this way points at hipMalloc'd memory, where generic and
 * global addresses coincide.
 */
template <typename T>
__device__ inline T gld(const T *p)
{
    return *__builtin_bit_cast(const __attribute__((address_space(1))) T *, p);
}

/* 16-byte global load (p 16-byte aligned). */
__device__ inline float4 gld4(const float *p)
{
    typedef float nf4 __attribute__((ext_vector_type(4)));
    const nf4 v = *__builtin_bit_cast(const __attribute__((address_space(1))) nf4 *, p);
    return make_float4(v.x, v.y, v.z, v.w);
}

/*
 * MFMA f64 16x16x4 (measured on gfx950):
 *   A (16x4):  lane l holds A[l%16][l/16]
 *   B (4x16):  lane l holds B[l/16][l%16]
 *   D (16x16): lane l, element v holds D[l/16 + 4v][l%16]
 */
__device__ inline v4d mfma_f64(double a, double b, v4d c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

/*
 * MFMA f32 16x16x4 (measured on gfx950): A and B as for f64;
 *   D (16x16): lane l, element v holds D[4(l/16) + v][l%16]
 */
__device__ inline v4f mfma_f32(float a, float b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

/*
 * p[idx(0)] + p[idx(1)] + ... + p[idx(cnt-1)], added left to right (the fixed order every
 * reduce kernel uses), with eight loads in flight: a plain loop waits for each load in turn.
 */
template <typename Idx>
__device__ inline double ordered_sum(const double *p, int cnt, Idx idx)
{
    double s = 0.0;
    int i = 0;
    for (; i + 8 <= cnt; i += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = gld(p + idx(i + u));
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; i < cnt; ++i) s += gld(p + idx(i));
    return s;
}

/* Sum of p[idx(i)] for i in [lo, hi), left to right, eight loads in flight. */
template <typename Idx>
__device__ inline double range_sum(const double *p, int lo, int hi, Idx idx)
{
    double s = 0.0;
    int i = lo;
    for (; i + 8 <= hi; i += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = gld(p + idx(i + u));
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; i < hi; ++i) s += gld(p + idx(i));
    return s;
}

/* Fixed-order butterfly sum over the aligned group of L lanes of a wave (L a power of two, at
 * most 64); every lane of the group gets the result. */
__device__ inline double group_sum(double v, int L)
{
    for (int o = L >> 1; o > 0; o >>= 1) v = v + __shfl_xor(v, o, kWave);
    return v;
}

/* Fixed-order block sum of one double per thread (blockDim.x == 256). */
__device__ inline double block_sum_256(double v, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) sh[t] = sh[t] + sh[t + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

/* Fixed-order wave64 butterfly sum (every lane gets the same result). */
__device__ inline double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, kWave);
    return v;
}

/* Sum of `nb` (<= 256) partials, identical in every block that calls it. */
__device__ inline double sum_partials(const double *part, int nb, double *sh)
{
    return block_sum_256((int)threadIdx.x < nb ? part[threadIdx.x] : 0.0, sh);
}

} // namespace mh
