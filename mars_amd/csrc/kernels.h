// Device helpers shared by the libmars_hip translation units (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

namespace mh {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;
constexpr int kRedBlocks = 256;   // blocks of the fixed-shape grid-stride reductions
constexpr int kRedThreads = 256;

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

/* Sum of kRedBlocks partials, identical in every block that calls it. */
__device__ inline double sum_partials(const double *part, double *sh)
{
    return block_sum_256(part[threadIdx.x], sh);
}

/* f32 candidate value B[row, par] * f32(X[row, xc] * s[xc]) (marsalgo.cc:436-441, kernels.cc:298-312). */
__device__ inline float cand_value(const float *B, size_t ldb, const float *X, size_t ldx,
                                   const float *s, size_t row, int xc, int par)
{
    const float xs = __fmul_rn(X[(size_t)xc * ldx + row], s[xc]);
    return __fmul_rn(B[row * ldb + par], xs);
}

} // namespace mh
