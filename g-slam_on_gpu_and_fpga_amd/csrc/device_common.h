// Device helpers shared by the libgslam kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gslam {

constexpr int kWave = 64;
constexpr int kThreads = 256;  // threads per workgroup for the streaming kernels

// Sum over the 64 lanes of a wavefront (butterfly; every lane gets the total).
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Sum over groups of G consecutive lanes (G a power of two <= 64).
template <int G>
__device__ inline double group_sum(double v)
{
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}

// Deterministic workgroup sum of K values: wave butterflies, then wave 0
// adds the per-wave totals in wave order.  Result valid in thread 0.
template <int K, int NT>
__device__ inline void block_sum(double (&v)[K], double (*lds)[NT / kWave])
{
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = wave_sum(v[k]);
        if (l == 0) lds[k][w] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
            for (int i = 0; i < NT / kWave; ++i) s += lds[k][i];
            v[k] = s;
        }
    }
    __syncthreads();
}

// As block_sum, for the last use of lds in a kernel: no barrier after
// thread 0 has read the per-wave totals.
template <int K, int NT>
__device__ inline void block_sum_last(double (&v)[K], double (*lds)[NT / kWave])
{
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = wave_sum(v[k]);
        if (l == 0) lds[k][w] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
            for (int i = 0; i < NT / kWave; ++i) s += lds[k][i];
            v[k] = s;
        }
    }
}

// Workgroup sum of K values that every thread receives, with one barrier:
// wave butterflies, per-wave totals to LDS, then each thread adds the totals
// in wave order (the same bits in every thread and every workgroup that sums
// the same inputs).  lds must not be reused before another barrier.
template <int K, int NT>
__device__ inline void block_allsum(double (&v)[K], double (*lds)[NT / kWave])
{
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double s = wave_sum(v[k]);
        if (l == 0) lds[k][w] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NT / kWave; ++i) s += lds[k][i];
        v[k] = s;
    }
}

// Angle wrap to [-pi, pi): a - 2 pi floor((a + pi) / 2 pi).
__device__ inline double wrap_pi(double a)
{
    const double two_pi = 6.283185307179586476925286766559;
    return a - two_pi * floor((a + 3.1415926535897932384626433832795) / two_pi);
}

}  // namespace gslam
