// Device helpers shared by the libgslam kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gslam {

constexpr int kWave = 64;
constexpr int kThreads = 256;  // threads per workgroup for the streaming kernels

// v of another lane of the same row of 16 through the DPP crossbar (two
// 32-bit moves; no LDS traffic, unlike a shuffle): kCtrl 0xB1 lane i ^ 1,
// 0x4E lane i ^ 2 (quad permutations), 0x141 lane 7 - i of the half row,
// 0x140 lane 15 - i of the row.
template <int kCtrl>
__device__ inline double dpp_get(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), kCtrl, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), kCtrl, 0xF, 0xF, true);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

__device__ inline double read_lane(double v, int lane)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Sum over groups of G consecutive lanes (G a power of two <= 64), every lane
// of a group getting the same bits: butterflies in which the two partner
// lanes add the same two values (commutative, so both round alike) -- within
// a row of 16 through DPP (i ^ 1, i ^ 2, then the half-row and row mirrors,
// which pair each 4- and 8-lane group with its neighbour), beyond it through
// shuffles.  Whole waves only (every lane of the wave active).
template <int G>
__device__ inline double group_sum(double v)
{
    static_assert(G >= 1 && G <= kWave && (G & (G - 1)) == 0, "a power of two <= 64");
    if constexpr (G >= 2) v += dpp_get<0xB1>(v);
    if constexpr (G >= 4) v += dpp_get<0x4E>(v);
    if constexpr (G >= 8) v += dpp_get<0x141>(v);
    if constexpr (G >= 16) v += dpp_get<0x140>(v);
#pragma unroll
    for (int o = 16; o < G; o <<= 1) v += __shfl_xor(v, o, G);
    return v;
}

// Sum over the 64 lanes of a wavefront, every lane getting the same bits:
// row sums (group_sum<16>), then the four rows' sums read as scalars and
// added in row order.  Whole waves only.
__device__ inline double wave_sum(double v)
{
    v = group_sum<16>(v);
    return ((read_lane(v, 0) + read_lane(v, 16)) + read_lane(v, 32)) + read_lane(v, 48);
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
// the same inputs).  lds must not be reused before another barrier.  The two
// halves are also usable apart (wave_totals as soon as the values are there,
// sum_wave_totals after a barrier the kernel needs anyway).
template <int K, int NT>
__device__ inline void wave_totals(const double (&v)[K], double (*lds)[NT / kWave])
{
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double s = wave_sum(v[k]);
        if (l == 0) lds[k][w] = s;
    }
}

template <int K, int NT>
__device__ inline void sum_wave_totals(double (&v)[K], const double (*lds)[NT / kWave])
{
    if constexpr (NT / kWave <= 4) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < NT / kWave; ++i) s += lds[k][i];
            v[k] = s;
        }
    } else {
        // wider workgroups: lane l reads the total of wave l mod W and W-lane
        // butterflies add them (K reads per thread instead of K W; the xor
        // partners add the same two values, so every lane ends with the same
        // bits)
        constexpr int W = NT / kWave;
        static_assert((W & (W - 1)) == 0, "a power-of-two number of waves");
        const int l = threadIdx.x % kWave;
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = group_sum<W>(lds[k][l % W]);
    }
}

template <int K, int NT>
__device__ inline void block_allsum(double (&v)[K], double (*lds)[NT / kWave])
{
    wave_totals<K, NT>(v, lds);
    __syncthreads();
    sum_wave_totals<K, NT>(v, lds);
}

// Angle wrap to [-pi, pi): a - 2 pi floor((a + pi) / 2 pi).
__device__ inline double wrap_pi(double a)
{
    const double two_pi = 6.283185307179586476925286766559;
    return a - two_pi * floor((a + 3.1415926535897932384626433832795) / two_pi);
}

}  // namespace gslam
