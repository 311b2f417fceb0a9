// Pieces shared by the PCG solvers: the host-visible status record, the
// partial-sum layout and the redundant (atomic-free) grid reductions.
#pragma once

#include <functional>

#include "device_common.h"
#include "gslam_internal.h"

namespace gslam {

// Streaming kernels use at most kMaxGrid workgroups (grid-stride loops), so a
// consumer re-reads at most kMaxGrid partials per reduced quantity.
constexpr int kMaxGrid = 512;
constexpr int kPartStride = kMaxGrid;
// Offsets (doubles) into DeviceGraph::part.
constexpr int kPartOffPR0 = 0 * kPartStride;    // r.r, r.z of even iterations (2 arrays)
constexpr int kPartOffPR1 = 2 * kPartStride;    // r.r, r.z of odd iterations (2 arrays)
constexpr int kPartOffPQ = 4 * kPartStride;     // p.q (1 array)
constexpr int kPartOffResid = 6 * kPartStride;  // residual / update norms (2 arrays)
// The SpMV runs one workgroup per row tile (up to kSpmvMaxGrid, grid-stride
// beyond), so its p.q partials get an array of their own.
constexpr int kSpmvMaxGrid = 4096;
constexpr int kPartOffPQT = 8 * kPartStride;    // p.q of the tiled SpMV
constexpr int kPartOffRz = 8 * kPartStride + kSpmvMaxGrid;  // reduced r.z by parity (2)
constexpr int kPartOffAlpha = kPartOffRz + 2;                // alpha by parity (2)
constexpr int kPartTotal = kPartOffRz + 8;

constexpr int kUpdThreads = 192;  // entry-parallel update: 64 poses per pass

// The FP64 PCG of graphs with at most kFusedMaxTiles SpMV tiles (N up to
// ~120k) runs one kernel per iteration: every workgroup re-reads all
// kFusedParts partials of the previous launch (one per workgroup, at most
// kFusedMaxGroups; 10 loads per thread at most), so the update of r, z, p
// and x needs no launch of its own.  A workgroup takes one tile up to
// kFusedMaxGroups tiles, two consecutive tiles up to twice that, four
// beyond (kFusedMaxTiles).
constexpr int kFusedMaxGroups = 512;
constexpr int kFusedMaxTiles = 4 * kFusedMaxGroups;
constexpr int kFusedParts = 5;  // p.q, r.z, q.z, q.D^-1 q, r.r

// Arrays that a PCG kernel reads once per iteration (the matrix values and
// column indices, the FP32 factors, D^-1) are loaded nontemporally when they
// are larger than the 256 MB MI355X Infinity Cache: they cannot stay there
// from one iteration to the next anyway, and nontemporal loads keep them from
// evicting the vectors that can (FP64 SpMV at N = 1M: 109 -> 96.5 us; FP32
// SpMV at N = 4M: 202 -> 184 us).  Smaller arrays use ordinary loads, which
// keep them cached across iterations (N = 100k: 10.4 us against 11.1 us
// nontemporal); the FP32 operator, which shares the cache with the mixed
// iteration's FP64 x, FP32 vectors and D^-1, switches at half the cache.
// Once the FP64 solver's six vectors no longer fit either, the
// ones that only their own rows' threads read and write (x, q, r) are
// streamed nontemporally too, which leaves the cache to z and p, whose
// entries the SpMV gathers (FP64 iteration at N = 4M: 589 -> 561 us).
constexpr int64_t kInfinityCacheBytes = int64_t{256} << 20;

inline bool streams_past_cache(int64_t bytes) { return bytes > kInfinityCacheBytes; }

template <bool kNt, typename T>
__device__ inline T stream_load(const T *p)
{
    if constexpr (kNt) return __builtin_nontemporal_load(p);
    else return *p;
}

template <bool kNt, typename T>
__device__ inline void stream_store(T v, T *p)
{
    if constexpr (kNt) __builtin_nontemporal_store(v, p);
    else *p = v;
}

inline unsigned int grid_for(int64_t rows, int lanes)
{
    int64_t g = (rows * lanes + kThreads - 1) / kThreads;
    if (g < 1) g = 1;
    if (g > kMaxGrid) g = kMaxGrid;
    return static_cast<unsigned int>(g);
}

inline unsigned int grid_entries(int64_t n_poses)
{
    int64_t g = (3 * n_poses + kUpdThreads - 1) / kUpdThreads;
    if (g < 1) g = 1;
    if (g > kMaxGrid) g = kMaxGrid;
    return static_cast<unsigned int>(g);
}

// Block sums of K values, written by thread 0 to out[k * stride + slot]
// (slot = blockIdx.x unless given).  Called at the end of a kernel (its LDS
// is not used again).
template <int K, int NT = kThreads>
__device__ inline void write_partials(double (&v)[K], double *out, int stride,
                                      int slot = -1)
{
    __shared__ double lds[K][NT / kWave];
    block_sum_last<K, NT>(v, lds);
    if (threadIdx.x == 0) {
        const int at = slot < 0 ? static_cast<int>(blockIdx.x) : slot;
#pragma unroll
        for (int k = 0; k < K; ++k) out[k * stride + at] = v[k];
    }
}

// The dispatcher deals the blocks of a grid of g round-robin over the XCDs
// (block b on XCD b % kXcdCount).  Returns b's position when each XCD takes
// one contiguous range of [0, g) instead, so that neighbouring rows, and the
// vector lines their gathers share, stay in one XCD's L2.  A bijection on
// [0, g).
constexpr int kXcdCount = 8;  // MI355X
__device__ inline int32_t xcd_contiguous(int32_t b, int32_t g)
{
    const int32_t x = b % kXcdCount, k = b / kXcdCount;
    const int32_t q = g / kXcdCount, r = g % kXcdCount;  // XCDs x < r hold q + 1
    return x * q + min(x, r) + k;
}

// acc += part[i] for i = threadIdx.x, +NT, ... < np, in that order, with
// the loads of M consecutive terms issued together (predicated, absent terms
// add 0.0); a plain loop waits for each load before the next add, one memory
// round trip per term.
template <int NT, int M = 8>
__device__ inline double strided_sum(const double *__restrict__ part, int np, double acc)
{
    for (int base = threadIdx.x; base < np; base += M * NT) {
        double t[M];
#pragma unroll
        for (int u = 0; u < M; ++u) t[u] = (base + u * NT < np) ? part[base + u * NT] : 0.0;
#pragma unroll
        for (int u = 0; u < M; ++u) acc += t[u];
    }
    return acc;
}

// Sum of np partials per quantity in a fixed order; every thread of every
// workgroup obtains the same bits.
template <int K, int NT = kThreads>
__device__ inline void sum_partials(const double *part, int np, int stride, double (&out)[K])
{
    __shared__ double lds[K][NT / kWave];
    __shared__ double res[K];
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = strided_sum<NT>(part + k * stride, np, 0.0);
    block_sum<K, NT>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) res[k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = res[k];
}

int run_chunked(DeviceGraph &g, const gslam_solve_opts &o,
                const std::function<void(int32_t)> &enqueue, int32_t &k);
int pcg_solve_fp64(DeviceGraph &g, const gslam_solve_opts &o, const double *b, double *x_out,
                   gslam_solve_stats &s);
int pcg_solve_mixed(DeviceGraph &g, const gslam_solve_opts &o, const double *b, double *x_out,
                    gslam_solve_stats &s);
// Enqueues ||b - A x|| / ||b|| (FP64) into the device scalar dst.
int enqueue_true_residual(DeviceGraph &g, const double *b, const double *x, double *dst);

}  // namespace gslam
