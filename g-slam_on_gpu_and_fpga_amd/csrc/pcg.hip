// Block-Jacobi preconditioned CG on the BSR Omega, FP64, fully device
// resident.  Algorithm and stopping rule of pcg_solve_sparse
// (src/cpu_pcg_fp64/MatrixOps.c:281-391): x0 = 0, r = b, z = D^-1 r, p = z;
// per iteration q = A p, alpha = rz/pq (pq == 0 -> stop before updating),
// x += alpha p, r -= alpha q, stop when ||r|| < tol ||b|| (checked after the
// update), z = D^-1 r, beta = rz_new/rz_old, p = z + beta p.
//
// Two kernels per iteration (one on graphs of at most kFusedMaxTiles SpMV
// tiles: pcg_fused_kernel below), no host round trip and no atomics:
//   A(k) "spmv":   sums the r.r / r.z partials B(k-1) left behind (every
//                  workgroup redundantly, in the same fixed order, so all
//                  agree bit for bit), applies the stopping test, forms
//                  p = z + beta p_old on the fly for every gathered column
//                  (double-buffered p), q = A p, and leaves p.q partials;
//   B(k) "update": sums the p.q partials, alpha = rz/pq (rz as A(k) reduced
//                  it), updates r, z for its own 3x3 block rows, leaves
//                  r.r / r.z partials and alpha; x_(k+1) = x_k + alpha_k p_k
//                  is applied by A(k+1), which holds p_k for its output rows
//                  anyway (or after the loop), the same arithmetic one launch
//                  later.
// Kernel boundaries order the partials; a grid-wide reduction through one
// atomic counter costs ~15-35 us per iteration on MI355X at 1-2k
// workgroups, the redundant re-read of <= 512 partials ~1 us (hence at most
// kFusedMaxGroups workgroups in the one-launch form).  Only workgroup 0 (or
// tile group 0) writes the status record the host polls.
#include <hip/hip_ext.h>

#include "device_common.h"
#include "pcg_common.h"

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <type_traits>

namespace gslam {

constexpr int kLanesPerRow = 4;  // lanes per block row in the residual SpMV

__device__ inline void inv3_cofactor(const double *a, double *inv)
{
    // MatrixOps.c:301-314: det == 0 is replaced by 1e-30
    double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                 a[2] * (a[3] * a[7] - a[4] * a[6]);
    if (det == 0.0) det = 1e-30;
    inv[0] = (a[4] * a[8] - a[5] * a[7]) / det;
    inv[1] = (a[2] * a[7] - a[1] * a[8]) / det;
    inv[2] = (a[1] * a[5] - a[2] * a[4]) / det;
    inv[3] = (a[5] * a[6] - a[3] * a[8]) / det;
    inv[4] = (a[0] * a[8] - a[2] * a[6]) / det;
    inv[5] = (a[2] * a[3] - a[0] * a[5]) / det;
    inv[6] = (a[3] * a[7] - a[4] * a[6]) / det;
    inv[7] = (a[1] * a[6] - a[0] * a[7]) / det;
    inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
}

__global__ void __launch_bounds__(kThreads)
pcg_setup_kernel(const int32_t *__restrict__ bdiag, const double *__restrict__ bval,
                 const double *__restrict__ b, int32_t n, int32_t max_iter,
                 double *__restrict__ dinv, float *__restrict__ dinv32, double *__restrict__ x,
                 double *__restrict__ r, double *__restrict__ z, double *__restrict__ p0,
                 double *__restrict__ p1, double *__restrict__ part_out, PcgStatus *status)
{
    double v[2] = {0.0, 0.0};
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double a[9], inv[9];
        const double *blk = bval + 9 * static_cast<int64_t>(bdiag[i]);
#pragma unroll
        for (int k = 0; k < 9; ++k) a[k] = blk[k];
        inv3_cofactor(a, inv);
        const int64_t o = 3 * static_cast<int64_t>(i), dim = 3 * static_cast<int64_t>(n);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) dinv[c * dim + o + a] = inv[3 * a + c];
        if (dinv32) {
#pragma unroll
            for (int k = 0; k < 9; ++k) dinv32[9 * static_cast<int64_t>(i) + k] = static_cast<float>(inv[k]);
        }
        const double b0 = b[o], b1 = b[o + 1], b2 = b[o + 2];
        const double z0 = inv[0] * b0 + inv[1] * b1 + inv[2] * b2;
        const double z1 = inv[3] * b0 + inv[4] * b1 + inv[5] * b2;
        const double z2 = inv[6] * b0 + inv[7] * b1 + inv[8] * b2;
        if (x) { x[o] = 0.0; x[o + 1] = 0.0; x[o + 2] = 0.0; }
        if (r) { r[o] = b0; r[o + 1] = b1; r[o + 2] = b2; }
        if (z) { z[o] = z0; z[o + 1] = z1; z[o + 2] = z2; }
        if (p0) { p0[o] = 0.0; p0[o + 1] = 0.0; p0[o + 2] = 0.0; }
        if (p1) { p1[o] = 0.0; p1[o + 1] = 0.0; p1[o + 2] = 0.0; }
        v[0] += b0 * b0 + b1 * b1 + b2 * b2;   // ||r0||^2 = ||b||^2
        v[1] += b0 * z0 + b1 * z1 + b2 * z2;
    }
    write_partials<2>(v, part_out, kPartStride);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        status->done = max_iter <= 0 ? 1 : 0;
        status->converged = 0;
        status->iters = 0;
        status->flags = 0;
        status->refreshes = 0;
        status->x_iter = -1;
        status->rnorm = 0.0;
        status->bnorm = 0.0;
        status->tol = 0.0;
    }
}

// tval = the blocks of each tile, entry k of slot l at (9 t + k) kTileBlocks
// + l (plane-major within the tile, so that each of an SpMV lane's nine loads
// is one contiguous 512-byte run per wave; zero in the padding slots of a
// tile with fewer blocks).
__global__ void __launch_bounds__(kThreads)
tile_pack_kernel(const int4 *__restrict__ tile_info, const double *__restrict__ bval,
                 int32_t n_tiles, double *__restrict__ tval)
{
    const int64_t n = static_cast<int64_t>(n_tiles) * kTileBlocks * 9;
    for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < n;
         e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t plane = e / kTileBlocks;  // 9 t + k
        const int64_t t = plane / 9;
        const int k = static_cast<int>(plane - 9 * t);
        const int4 ti = tile_info[t];
        const int64_t b = ti.z + (e - plane * kTileBlocks);
        tval[e] = b < ti.w ? bval[9 * b + k] : 0.0;
    }
}

// A(k): stopping test on the partials of B(k-1), then q = A p with
// p = z + beta p_old formed per gathered column; p.q partials.
//
// One workgroup per tile of whole block rows (<= 256 blocks, <= 85 rows;
// tile_info[t] = {first row, end row, first block, end block}): one lane per
// 3x3 block (its 9 values and column index from the tile-padded copy, slot
// 256 t + lane, and the gathered p entries), the block's 3 products staged in
// LDS, then one thread per output row (3 per block row) sums its row's
// products in block order.
//
// At N = 100k the matrix and vectors sit in the MI355X Infinity Cache and the
// kernel is bound by its chain of dependent loads, not by bandwidth.  Issue
// order: the partials that give beta, the lane's slot (column index and
// values), the tile record and the output row's range and z / p_old, then
// the gathered z / p_old -- two dependent loads deep; the partials are
// reduced (one barrier) once the products are formed.  To hold fewer
// registers meanwhile, each lane keeps V z_j and V p_old_j (6 values instead
// of 15) and forms V p_j = V z_j + beta V p_old_j afterwards.
struct SpmvTileRegs {
    int32_t r0, r1, b0, b1, ob0, ob1;
    bool has_blk, has_out;
    double yz[3], yp[3], zo, po, xo;
};

template <bool kNt, bool kNtVec>
__device__ inline void spmv_tile_load(int32_t t, const int4 *__restrict__ tile_info,
                                      const int32_t *__restrict__ brow_ptr,
                                      const int32_t *__restrict__ tcol,
                                      const double *__restrict__ tval,
                                      const double *__restrict__ z,
                                      const double *__restrict__ p_old,
                                      const double *__restrict__ x, SpmvTileRegs &T)
{
    static_assert(kTileBlocks == kThreads, "one lane per padded block slot");
    // the block of this lane from its padded slot, loaded without waiting for
    // the tile record (a padding slot holds zeros and a valid column, so its
    // products are 0); then the tile record and the output row's range and z /
    // p_old (from a row clamped into the tile: unconditional, so that no
    // branch waits on them); then the gathered z / p_old and the products
    const int64_t slot = static_cast<int64_t>(t) * kTileBlocks + threadIdx.x;
    const int64_t j = 3 * static_cast<int64_t>(stream_load<kNt>(tcol + slot));
    const double *tv = tval + 9 * static_cast<int64_t>(t) * kTileBlocks + threadIdx.x;
    double v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = stream_load<kNt>(tv + k * kTileBlocks);
    const int4 ti = tile_info[t];
    T.r0 = ti.x;
    T.r1 = ti.y;
    T.b0 = ti.z;
    T.b1 = ti.w;
    T.has_blk = T.b0 + static_cast<int32_t>(threadIdx.x) < T.b1;
    T.has_out = static_cast<int32_t>(threadIdx.x) < 3 * (T.r1 - T.r0);
    const int32_t orow = min(T.r0 + static_cast<int32_t>(threadIdx.x) / 3, T.r1 - 1);
    T.ob0 = brow_ptr[orow];
    T.ob1 = brow_ptr[orow + 1];
    const int64_t o = 3 * static_cast<int64_t>(orow) + static_cast<int>(threadIdx.x) % 3;
    T.zo = z[o];
    T.po = p_old[o];
    T.xo = stream_load<kNtVec>(x + o);
    const double z0 = z[j], z1 = z[j + 1], z2 = z[j + 2];
    const double q0 = p_old[j], q1 = p_old[j + 1], q2 = p_old[j + 2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        T.yz[a] = v[3 * a] * z0 + v[3 * a + 1] * z1 + v[3 * a + 2] * z2;
        T.yp[a] = v[3 * a] * q0 + v[3 * a + 1] * q1 + v[3 * a + 2] * q2;
    }
}

template <bool kNtVec>
__device__ inline double spmv_tile_run(const SpmvTileRegs &T, double beta,
                                       const int32_t *__restrict__ bcol,
                                       const double *__restrict__ bval,
                                       const double *__restrict__ z,
                                       const double *__restrict__ p_old,
                                       double *__restrict__ p_new, double *__restrict__ q,
                                       double *__restrict__ x, bool apply_x, double alpha_prev,
                                       double *ys)
{
    const int a = static_cast<int>(threadIdx.x) % 3;
    double acc = 0.0;
    if (T.has_blk) {
#pragma unroll
        for (int r = 0; r < 3; ++r) ys[r * kTileBlocks + threadIdx.x] = T.yz[r] + beta * T.yp[r];
    }
    __syncthreads();
    if (T.has_out) {
        const int32_t hi = min(T.ob1, T.b0 + kTileBlocks);
        for (int32_t c = max(T.ob0, T.b0); c < hi; ++c) acc += ys[a * kTileBlocks + (c - T.b0)];
    }
    // a single block row longer than one tile is walked in further chunks,
    // after every thread has summed the first (the condition is the same in
    // the whole workgroup)
    if (T.b1 > T.b0 + kTileBlocks) __syncthreads();
    for (int32_t cb = T.b0 + kTileBlocks; cb < T.b1; cb += kTileBlocks) {
        const int32_t bb = cb + static_cast<int32_t>(threadIdx.x);
        if (bb < T.b1) {
            const int64_t j = 3 * static_cast<int64_t>(bcol[bb]);
            const double p0 = z[j] + beta * p_old[j];
            const double p1 = z[j + 1] + beta * p_old[j + 1];
            const double p2 = z[j + 2] + beta * p_old[j + 2];
            const double *v = bval + 9 * static_cast<int64_t>(bb);
            ys[threadIdx.x] = v[0] * p0 + v[1] * p1 + v[2] * p2;
            ys[kTileBlocks + threadIdx.x] = v[3] * p0 + v[4] * p1 + v[5] * p2;
            ys[2 * kTileBlocks + threadIdx.x] = v[6] * p0 + v[7] * p1 + v[8] * p2;
        }
        __syncthreads();
        if (T.has_out) {
            const int32_t lo = max(T.ob0, cb), hi = min(T.ob1, cb + kTileBlocks);
            for (int32_t c = lo; c < hi; ++c) acc += ys[a * kTileBlocks + (c - cb)];
        }
        __syncthreads();
    }
    double pq = 0.0;
    if (T.has_out) {
        const int64_t o = 3 * static_cast<int64_t>(T.r0 + static_cast<int32_t>(threadIdx.x) / 3) + a;
        const double pi = T.zo + beta * T.po;
        p_new[o] = pi;
        stream_store<kNtVec>(acc, q + o);
        if (apply_x) stream_store<kNtVec>(T.xo + alpha_prev * T.po, x + o);
        pq = pi * acc;
    }
    return pq;
}

// (at most 72 VGPRs: 7 workgroups per CU, 1792 on the chip, keep the 1697
// tiles of the N = 100k grid city in one round; at 92 VGPRs, 5 per CU, the
// kernel measured 16 us in a solve.  A 64-VGPR cap spills.)  kNt: the tile
// copy is loaded nontemporally (it exceeds the Infinity Cache); kNtVec: x
// and q as well (the solver's vectors exceed it).
template <bool kNt, bool kNtVec>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(7, 8)))
pcg_spmv_kernel(const int32_t *__restrict__ brow_ptr, const int32_t *__restrict__ bcol,
                const double *__restrict__ bval, const int4 *__restrict__ tile_info,
                const int32_t *__restrict__ tcol, const double *__restrict__ tval,
                int32_t n_tiles, const double *__restrict__ z,
                const double *__restrict__ p_old, double *__restrict__ p_new,
                double *__restrict__ q, int32_t k, double tol_rel,
                const double *__restrict__ pr_cur, int32_t np_cur,
                double *__restrict__ rz_hist, double *__restrict__ pq_out,
                double *__restrict__ x, const double *__restrict__ alpha_hist, PcgStatus *status)
{
    __shared__ double ys[3 * kTileBlocks];
    __shared__ double red[2][kThreads / kWave];
    static_assert(kMaxGrid <= 2 * kThreads, "two r.r / r.z partials per thread");
    SpmvTileRegs T;
    // first tile: with a tile per workgroup (the matrix then fits the
    // Infinity Cache and the kernel is latency-bound) each XCD takes a
    // contiguous run of tiles, so that gathers of neighbouring rows hit its
    // L2 (N = 100k: 12.06 -> 10.89 us in the solve); workgroups that walk
    // several tiles keep the dispatcher's round robin, which streams HBM
    // faster (N = 4M: 409 against 434 us contiguous)
    const int32_t g = static_cast<int32_t>(gridDim.x), b = static_cast<int32_t>(blockIdx.x);
    const int32_t t0 = n_tiles <= g ? xcd_contiguous(b, g) : b;
    // Loaded first, so that their round trips overlap the tile's loads: the
    // stop flag and tolerance as earlier launches left them, rz_{k-1} as
    // A(k-1) reduced it, and the r.r_k / r.z_k partials (np_cur <= 2 x
    // kThreads: two per thread and array, summed below in strided_sum's order)
    const int32_t done_in = status->done;
    const double tol_in = status->tol;
    const double rz_prev = k > 0 ? rz_hist[(k + 1) & 1] : 1.0;
    // x_k = x_(k-1) + alpha_(k-1) p_(k-1) is this launch's to apply when B(k-1)
    // computed alpha (x_iter == k), whether or not the solve stops here
    const bool apply_x = status->x_iter == k;
    const double alpha_prev = k > 0 ? alpha_hist[(k + 1) & 1] : 0.0;
    // (unconditional loads from clamped, in-range indices: a predicated load
    // would be waited for where its branch joins)
    const int tid = static_cast<int>(threadIdx.x);
    const int ia = min(tid, np_cur - 1), ib = min(tid + kThreads, np_cur - 1);
    const double rr_a = pr_cur[ia], rr_b = pr_cur[ib];
    const double rz_a = pr_cur[kPartStride + ia], rz_b = pr_cur[kPartStride + ib];
    asm volatile("" ::: "memory");  // keeps the loads above here (not sunk to their use)
    if (t0 < n_tiles) spmv_tile_load<kNt, kNtVec>(t0, tile_info, brow_ptr, tcol, tval, z, p_old, x, T);
    const bool in_a = tid < np_cur, in_b = tid + kThreads < np_cur;
    double red_v[2];
    red_v[0] = (0.0 + (in_a ? rr_a : 0.0)) + (in_b ? rr_b : 0.0);
    red_v[1] = (0.0 + (in_a ? rz_a : 0.0)) + (in_b ? rz_b : 0.0);
    block_allsum<2, kThreads>(red_v, red);
    // the pending x update of this workgroup's rows, for launches that stop
    auto apply_x_only = [&]() {
        for (int32_t t = t0; t < n_tiles; t += gridDim.x) {
            const int4 ti = tile_info[t];
            const int64_t o = 3 * static_cast<int64_t>(ti.x) + threadIdx.x;
            if (o < 3 * static_cast<int64_t>(ti.y)) x[o] = x[o] + alpha_prev * p_old[o];
        }
    };
    if (done_in) {
        if (apply_x) apply_x_only();
        return;
    }
    const double rr = red_v[0], rz = red_v[1];
    if (!isfinite(rr) || !isfinite(rz)) {  // NaN / Inf in the residual (flag bit 1)
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            status->done = 1;
            status->flags |= 2;
            status->iters = k;
        }
        if (apply_x) apply_x_only();
        return;
    }
    double tol;
    if (k == 0) {
        tol = tol_rel * sqrt(rr);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            status->bnorm = sqrt(rr);
            status->tol = tol;
            status->rnorm = sqrt(rr);
        }
    } else {
        tol = tol_in;
        if (sqrt(rr) < tol) {  // MatrixOps.c:366: checked after the update
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                status->done = 1;
                status->converged = 1;
                status->iters = k;
                status->rnorm = sqrt(rr);
            }
            if (apply_x) apply_x_only();
            return;
        }
    }
    const double beta = k > 0 ? rz / rz_prev : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) rz_hist[k & 1] = rz;
    double pq[1] = {0.0};
    for (int32_t t = t0; t < n_tiles; t += gridDim.x) {
        if (t != t0) {
            __syncthreads();  // the previous tile is summed before ys is rewritten
            spmv_tile_load<kNt, kNtVec>(t, tile_info, brow_ptr, tcol, tval, z, p_old, x, T);
        }
        pq[0] += spmv_tile_run<kNtVec>(T, beta, bcol, bval, z, p_old, p_new, q, x, apply_x, alpha_prev, ys);
    }
    write_partials<1>(pq, pq_out, kSpmvMaxGrid, t0);  // slot t0: the sum of tiles t0, t0 + grid, ...
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        status->iters = k;
        status->rnorm = sqrt(rr);
    }
}

// B(k): alpha from the p.q partials and the r.z that A(k) reduced from the
// same partials (the same bits as summing them again here, without the
// loads), x/r/z update, r.r / r.z partials.  One thread per pose: its 21
// loads (x, r, p, q, D^-1) are issued before the p.q partials are loaded and
// reduced; D^-1 is stored column-major per entry ([c][3i+a]).  kNt: D^-1
// and q are loaded nontemporally (D^-1 exceeds the Infinity Cache); kNtVec:
// q and r (the solver's vectors exceed it).
template <bool kNt, bool kNtVec>
__global__ void __launch_bounds__(kThreads)
pcg_update_kernel(const double *__restrict__ q, const double *__restrict__ dinv,
                  double *__restrict__ r, double *__restrict__ z, int32_t n, int32_t k,
                  const double *__restrict__ pq_in, int32_t np_pq,
                  const double *__restrict__ rz_cur, double *__restrict__ pr_out,
                  double *__restrict__ alpha_out, PcgStatus *status)
{
    __shared__ double red[1][kThreads / kWave];
    const int64_t dim = 3 * static_cast<int64_t>(n);
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double rp[3], qp[3], di[9];
    auto load = [&](int32_t ii) {
        const int64_t o = 3 * static_cast<int64_t>(ii);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            rp[c] = stream_load<kNtVec>(r + o + c);
            qp[c] = stream_load<kNt || kNtVec>(q + o + c);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int a = 0; a < 3; ++a) di[3 * c + a] = stream_load<kNt>(dinv + c * dim + o + a);
    };
    const int32_t done_in = status->done;  // set by earlier launches only
    if (i < n) load(i);
    const double rz = *rz_cur;
    double red_v[1];
    red_v[0] = strided_sum<kThreads>(pq_in, np_pq, 0.0);
    block_allsum<1, kThreads>(red_v, red);
    if (done_in) return;
    // p.q == 0: the reference's break (MatrixOps.c:357); not finite: a NaN or
    // Inf in the system or the right-hand side (flag bit 1)
    if (red_v[0] == 0.0 || !isfinite(red_v[0])) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            status->done = 1;
            status->flags |= red_v[0] == 0.0 ? 1 : 2;
        }
        return;
    }
    const double alpha = rz / red_v[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *alpha_out = alpha;
        status->x_iter = k + 1;
    }
    double v[2] = {0.0, 0.0};
    for (; i < n; i += gridDim.x * blockDim.x) {
        if (i != static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x)) load(i);
        const int64_t o = 3 * static_cast<int64_t>(i);
        double rr[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            rr[c] = rp[c] - alpha * qp[c];
            stream_store<kNtVec>(rr[c], r + o + c);
        }
        double zz[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) zz[a] = di[a] * rr[0] + di[3 + a] * rr[1] + di[6 + a] * rr[2];
        z[o] = zz[0];
        z[o + 1] = zz[1];
        z[o + 2] = zz[2];
        v[0] += rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2];
        v[1] += rr[0] * zz[0] + rr[1] * zz[1] + rr[2] * zz[2];
    }
    write_partials<2>(v, pr_out, kPartStride);
}

// One launch L per PCG iteration for graphs of at most kFusedMaxTiles SpMV
// tiles (one, two or four tiles per workgroup), where the dependent
// launches of the two-launch form, not the bytes, bound the iteration.
// Launch L completes iteration L-1 and starts iteration L of the recurrence
// above:
//   from the partials of launch L-1 -- p.q, r.z, q.z, q.w and r.r of
//   p_(L-1), q_(L-1) = A p_(L-1), r_(L-1), z_(L-1) = D^-1 r_(L-1) and
//   w_(L-1) = D^-1 q_(L-1) --, reduced by every workgroup in the same order:
//     stop if ||r_(L-1)|| < tol (L >= 2: the test after each update,
//     MatrixOps.c:366), or if p.q == 0 (MatrixOps.c:357);
//     alpha = r.z / p.q;
//     r.z_L = r.z - 2 alpha q.z + alpha^2 q.w, which is r_L . D^-1 r_L for
//       r_L = r_(L-1) - alpha q_(L-1) (D^-1 is symmetric); beta = r.z_L / r.z;
//   then, for every block, V p_L = V z_(L-1) - alpha V w_(L-1) +
//   beta V p_(L-1) from the record {z, w, p} of launch L-1 of its column
//   (one 72-byte run per pose; the three products formed as the loads
//   arrive), summed per row into q_L = A p_L; for its own rows
//   p_L = (z_(L-1) - alpha w_(L-1)) + beta p_(L-1), r_L,
//   x_L = x_(L-1) + alpha p_(L-1), z_L = D^-1 r_L and w_L = D^-1 q_L (the
//   pose's three entries through LDS), the record of launch L and its
//   partials.  (q_L is A p_L up to the rounding of the products' order, as
//   in the two-launch SpMV, which forms V z + beta V p_old.)
// z_(L-1) - alpha w_(L-1) is D^-1 r_L up to the rounding of this one step
// (z is recomputed from r by every launch, so nothing accumulates); r.z_L is
// reduced from the recomputed z_L.  Launch 0 starts from r_(-1) = b,
// q_(-1) = 0, the record {D^-1 b, 0, 0} and x = 0 (pcg_fused_setup_kernel)
// with alpha = beta = 0.  r, q and the records are double-buffered by the
// parity of L (other workgroups gather launch L-1's while this one writes
// launch L's).
//
// beta comes from the one-step identity instead of a reduction of r_L . z_L,
// which would need a second grid-wide synchronisation; anchored on the reduced
// r.z_(L-1) every iteration, its rounding does not accumulate.  CPU study (the
// same block-Jacobi PCG in FP64, iterations to the tolerance and rel dx error
// against the refined direct solve, two-launch recurrence in brackets):
// grid-city N=10k at tol 1e-13 11,889 its, 6.5e-10 (11,967, 7.6e-10); park
// N=7120 1,500, 3.7e-13 (1,499, 2.6e-13); grid-city N=100k at 3e-13 153,799,
// 7.3e-9 (153,778, 5.7e-9).
__global__ void __launch_bounds__(kThreads)
pcg_fused_setup_kernel(const int32_t *__restrict__ bdiag, const double *__restrict__ bval,
                       const double *__restrict__ b, int32_t n, int32_t max_iter,
                       double *__restrict__ dinv, double *__restrict__ x,
                       double *__restrict__ r, double *__restrict__ q,
                       double *__restrict__ rec, double *__restrict__ part_out,
                       PcgStatus *status)
{
    double v[1] = {0.0};
    const int64_t dim = 3 * static_cast<int64_t>(n);
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double a[9], inv[9];
        const double *blk = bval + 9 * static_cast<int64_t>(bdiag[i]);
#pragma unroll
        for (int k = 0; k < 9; ++k) a[k] = blk[k];
        inv3_cofactor(a, inv);
        const int64_t o = 3 * static_cast<int64_t>(i);
#pragma unroll
        for (int aa = 0; aa < 3; ++aa)
#pragma unroll
            for (int c = 0; c < 3; ++c) dinv[c * dim + o + aa] = inv[3 * aa + c];
        const double b0 = b[o], b1 = b[o + 1], b2 = b[o + 2];
        double *e = rec + 9 * static_cast<int64_t>(i);
        e[0] = fma(inv[2], b2, fma(inv[1], b1, inv[0] * b0));
        e[1] = fma(inv[5], b2, fma(inv[4], b1, inv[3] * b0));
        e[2] = fma(inv[8], b2, fma(inv[7], b1, inv[6] * b0));
#pragma unroll
        for (int k = 3; k < 9; ++k) e[k] = 0.0;
        x[o] = 0.0; x[o + 1] = 0.0; x[o + 2] = 0.0;
        r[o] = b0; r[o + 1] = b1; r[o + 2] = b2;
        q[o] = 0.0; q[o + 1] = 0.0; q[o + 2] = 0.0;
        v[0] += b0 * b0 + b1 * b1 + b2 * b2;   // ||r0||^2 = ||b||^2
    }
    write_partials<1>(v, part_out, kPartStride);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        status->done = max_iter <= 0 ? 1 : 0;
        status->converged = 0;
        status->iters = 0;
        status->flags = 0;
        status->refreshes = 0;
        status->x_iter = -1;
        status->rnorm = 0.0;
        status->bnorm = 0.0;
        status->tol = 0.0;
    }
}

// p_L of one entry from its record {z, w, p} of launch L-1 (explicit fma, so
// that every workgroup that forms the entry rounds alike).
__device__ inline double fused_p(double z, double w, double p_old, double alpha, double beta)
{
    return fma(beta, p_old, fma(-alpha, w, z));
}

// base + byte offset, the offset a 32-bit value: the form the compiler turns
// into a scalar base register plus one 32-bit offset register (instead of a
// 64-bit address pair per load).
template <typename T>
__device__ inline T *at_byte(T *base, uint32_t off)
{
    using B = std::conditional_t<std::is_const_v<T>, const char, char>;
    return reinterpret_cast<T *>(reinterpret_cast<B *>(base) + off);
}

// A 3x3 block V (row-major) times the three parts of its column's record
// e = {z, w, p}: y = {V z, V w, V p}, formed as soon as the loads arrive,
// before alpha and beta are known (nine values held instead of eighteen).
__device__ inline void fused_products(const double (&v)[9], const double (&e)[9], double (&y)[9])
{
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            y[3 * s + r] = v[3 * r] * e[3 * s] + v[3 * r + 1] * e[3 * s + 1] + v[3 * r + 2] * e[3 * s + 2];
}

// Row r of V p_L = V z - alpha V w + beta V p.
__device__ inline double fused_combine(const double (&y)[9], int r, double alpha, double beta)
{
    return fma(beta, y[6 + r], fma(-alpha, y[3 + r], y[r]));
}

// kTpw consecutive tiles per workgroup of kTpw x kThreads threads: part
// w = tid / kThreads of the workgroup takes tile kTpw g + w of tile group g,
// with the tile kernel's lanes and LDS layout.  kTpw = 2 and 4 keep the
// partials at one per group and thread (the N = 100k grid city: 1,697
// tiles, 425 groups of four); held to 64 VGPRs, four groups of two or two
// of four share a CU, so all of them run in one round.  Groups map onto the
// XCDs contiguously (neighbouring rows in one L2; the dispatcher's round
// robin measured 16.5 against 15.8 us per iteration at N = 100k), and the
// partials are indexed by group.
template <int kTpw>
__global__ void __launch_bounds__(kTpw * kThreads)
__attribute__((amdgpu_waves_per_eu(kTpw == 1 ? 1 : 8, 8)))
pcg_fused_kernel(const int32_t *__restrict__ brow_ptr, const int32_t *__restrict__ bcol,
                 const double *__restrict__ bval, const int4 *__restrict__ tile_info,
                 const int32_t *__restrict__ tcol, const double *__restrict__ tval,
                 int32_t n_tiles, const double *__restrict__ dinv, int32_t n,
                 const double *__restrict__ r_in, const double *__restrict__ q_in,
                 const double *__restrict__ rec_in, double *__restrict__ r_out,
                 double *__restrict__ q_out, double *__restrict__ rec_out, double *__restrict__ x,
                 int32_t L, double tol_rel, const double *__restrict__ part_in, int32_t np_in,
                 double *__restrict__ part_out, PcgStatus *status)
{
    constexpr int NT = kTpw * kThreads;
    constexpr int kLoads = (kFusedMaxGroups + NT - 1) / NT;  // partials per thread and quantity
    static_assert(kTileBlocks == kThreads, "one lane per padded block slot");
    __shared__ double ys[kTpw][3 * kTileBlocks];
    __shared__ double xs[2][NT];  // the pose's r_L, then q_L entries
    __shared__ double red[kFusedParts][NT / kWave];
    __shared__ int32_t more[kTpw];  // chunks of a block row longer than a tile, per part
    const int tid = static_cast<int>(threadIdx.x);
    const int w = tid / kThreads, lt = tid % kThreads;
    const int32_t grp = xcd_contiguous(static_cast<int32_t>(blockIdx.x), static_cast<int32_t>(gridDim.x));
    const int32_t tw = kTpw * grp + w;
    const bool live = tw < n_tiles;  // the last group may hold fewer than kTpw tiles
    const int32_t t = live ? tw : n_tiles - 1;
    const int64_t dim = 3 * static_cast<int64_t>(n);
    const int32_t done_in = status->done;
    const double tol_in = status->tol;
    // the partials of launch L-1 first (launch 0: ||b||^2 of the setup
    // kernel, np_in <= kFusedMaxGroups), as unconditional loads from clamped
    // indices, summed below in the same order in every workgroup
    double pa[kLoads][kFusedParts];
    // (four tiles per workgroup: waves past the partials load none -- 9 of 16
    // at N = 100k, 15.88 -> 15.50 us per iteration; with one or two tiles
    // the branch costs more than the loads it saves, 5.47 -> 5.54 and 8.66
    // -> 9.02 us at N = 10k and 40k)
    const int wave_first = __builtin_amdgcn_readfirstlane(tid);
#pragma unroll
    for (int u = 0; u < kLoads; ++u) {
        const uint32_t iu = 8u * static_cast<uint32_t>(min(tid + u * NT, np_in - 1));
        if (kTpw < 4 || wave_first + u * NT < np_in) {
#pragma unroll
            for (int k = 0; k < kFusedParts; ++k)
                pa[u][k] = *at_byte(part_in + (L == 0 ? 0 : k * kFusedMaxGroups), iu);
        } else {
#pragma unroll
            for (int k = 0; k < kFusedParts; ++k) pa[u][k] = 0.0;
        }
    }
    asm volatile("" ::: "memory");  // keeps the loads above here (not sunk to their use)
    // this lane's block (tile-padded slot: zeros and a valid column in padding)
    // and the gathered record of its column, multiplied out on arrival
    const int4 ti = tile_info[t];
    double y[9];
    {
        // (32-bit byte offsets: the tile copy holds at most kFusedMaxTiles x
        // 18 KB, the records at most 72 B x ~120k poses)
        const uint32_t slot = static_cast<uint32_t>(t) * kTileBlocks + lt;
        const uint32_t col = static_cast<uint32_t>(*at_byte(tcol, 4u * slot));
        // the partials return before the column index (loads complete in
        // issue order): their wave totals go to LDS meanwhile, which frees
        // their registers; the block's values are loaded with the gathered
        // record, whose address waits for the column index anyway
        double pv[kFusedParts];
#pragma unroll
        for (int k = 0; k < kFusedParts; ++k) {
            pv[k] = 0.0;
#pragma unroll
            for (int u = 0; u < kLoads; ++u) pv[k] += tid + u * NT < np_in ? pa[u][k] : 0.0;
        }
        wave_totals<kFusedParts, NT>(pv, red);
        asm volatile("" ::: "memory");
        const uint32_t tv = 8u * (static_cast<uint32_t>(t) * (9u * kTileBlocks) + lt);
        double v[9], e[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) v[k] = *at_byte(tval + k * kTileBlocks, tv);
#pragma unroll
        for (int k = 0; k < 9; ++k) e[k] = *at_byte(rec_in + k, 72u * col);
        fused_products(v, e, y);
    }
    if constexpr (kTpw > 1) {
        // the products are formed here (not sunk to their use after the
        // reduction), so that their 18 inputs are dead before the own-row
        // loads
#pragma unroll
        for (int k = 0; k < 9; ++k) asm volatile("" ::"v"(y[k]));
    }
    // this thread's output entry (row a of pose orow, clamped into the tile);
    // its row range first, its values (used after the row sums) early in the
    // one-tile form and after the block products in the wide one, which
    // keeps the products' 18 inputs and the partials out of the registers
    // meanwhile (64 VGPRs, no spills)
    const int a = lt % 3;
    const bool has_blk = live && ti.z + lt < ti.w, has_out = live && lt < 3 * (ti.y - ti.x);
    const int32_t orow = min(ti.x + lt / 3, ti.y - 1);
    const int32_t ob0 = brow_ptr[orow], ob1 = brow_ptr[orow + 1];
    const int64_t o = 3 * static_cast<int64_t>(orow) + a;
    double oz, ow, op, orr, oq, ox, od0, od1, od2;
    // (32-bit byte offsets into the vectors: 3N <= 3 x ~120k entries)
    const uint32_t ob = 8u * static_cast<uint32_t>(o);
    const uint32_t oeb = 8u * (9u * static_cast<uint32_t>(orow) + static_cast<uint32_t>(a));
    auto load_own = [&]() {
        oz = *at_byte(rec_in, oeb);
        ow = *at_byte(rec_in + 3, oeb);
        op = *at_byte(rec_in + 6, oeb);
        orr = *at_byte(r_in, ob);
        oq = *at_byte(q_in, ob);
        ox = *at_byte(static_cast<const double *>(x), ob);
        od0 = *at_byte(dinv, ob);  // row a of D^-1
        od1 = *at_byte(dinv + dim, ob);
        od2 = *at_byte(dinv + 2 * dim, ob);
    };
    if constexpr (kTpw == 1) load_own();
    if (lt == 0) more[w] = live ? max(0, (ti.w - ti.z - 1) / kTileBlocks) : 0;
    double pv[kFusedParts];
    __syncthreads();
    sum_wave_totals<kFusedParts, NT>(pv, red);
    if (done_in) return;
    const double pq = pv[0], rz = pv[1], qz = pv[2], qw = pv[3], rr = L == 0 ? pv[0] : pv[4];
    double alpha = 0.0, beta = 0.0;
    const bool lead = grp == 0 && tid == 0;
    if (L == 0) {
        if (lead) {
            status->bnorm = sqrt(rr);
            status->tol = tol_rel * sqrt(rr);
            status->rnorm = sqrt(rr);
        }
    } else {
        if (!isfinite(rr) || !isfinite(rz)) {  // NaN / Inf in the residual (flag bit 1)
            if (lead) {
                status->done = 1;
                status->flags |= 2;
                status->iters = L - 1;
            }
            return;
        }
        if (L >= 2 && sqrt(rr) < tol_in) {  // the test after update L-1
            if (lead) {
                status->done = 1;
                status->converged = 1;
                status->iters = L - 1;
                status->rnorm = sqrt(rr);
            }
            return;
        }
        if (pq == 0.0 || !isfinite(pq)) {  // MatrixOps.c:357 (flag bit 0), or NaN / Inf
            if (lead) {
                status->done = 1;
                status->flags |= pq == 0.0 ? 1 : 2;
                status->iters = L - 1;
            }
            return;
        }
        alpha = rz / pq;
        beta = fma(alpha * alpha, qw, fma(-2.0 * alpha, qz, rz)) / rz;
    }
    // V p_L of this lane's block into LDS
    double *yl = ys[w];
    if (has_blk) {
#pragma unroll
        for (int r = 0; r < 3; ++r) yl[r * kTileBlocks + lt] = fused_combine(y, r, alpha, beta);
    }
    if constexpr (kTpw > 1) load_own();
    __syncthreads();
    double acc = 0.0;
    if (has_out) {
        const int32_t hi = min(ob1, ti.z + kTileBlocks);
        for (int32_t c = max(ob0, ti.z); c < hi; ++c) acc += yl[a * kTileBlocks + (c - ti.z)];
    }
    // a single block row longer than one tile: further chunks from the BSR
    // arrays, as many rounds as the longest of the workgroup's tiles needs
    // (the same count, hence the same barriers, in every thread)
    int32_t n_more = 0;
#pragma unroll
    for (int i = 0; i < kTpw; ++i) n_more = max(n_more, more[i]);
    if (n_more > 0) __syncthreads();
    for (int32_t m = 1; m <= n_more; ++m) {
        const int32_t cb = ti.z + m * kTileBlocks;
        const int32_t bb = cb + lt;
        if (live && bb < ti.w) {
            // (rows this long are rare -- a pose with more than kTileBlocks
            // edge ends --: three values per round trip, so that this path
            // needs few registers beside the own-row values; p_L as fused_p
            // rounds it)
            const double *ee = rec_in + 9 * static_cast<int64_t>(bcol[bb]);
            double pj[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) pj[c] = ee[c];
            asm volatile("" ::: "memory");
#pragma unroll
            for (int c = 0; c < 3; ++c) pj[c] = fma(-alpha, ee[3 + c], pj[c]);
            asm volatile("" ::: "memory");
#pragma unroll
            for (int c = 0; c < 3; ++c) pj[c] = fma(beta, ee[6 + c], pj[c]);
            const double *bv = bval + 9 * static_cast<int64_t>(bb);
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                asm volatile("" ::: "memory");
                yl[r * kTileBlocks + lt] = bv[3 * r] * pj[0] + bv[3 * r + 1] * pj[1] + bv[3 * r + 2] * pj[2];
            }
        }
        __syncthreads();
        if (has_out) {
            const int32_t lo = max(ob0, cb), hi = min(ob1, cb + kTileBlocks);
            for (int32_t c = lo; c < hi; ++c) acc += yl[a * kTileBlocks + (c - cb)];
        }
        __syncthreads();
    }
    // own entry: r_L and p_L; r_L and q_L to LDS for z_L = D^-1 r_L and
    // w_L = D^-1 q_L
    const double r_a = fma(-alpha, oq, orr);
    const double p_a = fused_p(oz, ow, op, alpha, beta);
    xs[0][tid] = r_a;
    // (both D^-1 products from the pose's three entries: tid - a is the
    // pose's first thread, in the same tile's part)
    xs[1][tid] = acc;
    __syncthreads();
    double part[kFusedParts] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (has_out) {
        const int base = tid - a;
        const double z_a = fma(od2, xs[0][base + 2], fma(od1, xs[0][base + 1], od0 * xs[0][base]));
        const double w_a = fma(od2, xs[1][base + 2], fma(od1, xs[1][base + 1], od0 * xs[1][base]));
        *at_byte(r_out, ob) = r_a;
        *at_byte(q_out, ob) = acc;
        *at_byte(rec_out, oeb) = z_a;
        *at_byte(rec_out + 3, oeb) = w_a;
        *at_byte(rec_out + 6, oeb) = p_a;
        if (L > 0) *at_byte(x, ob) = fma(alpha, op, ox);
        part[0] = p_a * acc;
        part[1] = r_a * z_a;
        part[2] = acc * z_a;
        part[3] = acc * w_a;
        part[4] = r_a * r_a;
    }
    write_partials<kFusedParts, NT>(part, part_out, kFusedMaxGroups, grp);
    if (lead) status->iters = L;
}

// After the last launch of a chunk (launches = L + 1): the stopping test on
// r_L for the host, and the iteration cap.
__global__ void __launch_bounds__(kThreads)
pcg_fused_finalize_kernel(const double *__restrict__ part_last, int32_t np, int32_t launches,
                          int32_t max_iter, PcgStatus *status)
{
    if (status->done || launches < 2) return;  // launch 0's partials hold r_0 = b
    double cur[1];
    sum_partials<1>(part_last + 4 * kFusedMaxGroups, np, kFusedMaxGroups, cur);
    if (threadIdx.x == 0) {
        const int32_t it = launches - 1;
        status->iters = it;
        status->rnorm = sqrt(cur[0]);
        if (sqrt(cur[0]) < status->tol) {
            status->done = 1;
            status->converged = 1;
        } else if (it >= max_iter) {
            status->done = 1;
        }
    }
}

// After the last launched iteration (or at a poll): the stopping test on the
// newest partials, for the host.
__global__ void __launch_bounds__(kThreads)
pcg_finalize_kernel(const double *__restrict__ pr_cur, int32_t np_cur, int32_t k,
                    int32_t max_iter, PcgStatus *status)
{
    if (status->done || k == 0) return;
    double cur[2];
    sum_partials<2>(pr_cur, np_cur, kPartStride, cur);
    if (threadIdx.x == 0) {
        status->iters = k;
        status->rnorm = sqrt(cur[0]);
        if (sqrt(cur[0]) < status->tol) {
            status->done = 1;
            status->converged = 1;
        } else if (k >= max_iter) {
            status->done = 1;
        }
    }
}

// After the last launched iteration K: x_K = x_(K-1) + alpha_(K-1) p_(K-1)
// when no A(K) ran to apply it (cap, or a stop found by the finalize test).
__global__ void __launch_bounds__(kThreads)
pcg_xfinish_kernel(double *__restrict__ x, const double *__restrict__ p_last,
                   const double *__restrict__ alpha_last, int64_t m, int32_t k,
                   const PcgStatus *status)
{
    if (status->x_iter != k) return;
    const double alpha = *alpha_last;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < m;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        x[i] = x[i] + alpha * p_last[i];
}

// y = A x (FP64) and, if b != NULL, ||b - y||^2 / ||b||^2 partials.
__global__ void __launch_bounds__(kThreads)
spmv64_kernel(const int32_t *__restrict__ brow_ptr, const int32_t *__restrict__ bcol,
              const double *__restrict__ bval, const double *__restrict__ x,
              const double *__restrict__ b, double *__restrict__ y, int32_t n,
              double *__restrict__ part_out)
{
    constexpr int G = kLanesPerRow, kRowsPerTile = kThreads / G;
    const int lane = threadIdx.x % G;
    double s[2] = {0.0, 0.0};
    for (int64_t row0 = static_cast<int64_t>(blockIdx.x) * kRowsPerTile; row0 < n;
         row0 += static_cast<int64_t>(gridDim.x) * kRowsPerTile) {
        const int64_t row = row0 + threadIdx.x / G;
        double y0 = 0.0, y1 = 0.0, y2 = 0.0;
        if (row < n) {
            const int32_t b1 = brow_ptr[row + 1];
            for (int32_t bb = brow_ptr[row] + lane; bb < b1; bb += G) {
                const int64_t j = 3 * static_cast<int64_t>(bcol[bb]);
                const double *v = bval + 9 * static_cast<int64_t>(bb);
                y0 += v[0] * x[j] + v[1] * x[j + 1] + v[2] * x[j + 2];
                y1 += v[3] * x[j] + v[4] * x[j + 1] + v[5] * x[j + 2];
                y2 += v[6] * x[j] + v[7] * x[j + 1] + v[8] * x[j + 2];
            }
        }
        y0 = group_sum<G>(y0);
        y1 = group_sum<G>(y1);
        y2 = group_sum<G>(y2);
        if (row < n && lane == 0) {
            const int64_t o = 3 * row;
            if (y) { y[o] = y0; y[o + 1] = y1; y[o + 2] = y2; }
            if (b) {
                const double d0 = b[o] - y0, d1 = b[o + 1] - y1, d2 = b[o + 2] - y2;
                s[0] += d0 * d0 + d1 * d1 + d2 * d2;
                s[1] += b[o] * b[o] + b[o + 1] * b[o + 1] + b[o + 2] * b[o + 2];
            }
        }
    }
    if (b) write_partials<2>(s, part_out, kPartStride);
}

__global__ void __launch_bounds__(kThreads)
ratio_kernel(const double *__restrict__ part, int32_t np, double *__restrict__ out)
{
    double t[2];
    sum_partials<2>(part, np, kPartStride, t);
    if (threadIdx.x == 0) *out = t[1] > 0.0 ? sqrt(t[0]) / sqrt(t[1]) : sqrt(t[0]);
}

// poses += dx (GSLAM.c:98-102), optional theta wrap; ||dx||^2 partials.
__global__ void __launch_bounds__(kThreads)
pose_update_kernel(double *__restrict__ poses, const double *__restrict__ dx, int32_t n,
                   int wrap, double *__restrict__ part_out)
{
    double s[1] = {0.0};
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int64_t o = 3 * static_cast<int64_t>(i);
        const double d0 = dx[o], d1 = dx[o + 1], d2 = dx[o + 2];
        poses[o] = poses[o] + d0;
        poses[o + 1] = poses[o + 1] + d1;
        double th = poses[o + 2] + d2;
        if (wrap) th = wrap_pi(th);
        poses[o + 2] = th;
        s[0] += d0 * d0 + d1 * d1 + d2 * d2;
    }
    write_partials<1>(s, part_out, kPartStride);
}

__global__ void __launch_bounds__(kThreads)
diff_kernel(const double *__restrict__ part, int32_t np, int32_t n, double *__restrict__ out)
{
    double t[1];
    sum_partials<1>(part, np, kPartStride, t);
    if (threadIdx.x == 0) *out = sqrt(t[0]) / sqrt(3.0 * static_cast<double>(n));
}

// The SpMV and update kernels of this graph: nontemporal variants where the
// tile copy (76 B per slot) or D^-1 (72 B per pose) exceeds the Infinity
// Cache, and where the six FP64 solver vectors (144 B per pose) do.
using SpmvKernel = decltype(&pcg_spmv_kernel<false, false>);
using UpdateKernel = decltype(&pcg_update_kernel<false, false>);

static bool vectors_stream(const DeviceGraph &g)
{
    return streams_past_cache(144 * static_cast<int64_t>(g.h.n));
}

static SpmvKernel spmv_kernel_for(const DeviceGraph &g)
{
    const bool mat = streams_past_cache(76 * int64_t{kTileBlocks} * g.n_tiles);
    if (vectors_stream(g)) return mat ? pcg_spmv_kernel<true, true> : pcg_spmv_kernel<false, true>;
    return mat ? pcg_spmv_kernel<true, false> : pcg_spmv_kernel<false, false>;
}

static UpdateKernel update_kernel_for(const DeviceGraph &g)
{
    const bool dinv = streams_past_cache(72 * static_cast<int64_t>(g.h.n));
    if (vectors_stream(g)) return dinv ? pcg_update_kernel<true, true> : pcg_update_kernel<false, true>;
    return dinv ? pcg_update_kernel<true, false> : pcg_update_kernel<false, false>;
}

// The padded tile copy of the current block values (before each FP64 solve).
static int launch_tile_pack(DeviceGraph &g)
{
    GSLAM_TRY(g.tval_reserve());
    const int64_t n = static_cast<int64_t>(g.n_tiles) * kTileBlocks * 9;
    const unsigned int grid = static_cast<unsigned int>(std::min<int64_t>(4096, (n + kThreads - 1) / kThreads));
    hipLaunchKernelGGL(tile_pack_kernel, dim3(grid), dim3(kThreads), 0, g.stream, g.tile_info, g.bval,
                       g.n_tiles, g.tval);
    return GSLAM_OK;
}

int launch_spmv64(DeviceGraph &g, const double *x, double *y)
{
    hipLaunchKernelGGL(spmv64_kernel, dim3(grid_for(g.h.n, kLanesPerRow)), dim3(kThreads), 0,
                       g.stream, g.brow_ptr, g.bcol, g.bval, x, nullptr, y, g.h.n, g.part);
    GSLAM_HIP(hipGetLastError());
    return GSLAM_OK;
}

int enqueue_true_residual(DeviceGraph &g, const double *b, const double *x, double *dst)
{
    const unsigned int gr = grid_for(g.h.n, kLanesPerRow);
    hipLaunchKernelGGL(spmv64_kernel, dim3(gr), dim3(kThreads), 0, g.stream, g.brow_ptr, g.bcol,
                       g.bval, x, b, nullptr, g.h.n, g.part + kPartOffResid);
    hipLaunchKernelGGL(ratio_kernel, dim3(1), dim3(kThreads), 0, g.stream,
                       g.part + kPartOffResid, static_cast<int32_t>(gr), dst);
    GSLAM_HIP(hipGetLastError());
    return GSLAM_OK;
}

int true_rel_residual(DeviceGraph &g, const double *b, const double *x, double *out)
{
    GSLAM_TRY(enqueue_true_residual(g, b, x, g.scal + 2));
    GSLAM_HIP(hipMemcpyAsync(g.scal_host + 2, g.scal + 2, sizeof(double), hipMemcpyDeviceToHost,
                             g.stream));
    GSLAM_HIP(hipStreamSynchronize(g.stream));
    *out = g.scal_host[2];
    return GSLAM_OK;
}

int pose_update(DeviceGraph &g, double *diff_out)
{
    const int32_t n = g.h.n;
    const unsigned int gr = grid_for(n, 1);
    hipLaunchKernelGGL(pose_update_kernel, dim3(gr), dim3(kThreads), 0, g.stream, g.poses, g.dx,
                       n, g.wrap, g.part + kPartOffResid);
    hipLaunchKernelGGL(diff_kernel, dim3(1), dim3(kThreads), 0, g.stream, g.part + kPartOffResid,
                       static_cast<int32_t>(gr), n, g.scal + 1);
    GSLAM_HIP(hipGetLastError());
    GSLAM_HIP(hipMemcpyAsync(g.scal_host + 1, g.scal + 1, sizeof(double), hipMemcpyDeviceToHost,
                             g.stream));
    GSLAM_HIP(hipStreamSynchronize(g.stream));
    if (diff_out) *diff_out = g.scal_host[1];
    return GSLAM_OK;
}

// One kernel per iteration (pcg_fused_kernel) for graphs of at most
// kFusedMaxTiles tiles: one tile per workgroup up to kFusedMaxGroups tiles,
// two up to twice that, four beyond -- the fewest tiles per workgroup that
// keep the partials at one per thread and quantity (a second partial load
// per thread cost ~1 us per iteration at N = 40k, more than the larger
// workgroups' barriers: 8.65 us against 9.42 with four tiles and 9.69 with
// two tiles and two loads).
static int pcg_solve_fused(DeviceGraph &g, const gslam_solve_opts &o, const double *b,
                           double *x_out, gslam_solve_stats &s)
{
    const int32_t n = g.h.n;
    const unsigned int ge = grid_for(n, 1);
    const int tpw = g.n_tiles <= kFusedMaxGroups ? 1 : (g.n_tiles <= 2 * kFusedMaxGroups ? 2 : 4);
    const unsigned int gt = static_cast<unsigned int>((g.n_tiles + tpw - 1) / tpw);
    const dim3 blk(tpw * kThreads);
    auto kern = tpw == 1 ? pcg_fused_kernel<1> : (tpw == 2 ? pcg_fused_kernel<2> : pcg_fused_kernel<4>);
    PcgStatus *st = g.status;
    GSLAM_TRY(g.tval_reserve());  // first solve: outside the timed region
    GSLAM_TRY(g.fused_reserve());
    double *R[2] = {g.r, g.r_alt}, *Q[2] = {g.q, g.q_alt};
    double *E[2] = {g.zwp, g.zwp + 3 * g.dim()};
    double *FP[2] = {g.fpart, g.fpart + kFusedParts * kFusedMaxGroups};
    GSLAM_HIP(hipEventRecord(g.ev0, g.stream));
    GSLAM_TRY(launch_tile_pack(g));
    // launch 0 reads r_(-1) = b, q_(-1) = 0, the record {D^-1 b, 0, 0} (parity
    // 1) and x = 0, and ||b||^2 from the partials of the setup kernel
    hipLaunchKernelGGL(pcg_fused_setup_kernel, dim3(ge), dim3(kThreads), 0, g.stream, g.bdiag,
                       g.bval, b, n, o.max_iter, g.dinv, x_out, R[1], Q[1], E[1],
                       g.part + kPartOffPR0, st);
    GSLAM_HIP(hipGetLastError());
    GSLAM_HIP(hipEventRecord(g.ev2, g.stream));  // end of the setup phase
    constexpr int kMaxSamples = 128;
    int n_samp = 0;
    if (o.profile_every > 0 && g.prof_ev.empty()) {
        g.prof_ev.resize(4 * kMaxSamples, nullptr);
        for (hipEvent_t &e : g.prof_ev) GSLAM_HIP(hipEventCreate(&e));
    }
    int32_t k = 0;  // launches enqueued: launch L performs update L (L >= 1)
    auto enqueue = [&](int32_t m) {
        for (int32_t c = 0; c < m; ++c, ++k) {
            const int cur = k & 1, prev = cur ^ 1;
            const double *pin = k == 0 ? g.part + kPartOffPR0 : FP[prev];
            const int32_t npin = k == 0 ? static_cast<int32_t>(ge) : static_cast<int32_t>(gt);
            const bool samp = o.profile_every > 0 && k % o.profile_every == 0 && n_samp < kMaxSamples;
            auto launch = [&](auto... a) {
                if (samp) {
                    hipEvent_t *ev = &g.prof_ev[4 * n_samp++];
                    hipExtLaunchKernelGGL(kern, dim3(gt), blk, 0, g.stream, ev[0], ev[1], 0, a...);
                } else {
                    hipLaunchKernelGGL(kern, dim3(gt), blk, 0, g.stream, a...);
                }
            };
            launch(g.brow_ptr, g.bcol, g.bval, g.tile_info, g.tcol, g.tval, g.n_tiles, g.dinv, n,
                   R[prev], Q[prev], E[prev], R[cur], Q[cur], E[cur], x_out, k, o.tol, pin, npin,
                   FP[cur], st);
        }
        hipLaunchKernelGGL(pcg_fused_finalize_kernel, dim3(1), dim3(kThreads), 0, g.stream,
                           FP[(k - 1) & 1], static_cast<int32_t>(gt), k, o.max_iter, st);
    };
    // max_iter updates take max_iter + 1 launches
    gslam_solve_opts oc = o;
    oc.max_iter = o.max_iter < INT32_MAX ? o.max_iter + 1 : o.max_iter;
    GSLAM_TRY(run_chunked(g, oc, enqueue, k));
    GSLAM_HIP(hipEventRecord(g.ev1, g.stream));
    GSLAM_HIP(hipMemcpyAsync(g.status_host, st, sizeof(PcgStatus), hipMemcpyDeviceToHost, g.stream));
    GSLAM_HIP(hipStreamSynchronize(g.stream));
    float ms = 0.f, ms_setup = 0.f;
    GSLAM_HIP(hipEventElapsedTime(&ms, g.ev0, g.ev1));
    GSLAM_HIP(hipEventElapsedTime(&ms_setup, g.ev0, g.ev2));
    s.ms_setup = ms_setup;
    const PcgStatus &ps = *g.status_host;
    s.iters = ps.iters;
    s.converged = ps.converged;
    s.flags = ps.flags;
    s.reliable_updates = 0;
    s.rel_res = ps.bnorm > 0 ? ps.rnorm / ps.bnorm : ps.rnorm;
    s.ms = ms;
    s.spmv_us = s.update_us = 0.0;
    if (n_samp > 0) {
        // samples past the converged iteration timed early-exit launches
        const int use = std::max(1, std::min(n_samp, (ps.iters + o.profile_every - 1) / o.profile_every));
        double a = 0.0;
        for (int i = 0; i < use; ++i) {
            float t0 = 0.f;
            GSLAM_HIP(hipEventElapsedTime(&t0, g.prof_ev[4 * i], g.prof_ev[4 * i + 1]));
            a += t0;
        }
        s.spmv_us = 1000.0 * a / use;  // the whole iteration's one launch
    }
    GSLAM_TRY(true_rel_residual(g, b, x_out, &s.true_rel_res));
    return GSLAM_OK;
}

int pcg_solve_fp64(DeviceGraph &g, const gslam_solve_opts &o, const double *b, double *x_out,
                   gslam_solve_stats &s)
{
    if (o.pcg_schedule == GSLAM_PCG_AUTO && g.n_tiles <= kFusedMaxTiles)
        return pcg_solve_fused(g, o, b, x_out, s);
    const int32_t n = g.h.n;
    const unsigned int ge = grid_for(n, 1);
    const unsigned int gt = static_cast<unsigned int>(std::min<int32_t>(g.n_tiles, kSpmvMaxGrid));
    double *PR[2] = {g.part + kPartOffPR0, g.part + kPartOffPR1};
    double *PQ = g.part + kPartOffPQT;
    PcgStatus *st = g.status;
    GSLAM_TRY(g.tval_reserve());  // first FP64 solve: outside the timed region
    GSLAM_HIP(hipEventRecord(g.ev0, g.stream));
    GSLAM_TRY(launch_tile_pack(g));
    hipLaunchKernelGGL(pcg_setup_kernel, dim3(ge), dim3(kThreads), 0, g.stream, g.bdiag, g.bval, b,
                       n, o.max_iter, g.dinv, static_cast<float *>(nullptr), x_out, g.r, g.z,
                       g.p[0], g.p[1], PR[0], st);
    GSLAM_HIP(hipGetLastError());
    GSLAM_HIP(hipEventRecord(g.ev2, g.stream));  // end of the setup phase
    int32_t k = 0;
    // optional in-solve timing: events around the SpMV and the update of
    // every profile_every-th iteration (4 events per sample)
    constexpr int kMaxSamples = 128;
    int n_samp = 0;
    if (o.profile_every > 0 && g.prof_ev.empty()) {
        g.prof_ev.resize(4 * kMaxSamples, nullptr);
        for (hipEvent_t &e : g.prof_ev) GSLAM_HIP(hipEventCreate(&e));
    }
    const SpmvKernel spmv_kern = spmv_kernel_for(g);
    const UpdateKernel update_kern = update_kernel_for(g);
    auto enqueue_iters = [&](int32_t m) {
        for (int32_t c = 0; c < m; ++c, ++k) {
            double *pold = g.p[k & 1], *pnew = g.p[(k + 1) & 1];
            const bool samp = o.profile_every > 0 && k % o.profile_every == 0 && n_samp < kMaxSamples;
            hipEvent_t *ev = samp ? &g.prof_ev[4 * n_samp++] : nullptr;
            auto spmv_args = [&](auto launch) {
                launch(spmv_kern, g.brow_ptr, g.bcol, g.bval, g.tile_info, g.tcol, g.tval,
                       g.n_tiles, g.z,
                       pold, pnew, g.q, k, o.tol, PR[k & 1], static_cast<int32_t>(ge),
                       g.part + kPartOffRz, PQ, x_out, g.part + kPartOffAlpha, st);
            };
            auto update_args = [&](auto launch) {
                launch(update_kern, g.q, g.dinv, g.r, g.z, n, k, PQ,
                       static_cast<int32_t>(gt), g.part + kPartOffRz + (k & 1),
                       PR[(k + 1) & 1], g.part + kPartOffAlpha + (k & 1), st);
            };
            const dim3 blk(kThreads);
            if (ev) {
                // the events are tied to the kernel's own dispatch packet
                spmv_args([&](auto kern, auto... a) {
                    hipExtLaunchKernelGGL(kern, dim3(gt), blk, 0, g.stream, ev[0], ev[1], 0, a...);
                });
                update_args([&](auto kern, auto... a) {
                    hipExtLaunchKernelGGL(kern, dim3(ge), blk, 0, g.stream, ev[2], ev[3], 0, a...);
                });
            } else {
                spmv_args([&](auto kern, auto... a) {
                    hipLaunchKernelGGL(kern, dim3(gt), blk, 0, g.stream, a...);
                });
                update_args([&](auto kern, auto... a) {
                    hipLaunchKernelGGL(kern, dim3(ge), blk, 0, g.stream, a...);
                });
            }
        }
        hipLaunchKernelGGL(pcg_finalize_kernel, dim3(1), dim3(kThreads), 0, g.stream, PR[k & 1],
                           static_cast<int32_t>(ge), k, o.max_iter, st);
    };
    GSLAM_TRY(run_chunked(g, o, enqueue_iters, k));
    if (k > 0)
        hipLaunchKernelGGL(pcg_xfinish_kernel, dim3(grid_for(3 * static_cast<int64_t>(n), 1)),
                           dim3(kThreads), 0, g.stream, x_out, g.p[k & 1],
                           g.part + kPartOffAlpha + ((k - 1) & 1), 3 * static_cast<int64_t>(n), k,
                           st);
    GSLAM_HIP(hipEventRecord(g.ev1, g.stream));
    GSLAM_HIP(hipMemcpyAsync(g.status_host, st, sizeof(PcgStatus), hipMemcpyDeviceToHost, g.stream));
    GSLAM_HIP(hipStreamSynchronize(g.stream));
    float ms = 0.f, ms_setup = 0.f;
    GSLAM_HIP(hipEventElapsedTime(&ms, g.ev0, g.ev1));
    GSLAM_HIP(hipEventElapsedTime(&ms_setup, g.ev0, g.ev2));
    s.ms_setup = ms_setup;
    const PcgStatus &ps = *g.status_host;
    s.iters = ps.iters;
    s.converged = ps.converged;
    s.flags = ps.flags;
    s.reliable_updates = 0;
    s.rel_res = ps.bnorm > 0 ? ps.rnorm / ps.bnorm : ps.rnorm;
    s.ms = ms;
    s.spmv_us = s.update_us = 0.0;
    if (n_samp > 0) {
        // samples past the converged iteration timed early-exit launches
        const int use = std::max(1, std::min(n_samp, (ps.iters + o.profile_every - 1) / o.profile_every));
        double a = 0.0, u = 0.0;
        for (int i = 0; i < use; ++i) {
            float t0 = 0.f, t1 = 0.f;
            GSLAM_HIP(hipEventElapsedTime(&t0, g.prof_ev[4 * i], g.prof_ev[4 * i + 1]));
            GSLAM_HIP(hipEventElapsedTime(&t1, g.prof_ev[4 * i + 2], g.prof_ev[4 * i + 3]));
            a += t0;
            u += t1;
        }
        s.spmv_us = 1000.0 * a / use;
        s.update_us = 1000.0 * u / use;
    }
    GSLAM_TRY(true_rel_residual(g, b, x_out, &s.true_rel_res));
    return GSLAM_OK;
}

// Size of the next chunk: doubling up to the cap, but, once the residual
// falls, no more than the iterations still predicted beyond the chunk in
// flight (at least 8), at the faster of the average rate since the start and
// the rate between the last two polls (CG speeds up towards the end).
// Launches after the stop exit at once but still cost ~2-3 us each, and up
// to two full chunks of them would otherwise follow a short solve (park N =
// 7120: ~300 of 1,475 FP64 iterations; the mixed solver's 128-iteration FP32
// phase: ~370).  The estimate only shapes the chunks: the device decides
// when to stop.
static int32_t next_chunk(const PcgStatus &ps, const PcgStatus *prev, int32_t chunk, int32_t cap,
                          int32_t in_flight)
{
    const int32_t next = std::min(chunk * 2, cap);
    if (ps.iters < 16 || !(ps.bnorm > 0.0) || !(ps.rnorm > ps.tol) || !(ps.tol > 0.0)) return next;
    double rate = std::log(ps.rnorm / ps.bnorm) / ps.iters;  // < 0 while converging
    if (prev && prev->iters < ps.iters && prev->rnorm > 0.0)
        rate = std::min(rate, std::log(ps.rnorm / prev->rnorm) / (ps.iters - prev->iters));
    if (!(rate < 0.0)) return next;
    const double left = std::log(ps.tol / ps.rnorm) / rate - in_flight;
    const double est = left > 0.0 ? std::ceil(1.25 * left) : 0.0;
    return std::max<int32_t>(8, std::min<double>(est, next));
}

// Chunked launch loop shared by the solvers: enqueue chunks of iterations,
// keep one chunk in flight ahead of the host's status check so the GPU never
// idles on a synchronisation, stop when the device reports done or the cap.
int run_chunked(DeviceGraph &g, const gslam_solve_opts &o,
                const std::function<void(int32_t)> &enqueue, int32_t &k)
{
    int32_t chunk = 8;
    const bool fixed = o.check_every > 0;
    const int32_t cap = std::max(1, fixed ? o.check_every : 256);
    int slot = 0;
    bool pending = false, have_prev = false;
    PcgStatus prev{};
    while (k < o.max_iter) {
        const int32_t m = std::min(chunk, o.max_iter - k);
        enqueue(m);
        GSLAM_HIP(hipGetLastError());
        GSLAM_HIP(hipMemcpyAsync(g.status_ring + slot, g.status, sizeof(PcgStatus),
                                 hipMemcpyDeviceToHost, g.stream));
        GSLAM_HIP(hipEventRecord(g.ev_poll[slot], g.stream));
        int32_t next = std::min(chunk * 2, cap);
        if (pending) {
            GSLAM_HIP(hipEventSynchronize(g.ev_poll[slot ^ 1]));
            const PcgStatus ps = g.status_ring[slot ^ 1];
            if (ps.done) break;
            if (!fixed) next = next_chunk(ps, have_prev ? &prev : nullptr, chunk, cap, m);
            prev = ps;
            have_prev = true;
        }
        pending = true;
        slot ^= 1;
        chunk = next;
    }
    return GSLAM_OK;
}

int time_mixed_spmv(DeviceGraph &g, int reps, double *us);

int time_kernel(DeviceGraph &g, int which, int reps, double *us)
{
    const int32_t n = g.h.n;
    const unsigned int ge = grid_for(n, 1);
    const unsigned int gt = static_cast<unsigned int>(std::min<int32_t>(g.n_tiles, kSpmvMaxGrid));
    double *PR[2] = {g.part + kPartOffPR0, g.part + kPartOffPR1};
    double *PQ = g.part + kPartOffPQT;
    if (which == GSLAM_KERNEL_SPMV_MIXED) return time_mixed_spmv(g, reps, us);
    if (which < GSLAM_KERNEL_SPMV_FP64 || which > GSLAM_KERNEL_ASSEMBLY) return GSLAM_ERR_ARG;
    const SpmvKernel spmv_kern = spmv_kernel_for(g);
    const UpdateKernel update_kern = update_kernel_for(g);
    auto spmv = [&]() {
        hipLaunchKernelGGL(spmv_kern, dim3(gt), dim3(kThreads), 0, g.stream, g.brow_ptr,
                           g.bcol, g.bval, g.tile_info, g.tcol, g.tval, g.n_tiles, g.z, g.p[0],
                           g.p[1], g.q, 0,
                           1e-30, PR[0], static_cast<int32_t>(ge), g.part + kPartOffRz, PQ,
                           g.x, g.part + kPartOffAlpha, g.status);
    };
    auto update = [&]() {
        hipLaunchKernelGGL(update_kern, dim3(ge), dim3(kThreads), 0, g.stream, g.q,
                           g.dinv, g.r, g.z, n, 0, PQ, static_cast<int32_t>(gt),
                           g.part + kPartOffRz, PR[1], g.part + kPartOffAlpha, g.status);
    };
    auto run = [&](int rr) {
        for (int i = 0; i < rr; ++i) {
            if (which == GSLAM_KERNEL_SPMV_FP64) spmv();
            else if (which == GSLAM_KERNEL_UPDATE_FP64) update();
            else (void)launch_assembly(g);
        }
    };
    // a consistent solver state: setup, one SpMV (valid p.q partials)
    GSLAM_TRY(launch_tile_pack(g));
    hipLaunchKernelGGL(pcg_setup_kernel, dim3(ge), dim3(kThreads), 0, g.stream, g.bdiag, g.bval,
                       g.xi, n, 1, g.dinv, static_cast<float *>(nullptr), g.x, g.r, g.z, g.p[0],
                       g.p[1], PR[0], g.status);
    // every timed SpMV launch applies the pending x update (with alpha 0), as
    // in a solve, so it moves the same bytes
    GSLAM_HIP(hipMemsetAsync(&g.status->x_iter, 0, sizeof(int32_t), g.stream));
    spmv();
    run(3);  // warm-up
    GSLAM_HIP(hipGetLastError());
    GSLAM_HIP(hipEventRecord(g.ev0, g.stream));
    run(reps);
    GSLAM_HIP(hipEventRecord(g.ev1, g.stream));
    GSLAM_HIP(hipEventSynchronize(g.ev1));
    float ms = 0.f;
    GSLAM_HIP(hipEventElapsedTime(&ms, g.ev0, g.ev1));
    *us = 1000.0 * ms / reps;
    g.solved = false;
    return GSLAM_OK;
}

int pcg_solve(DeviceGraph &g, const gslam_solve_opts &o, const double *b, double *x_out,
              gslam_solve_stats &s)
{
    if (o.solver == GSLAM_SOLVER_PCG_MIXED) return pcg_solve_mixed(g, o, b, x_out, s);
    return pcg_solve_fp64(g, o, b, x_out, s);
}

}  // namespace gslam
