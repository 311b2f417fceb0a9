// Mixed-precision block-Jacobi PCG: FP32 inner iterations, FP64 refinement
// on the true residual.
//
// Inner iterations: FP32 r, z, p, q, FP32 block-Jacobi inverses and an FP32
// operator; the iterate x is accumulated in FP64 and dot products are summed
// in FP64.  Every refresh_every iterations an FP64 pass recomputes the true
// residual r = b - Omega x with the FP64 BSR Omega, rounds it into the FP32
// recurrence and restarts z and beta from it while keeping the search
// direction p (reliable updates, Sleijpen & van der Vorst 1996).
// Convergence is declared only on that FP64 true residual.
//
// The FP32 operator is applied in factorized form,
//     (Omega p)_i = anchor_i p_i + sum_{edges e at i} J_a^T W_e (J_a p_i + J_b p_o),
// from 8 FP32 factors per edge end (frec) written by the assembly kernels.
// Forming the edge residual-like term J_a p_i + J_b p_o (a difference of
// neighbouring values) before scaling keeps the near-null rigid-motion and
// low-frequency modes of long pose chains representable in FP32; an FP32
// copy of the assembled blocks loses them (entry-wise rounding of ~1e3
// entries whose combinations cancel to ~1e-4) and stalls at a true relative
// residual of ~1e-8 on the N=10k grid-city graph.
//
// The reference's only reduced-precision PCG (src/gpu_pcg_fp32/
// device_graph.cu:611-692) casts the assembled system to FP32 and never
// refines.
#include "device_common.h"
#include "pcg_common.h"

#include <algorithm>
#include <cmath>

namespace gslam {

// Lanes per pose row of the FP64 residual kernels below.
constexpr int kLanesMixed = 4;

// One workgroup per operator tile, one pass (up to kSpmvMaxGrid workgroups,
// tile-stride beyond); p.q partials in the kSpmvMaxGrid array.
static inline unsigned int mixed_spmv_grid(const DeviceGraph &g)
{
    return static_cast<unsigned int>(std::max<int32_t>(1, std::min<int32_t>(g.mt_tiles, kSpmvMaxGrid)));
}

__global__ void __launch_bounds__(kThreads)
mixed_setup_kernel(const int32_t *__restrict__ bdiag, const double *__restrict__ bval,
                   const double *__restrict__ b, int32_t n, int32_t max_iter,
                   float *__restrict__ dinv32, double *__restrict__ x, float *__restrict__ r,
                   float *__restrict__ z, float *__restrict__ p0, float *__restrict__ p1,
                   double *__restrict__ part_out, PcgStatus *status)
{
    double v[2] = {0.0, 0.0};
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double *a = bval + 9 * static_cast<int64_t>(bdiag[i]);
        double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                     a[2] * (a[3] * a[7] - a[4] * a[6]);
        if (det == 0.0) det = 1e-30;
        float fi[9];
        fi[0] = static_cast<float>((a[4] * a[8] - a[5] * a[7]) / det);
        fi[1] = static_cast<float>((a[2] * a[7] - a[1] * a[8]) / det);
        fi[2] = static_cast<float>((a[1] * a[5] - a[2] * a[4]) / det);
        fi[3] = static_cast<float>((a[5] * a[6] - a[3] * a[8]) / det);
        fi[4] = static_cast<float>((a[0] * a[8] - a[2] * a[6]) / det);
        fi[5] = static_cast<float>((a[2] * a[3] - a[0] * a[5]) / det);
        fi[6] = static_cast<float>((a[3] * a[7] - a[4] * a[6]) / det);
        fi[7] = static_cast<float>((a[1] * a[6] - a[0] * a[7]) / det);
        fi[8] = static_cast<float>((a[0] * a[4] - a[1] * a[3]) / det);
#pragma unroll
        for (int k = 0; k < 9; ++k) dinv32[9 * static_cast<int64_t>(i) + k] = fi[k];
        const int64_t o = 3 * static_cast<int64_t>(i);
        const float b0 = static_cast<float>(b[o]), b1 = static_cast<float>(b[o + 1]),
                    b2 = static_cast<float>(b[o + 2]);
        const float z0 = fi[0] * b0 + fi[1] * b1 + fi[2] * b2;
        const float z1 = fi[3] * b0 + fi[4] * b1 + fi[5] * b2;
        const float z2 = fi[6] * b0 + fi[7] * b1 + fi[8] * b2;
        x[o] = 0.0; x[o + 1] = 0.0; x[o + 2] = 0.0;
        r[o] = b0; r[o + 1] = b1; r[o + 2] = b2;
        z[o] = z0; z[o + 1] = z1; z[o + 2] = z2;
        p0[o] = 0.f; p0[o + 1] = 0.f; p0[o + 2] = 0.f;
        p1[o] = 0.f; p1[o + 1] = 0.f; p1[o + 2] = 0.f;
        v[0] += b[o] * b[o] + b[o + 1] * b[o + 1] + b[o + 2] * b[o + 2];
        v[1] += static_cast<double>(b0) * z0 + static_cast<double>(b1) * z1 +
                static_cast<double>(b2) * z2;
    }
    write_partials<2>(v, part_out, kPartStride);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        status->done = max_iter <= 0 ? 1 : 0;
        status->converged = 0;
        status->iters = 0;
        status->flags = 0;
        status->refreshes = 0;
        status->rnorm = 0.0;
        status->bnorm = 0.0;
        status->tol = 0.0;
    }
}

// y_i contribution of one edge end; see the factor layout in assembly.hip.
// type 0: landmark pair (B1 = [-I | b1], B2 = [I | b2], W 2x2)
// type 1: motion (J1 = -F, J2 = I, W = lambda Om 3x3)
// type 2: SE(2) between (Ji, Jj of the rotation c, s and ga, gb; W diagonal)
__device__ inline void edge_apply(int type, int side, const float4 f0, const float4 f1,
                                  const float pi[3], const float po[3], float y[3])
{
    if (type == 2) {
        const float c = f0.x, s = f0.y, ga = f0.z, gb = f0.w;
        if (side == 0) {  // pose i is the "from" end
            const float d0 = po[0] - pi[0], d1 = po[1] - pi[1];
            const float u0 = c * d0 + s * d1 + ga * pi[2];
            const float u1 = -s * d0 + c * d1 + gb * pi[2];
            const float u2 = po[2] - pi[2];
            const float w0 = f1.x * u0, w1 = f1.y * u1, w2 = f1.z * u2;
            y[0] += -c * w0 + s * w1;
            y[1] += -s * w0 - c * w1;
            y[2] += ga * w0 + gb * w1 - w2;
        } else {          // pose i is the "to" end, o the "from" end
            const float d0 = pi[0] - po[0], d1 = pi[1] - po[1];
            const float u0 = c * d0 + s * d1 + ga * po[2];
            const float u1 = -s * d0 + c * d1 + gb * po[2];
            const float u2 = pi[2] - po[2];
            const float w0 = f1.x * u0, w1 = f1.y * u1, w2 = f1.z * u2;
            y[0] += c * w0 - s * w1;
            y[1] += s * w0 + c * w1;
            y[2] += w2;
        }
    } else if (type == 0) {
        const float b1x = f0.x, b1y = f0.y, b2x = f0.z, b2y = f0.w;
        const float w00 = f1.x, w01 = f1.y, w11 = f1.z;
        // u = B1 p_(obs1 pose) + B2 p_(obs2 pose)
        const float *p1 = side == 0 ? pi : po, *p2 = side == 0 ? po : pi;
        const float u0 = (p2[0] - p1[0]) + b1x * p1[2] + b2x * p2[2];
        const float u1 = (p2[1] - p1[1]) + b1y * p1[2] + b2y * p2[2];
        const float w0 = w00 * u0 + w01 * u1, w1 = w01 * u0 + w11 * u1;
        if (side == 0) {
            y[0] -= w0;
            y[1] -= w1;
            y[2] += b1x * w0 + b1y * w1;
        } else {
            y[0] += w0;
            y[1] += w1;
            y[2] += b2x * w0 + b2y * w1;
        }
    } else {
        const float F02 = f0.x, F12 = f0.y;
        const float w00 = f0.z, w01 = f0.w, w02 = f1.x, w11 = f1.y, w12 = f1.z, w22 = f1.w;
        // u = -F p_(t1) + p_(t2)
        const float *p1 = side == 0 ? pi : po, *p2 = side == 0 ? po : pi;
        const float u0 = (p2[0] - p1[0]) - F02 * p1[2];
        const float u1 = (p2[1] - p1[1]) - F12 * p1[2];
        const float u2 = p2[2] - p1[2];
        const float w0 = w00 * u0 + w01 * u1 + w02 * u2;
        const float w1 = w01 * u0 + w11 * u1 + w12 * u2;
        const float w2 = w02 * u0 + w12 * u1 + w22 * u2;
        if (side == 0) {
            y[0] -= w0;
            y[1] -= w1;
            y[2] -= F02 * w0 + F12 * w1 + w2;
        } else {
            y[0] += w0;
            y[1] += w1;
            y[2] += w2;
        }
    }
}

// A(k): stopping test (true residual slots only), beta, q = Omega_32 p with
// p = z + beta p_old formed on the fly, p.q partials.
//
// One workgroup per tile of whole poses (<= 256 incidence entries, <= 85
// poses; the tiles of build_mixed_tiles): one lane per incidence entry (its
// entry, own and other pose from slot 256 t + lane -- no wait for the tile
// record -- and its edge's 8 factors, plane-major, one 1 KB run per factor
// and workgroup), the entry's 3 contributions to its own pose staged in LDS,
// then one thread per output row (3 per pose) sums its pose's entries in
// entry order.  Issue order, as in the FP64 SpMV: the r.r / r.z partials
// that give beta and the stop flag, the slot, the tile record and the output
// row's range and z / p_old, then the gathered z / p_old of both poses -- two
// dependent loads deep; the partials are reduced (one barrier) while the
// gathers are in flight.  A pose with more than 256 entries is a tile of its
// own and continues in extra chunks, summed one chunk at a time.
struct MixedSlotRegs {
    int32_t ent;
    float4 f0, f1;
    float zi[3], qi[3], zo[3], qo[3];
};

template <bool kNt>
__device__ inline void mixed_slot_load(int64_t slot, const int32_t *__restrict__ mt_ent,
                                       const int32_t *__restrict__ mt_own,
                                       const int32_t *__restrict__ mt_oth,
                                       const float *__restrict__ frec, const float *__restrict__ z,
                                       const float *__restrict__ p_old, MixedSlotRegs &S)
{
    const int64_t oi = 3 * static_cast<int64_t>(stream_load<kNt>(mt_own + slot));
    const int64_t oo = 3 * static_cast<int64_t>(stream_load<kNt>(mt_oth + slot));
    S.ent = stream_load<kNt>(mt_ent + slot);
    const float *f = frec + (slot / kTileBlocks) * 8 * kTileBlocks + slot % kTileBlocks;
    float fv[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) fv[c] = stream_load<kNt>(f + c * kTileBlocks);
    S.f0 = make_float4(fv[0], fv[1], fv[2], fv[3]);
    S.f1 = make_float4(fv[4], fv[5], fv[6], fv[7]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        S.zi[c] = z[oi + c];
        S.qi[c] = p_old[oi + c];
        S.zo[c] = z[oo + c];
        S.qo[c] = p_old[oo + c];
    }
}

// The entry's contributions to its own pose, into LDS column `col` (zeros
// for a padding slot).
__device__ inline void mixed_slot_apply(const MixedSlotRegs &S, float beta, int64_t n_obs_e,
                                        int64_t n_mot_e, float *ys, int col)
{
    float y[3] = {0.f, 0.f, 0.f};
    if (S.ent >= 0) {
        const int64_t e = S.ent >> 1;
        const int side = S.ent & 1;
        const int type = e < n_obs_e ? 0 : (e < n_obs_e + n_mot_e ? 1 : 2);
        float pi[3], po[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            pi[c] = S.zi[c] + beta * S.qi[c];
            po[c] = S.zo[c] + beta * S.qo[c];
        }
        edge_apply(type, side, S.f0, S.f1, pi, po, y);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) ys[a * kTileBlocks + col] = y[a];
}

template <bool kNt>
__global__ void __launch_bounds__(kThreads)
mixed_spmv_kernel(const int4 *__restrict__ mt_info, const int32_t *__restrict__ mt_ext,
                  const int32_t *__restrict__ mt_ent, const int32_t *__restrict__ mt_own,
                  const int32_t *__restrict__ mt_oth, const float *__restrict__ frec,
                  int32_t n_tiles, const int32_t *__restrict__ xptr, int64_t n_obs_e,
                  int64_t n_mot_e, float anchor, const float *__restrict__ z,
                  const float *__restrict__ p_old, float *__restrict__ p_new,
                  float *__restrict__ q, int32_t k, int32_t is_true, double tol_rel,
                  const double *__restrict__ pr_cur, int32_t np_cur,
                  const double *__restrict__ pr_prev, int32_t np_prev,
                  double *__restrict__ pq_out, PcgStatus *status)
{
    static_assert(kTileBlocks == kThreads, "one lane per slot of a 256-slot chunk");
    __shared__ float ys[3 * kTileBlocks];
    __shared__ double red[3][kThreads / kWave];
    static_assert(kMaxGrid <= 2 * kThreads, "two partials per thread and array");
    const int tid = static_cast<int>(threadIdx.x);
    const int a = tid % 3;
    // tiles in contiguous runs per XCD when every tile has a workgroup
    // (neighbouring poses share an L2), else the round robin
    const int32_t g = static_cast<int32_t>(gridDim.x), b = static_cast<int32_t>(blockIdx.x);
    const int32_t t0 = n_tiles <= g ? xcd_contiguous(b, g) : b;
    // The r.r / r.z partials of this and the previous parity first
    // (np_cur, np_prev <= 2 x kThreads: two per thread and array, summed below
    // in strided_sum's order), as unconditional loads from clamped indices and
    // pinned here, so that their round trips overlap the tile's loads.
    const int ia = min(tid, np_cur - 1), ib = min(tid + kThreads, np_cur - 1);
    const int ja = min(tid, np_prev - 1), jb = min(tid + kThreads, np_prev - 1);
    const double rr_a = pr_cur[ia], rr_b = pr_cur[ib];
    const double rz_a = pr_cur[kPartStride + ia], rz_b = pr_cur[kPartStride + ib];
    const double zp_a = pr_prev[kPartStride + ja], zp_b = pr_prev[kPartStride + jb];
    const int32_t done_in = status->done;
    const double tol_in = status->tol;
    asm volatile("" ::: "memory");  // keeps the loads above here (not sunk to their use)
    MixedSlotRegs S;
    int4 ti = make_int4(0, 0, 0, 0);
    int32_t ob0 = 0, ob1 = 0;
    float zr = 0.f, pr = 0.f;
    auto tile_load = [&](int32_t t) {
        mixed_slot_load<kNt>(static_cast<int64_t>(t) * kTileBlocks + tid, mt_ent, mt_own, mt_oth, frec,
                        z, p_old, S);
        ti = mt_info[t];
        const int32_t orow = min(ti.x + tid / 3, ti.y - 1);
        ob0 = xptr[orow];
        ob1 = xptr[orow + 1];
        const int64_t o = 3 * static_cast<int64_t>(orow) + a;
        zr = z[o];
        pr = p_old[o];
    };
    if (t0 < n_tiles) tile_load(t0);
    const bool in_a = tid < np_cur, in_b = tid + kThreads < np_cur;
    const bool jn_a = tid < np_prev, jn_b = tid + kThreads < np_prev;
    double red_v[3];
    red_v[0] = (0.0 + (in_a ? rr_a : 0.0)) + (in_b ? rr_b : 0.0);
    red_v[1] = (0.0 + (in_a ? rz_a : 0.0)) + (in_b ? rz_b : 0.0);
    red_v[2] = k > 0 ? (0.0 + (jn_a ? zp_a : 0.0)) + (jn_b ? zp_b : 0.0) : 0.0;
    block_allsum<3, kThreads>(red_v, red);
    if (done_in) return;
    const double rr = red_v[0];
    if (!isfinite(rr)) {  // NaN / Inf in the residual (flag bit 1)
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            status->done = 1;
            status->flags |= 2;
            status->iters = k;
        }
        return;
    }
    if (k == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            status->bnorm = sqrt(rr);
            status->tol = tol_rel * sqrt(rr);
            status->rnorm = sqrt(rr);
        }
    } else if (is_true) {
        const double tol = tol_in;
        if (sqrt(rr) < tol) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                status->done = 1;
                status->converged = 1;
                status->iters = k;
                status->rnorm = sqrt(rr);
            }
            return;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) status->rnorm = sqrt(rr);
    }
    const float beta = k > 0 ? static_cast<float>(red_v[1] / red_v[2]) : 0.f;
    double pq[1] = {0.0};
    for (int32_t t = t0; t < n_tiles; t += g) {
        if (t != t0) {
            __syncthreads();  // the previous tile is summed before ys is rewritten
            tile_load(t);
        }
        const bool has_out = tid < 3 * (ti.y - ti.x);
        mixed_slot_apply(S, beta, n_obs_e, n_mot_e, ys, tid);
        __syncthreads();
        float acc = 0.f;
        if (has_out) {
            const int32_t hi = min(ob1, ti.z + kTileBlocks);
            for (int32_t c = max(ob0, ti.z); c < hi; ++c) acc += ys[a * kTileBlocks + (c - ti.z)];
        }
        // a pose with more than 256 entries (a tile of its own) continues in
        // extra chunks (the condition is the same in the whole workgroup)
        if (ti.w > ti.z + kTileBlocks) {
            const int32_t ext = mt_ext[t];
            for (int32_t cb = ti.z + kTileBlocks, h = ext; cb < ti.w; cb += kTileBlocks, ++h) {
                MixedSlotRegs E;
                mixed_slot_load<kNt>(static_cast<int64_t>(h) * kTileBlocks + tid, mt_ent, mt_own, mt_oth,
                                frec, z, p_old, E);
                __syncthreads();  // the previous chunk is summed
                mixed_slot_apply(E, beta, n_obs_e, n_mot_e, ys, tid);
                __syncthreads();
                if (has_out) {
                    const int32_t hi = min(ob1, cb + kTileBlocks);
                    for (int32_t c = max(ob0, cb); c < hi; ++c) acc += ys[a * kTileBlocks + (c - cb)];
                }
            }
        }
        if (has_out) {
            const int32_t row = ti.x + tid / 3;
            const float pi = zr + beta * pr;
            if (row == 0) acc += anchor * pi;
            const int64_t o = 3 * static_cast<int64_t>(row) + a;
            p_new[o] = pi;
            q[o] = acc;
            pq[0] += static_cast<double>(pi) * acc;
        }
    }
    write_partials<1>(pq, pq_out, kSpmvMaxGrid, t0);  // slot t0: the sum of tiles t0, t0 + grid, ...
    if (blockIdx.x == 0 && threadIdx.x == 0) status->iters = k;
}

// B(k): alpha from the p.q partials, x / r / z update, r.r and r.z partials;
// the per-pose loads are issued before the partials are reduced.
__global__ void __launch_bounds__(kThreads)
mixed_update_kernel(const float *__restrict__ p, const float *__restrict__ q,
                    const float *__restrict__ dinv32, double *__restrict__ x,
                    float *__restrict__ r, float *__restrict__ z, int32_t n,
                    const double *__restrict__ pq_in, int32_t np_pq,
                    const double *__restrict__ pr_cur, int32_t np_cur,
                    double *__restrict__ pr_out, PcgStatus *status)
{
    __shared__ double red[2][kThreads / kWave];
    const int32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
    double xp[3];
    float pp[3], qp[3], rp[3], d[9];
    auto load = [&](int32_t ii) {
        const int64_t o = 3 * static_cast<int64_t>(ii);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            xp[c] = x[o + c];
            pp[c] = p[o + c];
            qp[c] = q[o + c];
            rp[c] = r[o + c];
        }
#pragma unroll
        for (int c = 0; c < 9; ++c) d[c] = dinv32[9 * static_cast<int64_t>(ii) + c];
    };
    const int32_t done_in = status->done;  // set by earlier launches only
    if (i0 < n) load(i0);
    double red_v[2];
    red_v[0] = strided_sum<kThreads, 8>(pq_in, np_pq, 0.0);
    red_v[1] = strided_sum<kThreads, 2>(pr_cur + kPartStride, np_cur, 0.0);
    block_allsum<2, kThreads>(red_v, red);
    if (done_in) return;
    if (red_v[0] == 0.0 || !isfinite(red_v[0])) {  // p.q == 0, or NaN / Inf (bit 1)
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            status->done = 1;
            status->flags |= red_v[0] == 0.0 ? 1 : 2;
        }
        return;
    }
    const double alpha = red_v[1] / red_v[0];
    const float af = static_cast<float>(alpha);
    double v[2] = {0.0, 0.0};
    for (int32_t i = i0; i < n; i += gridDim.x * blockDim.x) {
        if (i != i0) load(i);
        const int64_t o = 3 * static_cast<int64_t>(i);
        float rr[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x[o + c] = xp[c] + alpha * static_cast<double>(pp[c]);
            rr[c] = rp[c] - af * qp[c];
            r[o + c] = rr[c];
        }
        float zz[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) zz[c] = d[3 * c] * rr[0] + d[3 * c + 1] * rr[1] + d[3 * c + 2] * rr[2];
        z[o] = zz[0]; z[o + 1] = zz[1]; z[o + 2] = zz[2];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[0] += static_cast<double>(rr[c]) * rr[c];
            v[1] += static_cast<double>(rr[c]) * zz[c];
        }
    }
    write_partials<2>(v, pr_out, kPartStride);
}

// Refresh: r = b - Omega x in FP64 (BSR), rounded into the FP32 recurrence,
// z = D^-1 r; overwrites the r.r / r.z partials the next A(k) reads.
__global__ void __launch_bounds__(kThreads)
mixed_refresh_kernel(const int32_t *__restrict__ brow_ptr, const int32_t *__restrict__ bcol,
                     const double *__restrict__ bval, const float *__restrict__ dinv32,
                     const double *__restrict__ b, const double *__restrict__ x,
                     float *__restrict__ r, float *__restrict__ z, int32_t n,
                     double *__restrict__ pr_out, PcgStatus *status)
{
    if (status->done) return;
    constexpr int G = kLanesMixed, kRowsPerTile = kThreads / G;
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
            const double r0 = b[o] - y0, r1 = b[o + 1] - y1, r2 = b[o + 2] - y2;
            const float f0 = static_cast<float>(r0), f1 = static_cast<float>(r1),
                        f2 = static_cast<float>(r2);
            const float *d = dinv32 + 9 * row;
            const float z0 = d[0] * f0 + d[1] * f1 + d[2] * f2;
            const float z1 = d[3] * f0 + d[4] * f1 + d[5] * f2;
            const float z2 = d[6] * f0 + d[7] * f1 + d[8] * f2;
            r[o] = f0; r[o + 1] = f1; r[o + 2] = f2;
            z[o] = z0; z[o + 1] = z1; z[o + 2] = z2;
            s[0] += r0 * r0 + r1 * r1 + r2 * r2;
            s[1] += static_cast<double>(f0) * z0 + static_cast<double>(f1) * z1 +
                    static_cast<double>(f2) * z2;
        }
    }
    write_partials<2>(s, pr_out, kPartStride);
    if (blockIdx.x == 0 && threadIdx.x == 0) status->refreshes += 1;
}

__global__ void __launch_bounds__(kThreads)
mixed_finalize_kernel(const double *__restrict__ pr_cur, int32_t np_cur, int32_t k,
                      PcgStatus *status)
{
    double cur[2];
    sum_partials<2>(pr_cur, np_cur, kPartStride, cur);
    if (threadIdx.x == 0 && !status->done) {
        status->iters = k;
        status->rnorm = sqrt(cur[0]);
        status->done = 1;
        status->converged = sqrt(cur[0]) < status->tol ? 1 : 0;
    }
}

// r = b - Omega x (FP64 BSR), for the FP64 correction after the FP32 phase.
__global__ void __launch_bounds__(kThreads)
resid64_kernel(const int32_t *__restrict__ brow_ptr, const int32_t *__restrict__ bcol,
               const double *__restrict__ bval, const double *__restrict__ b,
               const double *__restrict__ x, double *__restrict__ r, int32_t n)
{
    constexpr int G = kLanesMixed, kRowsPerTile = kThreads / G;
    const int lane = threadIdx.x % G;
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
            r[o] = b[o] - y0;
            r[o + 1] = b[o + 1] - y1;
            r[o + 2] = b[o + 2] - y2;
        }
    }
}

// x += d
__global__ void __launch_bounds__(kThreads)
add64_kernel(double *__restrict__ x, const double *__restrict__ d, int64_t m)
{
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < m;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        x[i] += d[i];
}

static int mixed_phase(DeviceGraph &g, const gslam_solve_opts &o, const double *b, double *x_out,
                       gslam_solve_stats &s);

// FP32 phase, then (mixed_handover > 0) FP64 iterative refinement: once the
// FP64 true residual of the FP32 phase is below mixed_handover ||b||, FP64
// PCG solves Omega d = b - Omega x to the remaining tolerance and x += d.
// The FP32 recurrence cannot resolve the modes of the block-Jacobi
// preconditioned operator whose eigenvalues lie below ~eps32 of the largest
// (kappa up to ~1e13 on long grid-city chains), so the FP64 correction is
// what certifies tol on such graphs.
int pcg_solve_mixed(DeviceGraph &g, const gslam_solve_opts &o, const double *b, double *x_out,
                    gslam_solve_stats &s)
{
    const double handover = o.mixed_handover;
    gslam_solve_opts o1 = o;
    if (handover > o.tol) o1.tol = handover;
    GSLAM_TRY(mixed_phase(g, o1, b, x_out, s));
    s.iters_fp64 = 0;
    if (!(handover > 0.0) || s.true_rel_res < o.tol || (s.flags & 3)) return GSLAM_OK;
    GSLAM_TRY(g.vecs_reserve());
    const int64_t dim = g.dim();
    double *r2 = g.vecs + 2 * dim, *d = g.vecs + 3 * dim;
    const int32_t n = g.h.n;
    const unsigned int gr = grid_for(n, kLanesMixed);
    hipLaunchKernelGGL(resid64_kernel, dim3(gr), dim3(kThreads), 0, g.stream, g.brow_ptr, g.bcol,
                       g.bval, b, x_out, r2, n);
    GSLAM_HIP(hipGetLastError());
    gslam_solve_opts o2 = o;
    o2.solver = GSLAM_SOLVER_PCG_FP64;
    o2.tol = s.true_rel_res > 0.0 ? o.tol / s.true_rel_res : o.tol;
    o2.max_iter = std::max(0, o.max_iter - s.iters);
    o2.profile_every = 0;
    gslam_solve_stats s2{};
    GSLAM_TRY(pcg_solve_fp64(g, o2, r2, d, s2));
    hipLaunchKernelGGL(add64_kernel, dim3(grid_for(dim, 1)), dim3(kThreads), 0, g.stream, x_out, d,
                       dim);
    GSLAM_HIP(hipGetLastError());
    s.iters += s2.iters;
    s.iters_fp64 = s2.iters;
    s.ms += s2.ms;
    s.flags |= s2.flags;
    GSLAM_TRY(true_rel_residual(g, b, x_out, &s.true_rel_res));
    s.converged = s2.converged;  // the FP64 recurrence's test, as for PCG_FP64
    s.rel_res = s.true_rel_res;
    return GSLAM_OK;
}

// The operator kernel of this graph: the nontemporal variant where its slots
// (44 B each: factors, entry and both poses) exceed half the Infinity Cache,
// whose other half the mixed iteration's FP64 x, FP32 vectors and D^-1 fill
// (FP32 iteration in the solve at N = 1.2M: 91.3 -> 88.4 us; at 0.8M-2M
// within 1%; at 500k, 84 MB of slots, ordinary loads are faster: 39.3
// against 40.9 us).
using MixedSpmvKernel = decltype(&mixed_spmv_kernel<false>);

static MixedSpmvKernel mixed_spmv_kernel_for(const DeviceGraph &g)
{
    return streams_past_cache(2 * 44 * g.mt_slots) ? mixed_spmv_kernel<true> : mixed_spmv_kernel<false>;
}

static int mixed_phase(DeviceGraph &g, const gslam_solve_opts &o, const double *b, double *x_out,
                       gslam_solve_stats &s)
{
    const int32_t n = g.h.n;
    const unsigned int gs = mixed_spmv_grid(g), gu = grid_for(n, 1);
    const MixedSpmvKernel spmv_kern = mixed_spmv_kernel_for(g);
    const int32_t every = o.refresh_every > 0 ? o.refresh_every : 64;
    double *PR[2] = {g.part + kPartOffPR0, g.part + kPartOffPR1};
    double *PQ = g.part + kPartOffPQT;  // stride kSpmvMaxGrid
    PcgStatus *st = g.status;
    GSLAM_HIP(hipEventRecord(g.ev0, g.stream));
    hipLaunchKernelGGL(mixed_setup_kernel, dim3(gu), dim3(kThreads), 0, g.stream, g.bdiag, g.bval,
                       b, n, o.max_iter, g.dinv32, x_out, g.r32, g.z32, g.p32[0], g.p32[1], PR[0],
                       st);
    GSLAM_HIP(hipGetLastError());
    GSLAM_HIP(hipEventRecord(g.ev2, g.stream));  // end of the setup phase
    int32_t k = 0;
    auto refresh = [&]() {
        hipLaunchKernelGGL(mixed_refresh_kernel, dim3(gu), dim3(kThreads), 0, g.stream, g.brow_ptr,
                           g.bcol, g.bval, g.dinv32, b, x_out, g.r32, g.z32, n, PR[k & 1], st);
    };
    auto enqueue_iters = [&](int32_t m) {
        for (int32_t c = 0; c < m; ++c, ++k) {
            int32_t is_true = (k == 0) ? 1 : 0;
            if (k > 0 && k % every == 0) {
                refresh();
                is_true = 1;
            }
            float *pold = g.p32[k & 1], *pnew = g.p32[(k + 1) & 1];
            hipLaunchKernelGGL(spmv_kern, dim3(gs), dim3(kThreads), 0, g.stream, g.mt_info,
                               g.mt_ext, g.mt_ent, g.mt_own, g.mt_oth, g.frec, g.mt_tiles, g.xptr,
                               g.h.n_obs_e, g.h.n_mot_e, static_cast<float>(g.anchor), g.z32, pold,
                               pnew, g.q32, k, is_true, o.tol, PR[k & 1],
                               static_cast<int32_t>(gu), PR[(k + 1) & 1],
                               static_cast<int32_t>(gu), PQ, st);
            hipLaunchKernelGGL(mixed_update_kernel, dim3(gu), dim3(kThreads), 0, g.stream, pnew,
                               g.q32, g.dinv32, x_out, g.r32, g.z32, n, PQ,
                               static_cast<int32_t>(gs), PR[k & 1], static_cast<int32_t>(gu),
                               PR[(k + 1) & 1], st);
        }
    };
    GSLAM_TRY(run_chunked(g, o, enqueue_iters, k));
    // final true residual (no-op if an earlier refresh already converged)
    refresh();
    hipLaunchKernelGGL(mixed_finalize_kernel, dim3(1), dim3(kThreads), 0, g.stream, PR[k & 1],
                       static_cast<int32_t>(gu), k, st);
    GSLAM_HIP(hipGetLastError());
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
    s.reliable_updates = ps.refreshes;
    s.rel_res = ps.bnorm > 0 ? ps.rnorm / ps.bnorm : ps.rnorm;
    s.ms = ms;
    GSLAM_TRY(true_rel_residual(g, b, x_out, &s.true_rel_res));
    return GSLAM_OK;
}

int time_mixed_spmv(DeviceGraph &g, int reps, double *us)
{
    GSLAM_TRY(g.mixed_factors());
    const int32_t n = g.h.n;
    const unsigned int gs = mixed_spmv_grid(g), gu = grid_for(n, 1);
    const MixedSpmvKernel spmv_kern = mixed_spmv_kernel_for(g);
    double *PR[2] = {g.part + kPartOffPR0, g.part + kPartOffPR1};
    double *PQ = g.part + kPartOffPQT;  // stride kSpmvMaxGrid
    hipLaunchKernelGGL(mixed_setup_kernel, dim3(gu), dim3(kThreads), 0, g.stream, g.bdiag, g.bval,
                       g.xi, n, 1, g.dinv32, g.x, g.r32, g.z32, g.p32[0], g.p32[1], PR[0],
                       g.status);
    auto spmv = [&]() {
        hipLaunchKernelGGL(spmv_kern, dim3(gs), dim3(kThreads), 0, g.stream, g.mt_info,
                           g.mt_ext, g.mt_ent, g.mt_own, g.mt_oth, g.frec, g.mt_tiles, g.xptr,
                           g.h.n_obs_e, g.h.n_mot_e, static_cast<float>(g.anchor), g.z32, g.p32[0],
                           g.p32[1], g.q32, 0, 1, 1e-30, PR[0], static_cast<int32_t>(gu), PR[1],
                           static_cast<int32_t>(gu), PQ, g.status);
    };
    for (int i = 0; i < 3; ++i) spmv();
    GSLAM_HIP(hipGetLastError());
    GSLAM_HIP(hipEventRecord(g.ev0, g.stream));
    for (int i = 0; i < reps; ++i) spmv();
    GSLAM_HIP(hipEventRecord(g.ev1, g.stream));
    GSLAM_HIP(hipEventSynchronize(g.ev1));
    float ms = 0.f;
    GSLAM_HIP(hipEventElapsedTime(&ms, g.ev0, g.ev1));
    *us = 1000.0 * ms / reps;
    g.solved = false;
    return GSLAM_OK;
}

}  // namespace gslam
