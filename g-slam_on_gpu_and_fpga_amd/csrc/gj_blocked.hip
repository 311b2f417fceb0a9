// Blocked Gauss-Jordan inverse: the GJ_MIXED dense solve (FP32 inverse of the
// Jacobi-scaled Omega with FP32 MFMA trailing updates, then FP64 refinement
// on the true residual xi - Omega dx) and GJ_BLOCKED64 (the same block steps
// in FP64 with FP64 MFMA trailing updates, no refinement).
//
// The reference inverts Omega with an unblocked, unpivoted Gauss-Jordan
// (inv, src/cpu_pcg_fp64/MatrixOps.c:29-91; inv_float,
// src/cpu_gj_mixed/MatrixOps.c:174-212; inv_cuda, src/gpu_gj_fp64/inv.cu:
// 221-354): one pass over the whole matrix per pivot.  Here the pivots are
// taken kNb = 128 at a time (still unpivoted, in place):
//
//   P    = A_KK^-1                      FP64, one workgroup (pivot_block)
//   R_J  = P A_KJ            (J != K)   FP64 MFMA (row panel)
//   CT   = -A_IK^T           (I != K)   transposed copy of the column panel
//   A_IJ = A_IJ + CT^T R_J   (I, J != K)  MFMA trailing update
//   A_IK = CT^T P            (I != K)   same kernel, R_K = P
//   A_KJ = R_J, A_KK = P
//
// which is 2 n^3 flops in total, almost all in the trailing update.  The
// matrix is padded to a multiple of kNb with an identity block (its inverse
// is the identity, and exact zeros stay exact, so the padding never mixes
// into the result).  In GJ_MIXED the trailing launch of step K also updates
// and inverts the pivot block of step K+1 (its workgroup 0), so only the
// first pivot block has a launch of its own.
//
// Precision: Omega is scaled to S = D^-1/2 Omega D^-1/2 (unit diagonal) in
// FP64 before rounding to the storage type; the pivot blocks are inverted
// and the pivot rows scaled in FP64; GJ_MIXED stores and updates in FP32,
// GJ_BLOCKED64 in FP64.  The FP32 inverse M preconditions FP64 refinement
// on Omega itself:
//   GSLAM_REFINE_IR: x += D^-1/2 M D^-1/2 (xi - Omega x), refine_sweeps times;
//   GSLAM_REFINE_CG: conjugate gradients preconditioned by the same operator,
//     flexible (Polak-Ribiere) beta because M is not exactly symmetric,
//     residual recomputed in FP64 every iteration.
// Both stop early when ||xi - Omega x|| < tol ||xi||.
#include "device_common.h"
#include "pcg_common.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <utility>

namespace gslam {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kNb = 128;          // panel width, trailing tile edge and GEMM depth
constexpr int kPb = 8;            // pivot kernel: pivots per inner block step
constexpr int kPivThreads = 8 * 64;  // pivot kernel: six MFMA waves, one pivot-chain wave
constexpr int kXcd = 8;           // MI355X: 8 XCDs, workgroups dealt round-robin

// ---------------------------------------------------------------- loading

// dsc[i] = 1/sqrt(|d_i|) (1 if d_i == 0) of the diagonal of Omega (BSR) for
// i < n, 1 on the padding; u (padding part) = 0.
__global__ void __launch_bounds__(kThreads)
bgj_scale_bsr_kernel(const int32_t *__restrict__ bdiag, const double *__restrict__ bval,
                     int64_t n, int64_t np, double *__restrict__ dsc)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= np) return;
    double s = 1.0;
    if (i < n) {
        const int64_t row = i / 3;
        const int a = static_cast<int>(i - 3 * row);
        const double d = fabs(bval[9 * static_cast<int64_t>(bdiag[row]) + 4 * a]);
        if (d > 0.0) s = 1.0 / sqrt(d);
    }
    dsc[i] = s;
}

__global__ void __launch_bounds__(kThreads)
bgj_scale_dense_kernel(const double *__restrict__ m, int64_t n, int64_t np,
                       double *__restrict__ dsc)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= np) return;
    double s = 1.0;
    if (i < n) {
        const double d = fabs(m[i * n + i]);
        if (d > 0.0) s = 1.0 / sqrt(d);
    }
    dsc[i] = s;
}

// a = identity on the padding, zero elsewhere (np x np FP32, 4 per thread).
__global__ void __launch_bounds__(kThreads)
bgj_clear_kernel(float *__restrict__ a, int64_t n, int64_t np)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    const int64_t e = 4 * t;
    if (e >= np * np) return;
    const int64_t i = e / np, j = e - i * np;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i >= n && i >= j && i < j + 4) {
        const int d = static_cast<int>(i - j);
        if (d == 0) v.x = 1.f; else if (d == 1) v.y = 1.f; else if (d == 2) v.z = 1.f; else v.w = 1.f;
    }
    *reinterpret_cast<float4 *>(a + e) = v;
}

// Scaled BSR blocks into the dense FP32 matrix: one thread per block entry.
__global__ void __launch_bounds__(kThreads)
bgj_scatter_bsr_kernel(const int32_t *__restrict__ brow_ptr, const int32_t *__restrict__ bcol,
                       const double *__restrict__ bval, int64_t n_blocks, int32_t n_rows,
                       const double *__restrict__ dsc, int64_t np, float *__restrict__ a)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= 9 * n_blocks) return;
    const int64_t b = t / 9;
    const int k = static_cast<int>(t - 9 * b);
    int32_t lo = 0, hi = n_rows - 1;  // block row of b
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (brow_ptr[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const int64_t i = 3 * static_cast<int64_t>(lo) + k / 3;
    const int64_t j = 3 * static_cast<int64_t>(bcol[b]) + k % 3;
    a[i * np + j] = static_cast<float>(bval[t] * dsc[i] * dsc[j]);
}

// Scaled dense FP64 matrix (n x n) into the padded FP32 matrix.
__global__ void __launch_bounds__(kThreads)
bgj_load_dense_kernel(const double *__restrict__ m, int64_t n, int64_t np,
                      const double *__restrict__ dsc, float *__restrict__ a)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= np * np) return;
    const int64_t i = t / np, j = t - i * np;
    float v;
    if (i < n && j < n) v = static_cast<float>(m[i * n + j] * dsc[i] * dsc[j]);
    else v = (i == j) ? 1.f : 0.f;
    a[t] = v;
}

// inv = D^-1/2 M D^-1/2 (FP64 out, n x n); zeros if the inversion failed.
__global__ void __launch_bounds__(kThreads)
bgj_store_dense_kernel(const float *__restrict__ a, int64_t n, int64_t np,
                       const double *__restrict__ dsc, const int32_t *__restrict__ singular,
                       double *__restrict__ m)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n * n) return;
    const int64_t i = t / n, j = t - i * n;
    m[t] = *singular ? 0.0 : static_cast<double>(a[i * np + j]) * dsc[i] * dsc[j];
}

// ---------------------------------------------------------------- block step

// 1/d to within an ulp: the hardware reciprocal refined by two
// Newton-Raphson steps r <- r + r (1 - d r).
__device__ inline double recip_nr(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(r, fma(-d, r, 1.0), r);
}

// v of lane `src` in every lane (scalar broadcast of both halves).
__device__ inline double readlane_f64(double v, int src)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), src);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), src);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) |
                                          static_cast<unsigned int>(lo));
}

// In-place Gauss-Jordan of an 8 x 8 block held one entry per lane (lane
// 8i + j holds entry (i, j)) of a full wave: for each pivot p, row p is
// scaled by 1/piv, column p becomes -a_ip/piv, the pivot 1/piv, and every
// other entry a_ij - (a_ip/piv) a_pj.  A zero pivot sets *zero and is taken
// as 1 (the caller then reports the whole inverse as zero).
__device__ inline double gj8_wave(double v, int i, int j, bool &zero)
{
#pragma unroll
    for (int p = 0; p < kPb; ++p) {
        const double rowv = __shfl(v, 8 * p + j, kWave);
        const double colv = __shfl(v, 8 * i + p, kWave);
        double piv = readlane_f64(v, 9 * p);
        if (piv == 0.0) {
            zero = true;
            piv = 1.0;
        }
        const double ip = recip_nr(piv);
        if (i == p) v = (j == p) ? ip : v * ip;
        else if (j == p) v = -(colv * ip);
        else v = fma(-(colv * ip), rowv, v);
    }
    return v;
}

template <int... I, class F>
__device__ inline void static_for(std::integer_sequence<int, I...>, F &&f)
{
    (f(std::integral_constant<int, I>{}), ...);
}

// P = A_KK^-1 (128 x 128) in FP64 by Gauss-Jordan in blocks of 8 pivots,
// one workgroup of eight waves.  The MFMA waves hold the block in FP64 MFMA
// accumulator layout as 16-row tiles (one or two per wave, 8 x 16 x 16 each;
// register e of lane l in a tile holds (l/16 + 4e, l%16)).  Step pb
// (pivots 8pb..8pb+7, "K"; P8 = A_KK^-1 made by the chain wave one step
// ahead):
//   1. the MFMA waves publish column panel A[:, K], row panel A[K, :] with I
//      added to its A_KK part, and the next pivot block to LDS; barrier;
//   2. the MFMA waves: A += C B with FP64 MFMA 16x16x4 (rank 8), where
//      C[i, :] = -A[i, K] P8 (i not in K), C[K, :] = P8 - I and
//      B = A[K, :] + [0 I 0], which in one update gives the trailing
//      entries A_ij - A_iK P8 A_Kj, row K = P8 A_Kj, column K = -A_iK P8 and
//      block KK = P8; meanwhile the chain wave applies the step to the next
//      pivot block and inverts it (gj8_wave), so the serial pivot chain runs
//      beside the MFMAs.
// Roles by wave: a CU places wave w of a workgroup on SIMD w % 4, and FP64
// MFMA and FP64 VALU work on one SIMD do not overlap, so the chain wave (3)
// has SIMD 3 to itself while waves 0, 1 (two tiles each), 2, 4, 5, 6 (one
// each) keep 3 / 3 / 2 tiles on SIMDs 0 / 1 / 2.  Wave 7 and waves beyond the
// eighth take part in the barriers only.
// LDS buffers alternate with the parity of pb: one barrier per step.  Writes
// A_KK = P and R_K = P (in T) and PT = P^T (FP64) for the row-panel kernel;
// a zero pivot sets *singular (the whole inverse is then reported as zero,
// MatrixOps.c:60-68).
struct PivLds {
    double colraw[2][kNb][kPb];  // A[:, K]
    double rowraw[2][kPb][kNb];  // A[K, :] + [0 I 0]
    double pblk[2][kPb][kPb];    // P8 of step pb (parity pb)
    double nextblk[2][kPb][kPb]; // A[K+1, K+1] before step pb
    double rnext[kPb][kPb];      // chain wave: (P8 A[K, K+1])
};

constexpr int kChainWave = 3;
// 16-row tiles of wave w: count and first tile; owner wave and slot of tile rt
__device__ constexpr int piv_tiles(int w) { return w < 2 ? 2 : (w == 2 || (w >= 4 && w <= 6)) ? 1 : 0; }
__device__ constexpr int piv_first_tile(int w) { return w < 3 ? 2 * w : w + 1; }
__device__ constexpr int piv_owner(int rt) { return rt < 4 ? rt >> 1 : (rt == 4 ? 2 : rt - 1); }
__device__ constexpr int piv_slot(int rt) { return rt < 4 ? (rt & 1) : 0; }

// T is the storage type of the matrix and of R (float for GJ_MIXED, double
// for GJ_BLOCKED64); the block itself is inverted in FP64 either way.
template <class T>
__device__ __attribute__((always_inline)) inline void pivot_block(T *__restrict__ a, int64_t np, int32_t kp,
                                   double *__restrict__ pt, T *__restrict__ rp,
                                   int32_t *__restrict__ singular, PivLds &L)
{
    auto &colraw = L.colraw;
    auto &rowraw = L.rowraw;
    auto &pblk = L.pblk;
    auto &nextblk = L.nextblk;
    auto &rnext = L.rnext;
    constexpr int kSteps = kNb / kPb;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int l16 = lane & 15, h = lane >> 4;
    const int pi = lane >> 3, pj = lane & 7;  // chain wave: entry of the 8 x 8 block
    const int64_t k0 = static_cast<int64_t>(kp) * kNb;
    const int ntile = wave < 8 ? piv_tiles(wave) : 0;
    const int rt0 = piv_first_tile(wave);
    f64x4 acc[2][8];
    bool zero = false;
    if (ntile > 0) {
#pragma unroll
        for (int x = 0; x < 2; ++x)
            if (x < ntile) {
#pragma unroll
                for (int y = 0; y < 8; ++y)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int64_t row = 16 * (rt0 + x) + h + 4 * e;
                        acc[x][y][e] = static_cast<double>(a[(k0 + row) * np + k0 + 16 * y + l16]);
                    }
            }
    } else if (wave == kChainWave) {
        const double v = static_cast<double>(a[(k0 + pi) * np + k0 + pj]);
        pblk[0][pi][pj] = gj8_wave(v, pi, pj, zero);
    }
    __syncthreads();
    static_for(std::make_integer_sequence<int, kSteps>{}, [&](auto pbc) {
        constexpr int pb = decltype(pbc)::value;
        constexpr int buf = pb & 1;
        constexpr int yk = pb >> 1, half = pb & 1;  // column tile / lane half of K
        constexpr int wk = piv_owner(pb >> 1), xk = piv_slot(pb >> 1);  // row tile of K
        const bool lane_in_k = (l16 >> 3) == half;
        // 1. publish
        if (ntile > 0) {
            if (lane_in_k) {
#pragma unroll
                for (int x = 0; x < 2; ++x)
                    if (x < ntile) {
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            colraw[buf][16 * (rt0 + x) + h + 4 * e][l16 - 8 * half] = acc[x][yk][e];
                    }
            }
            if (wave == wk) {
#pragma unroll
                for (int y = 0; y < 8; ++y)
#pragma unroll
                    for (int ee = 0; ee < 2; ++ee) {
                        const int e = 2 * half + ee, r = h + 4 * ee;  // row of K
                        double v = acc[xk][y][e];
                        if (y == yk && lane_in_k && r == l16 - 8 * half) v += 1.0;
                        rowraw[buf][r][16 * y + l16] = v;
                    }
            }
            if constexpr (pb + 1 < kSteps) {
                constexpr int q = pb + 1;
                constexpr int yq = q >> 1, hq = q & 1;
                constexpr int wq = piv_owner(q >> 1), xq = piv_slot(q >> 1);
                if (wave == wq && (l16 >> 3) == hq) {
#pragma unroll
                    for (int ee = 0; ee < 2; ++ee)
                        nextblk[buf][h + 4 * ee][l16 - 8 * hq] = acc[xq][yq][2 * hq + ee];
                }
            }
        }
        __syncthreads();
        if (ntile > 0) {
            // 2. C in the MFMA A layout (lane l: row 16 rt + l%16, k = 4c + l/16)
            double pcol[2][kPb];
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int m = 0; m < kPb; ++m) pcol[c][m] = pblk[buf][m][4 * c + h];
            double av[2][2], bv[8][2];
#pragma unroll
            for (int x = 0; x < 2; ++x)
                if (x < ntile) {
                    const int r = 16 * (rt0 + x) + l16;
                    double cr[kPb];
#pragma unroll
                    for (int m = 0; m < kPb; ++m) cr[m] = colraw[buf][r][m];
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        double s = 0.0;
#pragma unroll
                        for (int m = 0; m < kPb; ++m) s = fma(cr[m], pcol[c][m], s);
                        av[x][c] = -s;
                        if (wave == wk && x == xk && lane_in_k)  // row r of K: (P8 - I)
                            av[x][c] = pblk[buf][r & 7][4 * c + h] - ((r & 7) == 4 * c + h ? 1.0 : 0.0);
                    }
                }
#pragma unroll
            for (int y = 0; y < 8; ++y)
#pragma unroll
                for (int c = 0; c < 2; ++c) bv[y][c] = rowraw[buf][4 * c + h][16 * y + l16];
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int x = 0; x < 2; ++x)
                    if (x < ntile) {
#pragma unroll
                        for (int y = 0; y < 8; ++y)
                            acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x][c], bv[y][c],
                                                                           acc[x][y], 0, 0, 0);
                    }
        } else if (wave == kChainWave) {
            if constexpr (pb + 1 < kSteps) {
                // next pivot block: A[K+1, K+1] - A[K+1, K] P8 A[K, K+1], inverted
                constexpr int n0 = kPb * (pb + 1);
                double r = 0.0;
#pragma unroll
                for (int k = 0; k < kPb; ++k) r = fma(pblk[buf][pi][k], rowraw[buf][k][n0 + pj], r);
                rnext[pi][pj] = r;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                double v = nextblk[buf][pi][pj];
#pragma unroll
                for (int m = 0; m < kPb; ++m) v = fma(-colraw[buf][n0 + pi][m], rnext[m][pj], v);
                pblk[buf ^ 1][pi][pj] = gj8_wave(v, pi, pj, zero);
            }
        }
    });
    if (ntile == 0) {
        if (wave == kChainWave && zero && lane == 0) *singular = 1;
        return;
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
        if (x < ntile) {
#pragma unroll
            for (int y = 0; y < 8; ++y)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t row = 16 * (rt0 + x) + h + 4 * e;
                    const int64_t col = 16 * y + l16;
                    const double v = acc[x][y][e];
                    a[(k0 + row) * np + k0 + col] = static_cast<T>(v);
                    rp[row * np + k0 + col] = static_cast<T>(v);
                    pt[col * kNb + row] = v;
                }
        }
}

template <class T>
__global__ void __launch_bounds__(kPivThreads)
bgj_pivot_kernel(T *__restrict__ a, int64_t np, int32_t kp, double *__restrict__ pt,
                 T *__restrict__ rp, int32_t *__restrict__ singular)
{
    __shared__ PivLds lds;
    pivot_block<T>(a, np, kp, pt, rp, singular, lds);
}

// Row panel and column panel of block step kp.
//  blockIdx.x < np/32: R[:, j0:j0+32] = P A[K, j0:j0+32] with FP64 MFMA
//    16x16x4 (wave w: rows 32w..32w+31 of R, 2 x 2 blocks), written to R
//    and to row panel K of A (FP32); column tiles inside panel K skip.
//  otherwise: CT[:, i0:i0+64] = -A[i0:i0+64, K]^T through an LDS transpose,
//    row tiles inside panel K skip.
template <class T>
__global__ void __launch_bounds__(kThreads)
bgj_panel_kernel(T *__restrict__ a, int64_t np, int32_t kp, const double *__restrict__ pt,
                 T *__restrict__ rp, T *__restrict__ ct)
{
    const int64_t k0 = static_cast<int64_t>(kp) * kNb;
    const int32_t n_col_tiles = static_cast<int32_t>(np / 32);
    if (static_cast<int32_t>(blockIdx.x) < n_col_tiles) {
        const int64_t j0 = 32 * static_cast<int64_t>(blockIdx.x);
        if (j0 / kNb == kp) return;
        const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
        const int l16 = lane & 15, h = lane >> 4;
        f64x4 acc[2][2];
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) acc[x][y] = f64x4{0.0, 0.0, 0.0, 0.0};
        const double *pa = pt + static_cast<int64_t>(h) * kNb + 32 * wave + l16;
        const T *pb = a + (k0 + h) * np + j0 + l16;
#pragma unroll 4
        for (int s = 0; s < kNb / 4; ++s) {
            const double a0 = pa[4 * s * kNb], a1 = pa[4 * s * kNb + 16];
            const double b0 = static_cast<double>(pb[4 * s * np]);
            const double b1 = static_cast<double>(pb[4 * s * np + 16]);
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();  // every wave has read rows K of A before any is overwritten
        // 16x16x4 FP64 accumulator layout (measured on gfx950): register e of
        // lane l holds D[l/16 + 4e][l%16]
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t row = 32 * wave + 16 * x + h + 4 * e;
                    const int64_t col = j0 + 16 * y + l16;
                    const T v = static_cast<T>(acc[x][y][e]);
                    rp[row * np + col] = v;
                    a[(k0 + row) * np + col] = v;
                }
        return;
    }
    __shared__ T tile[kNb][65];
    const int64_t i0 = 64 * static_cast<int64_t>(blockIdx.x - n_col_tiles);
    if (i0 >= np || i0 / kNb == kp) return;
#pragma unroll 4
    for (int e = threadIdx.x; e < 64 * kNb; e += kThreads) {
        const int i = e / kNb, k = e % kNb;
        tile[k][i] = a[(i0 + i) * np + k0 + k];
    }
    __syncthreads();
#pragma unroll 4
    for (int e = threadIdx.x; e < 64 * kNb; e += kThreads) {
        const int k = e / 64, i = e % 64;
        ct[static_cast<int64_t>(k) * np + i0 + i] = -tile[k][i];
    }
}

// Trailing update of block step kp: every 128 x 128 tile (ti, tj), ti != kp,
//   A_tile = (tj == kp ? 0 : A_tile) + CT[:, ti]^T R[:, tj]
// with MFMA 32x32x2 FP32 (FP32 accumulate).  Persistent workgroups of eight
// waves, one per CU, each walking its tiles:
//  * waves 4-7 stream the operands -- the 128 x kTrKc chunks of the CT slab
//    (block row ti) and of the R slab (block column tj), k-major -- into
//    double-buffered LDS (global loads issued two chunks ahead into one of
//    two register sets, written to the buffer the MFMA waves just left);
//  * waves 0-3 (a 64 x 64 quadrant each, 2 x 2 MFMA blocks) issue the loads
//    of the old tile as the tile starts and accumulate the first chunk's
//    products from zero while they land (operands from LDS only); the old
//    tile is then added and the other chunks accumulate onto it, so the
//    HBM traffic of a tile overlaps its MFMAs.
// One barrier per chunk.  Tiles are grouped per XCD (workgroup b runs on XCD
// b % 8; the workgroups of one XCD share one contiguous range of the strip
// order of trailing_tile), so the CT and R slabs in use stay in one XCD's L2.
//
// With `fuse`, workgroup 0 instead updates tile (kp+1, kp+1) itself (it is
// taken out of the other workgroups' tile order) and then inverts that block
// as the pivot of step kp+1 (pivot_block, writing R_{kp+1} = P into rp_next):
// the serial pivot inversion runs beside the trailing tiles.
constexpr int kTrKc = 32;                 // k rows per operand chunk
constexpr int kTrChunks = kNb / kTrKc;    // chunks per tile
constexpr int kTrThreads = 8 * 64;        // 4 MFMA waves + 4 loader waves
constexpr int kTrLd = kNb + 32;           // LDS row stride (floats): the two
                                          // half-waves read different banks
constexpr int kTrWgPerXcd = 32;           // one workgroup per CU (256 CUs)

struct TrLds {
    float ctl[2][kTrKc][kTrLd];
    float rl[2][kTrKc][kTrLd];
};

union GjLds {
    TrLds t;
    PivLds p;
};

// Tile t of block step kp -> (ti, tj): the nt - 1 block rows other than kp
// are taken in strips of four, each strip column by column, so that the
// tiles an XCD works on at one time share few CT and R slabs.
__device__ inline void trailing_tile(int32_t t, int32_t nt, int32_t kp, int32_t &ti, int32_t &tj)
{
    const int32_t strip = t / (4 * nt);
    const int32_t hgt = min(4, nt - 1 - 4 * strip);
    const int32_t off = t - strip * 4 * nt;
    tj = off / hgt;
    ti = 4 * strip + (off - tj * hgt);
    if (ti >= kp) ++ti;
}

// Index in the tile order with tile t_skip taken out.
__device__ inline int32_t skip_tile(int32_t t, int32_t t_skip) { return t >= t_skip ? t + 1 : t; }

// Tiles of this workgroup: first, first + stride, ... (n_my of them).  XCD x
// has wpx workgroups (wpx - fuse for XCD 0, whose workgroup 0 is the pivot
// one) and the range [n cum(x) / total, n cum(x+1) / total) of the tiles,
// cum(x) = workgroups on the XCDs before x.
struct TrPlan {
    int32_t first, stride, n_my;
};

__device__ inline TrPlan trailing_plan(int32_t n_tiles, int32_t wpx, int32_t fuse)
{
    const int32_t b = static_cast<int32_t>(blockIdx.x), x = b % kXcd;
    const int32_t total = kXcd * wpx - fuse;
    const int32_t cum0 = x == 0 ? 0 : x * wpx - fuse, cum1 = (x + 1) * wpx - fuse;
    const int32_t lo = static_cast<int32_t>(static_cast<int64_t>(n_tiles) * cum0 / total);
    const int32_t hi = static_cast<int32_t>(static_cast<int64_t>(n_tiles) * cum1 / total);
    TrPlan pl;
    pl.stride = (x == 0) ? wpx - fuse : wpx;
    pl.first = lo + b / kXcd - ((x == 0) ? fuse : 0);
    pl.n_my = pl.first < hi ? (hi - pl.first + pl.stride - 1) / pl.stride : 0;
    return pl;
}

__device__ __attribute__((always_inline)) inline void trailing_tiles(float *__restrict__ a, const float *__restrict__ ct,
                                      const float *__restrict__ rp, int64_t np, int32_t nt,
                                      int32_t kp, int32_t t_skip, TrPlan pl, TrLds &L)
{
    auto &ctl = L.ctl;
    auto &rl = L.rl;
    const int32_t first = pl.first, stride = pl.stride, n_my = pl.n_my;
    const int32_t n_chunks = n_my * kTrChunks;
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (wave >= 4) {
        // loader: chunk g = (tile g / kTrChunks, rows kTrKc (g % kTrChunks)..)
        const int lt = threadIdx.x - 4 * kWave;
        constexpr int kPer = kTrKc * kNb / 4 / (4 * kWave);  // float4 per thread per slab
        static_assert(kPer == 4, "loader holds four float4 per slab");
        // two register sets: the loads of chunk g + 2 are issued while chunk g
        // is multiplied, and land in LDS as the MFMA waves move to chunk g + 1
        float4 ra0_0, ra0_1, ra0_2, ra0_3, rb0_0, rb0_1, rb0_2, rb0_3;
        float4 ra1_0, ra1_1, ra1_2, ra1_3, rb1_0, rb1_1, rb1_2, rb1_3;
        const int row0 = lt / (kNb / 4), col0 = 4 * (lt % (kNb / 4));  // q-th: row0 + 8q
        constexpr int kRowStep = 4 * kWave / (kNb / 4);
#define GSLAM_TR_LOAD(g, S)                                                                      \
    {                                                                                            \
        int32_t ti_, tj_;                                                                        \
        trailing_tile(skip_tile(first + ((g) / kTrChunks) * stride, t_skip), nt, kp, ti_, tj_);                     \
        const int64_t k0_ = static_cast<int64_t>((g) % kTrChunks) * kTrKc + row0;                 \
        const float *pct = ct + k0_ * np + static_cast<int64_t>(ti_) * kNb + col0;               \
        const float *prp = rp + k0_ * np + static_cast<int64_t>(tj_) * kNb + col0;               \
        ra##S##_0 = *reinterpret_cast<const float4 *>(pct);                                       \
        ra##S##_1 = *reinterpret_cast<const float4 *>(pct + kRowStep * np);                       \
        ra##S##_2 = *reinterpret_cast<const float4 *>(pct + 2 * kRowStep * np);                   \
        ra##S##_3 = *reinterpret_cast<const float4 *>(pct + 3 * kRowStep * np);                   \
        rb##S##_0 = *reinterpret_cast<const float4 *>(prp);                                       \
        rb##S##_1 = *reinterpret_cast<const float4 *>(prp + kRowStep * np);                       \
        rb##S##_2 = *reinterpret_cast<const float4 *>(prp + 2 * kRowStep * np);                   \
        rb##S##_3 = *reinterpret_cast<const float4 *>(prp + 3 * kRowStep * np);                   \
    }
#define GSLAM_TR_STORE(b, S)                                                                     \
    {                                                                                            \
        *reinterpret_cast<float4 *>(&ctl[(b)][row0][col0]) = ra##S##_0;                           \
        *reinterpret_cast<float4 *>(&ctl[(b)][row0 + kRowStep][col0]) = ra##S##_1;                \
        *reinterpret_cast<float4 *>(&ctl[(b)][row0 + 2 * kRowStep][col0]) = ra##S##_2;            \
        *reinterpret_cast<float4 *>(&ctl[(b)][row0 + 3 * kRowStep][col0]) = ra##S##_3;            \
        *reinterpret_cast<float4 *>(&rl[(b)][row0][col0]) = rb##S##_0;                            \
        *reinterpret_cast<float4 *>(&rl[(b)][row0 + kRowStep][col0]) = rb##S##_1;                 \
        *reinterpret_cast<float4 *>(&rl[(b)][row0 + 2 * kRowStep][col0]) = rb##S##_2;             \
        *reinterpret_cast<float4 *>(&rl[(b)][row0 + 3 * kRowStep][col0]) = rb##S##_3;             \
    }
        // chunk c sits in register set c & 1 and LDS buffer c & 1
        if (n_chunks > 0) {
            GSLAM_TR_LOAD(0, 0)
            GSLAM_TR_STORE(0, 0)
        }
        if (n_chunks > 1) GSLAM_TR_LOAD(1, 1)
        if (n_chunks > 2) GSLAM_TR_LOAD(2, 0)
        __syncthreads();
        for (int32_t g = 0; g < n_chunks; g += 2) {
            // g even: store chunk g + 1 (set 1), load chunk g + 3 (set 1)
            if (g + 1 < n_chunks) GSLAM_TR_STORE(1, 1)
            if (g + 3 < n_chunks) GSLAM_TR_LOAD(g + 3, 1)
            __syncthreads();
            if (g + 1 >= n_chunks) break;
            // g + 1 odd: store chunk g + 2 (set 0), load chunk g + 4 (set 0)
            if (g + 2 < n_chunks) GSLAM_TR_STORE(0, 0)
            if (g + 4 < n_chunks) GSLAM_TR_LOAD(g + 4, 0)
            __syncthreads();
        }
#undef GSLAM_TR_LOAD
#undef GSLAM_TR_STORE
        return;
    }
    const int l32 = lane & 31, h = lane >> 5;
    const int qi = 64 * (wave >> 1), qj = 64 * (wave & 1);
    __syncthreads();
    for (int32_t it = 0; it < n_my; ++it) {
        int32_t ti, tj;
        trailing_tile(skip_tile(first + it * stride, t_skip), nt, kp, ti, tj);
        float *cbase = a + (static_cast<int64_t>(ti) * kNb + qi + 4 * h) * np +
                       static_cast<int64_t>(tj) * kNb + qj + l32;
        f32x16 cold[2][2], acc[2][2];
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    cold[x][y][e] = tj != kp ? cbase[(32 * x + 8 * (e >> 2) + (e & 3)) * np + 32 * y] : 0.f;
                    acc[x][y][e] = 0.f;
                }
        for (int c = 0; c < kTrChunks; ++c) {
            const int b = (it * kTrChunks + c) & 1;
            const float *pa = &ctl[b][h][qi + l32];
            const float *pb = &rl[b][h][qj + l32];
#pragma unroll
            for (int s2 = 0; s2 < kTrKc / 2; ++s2) {
                const float a0 = pa[2 * s2 * kTrLd], a1 = pa[2 * s2 * kTrLd + 32];
                const float b0 = pb[2 * s2 * kTrLd], b1 = pb[2 * s2 * kTrLd + 32];
                acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[1][1], 0, 0, 0);
            }
            __syncthreads();
            if (c == 0) {
                // the old tile joins the sum after the first chunk, so the
                // remaining products accumulate onto it as in an in-place update
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y) acc[x][y] += cold[x][y];
            }
        }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int e = 0; e < 16; ++e)
                    cbase[(32 * x + 8 * (e >> 2) + (e & 3)) * np + 32 * y] = acc[x][y][e];
    }
}

__global__ void __launch_bounds__(kTrThreads)
bgj_trailing_kernel(float *__restrict__ a, const float *__restrict__ ct,
                    const float *__restrict__ rp, int64_t np, int32_t nt, int32_t kp,
                    int32_t wpx, int32_t fuse, int32_t t_skip, int32_t n_tail,
                    float *__restrict__ rp_next, double *__restrict__ pt,
                    int32_t *__restrict__ singular)
{
    __shared__ __attribute__((aligned(16))) GjLds lds;
    if (fuse && blockIdx.x == 0) {
        // tile (kp+1, kp+1) -- position t_skip of the full tile order -- by
        // the same pipeline as every other tile, then its inverse
        TrPlan one;
        one.first = t_skip;
        one.stride = 1;
        one.n_my = 1;
        trailing_tiles(a, ct, rp, np, nt, kp, (nt - 1) * nt + 1, one, lds.t);
        __syncthreads();  // the updated block is visible to the workgroup
        pivot_block<float>(a, np, kp + 1, pt, rp_next, singular, lds.p);
        if (n_tail > 0) {
            // then the last n_tail tiles of the others' order
            __syncthreads();  // lds.p and lds.t share memory
            TrPlan tail;
            tail.first = (nt - 1) * nt - 1 - n_tail;
            tail.stride = 1;
            tail.n_my = n_tail;
            trailing_tiles(a, ct, rp, np, nt, kp, t_skip, tail, lds.t);
        }
        return;
    }
    const int32_t n_tiles = (nt - 1) * nt - fuse - n_tail;
    trailing_tiles(a, ct, rp, np, nt, kp, fuse ? t_skip : n_tiles + 1,
                   trailing_plan(n_tiles, wpx, fuse), lds.t);
}

// ---------------------------------------------------------------- FP64 blocked

// GJ_BLOCKED64 trailing update of block step kp, FP64 throughout: every 128 x
// 128 tile (ti, tj), ti != kp, A_tile = (tj == kp ? 0 : A_tile) + CT^T R.
// A tile is worked by four waves (a 64 x 64 quadrant each, 4 x 4 FP64 MFMA
// 16x16x4 blocks accumulated onto the old tile, operands read as 128-byte
// rows from L2).  An FP64 MFMA takes about four times as long per tile as
// the tile's own HBM traffic, so the two tiles in flight per CU keep its
// matrix pipes busy without staging through LDS.
__device__ __attribute__((always_inline)) inline void trailing64_tile(
    double *__restrict__ a, const double *__restrict__ ct, const double *__restrict__ rp,
    int64_t np, int32_t nt, int32_t kp, int32_t t, int quad)
{
    int32_t ti, tj;
    trailing_tile(t, nt, kp, ti, tj);
    const int lane = threadIdx.x % kWave;
    const int l16 = lane & 15, h = lane >> 4;
    const int64_t i0 = static_cast<int64_t>(ti) * kNb + 64 * (quad >> 1);
    const int64_t j0 = static_cast<int64_t>(tj) * kNb + 64 * (quad & 1);
    double *cbase = a + (i0 + h) * np + j0 + l16;  // (x, y, e): row 16x + 4e, col 16y
    f64x4 acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                acc[x][y][e] = tj != kp ? cbase[(16 * x + 4 * e) * np + 16 * y] : 0.0;
    const double *pa = ct + static_cast<int64_t>(h) * np + i0 + l16;
    const double *pb = rp + static_cast<int64_t>(h) * np + j0 + l16;
#pragma unroll 2
    for (int s4 = 0; s4 < kNb / 4; ++s4) {
        const int64_t o = 4 * static_cast<int64_t>(s4) * np;
        double av[4], bv[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) av[x] = pa[o + 16 * x];
#pragma unroll
        for (int y = 0; y < 4; ++y) bv[y] = pb[o + 16 * y];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y)
                acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], acc[x][y], 0, 0, 0);
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int e = 0; e < 4; ++e) cbase[(16 * x + 4 * e) * np + 16 * y] = acc[x][y][e];
}

// Tiles of worker `half` (waves 4 half .. 4 half + 3) of this workgroup: as
// trailing_plan, with two workers per workgroup and, under `fuse`, both
// workers of workgroup 0 (the pivot one) taken out of XCD 0's share.
__device__ inline TrPlan trailing64_plan(int32_t n_tiles, int32_t wpx, int32_t fuse, int32_t half)
{
    const int32_t b = static_cast<int32_t>(blockIdx.x), x = b % kXcd;
    const int32_t f2 = 2 * fuse;
    const int32_t total = 2 * kXcd * wpx - f2;
    const int32_t cum0 = x == 0 ? 0 : 2 * x * wpx - f2, cum1 = 2 * (x + 1) * wpx - f2;
    const int32_t lo = static_cast<int32_t>(static_cast<int64_t>(n_tiles) * cum0 / total);
    const int32_t hi = static_cast<int32_t>(static_cast<int64_t>(n_tiles) * cum1 / total);
    TrPlan pl;
    pl.stride = (x == 0) ? 2 * wpx - f2 : 2 * wpx;
    pl.first = lo + 2 * (b / kXcd) + half - ((x == 0) ? f2 : 0);
    pl.n_my = pl.first < hi ? (hi - pl.first + pl.stride - 1) / pl.stride : 0;
    return pl;
}

// Persistent workgroups of eight waves, one per CU, each running two tile
// workers (waves 0-3 and 4-7) with no barrier between them.  With `fuse`,
// workgroup 0 instead updates tile (kp+1, kp+1) (taken out of the other
// workers' tile order) and inverts it as the pivot block of step kp+1
// (pivot_block, writing R_{kp+1} = P into rp_next), beside the other tiles,
// and then the last n_tail tiles of the others' order (trailing_tail).
__global__ void __launch_bounds__(kTrThreads)
bgj_trailing64_kernel(double *__restrict__ a, const double *__restrict__ ct,
                      const double *__restrict__ rp, int64_t np, int32_t nt, int32_t kp,
                      int32_t wpx, int32_t fuse, int32_t t_skip, int32_t n_tail,
                      double *__restrict__ rp_next, double *__restrict__ pt,
                      int32_t *__restrict__ singular)
{
    __shared__ PivLds lds;
    const int wave = threadIdx.x / kWave;
    if (fuse && blockIdx.x == 0) {
        if (wave < 4) trailing64_tile(a, ct, rp, np, nt, kp, t_skip, wave);
        __syncthreads();  // the updated block is visible to the workgroup
        pivot_block<double>(a, np, kp + 1, pt, rp_next, singular, lds);
        const int32_t end = (nt - 1) * nt - 1;
        for (int32_t t = end - n_tail + (wave >> 2); t < end; t += 2)
            trailing64_tile(a, ct, rp, np, nt, kp, skip_tile(t, t_skip), wave & 3);
        return;
    }
    const int32_t n_tiles = (nt - 1) * nt - fuse - n_tail;
    const int32_t skip = fuse ? t_skip : n_tiles + 1;
    const TrPlan pl = trailing64_plan(n_tiles, wpx, fuse, wave >> 2);
    for (int32_t i = 0; i < pl.n_my; ++i)
        trailing64_tile(a, ct, rp, np, nt, kp, skip_tile(pl.first + i * pl.stride, skip), wave & 3);
}

// a = identity on the padding, zero elsewhere (np x np FP64).
__global__ void __launch_bounds__(kThreads)
bgj_clear64_kernel(double *__restrict__ a, int64_t n, int64_t np)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= np * np) return;
    const int64_t i = t / np, j = t - i * np;
    a[t] = (i >= n && i == j) ? 1.0 : 0.0;
}

// Jacobi-scaled BSR blocks of Omega (dsc_i Omega_ij dsc_j) into the padded
// FP64 matrix.
__global__ void __launch_bounds__(kThreads)
bgj_scatter_bsr64_kernel(const int32_t *__restrict__ brow_ptr, const int32_t *__restrict__ bcol,
                         const double *__restrict__ bval, int64_t n_blocks, int32_t n_rows,
                         const double *__restrict__ dsc, int64_t np, double *__restrict__ a)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= 9 * n_blocks) return;
    const int64_t b = t / 9;
    const int k = static_cast<int>(t - 9 * b);
    int32_t lo = 0, hi = n_rows - 1;  // block row of b
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (brow_ptr[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const int64_t i = 3 * static_cast<int64_t>(lo) + k / 3;
    const int64_t j = 3 * static_cast<int64_t>(bcol[b]) + k % 3;
    a[i * np + j] = bval[t] * dsc[i] * dsc[j];
}

// n x n host-layout matrix into the padded matrix, Jacobi-scaled, and the
// inverse back out (unscaled; zeros if the inversion met a zero pivot).
__global__ void __launch_bounds__(kThreads)
bgj_load_dense64_kernel(const double *__restrict__ m, int64_t n, int64_t np,
                        const double *__restrict__ dsc, double *__restrict__ a)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= np * np) return;
    const int64_t i = t / np, j = t - i * np;
    a[t] = (i < n && j < n) ? m[i * n + j] * dsc[i] * dsc[j] : (i == j ? 1.0 : 0.0);
}

__global__ void __launch_bounds__(kThreads)
bgj_store_dense64_kernel(const double *__restrict__ a, int64_t n, int64_t np,
                         const double *__restrict__ dsc, const int32_t *__restrict__ singular,
                         double *__restrict__ m)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n * n) return;
    const int64_t i = t / n, j = t - i * n;
    m[t] = *singular ? 0.0 : dsc[i] * a[i * np + j] * dsc[j];
}

// x = D A[0:n, 0:n] D b (D = diag(dsc)) with the padded row stride np, one
// wave per row: the solve with the inverse of the scaled matrix.
__global__ void __launch_bounds__(kThreads)
bgj_matvec64_kernel(const double *__restrict__ a, int64_t np, int64_t n,
                    const double *__restrict__ dsc, const double *__restrict__ b,
                    double *__restrict__ x)
{
    const int lane = threadIdx.x % kWave;
    const int64_t wave0 = (blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x) / kWave;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    for (int64_t i = wave0; i < n; i += nwaves) {
        const double *row = a + i * np;
        double s = 0.0;
        for (int64_t j = lane; j < n; j += kWave) s = fma(row[j], dsc[j] * b[j], s);
        s = wave_sum(s);
        if (lane == 0) x[i] = dsc[i] * s;
    }
}

// ---------------------------------------------------------------- refinement

// MODE 0: out = Omega v, partial v.out.
// MODE 1: out = b - Omega v (v == NULL: v = 0), u = dsc * out, partial out.out
//         and (first call) b.b.
template <int MODE>
__global__ void __launch_bounds__(kThreads)
bgj_spmv_kernel(const int32_t *__restrict__ brow_ptr, const int32_t *__restrict__ bcol,
                const double *__restrict__ bval, const double *__restrict__ v,
                const double *__restrict__ b, const double *__restrict__ dsc,
                double *__restrict__ out, double *__restrict__ u, int32_t n,
                double *__restrict__ part, const PcgStatus *status)
{
    if (status && status->done) return;
    constexpr int G = 4, kRowsPerTile = kThreads / G;
    const int lane = threadIdx.x % G;
    double s[2] = {0.0, 0.0};
    for (int64_t row0 = static_cast<int64_t>(blockIdx.x) * kRowsPerTile; row0 < n;
         row0 += static_cast<int64_t>(gridDim.x) * kRowsPerTile) {
        const int64_t row = row0 + threadIdx.x / G;
        double y0 = 0.0, y1 = 0.0, y2 = 0.0;
        if (row < n && v) {
            const int32_t b1 = brow_ptr[row + 1];
            for (int32_t bb = brow_ptr[row] + lane; bb < b1; bb += G) {
                const int64_t j = 3 * static_cast<int64_t>(bcol[bb]);
                const double *m = bval + 9 * static_cast<int64_t>(bb);
                y0 += m[0] * v[j] + m[1] * v[j + 1] + m[2] * v[j + 2];
                y1 += m[3] * v[j] + m[4] * v[j + 1] + m[5] * v[j + 2];
                y2 += m[6] * v[j] + m[7] * v[j + 1] + m[8] * v[j + 2];
            }
        }
        y0 = group_sum<G>(y0);
        y1 = group_sum<G>(y1);
        y2 = group_sum<G>(y2);
        if (row < n && lane == 0) {
            const int64_t o = 3 * row;
            const double y[3] = {y0, y1, y2};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (MODE == 0) {
                    out[o + c] = y[c];
                    s[0] += v[o + c] * y[c];
                } else {
                    const double rr = b[o + c] - y[c];
                    out[o + c] = rr;
                    u[o + c] = dsc[o + c] * rr;
                    s[0] += rr * rr;
                    s[1] += b[o + c] * b[o + c];
                }
            }
        }
    }
    write_partials<2>(s, part, kPartStride);
}

// z = D^-1/2 M u over the n real rows, one wavefront per row (FP32 matrix
// entries, FP64 products and sums).  IR: x += z.  CG: z stored, partials of
// z.r_new and z.r_old.
template <bool kIR>
__global__ void __launch_bounds__(kThreads)
bgj_apply_kernel(const float *__restrict__ a, int64_t np, int64_t n,
                 const double *__restrict__ u, const double *__restrict__ dsc,
                 double *__restrict__ out, const double *__restrict__ r_new,
                 const double *__restrict__ r_old, double *__restrict__ part,
                 const PcgStatus *status)
{
    if (status->done) return;
    const int lane = threadIdx.x % kWave;
    const int64_t wave0 = (blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x) / kWave;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    double v[2] = {0.0, 0.0};
    for (int64_t i = wave0; i < n; i += nwaves) {
        const float *row = a + i * np;
        double s = 0.0;
        for (int64_t j = 4 * lane; j < np; j += 4 * kWave) {
            const float4 m4 = *reinterpret_cast<const float4 *>(row + j);
            s += static_cast<double>(m4.x) * u[j] + static_cast<double>(m4.y) * u[j + 1] +
                 static_cast<double>(m4.z) * u[j + 2] + static_cast<double>(m4.w) * u[j + 3];
        }
        s = wave_sum(s);
        if (lane == 0) {
            const double z = dsc[i] * s;
            if (kIR) {
                out[i] += z;
            } else {
                out[i] = z;
                v[0] += z * r_new[i];
                v[1] += z * r_old[i];
            }
        }
    }
    if (!kIR) write_partials<2>(v, part, kPartStride);
}

// Stopping test on ||r||^2 partials (and ||b||^2 at init) for the host.
__global__ void __launch_bounds__(kThreads)
bgj_check_kernel(const double *__restrict__ part, int32_t np, int32_t k, int32_t cap, int init,
                 double tol_rel, PcgStatus *status)
{
    if (!init && status->done) return;
    double v[2];
    sum_partials<2>(part, np, kPartStride, v);
    if (threadIdx.x != 0) return;
    if (init) {
        status->done = 0;
        status->converged = 0;
        status->flags = 0;
        status->refreshes = 0;
        status->bnorm = sqrt(v[1]);
        status->tol = tol_rel * sqrt(v[1]);
    }
    status->iters = k;
    status->rnorm = sqrt(v[0]);
    if (!init) {
        if (sqrt(v[0]) < status->tol) {
            status->done = 1;
            status->converged = 1;
        } else if (k >= cap) {
            status->done = 1;
        }
    }
}

// CG: alpha = rz / p.q; x += alpha p (p.q == 0 stops, MatrixOps.c:357).
__global__ void __launch_bounds__(kThreads)
bgj_cg_x_kernel(const double *__restrict__ p, double *__restrict__ x, int64_t n,
                const double *__restrict__ pq_part, int32_t np_pq,
                const double *__restrict__ rz, PcgStatus *status)
{
    if (status->done) return;
    double pq[2];
    sum_partials<2>(pq_part, np_pq, kPartStride, pq);
    if (pq[0] == 0.0 || !isfinite(pq[0])) {  // p.q == 0, or NaN / Inf (bit 1)
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            status->done = 1;
            status->flags |= pq[0] == 0.0 ? 1 : 2;
        }
        return;
    }
    const double alpha = *rz / pq[0];
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        x[i] += alpha * p[i];
}

// CG: beta = z.(r_new - r_old) / rz_old (0 at init), p = z + beta p,
// rz_new = z.r_new (workgroup 0 stores it).
__global__ void __launch_bounds__(kThreads)
bgj_cg_p_kernel(const double *__restrict__ z, double *__restrict__ p, int64_t n,
                const double *__restrict__ zr_part, int32_t np_zr, const double *__restrict__ rz_old,
                double *__restrict__ rz_new, int init, const PcgStatus *status)
{
    if (status->done) return;
    double zr[2];
    sum_partials<2>(zr_part, np_zr, kPartStride, zr);
    const double beta = init ? 0.0 : (zr[0] - zr[1]) / *rz_old;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[i] = init ? z[i] : z[i] + beta * p[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) *rz_new = zr[0];
}

// ---------------------------------------------------------------- drivers

static inline unsigned int nblk(int64_t count, int per = kThreads)
{
    return static_cast<unsigned int>((count + per - 1) / per);
}

int64_t bgj_padded(int64_t n) { return (n + kNb - 1) / kNb * kNb; }

int DenseWork::reserve(int64_t n_)
{
    const int64_t np_ = bgj_padded(n_);
    if (a && np_ <= np_cap) {
        n = n_;
        np = np_;
        return GSLAM_OK;
    }
    release();
    const size_t f = sizeof(float), d = sizeof(double);
    void *p[8] = {};
    const size_t sz[8] = {static_cast<size_t>(np_ * np_) * f, static_cast<size_t>(kNb * np_) * f,
                          static_cast<size_t>(2 * kNb * np_) * f, static_cast<size_t>(kNb * kNb) * d,
                          static_cast<size_t>(np_) * d, static_cast<size_t>(np_) * d,
                          static_cast<size_t>(6 * np_ + 8) * d, 256};
    for (int i = 0; i < 8; ++i) {
        if (hipMalloc(&p[i], sz[i]) != hipSuccess) {
            for (int j = 0; j < i; ++j) (void)hipFree(p[j]);
            set_error("blocked Gauss-Jordan: out of device memory for a " + std::to_string(np_) +
                      " x " + std::to_string(np_) + " FP32 matrix");
            return GSLAM_ERR_NOMEM;
        }
    }
    a = static_cast<float *>(p[0]);
    ct = static_cast<float *>(p[1]);
    rp = static_cast<float *>(p[2]);
    pt = static_cast<double *>(p[3]);
    dsc = static_cast<double *>(p[4]);
    u = static_cast<double *>(p[5]);
    vec = static_cast<double *>(p[6]);
    flag = static_cast<int32_t *>(p[7]);
    n = n_;
    np = np_cap = np_;
    bytes = 0;
    for (size_t s : sz) bytes += static_cast<int64_t>(s);
    return GSLAM_OK;
}

void DenseWork::release()
{
    void *p[8] = {a, ct, rp, pt, dsc, u, vec, flag};
    for (void *q : p)
        if (q) (void)hipFree(q);
    a = ct = rp = nullptr;
    pt = dsc = u = vec = nullptr;
    flag = nullptr;
    n = np = np_cap = bytes = 0;
}

// Trailing-update launch shape: persistent workgroups per XCD (at most one
// per CU, no more than a share of the tiles).
static int32_t trailing_wpx(int32_t nt)
{
    const int32_t per_xcd = ((nt - 1) * nt + kXcd - 1) / kXcd;
    return std::min(kTrWgPerXcd, per_xcd);
}

// Tiles a fused trailing launch leaves to workgroup 0 once it has made the
// next pivot block.  `workers` tile workers take R tiles each and workgroup
// 0's `own` workers R - lag each (lag = rounds of tiles its own tile and the
// pivot block take): R is the fewest rounds that cover n_tiles, and
// workgroup 0 takes what the others' R rounds leave.
static int32_t trailing_tail(int32_t n_tiles, int32_t workers, int32_t own, int32_t lag)
{
    int32_t r = (n_tiles + workers + own - 1) / (workers + own);
    while (static_cast<int64_t>(workers) * r + static_cast<int64_t>(own) * std::max(0, r - lag) <
           n_tiles)
        ++r;
    return std::max(0, n_tiles - workers * r);
}

// Position of tile (kp+1, kp+1) in the strip order of trailing_tile.
static int32_t trailing_pos_next_pivot(int32_t nt, int32_t kp)
{
    const int32_t r = kp;  // block row kp + 1 among the rows other than kp
    const int32_t strip = r / 4, hgt = std::min(4, nt - 1 - 4 * strip);
    return strip * 4 * nt + (kp + 1) * hgt + (r - 4 * strip);
}

// R of block step k (two buffers: step k + 1's pivot writes its R while
// step k's trailing update still reads R_k).
static float *rp_of(DenseWork &w, int32_t k)
{
    return w.rp + static_cast<int64_t>(k & 1) * kNb * w.np_cap;
}

// One block step: pivot (unless the previous step's trailing launch already
// made it), panels, trailing update -- which, when there is a next step and
// at least two workgroups per XCD, also makes the next pivot.  Returns
// whether it did.
static bool bgj_step(hipStream_t s, DenseWork &w, int32_t kp, bool pivot_done)
{
    const int32_t nt = static_cast<int32_t>(w.np / kNb);
    float *rpk = rp_of(w, kp);
    if (!pivot_done)
        hipLaunchKernelGGL(bgj_pivot_kernel<float>, dim3(1), dim3(kPivThreads), 0, s, w.a, w.np, kp, w.pt,
                           rpk, w.flag);
    if (nt == 1) return false;
    const unsigned int gp = static_cast<unsigned int>(w.np / 32 + w.np / 64);
    hipLaunchKernelGGL(bgj_panel_kernel<float>, dim3(gp), dim3(kThreads), 0, s, w.a, w.np, kp, w.pt,
                       rpk, w.ct);
    const int32_t wpx = trailing_wpx(nt);
    const int32_t fuse = (kp + 1 < nt && wpx >= 2) ? 1 : 0;
    // FP32 pivot block: about 52 us, 5 rounds of 11 us tiles (N=5000)
    const int32_t tail = fuse ? trailing_tail((nt - 1) * nt - 1, kXcd * wpx - 1, 1, 7) : 0;
    hipLaunchKernelGGL(bgj_trailing_kernel, dim3(kXcd * wpx), dim3(kTrThreads), 0, s, w.a, w.ct,
                       rpk, w.np, nt, kp, wpx, fuse, fuse ? trailing_pos_next_pivot(nt, kp) : 0,
                       tail, rp_of(w, kp + 1), w.pt, w.flag);
    return fuse != 0;
}

// In-place FP32 blocked inverse of w.a (np x np, padded).  *singular set on
// a zero pivot (flag cleared first).
int bgj_invert(hipStream_t s, DenseWork &w)
{
    GSLAM_HIP(hipMemsetAsync(w.flag, 0, sizeof(int32_t), s));
    const int32_t nt = static_cast<int32_t>(w.np / kNb);
    bool pivot_done = false;
    for (int32_t kp = 0; kp < nt; ++kp) pivot_done = bgj_step(s, w, kp, pivot_done);
    GSLAM_HIP(hipGetLastError());
    return GSLAM_OK;
}

int bgj_time_trailing(hipStream_t s, DenseWork &w, int reps, hipEvent_t e0, hipEvent_t e1,
                      double *us)
{
    const int32_t nt = static_cast<int32_t>(w.np / kNb);
    if (nt < 2) {
        set_error("trailing update timing needs 3N > 128");
        return GSLAM_ERR_STATE;
    }
    const int32_t wpx = trailing_wpx(nt);
    auto launch = [&]() {
        hipLaunchKernelGGL(bgj_trailing_kernel, dim3(kXcd * wpx), dim3(kTrThreads), 0, s, w.a, w.ct,
                           w.rp, w.np, nt, 0, wpx, 0, 0, 0, w.rp, w.pt, w.flag);
    };
    for (int i = 0; i < 3; ++i) launch();
    GSLAM_HIP(hipGetLastError());
    GSLAM_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) launch();
    GSLAM_HIP(hipEventRecord(e1, s));
    GSLAM_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GSLAM_HIP(hipEventElapsedTime(&ms, e0, e1));
    *us = 1000.0 * ms / reps;
    return GSLAM_OK;
}

int dense_solve_mixed(DeviceGraph &g, const gslam_solve_opts &o, const double *b, double *x,
                      gslam_solve_stats &st)
{
    const int64_t n = 3 * static_cast<int64_t>(g.h.n);
    DenseWork &w = g.dense;
    GSLAM_TRY(w.reserve(n));
    const int64_t np = w.np;
    hipStream_t s = g.stream;
    const int32_t nrows = g.h.n;
    double *r0 = w.vec, *r1 = w.vec + np, *z = w.vec + 2 * np, *p = w.vec + 3 * np,
           *q = w.vec + 4 * np, *rz = w.vec + 6 * np;  // rz[2]
    double *rbuf[2] = {r0, r1};
    double *PRs = g.part + kPartOffResid, *PZ = g.part + kPartOffPR0, *PQ = g.part + kPartOffPQ;
    const unsigned int gs = grid_for(nrows, 4);
    const unsigned int gv = grid_for(n, 1);
    const unsigned int ga = static_cast<unsigned int>(std::min<int64_t>(kMaxGrid, (n + 3) / 4));
    PcgStatus *stp = g.status;
    const bool cg = o.refine_mode == GSLAM_REFINE_CG;
    const int32_t cap = std::max(0, o.refine_sweeps);

    GSLAM_HIP(hipEventRecord(g.ev0, s));
    hipLaunchKernelGGL(bgj_clear_kernel, dim3(nblk(np * np / 4)), dim3(kThreads), 0, s, w.a, n, np);
    hipLaunchKernelGGL(bgj_scale_bsr_kernel, dim3(nblk(np)), dim3(kThreads), 0, s, g.bdiag, g.bval,
                       n, np, w.dsc);
    hipLaunchKernelGGL(bgj_scatter_bsr_kernel, dim3(nblk(9 * g.h.n_blocks)), dim3(kThreads), 0, s,
                       g.brow_ptr, g.bcol, g.bval, g.h.n_blocks, nrows, w.dsc, np, w.a);
    GSLAM_TRY(bgj_invert(s, w));
    GSLAM_HIP(hipEventRecord(g.ev2, s));

    // x = 0, r0 = b, u = D^-1/2 b (padding of u zero), ||b||
    GSLAM_HIP(hipMemsetAsync(w.u, 0, static_cast<size_t>(np) * sizeof(double), s));
    GSLAM_HIP(hipMemsetAsync(x, 0, static_cast<size_t>(n) * sizeof(double), s));
    hipLaunchKernelGGL(bgj_spmv_kernel<1>, dim3(gs), dim3(kThreads), 0, s, g.brow_ptr, g.bcol,
                       g.bval, static_cast<const double *>(nullptr), b, w.dsc, r0, w.u, nrows, PRs,
                       static_cast<const PcgStatus *>(nullptr));
    hipLaunchKernelGGL(bgj_check_kernel, dim3(1), dim3(kThreads), 0, s, PRs,
                       static_cast<int32_t>(gs), 0, cap, 1, o.tol, stp);
    int32_t k = 0;
    gslam_solve_opts lo = o;
    std::function<void(int32_t)> enqueue;
    if (!cg) {
        // sweep k = 0 is the FP32 solve x = M b, sweeps 1..cap refine
        lo.max_iter = cap + 1;
        enqueue = [&](int32_t m) {
            for (int32_t c = 0; c < m; ++c, ++k) {
                hipLaunchKernelGGL(bgj_apply_kernel<true>, dim3(ga), dim3(kThreads), 0, s, w.a, np,
                                   n, w.u, w.dsc, x, static_cast<const double *>(nullptr),
                                   static_cast<const double *>(nullptr), PZ, stp);
                hipLaunchKernelGGL(bgj_spmv_kernel<1>, dim3(gs), dim3(kThreads), 0, s, g.brow_ptr,
                                   g.bcol, g.bval, x, b, w.dsc, r0, w.u, nrows, PRs, stp);
                hipLaunchKernelGGL(bgj_check_kernel, dim3(1), dim3(kThreads), 0, s, PRs,
                                   static_cast<int32_t>(gs), k, cap, 0, o.tol, stp);
            }
        };
    } else {
        // z0 = M r0, p = z0, rz = z0.r0
        hipLaunchKernelGGL(bgj_apply_kernel<false>, dim3(ga), dim3(kThreads), 0, s, w.a, np, n,
                           w.u, w.dsc, z, r0, r0, PZ, stp);
        hipLaunchKernelGGL(bgj_cg_p_kernel, dim3(gv), dim3(kThreads), 0, s, z, p, n, PZ,
                           static_cast<int32_t>(ga), rz, rz + 1, 1, stp);
        lo.max_iter = cap;
        k = 0;
        enqueue = [&](int32_t m) {
            for (int32_t c = 0; c < m; ++c) {
                ++k;
                double *rnew = rbuf[k & 1], *rold = rbuf[(k + 1) & 1];
                hipLaunchKernelGGL(bgj_spmv_kernel<0>, dim3(gs), dim3(kThreads), 0, s, g.brow_ptr,
                                   g.bcol, g.bval, p, static_cast<const double *>(nullptr),
                                   w.dsc, q, static_cast<double *>(nullptr), nrows, PQ, stp);
                hipLaunchKernelGGL(bgj_cg_x_kernel, dim3(gv), dim3(kThreads), 0, s, p, x, n, PQ,
                                   static_cast<int32_t>(gs), rz + (k & 1), stp);
                hipLaunchKernelGGL(bgj_spmv_kernel<1>, dim3(gs), dim3(kThreads), 0, s, g.brow_ptr,
                                   g.bcol, g.bval, x, b, w.dsc, rnew, w.u, nrows, PRs, stp);
                hipLaunchKernelGGL(bgj_check_kernel, dim3(1), dim3(kThreads), 0, s, PRs,
                                   static_cast<int32_t>(gs), k, cap, 0, o.tol, stp);
                hipLaunchKernelGGL(bgj_apply_kernel<false>, dim3(ga), dim3(kThreads), 0, s, w.a,
                                   np, n, w.u, w.dsc, z, rnew, rold, PZ, stp);
                hipLaunchKernelGGL(bgj_cg_p_kernel, dim3(gv), dim3(kThreads), 0, s, z, p, n, PZ,
                                   static_cast<int32_t>(ga), rz + (k & 1), rz + ((k + 1) & 1), 0,
                                   stp);
            }
        };
    }
    if (cg && cap == 0) {
        // no refinement: x = M b
        hipLaunchKernelGGL(bgj_apply_kernel<true>, dim3(ga), dim3(kThreads), 0, s, w.a, np, n, w.u,
                           w.dsc, x, static_cast<const double *>(nullptr),
                           static_cast<const double *>(nullptr), PZ, stp);
    } else {
        GSLAM_TRY(run_chunked(g, lo, enqueue, k));
    }
    GSLAM_HIP(hipGetLastError());
    GSLAM_HIP(hipEventRecord(g.ev1, s));
    GSLAM_HIP(hipMemcpyAsync(g.status_host, stp, sizeof(PcgStatus), hipMemcpyDeviceToHost, s));
    int32_t singular = 0;
    GSLAM_HIP(hipMemcpyAsync(g.scal_host + 3, w.flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GSLAM_HIP(hipStreamSynchronize(s));
    std::memcpy(&singular, g.scal_host + 3, sizeof(int32_t));
    float ms = 0.f, msf = 0.f;
    GSLAM_HIP(hipEventElapsedTime(&ms, g.ev0, g.ev1));
    GSLAM_HIP(hipEventElapsedTime(&msf, g.ev0, g.ev2));
    const PcgStatus &ps = *g.status_host;
    st.ms = ms;
    st.ms_factor = msf;
    st.iters = ps.iters;
    st.converged = ps.converged;
    st.flags = ps.flags;
    st.reliable_updates = 0;
    GSLAM_TRY(true_rel_residual(g, b, x, &st.true_rel_res));
    st.rel_res = st.true_rel_res;
    if (singular) {
        set_error("blocked Gauss-Jordan: zero pivot");
        GSLAM_HIP(hipMemsetAsync(x, 0, static_cast<size_t>(n) * sizeof(double), s));
        GSLAM_HIP(hipStreamSynchronize(s));
        return GSLAM_ERR_SINGULAR;
    }
    return GSLAM_OK;
}

// Host matrix in, host FP64 inverse out (FP32 accuracy, no refinement).
int bgj_invert_host(const double *m_host, double *inv_host, int64_t n)
{
    hipStream_t s = nullptr;
    GSLAM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DenseWork w;
    double *m = nullptr;
    int rc = w.reserve(n);
    if (rc == GSLAM_OK && hipMalloc(reinterpret_cast<void **>(&m), static_cast<size_t>(n * n) * 8) != hipSuccess) {
        set_error("gslam_gj_invert: out of device memory");
        rc = GSLAM_ERR_NOMEM;
    }
    int32_t singular = 0;
    if (rc == GSLAM_OK) {
        const int64_t np = w.np;
        if (hipMemcpyAsync(m, m_host, static_cast<size_t>(n * n) * 8, hipMemcpyHostToDevice, s) != hipSuccess) {
            rc = GSLAM_ERR_HIP;
        } else {
            hipLaunchKernelGGL(bgj_scale_dense_kernel, dim3(nblk(np)), dim3(kThreads), 0, s, m, n,
                               np, w.dsc);
            hipLaunchKernelGGL(bgj_load_dense_kernel, dim3(nblk(np * np)), dim3(kThreads), 0, s, m,
                               n, np, w.dsc, w.a);
            rc = bgj_invert(s, w);
        }
        if (rc == GSLAM_OK) {
            hipLaunchKernelGGL(bgj_store_dense_kernel, dim3(nblk(n * n)), dim3(kThreads), 0, s, w.a,
                               n, np, w.dsc, w.flag, m);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(inv_host, m, static_cast<size_t>(n * n) * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipMemcpyAsync(&singular, w.flag, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                set_error("gslam_gj_invert: HIP error");
                rc = GSLAM_ERR_HIP;
            }
        }
    }
    (void)hipFree(m);
    w.release();
    (void)hipStreamDestroy(s);
    if (rc == GSLAM_OK && singular) {
        set_error("Gauss-Jordan: zero pivot (the inverse is the zero matrix)");
        return GSLAM_ERR_SINGULAR;
    }
    return rc;
}


// ---------------------------------------------------------------- FP64 blocked, host

int DenseWork64::reserve(int64_t n_)
{
    const int64_t np_ = bgj_padded(n_);
    if (a && np_ <= np_cap) {
        n = n_;
        np = np_;
        return GSLAM_OK;
    }
    release();
    const size_t d = sizeof(double);
    void *p[6] = {};
    const size_t sz[6] = {static_cast<size_t>(np_ * np_) * d, static_cast<size_t>(kNb * np_) * d,
                          static_cast<size_t>(2 * kNb * np_) * d, static_cast<size_t>(kNb * kNb) * d,
                          static_cast<size_t>(np_) * d, 256};
    for (int i = 0; i < 6; ++i) {
        if (hipMalloc(&p[i], sz[i]) != hipSuccess) {
            for (int j = 0; j < i; ++j) (void)hipFree(p[j]);
            set_error("blocked FP64 Gauss-Jordan: out of device memory for a " +
                      std::to_string(np_) + " x " + std::to_string(np_) + " FP64 matrix");
            return GSLAM_ERR_NOMEM;
        }
    }
    a = static_cast<double *>(p[0]);
    ct = static_cast<double *>(p[1]);
    rp = static_cast<double *>(p[2]);
    pt = static_cast<double *>(p[3]);
    dsc = static_cast<double *>(p[4]);
    flag = static_cast<int32_t *>(p[5]);
    n = n_;
    np = np_cap = np_;
    bytes = 0;
    for (size_t b : sz) bytes += static_cast<int64_t>(b);
    return GSLAM_OK;
}

void DenseWork64::release()
{
    void *p[6] = {a, ct, rp, pt, dsc, flag};
    for (void *q : p)
        if (q) (void)hipFree(q);
    a = ct = rp = pt = dsc = nullptr;
    flag = nullptr;
    n = np = np_cap = bytes = 0;
}

// Trailing-update launch shape of GJ_BLOCKED64: workgroups per XCD, two
// tile workers each, at most one workgroup per CU.
static int32_t trailing64_wpx(int32_t nt)
{
    const int32_t per_xcd = ((nt - 1) * nt + 2 * kXcd - 1) / (2 * kXcd);
    return std::max(1, std::min(kTrWgPerXcd, per_xcd));
}

// R of block step k (two buffers, as rp_of).
static double *rp64_of(DenseWork64 &w, int32_t k)
{
    return w.rp + static_cast<int64_t>(k & 1) * kNb * w.np_cap;
}

// In-place FP64 blocked inverse of w.a (np x np, padded): per block step the
// FP64 pivot block, the FP64 MFMA row panel and the FP64 MFMA trailing update,
// whose workgroup 0 also makes the next step's pivot block.
// The matrix is Jacobi-scaled (unit diagonal) like GJ_MIXED's: the pivot
// block's folded rank-8 update (pivot_block) forms P8 from terms of the size
// of A_KK, which costs accuracy when |A_KK| is far from 1.
static int bgj64_invert(hipStream_t s, DenseWork64 &w)
{
    GSLAM_HIP(hipMemsetAsync(w.flag, 0, sizeof(int32_t), s));
    const int32_t nt = static_cast<int32_t>(w.np / kNb);
    const unsigned int gp = static_cast<unsigned int>(w.np / 32 + w.np / 64);
    const int32_t wpx = trailing64_wpx(nt);
    bool pivot_done = false;
    for (int32_t kp = 0; kp < nt; ++kp) {
        double *rpk = rp64_of(w, kp);
        if (!pivot_done)
            hipLaunchKernelGGL(bgj_pivot_kernel<double>, dim3(1), dim3(kPivThreads), 0, s, w.a, w.np,
                               kp, w.pt, rpk, w.flag);
        if (nt == 1) break;
        hipLaunchKernelGGL(bgj_panel_kernel<double>, dim3(gp), dim3(kThreads), 0, s, w.a, w.np, kp,
                           w.pt, rpk, w.ct);
        const int32_t fuse = kp + 1 < nt ? 1 : 0;
        // FP64 tile then pivot block: about 45 + 56 us, under 3 rounds of tiles
        const int32_t tail =
            fuse ? trailing_tail((nt - 1) * nt - 1, 2 * kXcd * wpx - 2, 2, 3) : 0;
        hipLaunchKernelGGL(bgj_trailing64_kernel, dim3(kXcd * wpx), dim3(kTrThreads), 0, s, w.a,
                           w.ct, rpk, w.np, nt, kp, wpx, fuse,
                           fuse ? trailing_pos_next_pivot(nt, kp) : 0, tail, rp64_of(w, kp + 1),
                           w.pt, w.flag);
        pivot_done = fuse != 0;
    }
    GSLAM_HIP(hipGetLastError());
    return GSLAM_OK;
}

int dense_solve_blocked64(DeviceGraph &g, const gslam_solve_opts &o, const double *b, double *x,
                          gslam_solve_stats &st)
{
    (void)o;
    const int64_t n = 3 * static_cast<int64_t>(g.h.n);
    DenseWork64 &w = g.dense64;
    GSLAM_TRY(w.reserve(n));
    const int64_t np = w.np;
    hipStream_t s = g.stream;
    GSLAM_HIP(hipEventRecord(g.ev0, s));
    hipLaunchKernelGGL(bgj_clear64_kernel, dim3(nblk(np * np)), dim3(kThreads), 0, s, w.a, n, np);
    hipLaunchKernelGGL(bgj_scale_bsr_kernel, dim3(nblk(np)), dim3(kThreads), 0, s, g.bdiag, g.bval,
                       n, np, w.dsc);
    hipLaunchKernelGGL(bgj_scatter_bsr64_kernel, dim3(nblk(9 * g.h.n_blocks)), dim3(kThreads), 0, s,
                       g.brow_ptr, g.bcol, g.bval, g.h.n_blocks, g.h.n, w.dsc, np, w.a);
    GSLAM_TRY(bgj64_invert(s, w));
    GSLAM_HIP(hipEventRecord(g.ev2, s));
    const unsigned int gm = static_cast<unsigned int>(std::min<int64_t>(kMaxGrid, (n + 3) / 4));
    hipLaunchKernelGGL(bgj_matvec64_kernel, dim3(gm), dim3(kThreads), 0, s, w.a, np, n, w.dsc, b, x);
    GSLAM_HIP(hipGetLastError());
    GSLAM_HIP(hipEventRecord(g.ev1, s));
    int32_t singular = 0;
    GSLAM_HIP(hipMemcpyAsync(g.scal_host + 3, w.flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GSLAM_HIP(hipStreamSynchronize(s));
    std::memcpy(&singular, g.scal_host + 3, sizeof(int32_t));
    float ms = 0.f, msf = 0.f;
    GSLAM_HIP(hipEventElapsedTime(&ms, g.ev0, g.ev1));
    GSLAM_HIP(hipEventElapsedTime(&msf, g.ev0, g.ev2));
    st.ms = ms;
    st.ms_factor = msf;
    st.iters = 0;
    st.converged = singular ? 0 : 1;
    st.flags = 0;
    st.reliable_updates = 0;
    if (singular) {
        set_error("Gauss-Jordan: zero pivot (the solution is the zero vector)");
        GSLAM_HIP(hipMemsetAsync(x, 0, static_cast<size_t>(n) * sizeof(double), s));
        GSLAM_HIP(hipStreamSynchronize(s));
        st.true_rel_res = st.rel_res = 1.0;
        return GSLAM_ERR_SINGULAR;
    }
    GSLAM_TRY(true_rel_residual(g, b, x, &st.true_rel_res));
    st.rel_res = st.true_rel_res;
    return GSLAM_OK;
}

int bgj64_invert_host(const double *m_host, double *inv_host, int64_t n)
{
    hipStream_t s = nullptr;
    GSLAM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DenseWork64 w;
    double *m = nullptr;
    int rc = w.reserve(n);
    if (rc == GSLAM_OK && hipMalloc(reinterpret_cast<void **>(&m), static_cast<size_t>(n * n) * 8) != hipSuccess) {
        set_error("gslam_gj_invert: out of device memory");
        rc = GSLAM_ERR_NOMEM;
    }
    int32_t singular = 0;
    if (rc == GSLAM_OK) {
        const int64_t np = w.np;
        if (hipMemcpyAsync(m, m_host, static_cast<size_t>(n * n) * 8, hipMemcpyHostToDevice, s) != hipSuccess) {
            rc = GSLAM_ERR_HIP;
        } else {
            hipLaunchKernelGGL(bgj_scale_dense_kernel, dim3(nblk(np)), dim3(kThreads), 0, s, m, n,
                               np, w.dsc);
            hipLaunchKernelGGL(bgj_load_dense64_kernel, dim3(nblk(np * np)), dim3(kThreads), 0, s, m,
                               n, np, w.dsc, w.a);
            rc = bgj64_invert(s, w);
        }
        if (rc == GSLAM_OK) {
            hipLaunchKernelGGL(bgj_store_dense64_kernel, dim3(nblk(n * n)), dim3(kThreads), 0, s, w.a,
                               n, np, w.dsc, w.flag, m);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(inv_host, m, static_cast<size_t>(n * n) * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipMemcpyAsync(&singular, w.flag, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                set_error("gslam_gj_invert: HIP error");
                rc = GSLAM_ERR_HIP;
            }
        }
    }
    (void)hipFree(m);
    w.release();
    (void)hipStreamDestroy(s);
    if (rc == GSLAM_OK && singular) {
        set_error("Gauss-Jordan: zero pivot (the inverse is the zero matrix)");
        return GSLAM_ERR_SINGULAR;
    }
    return rc;
}

}  // namespace gslam
