// Internal state of a libgslam graph handle and the kernel launchers each
// translation unit provides.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "gslam.h"
#include "graph_build.h"

namespace gslam {

void set_error(const std::string &msg);

#define GSLAM_HIP(expr)                                                        \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            ::gslam::set_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
            return GSLAM_ERR_HIP;                                              \
        }                                                                      \
    } while (0)

#define GSLAM_TRY(expr)                                                        \
    do {                                                                       \
        int r_ = (expr);                                                       \
        if (r_ != GSLAM_OK) return r_;                                         \
    } while (0)

// Solver status written by workgroup 0 of the PCG kernels and polled by the
// host (device memory, mirrored into pinned host memory).
struct PcgStatus {
    int32_t done;        // 1: stop (converged, cap, or p.q == 0)
    int32_t converged;
    int32_t iters;       // iterations completed
    int32_t flags;       // bit0: p.q == 0 (MatrixOps.c:357)
    int32_t refreshes;   // mixed: FP64 true-residual refreshes
    int32_t x_iter;      // FP64 PCG: k when x_k = x_(k-1) + alpha_(k-1) p_(k-1) is
                         // still to be applied (by SpMV(k), or after the loop)
    double rnorm;        // ||r|| of the last test
    double bnorm;        // ||b||
    double tol;          // tol_rel * ||b||
};

// Work space of the blocked Gauss-Jordan (GJ_MIXED): the padded FP32 matrix
// (np x np, np = 3N rounded up to the 128-wide panel), the k-major column
// and row panels, the FP64 pivot block, the Jacobi scaling and refinement
// vectors.  Kept between solves; grown on demand.
struct DenseWork {
    float *a = nullptr, *ct = nullptr, *rp = nullptr;
    double *pt = nullptr, *dsc = nullptr, *u = nullptr, *vec = nullptr;
    int32_t *flag = nullptr;
    int64_t n = 0, np = 0, np_cap = 0, bytes = 0;
    int reserve(int64_t n_);
    void release();
    ~DenseWork() { release(); }
};

// GJ_BLOCKED64 work space: the padded, Jacobi-scaled FP64 matrix (inverted
// in place), the FP64 column panel, two row panels (block steps k and k + 1),
// P^T of the pivot block, the scale factors and a zero-pivot flag.
struct DenseWork64 {
    double *a = nullptr, *ct = nullptr, *rp = nullptr, *pt = nullptr, *dsc = nullptr;
    int32_t *flag = nullptr;
    int64_t n = 0, np = 0, np_cap = 0, bytes = 0;
    int reserve(int64_t n_);
    void release();
    ~DenseWork64() { release(); }
};

struct DeviceGraph {
    int dev = 0;
    hipStream_t stream = nullptr;
    HostGraph h;
    double lambda = 0.1, anchor = 1e6, delta = 1.0, mns[4] = {0.19, 0.001, 0.13, 0.2};
    double snr[2] = {0.14, 0.05};
    int wrap = 0;
    int anchor_prior = 0;
    double pose0_ref[3] = {0.0, 0.0, 0.0};  // anchor prior reference (pose 0 at create)
    int64_t bytes = 0;
    bool assembled = false, solved = false;
    bool factors_valid = false;  // frec matches the assembled system
    bool from_edges = false;     // the assembled system came from gslam_assemble

    // inputs / poses
    double *poses = nullptr;     // 3N
    double *dx = nullptr;        // 3N
    int32_t *e_t1 = nullptr, *e_t2 = nullptr;
    double *obs_param = nullptr, *mot_param = nullptr, *bet_param = nullptr;
    double *bet_w = nullptr;     // per between edge exp(w_e)

    // structure
    int32_t *brow_ptr = nullptr, *bcol = nullptr, *bdiag = nullptr;
    int32_t *cptr = nullptr, *clist = nullptr, *xptr = nullptr, *xlist = nullptr;
    // assembly schedule (assembly.hip): the pose walk reads xblk (per
    // incidence entry its off-diagonal block); the staged one reads cptr,
    // clist and the stage records
    bool walk = false;
    int32_t *xblk = nullptr;
    int32_t *csr_row_ptr = nullptr, *csr_col_idx = nullptr, *csr_pos = nullptr;
    int4 *tile_info = nullptr;    // per tile {first row, end row, first block, end block}
    int32_t n_tiles = 0;
    // The blocks of tile t padded to kTileBlocks slots from slot kTileBlocks t
    // (column index; a padding slot repeats the tile's first column), so an
    // SpMV lane finds its block without the tile record: tcol fixed, tval
    // packed from bval at the start of each FP64 solve (zero in padding),
    // entry k of slot l of tile t at (9 t + k) kTileBlocks + l.  tval is
    // allocated by the first FP64 PCG solve (tval_reserve).
    int32_t *tcol = nullptr;
    double *tval = nullptr;

    // assembled system
    double *stage = nullptr;     // kStage per edge
    double *bval = nullptr;      // 9 per block, FP64
    double *xi = nullptr;        // 3N

    // solver work space
    double *x = nullptr, *r = nullptr, *z = nullptr, *q = nullptr;
    double *p[2] = {nullptr, nullptr};
    double *dinv = nullptr;      // FP64 D^-1, column-major per entry: [c][3i+a]
    float *dinv32 = nullptr;
    float *r32 = nullptr, *z32 = nullptr, *q32 = nullptr, *p32[2] = {nullptr, nullptr};
    double *part = nullptr;      // reduction partials (pcg_common.h layout)
    PcgStatus *status = nullptr;
    PcgStatus *status_host = nullptr;  // pinned mirror
    PcgStatus *status_ring = nullptr;  // pinned, 2 slots for look-ahead polling
    double *scal = nullptr;      // small device scalars (diff etc.)
    double *scal_host = nullptr; // pinned
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev_poll[2] = {nullptr, nullptr};
    hipEvent_t ev_phase[2] = {nullptr, nullptr};  // around gslam_assemble / the pose update
    double ms_assembly = 0.0;    // device time of the last gslam_assemble

    // factorized FP32 operator of the mixed solver, in the tiles of
    // build_mixed_tiles (graph_build.h): per tile {first pose, end pose,
    // xptr of both} (mt_info) and its first extra chunk (mt_ext); per slot
    // the incidence entry e*2+side (-1 in padding), its own and other pose;
    // mt_pos[2e+side] = slot; the 8 FP32 factors of the slot's edge, stored
    // plane-major per 256-slot chunk: factor c of slot 256 h + l at
    // frec[(8 h + c) 256 + l].  These and the FP32 vectors above are
    // allocated by the first mixed solve (mixed_factors); until then the
    // assembly kernels skip frec.
    int4 *mt_info = nullptr;
    int32_t *mt_ext = nullptr, *mt_ent = nullptr, *mt_own = nullptr, *mt_oth = nullptr;
    int32_t *mt_pos = nullptr;
    int32_t mt_tiles = 0;
    int64_t mt_slots = 0;        // slots of all tiles and extra chunks
    float *frec = nullptr;

    // Vector work space, allocated on first use and kept: 4 x 3N + n_between
    // doubles (caller right-hand side and solution of gslam_solve_rhs and
    // gslam_adjoint, the mixed solver's FP64 correction r and d, adjoint
    // edge gradients), and an nnz-double staging buffer for the CSR copies.
    double *vecs = nullptr;
    double *csr_tmp = nullptr;
    int vecs_reserve();
    int csr_tmp_reserve();
    int tval_reserve();
    // One-launch-per-iteration FP64 PCG of small graphs (pcg.hip): second
    // buffers of r and q (3N each), per pose and parity the record {z, w =
    // D^-1 q, p} (9 doubles) the SpMV gathers, and the partials of two
    // launches; allocated by its first solve.
    double *r_alt = nullptr, *q_alt = nullptr, *zwp = nullptr, *fpart = nullptr;
    int fused_reserve();
    // Makes frec valid for the assembled system: allocates the mixed solver's
    // buffers on first use and, if the last assembly ran without them,
    // re-runs it (the assembly is deterministic: bval and xi keep their bits).
    int mixed_factors();
    int64_t dim() const { return 3 * static_cast<int64_t>(h.n); }

    DenseWork dense;             // GJ_MIXED work space (allocated on first use)
    DenseWork64 dense64;         // GJ_BLOCKED64 work space (allocated on first use)
    std::vector<hipEvent_t> prof_ev;  // in-solve kernel timing (created on first use)

    std::vector<void *> allocs;

    template <typename T>
    int alloc(T **p, int64_t count);
    int upload(void *dst, const void *src, int64_t bytes);
    int download(void *dst, const void *src, int64_t bytes);
    ~DeviceGraph();
};

template <typename T>
int DeviceGraph::alloc(T **ptr, int64_t count)
{
    size_t bytes_ = static_cast<size_t>(count > 0 ? count : 1) * sizeof(T);
    void *mem = nullptr;
    hipError_t e = hipMalloc(&mem, bytes_);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc(") + std::to_string(bytes_) + "): " +
                  hipGetErrorString(e));
        return GSLAM_ERR_NOMEM;
    }
    allocs.push_back(mem);
    bytes += static_cast<int64_t>(bytes_);
    *ptr = static_cast<T *>(mem);
    return GSLAM_OK;
}

// ---- assembly.hip ----
// Evaluates every edge at the current poses and gathers Omega (BSR, FP64 and
// FP32 copy) and xi.  exact_hessian: between edges add their second-order
// terms (the exact Hessian of the cost instead of the Gauss-Newton matrix).
int launch_assembly(DeviceGraph &g, bool exact_hessian = false);
// BSR values -> reference CSR values (device to device).
int launch_bsr_to_csr(DeviceGraph &g, const double *bval, double *csr_vals);
int launch_csr_to_bsr(DeviceGraph &g, const double *csr_vals, double *bval);

// ---- adjoint.hip ----
// dl_dw[k] = -(J_k v)^T L_k e_k for every between edge k (device arrays).
int launch_adjoint_grad(DeviceGraph &g, const double *v, double *dl_dw);

// ---- gj.hip ----
int dense_solve(DeviceGraph &g, const gslam_solve_opts &o, const double *b, double *x,
                gslam_solve_stats &s);

// ---- gj_blocked.hip ----
int64_t bgj_padded(int64_t n);
int dense_solve_mixed(DeviceGraph &g, const gslam_solve_opts &o, const double *b, double *x,
                      gslam_solve_stats &s);
// FP32 blocked inverse of a host n x n FP64 matrix, returned as FP64.
int bgj_invert_host(const double *m_host, double *inv_host, int64_t n);
int dense_solve_blocked64(DeviceGraph &g, const gslam_solve_opts &o, const double *b, double *x,
                          gslam_solve_stats &st);
int bgj64_invert_host(const double *m_host, double *inv_host, int64_t n);
// Mean duration of the trailing-update kernel of block step 0 on w.
int bgj_time_trailing(hipStream_t s, DenseWork &w, int reps, hipEvent_t e0, hipEvent_t e1,
                      double *us);

// ---- pcg.hip ----
int pcg_solve(DeviceGraph &g, const gslam_solve_opts &o, const double *b,
              double *x_out, gslam_solve_stats &s);
int pose_update(DeviceGraph &g, double *diff_out);
// y = Omega x (FP64, BSR) on the handle stream.
int launch_spmv64(DeviceGraph &g, const double *x, double *y);
// Mean per-launch duration (us) of one solver kernel (GSLAM_KERNEL_*).
int time_kernel(DeviceGraph &g, int which, int reps, double *us);
// ||b - Omega x|| / ||b|| in FP64 (device reductions, one host sync).
int true_rel_residual(DeviceGraph &g, const double *b, const double *x, double *out);

}  // namespace gslam

struct gslam_graph {
    gslam::DeviceGraph g;
};
