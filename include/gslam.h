/*
 * gslam.h -- C-ABI of libgslam, the MI355X (gfx950) Graph-SLAM back-end.
 *
 * One Gauss-Newton (GN) step = edge evaluation + assembly of the 3N x 3N
 * information matrix Omega and vector xi, a solve Omega dx = xi, and the
 * pose update.  Everything runs on one GPU through hand-written HIP kernels;
 * no vendor sparse/BLAS library is used.
 *
 * The handle API below replaces the module-static device graph of the
 * reference GPU PCG (src/gpu_pcg_fp64/device_graph.cuh:23-91) and its dense
 * GJ entry (src/gpu_gj_fp64/inv.h:9).  The reference's own entry points are
 * exported unchanged by gslam_compat.h.
 *
 * Conventions: plain pointers and sizes, host arrays are caller-owned and
 * copied in/out, the handle owns every device allocation, every call returns
 * GSLAM_OK (0) or a negative status and never exits the process.  A handle
 * is not thread-safe; distinct handles are independent.  Calls are
 * synchronous at return unless stated otherwise.
 */
#ifndef GSLAM_H
#define GSLAM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSLAM_OK             0
#define GSLAM_ERR_ARG       -1  /* invalid argument / index out of range     */
#define GSLAM_ERR_HIP       -2  /* HIP runtime error (message: last_error)   */
#define GSLAM_ERR_NOMEM     -3  /* host or device allocation failed          */
#define GSLAM_ERR_SINGULAR  -4  /* zero pivot in Gauss-Jordan                */
#define GSLAM_ERR_STATE     -5  /* call out of order (e.g. solve before assemble) */
#define GSLAM_ERR_TOO_LARGE -6  /* size beyond a documented limit             */
#define GSLAM_ERR_NOT_CONVERGED -7 /* solver hit its iteration cap (stats valid) */
#define GSLAM_ERR_NONFINITE -8 /* NaN or Inf in the system, the right-hand side
                                  or the solution (stats valid, flags bit 1) */

typedef struct gslam_graph gslam_graph;

/* Graph description.  Edges are evaluated in this order, which is also the
 * order of summation into Omega/xi:
 *   1. landmark-pair observation edges (ObsEdge, src/cpu_pcg_fp64/ObsEdge.c),
 *      enumerated from the z records per pair_order;
 *   2. n_poses-1 velocity-model motion edges (i, i+1) when motion != 0
 *      (MotionEdge, src/cpu_pcg_fp64/MotionEdge.c), using controls u[i+1];
 *   3. SE(2) between edges (pose-graph odometry / loop closures; no
 *      reference counterpart).
 * Items 1+2 with anchor 1e6 and wrap 0 are exactly the reference .log.txt
 * model (src/cpu_pcg_fp64/GSLAM.c:55-139). */
typedef struct {
    int32_t n_poses;
    const double *poses;        /* 3*n_poses: x, y, theta                    */

    int32_t motion;             /* 1: add the n_poses-1 motion edges         */
    const double *u_nu;         /* n_poses controls (u records in file order) */
    const double *u_omega;
    double delta;               /* time step ("delta" record)               */
    double lambda;              /* scales motion and between edges          */
    double mns[4];              /* motion noise {0.19, 0.001, 0.13, 0.2}     */

    int32_t n_obs;              /* z records                                 */
    const int32_t *z_step;      /* pose index of each observation            */
    const int32_t *z_lm;        /* landmark id                               */
    const double *z_range;
    const double *z_bearing;
    int32_t pair_order;         /* 0: first-appearance landmark order (CPU
                                   reference, GSLAM.c:246-276); 1: ascending
                                   landmark id (GPU reference,
                                   device_graph.cu:505-551)                  */
    double snr[2];              /* sensor noise ratios {0.14, 0.05}          */

    int64_t n_between;          /* SE(2) between edges                       */
    const int32_t *be_i;        /* from pose                                 */
    const int32_t *be_j;        /* to pose                                   */
    const double *be_meas;      /* 3 per edge: dx, dy (frame of pose i), dth */
    const double *be_info;      /* 3 per edge: diagonal information 1/sigma^2 */

    double anchor;              /* added to Omega[d][d], d<3 (1e6)           */
    int32_t wrap;               /* 0: reference semantics (no angle wrap);
                                   1: wrap theta residuals and updated poses
                                   to [-pi, pi)                               */
    int32_t anchor_prior;       /* 0: the anchor enters Omega only (reference,
                                   MatrixOps.c:223-231); 1: it is a prior on
                                   pose 0 at its value in `poses` at
                                   gslam_create, adding anchor * (ref - pose0)
                                   to xi, so the converged poses are a smooth
                                   function of the edge weights (gslam_adjoint) */
    int32_t assembly;           /* GSLAM_ASSEMBLY_AUTO, _POSE_WALK or _STAGED */
} gslam_graph_desc;

/* Assembly of Omega and xi (the same values, summed in the same edge order):
 *   POSE_WALK: one thread per pose evaluates each incident edge and adds its
 *              blocks and xi half in edge order (no staging pass);
 *   STAGED:    one thread per edge stages its 4 blocks and 2 xi halves, then
 *              one thread per Omega / xi entry sums its contribution list;
 *   AUTO:      POSE_WALK when no pose has more than 64 incident edge ends and
 *              no edge joins a pose to itself, else STAGED. */
#define GSLAM_ASSEMBLY_AUTO      0
#define GSLAM_ASSEMBLY_POSE_WALK 1
#define GSLAM_ASSEMBLY_STAGED    2

typedef struct {
    int32_t n_poses;
    int64_t n_edges;            /* obs pairs + motion + between              */
    int64_t n_obs_edges, n_motion_edges, n_between_edges;
    int64_t n_blocks;           /* 3x3 blocks of Omega (BSR)                 */
    int64_t nnz;                /* scalar nonzeros of the reference CSR      */
    int64_t device_bytes;       /* device memory held by the handle          */
    int32_t n_tiles;            /* SpMV tiles of whole block rows (<= 256
                                   blocks each); the FP64 PCG runs one launch
                                   per iteration up to 2048 of them          */
} gslam_graph_info;

#define GSLAM_SOLVER_PCG_FP64   0  /* block-Jacobi PCG, FP64 throughout      */
#define GSLAM_SOLVER_PCG_MIXED  1  /* FP32 inner PCG, FP64 true-residual
                                      reliable updates                      */
#define GSLAM_SOLVER_GJ_FP64    2  /* dense unpivoted Gauss-Jordan, FP64     */
#define GSLAM_SOLVER_GJ_MIXED   3  /* blocked GJ, FP32 MFMA trailing updates,
                                      FP64 iterative refinement             */
#define GSLAM_SOLVER_GJ_BLOCKED64 4 /* blocked GJ in FP64 (FP64 MFMA trailing
                                      updates; same pivots as GJ_FP64)      */

typedef struct {
    int32_t solver;             /* GSLAM_SOLVER_*                            */
    double tol;                 /* PCG: stop when ||r|| < tol * ||xi||        */
    int32_t max_iter;           /* PCG iteration cap                         */
    int32_t refine_sweeps;      /* GJ_MIXED: FP64 refinement sweeps          */
    int32_t refresh_every;      /* PCG_MIXED: FP64 true-residual refresh and
                                   convergence test every k iterations (64)  */
    int32_t check_every;        /* PCG: host polls convergence every k iters */
    int32_t refine_mode;        /* GJ_MIXED: GSLAM_REFINE_IR or _CG          */
    int32_t profile_every;      /* PCG_FP64: bracket the SpMV and update
                                   launches of every k-th iteration (at most
                                   128) with HIP events; 0 = off             */
    double mixed_handover;      /* PCG_MIXED: FP64 true relative residual at
                                   which the FP32 iterations stop and FP64
                                   PCG solves the remaining correction
                                   Omega d = xi - Omega dx to tol (FP64
                                   iterative refinement); 0 = FP32 iterations
                                   with FP64 refreshes down to tol           */
    int32_t pcg_schedule;       /* PCG_FP64 (and the mixed solver's FP64
                                   correction): GSLAM_PCG_AUTO or
                                   GSLAM_PCG_TWO_LAUNCH                      */
} gslam_solve_opts;

/* FP64 PCG launch schedules (same block-Jacobi PCG, same stopping rule):
 *   AUTO: one kernel per iteration on graphs of at most 2048 SpMV tiles
 *         (about 120,000 poses; gslam_graph_info.n_tiles); beta then comes
 *         from the one-step identity r.z' = r.z - 2 alpha q.z +
 *         alpha^2 q.D^-1 q, so the iterates differ from the two-launch ones
 *         in the last bits; two kernels per iteration on larger graphs.
 *   TWO_LAUNCH: two kernels per iteration at every size. */
#define GSLAM_PCG_AUTO       0
#define GSLAM_PCG_TWO_LAUNCH 1

/* GJ_MIXED refinement on the true residual xi - Omega dx (FP64), with the
 * FP32 blocked inverse M of the Jacobi-scaled Omega as the approximate
 * solver; both stop early once ||xi - Omega dx|| < tol ||xi||.
 *   IR: dx += D^-1/2 M D^-1/2 (xi - Omega dx), refine_sweeps times after the
 *       first FP32 solve;
 *   CG: up to refine_sweeps conjugate-gradient iterations preconditioned by
 *       the same operator (flexible beta; for ill-conditioned Omega, where the
 *       FP32 inverse alone has no correct digit, this still converges). */
#define GSLAM_REFINE_IR 0
#define GSLAM_REFINE_CG 1

typedef struct {
    int32_t iters;              /* PCG iterations / GJ_MIXED refinement
                                   iterations executed                      */
    int32_t converged;          /* 1 if the tolerance was met                */
    int32_t reliable_updates;   /* PCG_MIXED: FP64 residual refreshes        */
    int32_t flags;              /* bit0: p.q == 0 break (MatrixOps.c:357);
                                   bit1: NaN / Inf met (GSLAM_ERR_NONFINITE)  */
    double rel_res;             /* final ||r||/||xi|| of the solver's residual */
    double true_rel_res;        /* ||xi - Omega dx|| / ||xi||, FP64          */
    double ms;                  /* device time of the solve (HIP events)     */
    double ms_factor;           /* GJ: device time of the dense inversion     */
    double spmv_us, update_us;  /* PCG_FP64 with profile_every: mean sampled
                                   per-launch durations inside the solve (us) */
    int32_t iters_fp64;         /* PCG_MIXED with mixed_handover: FP64 PCG
                                   iterations of the correction (included in
                                   iters)                                    */
    /* Phase timers (device time, HIP events), milliseconds: */
    double ms_assembly;         /* the handle's last gslam_assemble (edges,
                                   Omega / xi gather)                         */
    double ms_setup;            /* PCG: tile copy, block-Jacobi inverses,
                                   r0 / z0 (part of ms); GJ: = ms_factor      */
    double ms_update;           /* gslam_gn_step: pose update and diff; else 0 */
} gslam_solve_stats;

/* Fills opts with the reference defaults (PCG FP64, tol 1e-6, 500 iterations:
 * src/cpu_pcg_fp64/MatrixOps.c:345-347). */
void gslam_default_opts(gslam_solve_opts *opts);
void gslam_default_desc(gslam_graph_desc *desc);

/* Lifecycle: replaces device_graph_init / device_graph_free
 * (src/gpu_pcg_fp64/device_graph.cuh:23-33, 91).  Builds the 64-bit-safe
 * sort-based block structure once (no dim x dim marker) and uploads it. */
int gslam_create(const gslam_graph_desc *desc, gslam_graph **out);
int gslam_destroy(gslam_graph *g);
int gslam_get_info(const gslam_graph *g, gslam_graph_info *info);

/* Poses (host <-> device, 3*n_poses doubles). */
int gslam_set_poses(gslam_graph *g, const double *poses);
int gslam_get_poses(gslam_graph *g, double *poses);
int gslam_set_lambda(gslam_graph *g, double lambda);
/* Per-between-edge information multipliers exp(w_e) given as log-weights
 * w_e (n_between doubles; NULL resets to 0). */
int gslam_set_log_weights(gslam_graph *g, const double *log_w);

/* Assembly of Omega and xi at the current poses: replaces
 * device_graph_update (device_graph.cuh:47-52) minus its host upload.
 * Deterministic: each entry is summed in edge order starting from the
 * anchor, the order of build_csr_from_edges (MatrixOps.c:223-272). */
int gslam_assemble(gslam_graph *g);

/* Download of the assembled system.  CSR in the reference layout
 * (row_ptr[3N+1], col_idx[nnz], values[nnz]), BSR block rows
 * (brow_ptr[N+1], bcol[n_blocks], bval[9*n_blocks]); any pointer may be
 * NULL. */
int gslam_get_system_csr(gslam_graph *g, int32_t *row_ptr, int32_t *col_idx,
                         double *values, double *xi);
int gslam_get_system_bsr(gslam_graph *g, int32_t *brow_ptr, int32_t *bcol,
                         double *bval, double *xi);

/* Replaces the assembled system by caller values in the reference CSR layout
 * (values[nnz] on the pattern gslam_get_system_csr reports, xi[3N]), e.g. a
 * system assembled by the reference CPU code (build_csr_from_edges,
 * src/cpu_pcg_fp64/MatrixOps.c:173-273).  The mixed solver needs the edge
 * factors gslam_assemble writes and returns GSLAM_ERR_STATE until the next
 * gslam_assemble. */
int gslam_set_system_csr(gslam_graph *g, const double *values, const double *xi);

/* Solve Omega dx = xi (replaces device_pcg_solve, device_graph.cuh:83-84,
 * and the inv_cuda + host matvec of src/gpu_gj_fp64/GSLAM.cu:391-433).
 * dx stays on the device; GSLAM_ERR_NOT_CONVERGED if the cap was hit (dx
 * and stats are still valid, as the reference returns silently). */
int gslam_solve(gslam_graph *g, const gslam_solve_opts *opts,
                gslam_solve_stats *stats);
int gslam_get_dx(gslam_graph *g, double *dx);
/* Solve with a caller right-hand side instead of xi (adjoint solves). */
int gslam_solve_rhs(gslam_graph *g, const gslam_solve_opts *opts,
                    const double *rhs, double *x, gslam_solve_stats *stats);

/* poses += dx (src/cpu_pcg_fp64/GSLAM.c:98-102), wrapped when desc.wrap;
 * diff = ||dx|| / sqrt(3N) (GSLAM.c:104-105). */
int gslam_apply_update(gslam_graph *g, double *diff_out);

/* Cost of the reference Gauss-Jordan drivers' use_diff = 0 stopping test
 * (src/gpu_gj_fp64/GSLAM.cu:245-257; src/cpu_gj_mixed/GSLAM.c:74-86) at the
 * current poses: for every landmark-pair edge, in edge order, the stored
 * residual e and 2x2 information inv(Sigma) (ObsEdge.c:76-82, 126-140) read
 * through the drivers' 3x3 view, e^T M e, added to *cost (motion edges add
 * zero).  *cost is the running sum the caller keeps: the reference never
 * resets it between rounds.  Host arithmetic in the reference's order (the
 * same bits as the reference's printf("cost: %.17g")). */
int gslam_obs_cost(gslam_graph *g, double *cost);

/* assemble + solve + apply_update. */
int gslam_gn_step(gslam_graph *g, const gslam_solve_opts *opts,
                  gslam_solve_stats *stats, double *diff_out);

/* Dense inverse of a host matrix by unpivoted Gauss-Jordan
 * (src/cpu_pcg_fp64/MatrixOps.c:29-91; GPU twin src/gpu_gj_fp64/inv.cu).
 * solver: GSLAM_SOLVER_GJ_FP64, GSLAM_SOLVER_GJ_BLOCKED64 (FP64, 128 pivots
 * per block step) or GSLAM_SOLVER_GJ_MIXED (FP32 MFMA blocked, no refinement
 * of the inverse itself).  GSLAM_ERR_SINGULAR on a zero pivot
 * (inv is then all zeros, as in the reference). */
int gslam_gj_invert(const double *a, double *inv, int32_t n, int32_t solver);

/* FP64 Gauss-Jordan inverse of a device-resident n x n matrix into the
 * caller's device buffer inv (a == inv inverts in place), allocating
 * nothing: the same kernels as gslam_gj_invert(..., GSLAM_SOLVER_GJ_FP64)
 * with their scratch (3n + 8 doubles: pivot columns, diagonal, zero-pivot
 * flag) in a static device array of 3 x 16384 + 8 doubles per GPU.  No
 * counterpart in the reference (its inv_cuda, src/gpu_gj_fp64/inv.h:9,
 * allocates per call); for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  n <= 16384, else GSLAM_ERR_TOO_LARGE.  Runs on the null
 * stream of the current device and returns when the inverse is complete;
 * GSLAM_ERR_SINGULAR on a zero pivot (inv is then all zeros).  Calls that
 * overlap on one GPU share its scratch and race. */
int gslam_gj_invert_device(const double *a, double *inv, int32_t n);

/* Adjoint of the converged poses x* of a between-edge graph (no reference
 * counterpart; the reference tunes lambda by a manual grid search,
 * README.md:137-163).  With f = 1/2 sum_e e_e^T L_e e_e (+ anchor prior),
 * L_e = exp(w_e) lambda diag(info_e), and the caller's dL/dx* (3N doubles),
 * it assembles the exact Hessian H of f at the current poses (Gauss-Newton
 * blocks plus the second-order terms of the rotated translation residual,
 * same block pattern as Omega), solves H v = dL/dx* with the FP64 PCG
 * (opts->tol, opts->max_iter; the solver field is ignored), then an
 * edge-parallel kernel forms
 *   dL/dw_e    = -(J_e v)^T L_e e_e            (n_between doubles, dl_dw)
 *   dL/dlambda = sum_e dL/dw_e / lambda          (dl_dlambda)
 * Requires desc.anchor_prior = 1 and converged poses.
 *
 * Graphs of the reference's landmark-pair and motion edges (no between
 * edges; dl_dw is not written): xi is not the gradient of a cost there (the
 * edge information depends on the poses), so dL/dlambda comes from the
 * tangent u = dx* / dlambda: K u = d xi / d lambda with K = -d xi / dx at
 * x*, applied by central differences of the assembled xi, solved by
 * right-preconditioned GMRES(40) with the FP64 PCG on Omega(x*) as the
 * preconditioner (opts->tol: GMRES relative residual and PCG tolerance;
 * at most min(opts->max_iter, 400) GMRES iterations, one restart); the
 * result is accepted once the residual through a fresh central-difference
 * product (stats->true_rel_res, ~1e-5 from their rounding) is <= 1e-4;
 * stats->iters counts GMRES iterations, stats->iters_fp64 PCG iterations.  lambda scales the
 * motion edges only (MotionEdge.c:113-175), so d xi / d lambda =
 * (xi(lambda) - xi(0)) / lambda.
 * Leaves the handle unassembled (the next solve assembles again). */
int gslam_adjoint(gslam_graph *g, const double *dl_dx, const gslam_solve_opts *opts,
                  double *dl_dw, double *dl_dlambda, gslam_solve_stats *stats);

/* Kernel timing on the current system: runs one kernel `reps` times back to
 * back on the handle's stream and returns the mean duration per launch in
 * microseconds, measured with HIP events on that stream. */
#define GSLAM_KERNEL_SPMV_FP64   0  /* FP64 PCG: BSR SpMV + p.q partials       */
#define GSLAM_KERNEL_UPDATE_FP64 1  /* FP64 PCG: x, r, z update + partials      */
#define GSLAM_KERNEL_SPMV_MIXED  2  /* mixed PCG: factorized FP32 operator      */
#define GSLAM_KERNEL_ASSEMBLY    3  /* edge evaluation + Omega/xi gathers       */
#define GSLAM_KERNEL_GJ_TRAILING 4  /* GJ_MIXED: FP32 MFMA trailing update of
                                       the first block step (needs a GJ_MIXED
                                       solve first; overwrites its workspace)  */
int gslam_time_kernel(gslam_graph *g, int32_t which, int32_t reps, double *us_per_launch);

/* Last error message of the calling thread ("" if none). */
const char *gslam_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
