/*
 * gslam_oracle.h -- CPU FP64 restatement of the reference Graph-SLAM
 * Gauss-Newton step.  TEST INFRASTRUCTURE ONLY: this library is the checker
 * that the HIP product is compared against.  It is loaded by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py, and by
 * nothing else.  The product library never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Reference paths
 * are relative to the reference repository root.
 *
 * Flat arrays, plain pointers, int32/int64 sizes: the same conventions as the
 * product C-ABI, so Python tests can drive both through ctypes.
 */
#ifndef GSLAM_ORACLE_H
#define GSLAM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-edge output layout shared by every edge model (42 doubles):
 *   [ 0.. 8] UL  block (t1,t1), row-major
 *   [ 9..17] UR  block (t1,t2)
 *   [18..26] BL  block (t2,t1)
 *   [27..35] BR  block (t2,t2)
 *   [36..38] xi  at pose t1
 *   [39..41] xi  at pose t2
 */
#define ORC_EDGE_VALS 42

/* ---------------- .log.txt reader: src/cpu_pcg_fp64/GSLAM.c:415-493 ------ */
/* Counts records.  Returns 0, or -1 if the file cannot be opened, -2 if a
 * record is truncated (the reference would dereference a NULL token). */
int orc_log_count(const char *path, int32_t *n_u, int32_t *n_x, int32_t *n_z);
/* Reads the records into caller arrays sized by orc_log_count. */
int orc_log_read(const char *path, double *delta,
                 int32_t *u_step, double *u_nu, double *u_omega,
                 int32_t *x_step, double *x_pose /* 3 per record */,
                 int32_t *z_step, int32_t *z_lm, double *z_range,
                 double *z_bearing);

/* ---------------- landmark-pair enumeration ------------------------------ */
/* order 0: first-appearance landmark order, src/cpu_pcg_fp64/GSLAM.c:222-318
 * order 1: ascending landmark id, src/gpu_pcg_fp64/device_graph.cu:489-551
 * Writes observation indices (into the z arrays) of each pair j<k.  Pass
 * NULL outputs to count.  Returns the number of pairs. */
int64_t orc_obs_pairs(int32_t n_obs, const int32_t *z_lm, int32_t order,
                      int32_t *pair_a, int32_t *pair_b);

/* ---------------- edge models -------------------------------------------- */
/* MotionEdge_create, src/cpu_pcg_fp64/MotionEdge.c:9-193, with
 * state_transition, src/cpu_pcg_fp64/GSLAM.c:373-401.  omega_out (9) may be
 * NULL; it receives the edge information matrix (MotionEdge.c:100-102). */
void orc_motion_edge(const double x1[3], const double x2[3], double nu,
                     double omega, double delta, double lambda,
                     const double mns[4], double out[ORC_EDGE_VALS],
                     double *omega_out);

/* ObsEdge_create, src/cpu_pcg_fp64/ObsEdge.c:10-212.  hat_e (2) and om (4)
 * may be NULL. */
void orc_obs_edge(const double x1[3], const double x2[3], double r1,
                  double phi1, double r2, double phi2, const double snr[2],
                  double out[ORC_EDGE_VALS], double *hat_e, double *om);

/* SE(2) between edge (no reference counterpart): residual
 *   e = [R(th_i)^T (t_j - t_i) - z_t ; th_j - th_i - z_th]   (wrapped if wrap)
 * information  diag(info) * weight, blocks J^T L J, xi = -J^T L e.
 * e_out (3) may be NULL. */
void orc_between_edge(const double xi[3], const double xj[3],
                      const double z[3], const double info[3], double weight,
                      int32_t wrap, double out[ORC_EDGE_VALS], double *e_out);

/* ---------------- sparse assembly: src/cpu_pcg_fp64/MatrixOps.c:173-273 --- */
/* Builds the sorted-column CSR pattern of the reference (anchor diagonal of
 * pose 0 + the four 3x3 blocks of every edge), 64-bit safe, without the
 * dim x dim marker.  Call once with col_idx == NULL to get nnz. */
int64_t orc_csr_pattern(int32_t n_poses, int64_t n_edges, const int32_t *t1,
                        const int32_t *t2, int32_t *row_ptr, int32_t *col_idx);
/* Accumulates anchor then edge blocks in edge order into values (nnz) and xi
 * (3N), zeroing them first -- the summation order of MatrixOps.c:223-272. */
int orc_csr_values(int32_t n_poses, const int32_t *row_ptr,
                   const int32_t *col_idx, int64_t n_edges, const int32_t *t1,
                   const int32_t *t2, const double *edge_vals, double anchor,
                   double *values, double *xi);

/* Dense assembly (src/cpu_pcg_fp64/GSLAM.c:320-371 with the anchor of
 * src/gpu_gj_fp64/GSLAM.cu:315-328).  omega: dim*dim row-major. */
int orc_dense_values(int32_t n_poses, int64_t n_edges, const int32_t *t1,
                     const int32_t *t2, const double *edge_vals,
                     double anchor, double *omega, double *xi);

/* ---------------- block-Jacobi PCG: src/cpu_pcg_fp64/MatrixOps.c:281-391 -- */
/* tol_rel and max_iter replace the hard-coded 1e-6 and 500 (MatrixOps.c:345,
 * 347).  res_hist (max_iter, may be NULL) receives ||r|| after each update.
 * Returns the number of iterations executed (the loop index reached). */
/* Threads the PCG loops run on: 1 in liboracle.so, the OpenMP team size in
 * liboracle_omp.so (the same source built with -fopenmp). */
int32_t orc_threads(void);
int32_t orc_pcg(int32_t dim, const int32_t *row_ptr, const int32_t *col_idx,
                const double *values, const double *b, double tol_rel,
                int32_t max_iter, double *x, double *res_hist);
/* 3x3 block inverses of MatrixOps.c:287-315 (det == 0 -> 1e-30). */
void orc_block_jacobi(int32_t dim, const int32_t *row_ptr,
                      const int32_t *col_idx, const double *values,
                      double *diag_inv);
/* y = A x over CSR (MatrixOps.c:349-353). */
void orc_spmv(int32_t dim, const int32_t *row_ptr, const int32_t *col_idx,
              const double *values, const double *x, double *y);

/* ---------------- dense Gauss-Jordan: src/cpu_pcg_fp64/MatrixOps.c:29-91 -- */
/* Unpivoted augmented GJ.  A zero pivot yields the zero matrix and returns 1
 * (MatrixOps.c:60-68). */
int orc_gj_inv(int32_t n, const double *a, double *a_inv);
/* FP32 twin, src/cpu_gj_mixed/MatrixOps.c:174-212. */
int orc_gj_inv_f32(int32_t n, const float *a, float *a_inv);
/* y = A x, dense, src/cpu_pcg_fp64/MatrixOps.c:123-147 order. */
void orc_dense_matvec(int32_t n, const double *a, const double *x, double *y);

/* ---------------- the Gauss-Newton step on the .log.txt model ------------- */
/* Edges of one round in reference order: ObsEdges (pairs, orc_obs_pairs
 * order) then MotionEdges (GSLAM.c:62-72).  Pass NULL outputs to count. */
int64_t orc_log_edges(int32_t n_poses, const double *poses,
                      const double *u_nu, const double *u_omega, double delta,
                      double lambda, const double mns[4], int32_t n_obs,
                      const int32_t *z_step, const int32_t *z_lm,
                      const double *z_range, const double *z_bearing,
                      int32_t order, int32_t *t1, int32_t *t2,
                      double *edge_vals);

/* ||v||_2 / sqrt(dim): src/cpu_pcg_fp64/MatrixOps.c:4-16 + GSLAM.c:104-105 */
double orc_diff(int32_t dim, const double *dx);

/* Landmark estimates after the GN loop (GSLAM.c:141-196, MapEdge.c:7-69):
 * first-appearance landmark order, mean of x + r (cos, sin)(th + phi) over a
 * landmark's observations.  Returns the number of landmarks. */
int32_t orc_landmark_means(int32_t n_obs, const double *poses, const int32_t *z_step,
                           const int32_t *z_lm, const double *z_range,
                           const double *z_bearing, int32_t *ids_out, double *means_out);

/* Pose update hat_x += dx (GSLAM.c:98-102); wrap != 0 wraps theta to
 * (-pi, pi] afterwards (the build's optional angle-wrap mode). */
void orc_pose_update(int32_t n_poses, double *poses, const double *dx,
                     int32_t wrap);

/* Full reference GN loop on a .log.txt graph, use_diff = 1 path of
 * src/cpu_pcg_fp64/GSLAM.c:55-139 with the PCG of MatrixOps.c:281-391.
 * diffs (num_rounds) receives the printed diff per executed round.
 * Returns the number of executed rounds. */
int32_t orc_log_gn_run(int32_t n_poses, double *poses, const double *u_nu,
                       const double *u_omega, double delta, double lambda,
                       int32_t n_obs, const int32_t *z_step,
                       const int32_t *z_lm, const double *z_range,
                       const double *z_bearing, int32_t num_rounds,
                       double tol_rel, int32_t max_iter, double *diffs);

/* ---------------- between-edge pose graph (no reference counterpart) ------ */
/* Edges of one GN step for a pose graph of between edges, in list order. */
void orc_between_edges(int32_t n_poses, const double *poses, int64_t n_edges,
                       const int32_t *ei, const int32_t *ej,
                       const double *meas, const double *info,
                       const double *weight, double lambda, int32_t wrap,
                       double *edge_vals);

/* One GN step on a between-edge graph: edges, CSR assembly (anchor on pose
 * 0), PCG from x = 0 to tol_rel / max_iter, pose update (with wrap).  CSR
 * pattern arrays must come from orc_csr_pattern.  Scratch: 8 * nnz + 40 * N
 * doubles allocated inside.  Returns PCG iterations, or < 0 on error. */
int32_t orc_between_gn_step(int32_t n_poses, double *poses, int64_t n_edges,
                            const int32_t *ei, const int32_t *ej,
                            const double *meas, const double *info,
                            const double *weight, double lambda,
                            double anchor, int32_t wrap,
                            const int32_t *row_ptr, const int32_t *col_idx,
                            double tol_rel, int32_t max_iter, double *dx_out,
                            double *diff_out);

#ifdef __cplusplus
}
#endif
#endif
