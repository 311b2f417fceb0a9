/*
 * gslam_compat.h -- the reference's own C entry points for the GPU solve,
 * exported unchanged by libgslam so that the reference host drivers
 * (src/gpu_pcg_fp64/GSLAM.cu, src/gpu_gj_fp64/GSLAM.cu) link against it as a
 * drop-in.  Semantics, argument meaning and error behaviour follow the
 * reference: module-level state (one graph per process), device buffers for
 * CSR values and xi allocated by the caller, HIP errors print a message to
 * stderr and exit(1) as the reference's cuda_check does
 * (src/gpu_pcg_fp64/device_graph.cu:99-106).
 *
 * New code should use the handle API of gslam.h instead.
 */
#ifndef GSLAM_COMPAT_H
#define GSLAM_COMPAT_H

#ifdef __cplusplus
extern "C" {
#endif

/* Record types of the .log.txt reader, field for field as
 * src/gpu_pcg_fp64/U.h, HAT_X.h, Z.h (same layout as src/cpu_pcg_fp64). */
struct U {
    unsigned int step;
    double nu;
    double omega;
};

struct HAT_X {
    unsigned int step;
    double hat_x[3];
};

struct Z {
    unsigned int step;
    unsigned int landmark_id;
    double z[2];
};

/* src/gpu_pcg_fp64/device_graph.cuh:23-33.  Builds the CSR structure of all
 * ObsEdges (ascending landmark id) and MotionEdges, prints the reference's
 * "device_graph_init: dim=... nnz=... obs_edges=... lm_active=..." line and
 * returns nnz.  mns = {0.19, 0.001, 0.13, 0.2}. */
void device_graph_init(struct Z *zlist, unsigned int n_obs, struct U *us,
                       unsigned int n_poses, unsigned int dim, double delta,
                       double lambda, const double *mns, int *nnz_out);

/* device_graph.cuh:47-52.  Uploads the poses, evaluates every edge and writes
 * the CSR values (nnz, reference layout) and xi (dim) into the caller's
 * device buffers. */
void device_graph_update(struct HAT_X *hat_xs, unsigned int n_poses,
                         double *d_csr_vals, double *d_xi);

/* device_graph.cuh:59-60: device pointers of the CSR structure. */
int *device_graph_get_row_ptr(void);
int *device_graph_get_col_idx(void);

/* device_graph.cuh:69: a no-op in the reference (device_graph.cu:834-837);
 * the preconditioner is built inside device_pcg_solve. */
void device_build_diag_inv(int dim);

/* device_graph.cuh:83-84.  Block-Jacobi PCG on the given CSR values and xi
 * (tol 1e-6 * ||xi||, max_iter iterations), solution copied to host x_h. */
void device_pcg_solve(double *x_h, double *d_csr_vals, double *d_xi, int dim,
                      int nnz, int max_iter);

/* device_graph.cuh:91. */
void device_graph_free(void);

/* src/gpu_gj_fp64/inv.h:9.  Inverts the host dim x dim matrix mat_h by
 * unpivoted Gauss-Jordan into the caller-allocated device buffer result_d
 * (dim*dim doubles); synchronous; a zero pivot gives the zero matrix. */
void inv_cuda(double *result_d, double *mat_h, unsigned int dim);

#ifdef __cplusplus
}
#endif
#endif
