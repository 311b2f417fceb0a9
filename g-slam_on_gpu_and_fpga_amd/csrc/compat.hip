// The reference's GPU entry points (include/gslam_compat.h) on top of the
// libgslam handle: one module-level graph, CSR values exchanged through the
// caller's device buffers, errors reported and exit(1) like the reference's
// cuda_check (src/gpu_pcg_fp64/device_graph.cu:99-106).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <unordered_map>
#include <vector>

#include "gslam.h"
#include "gslam_compat.h"
#include "gslam_internal.h"

namespace gslam {
int gj_invert_fp64_device(hipStream_t s, double *m, int64_t n, double *work, int32_t *flag,
                          int32_t *singular_out);
}

static gslam_graph *s_graph = nullptr;
static int s_dim = 0;

static void compat_check(int rc, const char *where)
{
    if (rc != GSLAM_OK) {
        std::fprintf(stderr, "HIP error at %s: %s\n", where, gslam_last_error());
        std::exit(1);
    }
}

static void hip_check(hipError_t e, const char *where)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "HIP error at %s: %s\n", where, hipGetErrorString(e));
        std::exit(1);
    }
}

extern "C" {

void device_graph_init(struct Z *zlist, unsigned int n_obs, struct U *us, unsigned int n_poses,
                       unsigned int dim, double delta, double lambda, const double *mns,
                       int *nnz_out)
{
    if (s_graph) device_graph_free();
    std::vector<double> nu(n_poses), om(n_poses), poses(3 * static_cast<size_t>(n_poses), 0.0);
    for (unsigned int i = 0; i < n_poses; ++i) {
        nu[i] = us[i].nu;
        om[i] = us[i].omega;
    }
    std::vector<int32_t> zs(n_obs), zl(n_obs);
    std::vector<double> zr(n_obs), zb(n_obs);
    std::unordered_map<int32_t, int> lm_count;
    for (unsigned int i = 0; i < n_obs; ++i) {
        zs[i] = static_cast<int32_t>(zlist[i].step);
        zl[i] = static_cast<int32_t>(zlist[i].landmark_id);
        zr[i] = zlist[i].z[0];
        zb[i] = zlist[i].z[1];
        lm_count[zl[i]]++;
    }
    gslam_graph_desc d;
    gslam_default_desc(&d);
    d.n_poses = static_cast<int32_t>(n_poses);
    d.poses = poses.data();
    d.motion = 1;
    d.u_nu = nu.data();
    d.u_omega = om.data();
    d.delta = delta;
    d.lambda = lambda;
    for (int k = 0; k < 4; ++k) d.mns[k] = mns[k];
    d.n_obs = static_cast<int32_t>(n_obs);
    d.z_step = zs.data();
    d.z_lm = zl.data();
    d.z_range = zr.data();
    d.z_bearing = zb.data();
    d.pair_order = 1;  // ascending landmark id, as build_csr_structure
    compat_check(gslam_create(&d, &s_graph), "device_graph_init");
    gslam_graph_info info;
    gslam_get_info(s_graph, &info);
    int lm_active = 0;
    for (auto &kv : lm_count) lm_active += kv.second >= 2;
    s_dim = static_cast<int>(dim);
    std::printf("device_graph_init: dim=%d nnz=%d obs_edges=%d lm_active=%d\n", s_dim,
                static_cast<int>(info.nnz), static_cast<int>(info.n_obs_edges), lm_active);
    *nnz_out = static_cast<int>(info.nnz);
}

void device_graph_update(struct HAT_X *hat_xs, unsigned int n_poses, double *d_csr_vals,
                         double *d_xi)
{
    std::vector<double> flat(3 * static_cast<size_t>(n_poses));
    for (unsigned int i = 0; i < n_poses; ++i)
        for (int k = 0; k < 3; ++k) flat[3 * i + k] = hat_xs[i].hat_x[k];
    compat_check(gslam_set_poses(s_graph, flat.data()), "device_graph_update: upload hat_x");
    compat_check(gslam_assemble(s_graph), "device_graph_update: assembly");
    gslam::DeviceGraph &g = s_graph->g;
    compat_check(gslam::launch_bsr_to_csr(g, g.bval, d_csr_vals), "device_graph_update: csr");
    hip_check(hipMemcpyAsync(d_xi, g.xi, 3 * static_cast<size_t>(g.h.n) * 8,
                             hipMemcpyDeviceToDevice, g.stream), "device_graph_update: xi");
    hip_check(hipStreamSynchronize(g.stream), "kernel sync");
}

int *device_graph_get_row_ptr(void) { return s_graph ? s_graph->g.csr_row_ptr : nullptr; }
int *device_graph_get_col_idx(void) { return s_graph ? s_graph->g.csr_col_idx : nullptr; }

void device_build_diag_inv(int dim) { (void)dim; }

void device_pcg_solve(double *x_h, double *d_csr_vals, double *d_xi, int dim, int nnz,
                      int max_iter)
{
    (void)nnz;
    gslam::DeviceGraph &g = s_graph->g;
    compat_check(gslam::launch_csr_to_bsr(g, d_csr_vals, g.bval), "device_pcg_solve: csr");
    gslam_solve_opts o;
    gslam_default_opts(&o);
    o.max_iter = max_iter;
    gslam_solve_stats s;
    compat_check(gslam::pcg_solve(g, o, d_xi, g.dx, s), "device_pcg_solve");
    hip_check(hipMemcpyAsync(x_h, g.dx, static_cast<size_t>(dim) * 8, hipMemcpyDeviceToHost,
                             g.stream), "copy result to host");
    hip_check(hipStreamSynchronize(g.stream), "copy result to host");
    g.assembled = true;
    g.from_edges = false;
    g.factors_valid = false;
    g.solved = true;
}

void device_graph_free(void)
{
    if (s_graph) gslam_destroy(s_graph);
    s_graph = nullptr;
}

void inv_cuda(double *result_d, double *mat_h, unsigned int dim)
{
    const int64_t n = dim;
    hipStream_t s = nullptr;
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "inv_cuda: stream");
    double *work = nullptr;
    int32_t *flag = nullptr;
    hip_check(hipMalloc(reinterpret_cast<void **>(&work), static_cast<size_t>(3 * n + 8) * 8),
              "inv_cuda: malloc");
    hip_check(hipMalloc(reinterpret_cast<void **>(&flag), 64), "inv_cuda: malloc");
    hip_check(hipMemcpyAsync(result_d, mat_h, static_cast<size_t>(n * n) * 8,
                             hipMemcpyHostToDevice, s), "inv_cuda: copy");
    int32_t singular = 0;
    compat_check(gslam::gj_invert_fp64_device(s, result_d, n, work, flag, &singular), "inv_cuda");
    hip_check(hipFree(work), "inv_cuda: free");
    hip_check(hipFree(flag), "inv_cuda: free");
    hip_check(hipStreamDestroy(s), "inv_cuda: stream");
}

}  // extern "C"
