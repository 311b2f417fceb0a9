// Adjoint of the converged poses of a between-edge graph.
//
// No reference counterpart: the reference picks lambda by a manual grid
// search over whole runs (README.md:137-163).  The converged poses x* are
// the stationary point of f(x) = 1/2 sum_e e_e^T L_e e_e + 1/2 a |x_0 - ref|^2
// with L_e = exp(w_e) lambda diag(info_e), so F(x*, theta) = grad f = 0 and
//   dL/dtheta = -v^T dF/dtheta,   H v = dL/dx*,   H = dF/dx (exact Hessian).
// dF/dw_e = J_e^T L_e e_e (L_e is proportional to exp(w_e)), hence
//   dL/dw_e = -(J_e v)^T L_e e_e,
// and every between edge scales with lambda (the anchor does not), hence
//   dL/dlambda = sum_e dL/dw_e / lambda.
// H is assembled by the same edge and gather kernels as Omega with the
// second-order terms switched on, and solved by the same FP64 PCG.
#include "device_common.h"
#include "pcg_common.h"

#include <cmath>
#include <vector>

namespace gslam {

// One thread per between edge: residual and Jacobians as in the assembly,
// u = J_i v_i + J_j v_j, dl_dw = -sum_k u_k L_k e_k.
__global__ void __launch_bounds__(kThreads)
adjoint_grad_kernel(const double *__restrict__ poses, const int32_t *__restrict__ t1,
                    const int32_t *__restrict__ t2, const double *__restrict__ bp,
                    const double *__restrict__ bw, int64_t n_bet, int64_t e_off, double lambda,
                    int wrap, const double *__restrict__ v, double *__restrict__ dl_dw)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n_bet) return;
    const int64_t e = e_off + k;
    const int64_t a = t1[e], b = t2[e];
    double s, c;
    sincos(poses[3 * a + 2], &s, &c);
    const double dx = poses[3 * b] - poses[3 * a], dy = poses[3 * b + 1] - poses[3 * a + 1];
    const double w = lambda * bw[k];
    const double L0 = w * bp[6 * k + 3], L1 = w * bp[6 * k + 4], L2 = w * bp[6 * k + 5];
    const double e0 = c * dx + s * dy - bp[6 * k];
    const double e1 = -s * dx + c * dy - bp[6 * k + 1];
    double e2 = poses[3 * b + 2] - poses[3 * a + 2] - bp[6 * k + 2];
    if (wrap) e2 = wrap_pi(e2);
    const double ga = -s * dx + c * dy, gb = -c * dx - s * dy;
    const double vi0 = v[3 * a], vi1 = v[3 * a + 1], vi2 = v[3 * a + 2];
    const double vj0 = v[3 * b], vj1 = v[3 * b + 1], vj2 = v[3 * b + 2];
    // J_i = [[-c,-s,ga],[s,-c,gb],[0,0,-1]], J_j = [[c,s,0],[-s,c,0],[0,0,1]]
    const double u0 = -c * vi0 - s * vi1 + ga * vi2 + c * vj0 + s * vj1;
    const double u1 = s * vi0 - c * vi1 + gb * vi2 - s * vj0 + c * vj1;
    const double u2 = vj2 - vi2;
    dl_dw[k] = -(u0 * L0 * e0 + u1 * L1 * e1 + u2 * L2 * e2);
}

int launch_adjoint_grad(DeviceGraph &g, const double *v, double *dl_dw)
{
    const HostGraph &h = g.h;
    if (h.n_bet_e == 0) return GSLAM_OK;
    hipLaunchKernelGGL(adjoint_grad_kernel,
                       dim3(static_cast<unsigned int>((h.n_bet_e + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, g.stream, g.poses, g.e_t1, g.e_t2, g.bet_param, g.bet_w,
                       h.n_bet_e, h.n_obs_e + h.n_mot_e, g.lambda, g.wrap, v, dl_dw);
    GSLAM_HIP(hipGetLastError());
    return GSLAM_OK;
}

}  // namespace gslam

using gslam::DeviceGraph;
using gslam::set_error;

extern "C" int gslam_adjoint(gslam_graph *gh, const double *dl_dx, const gslam_solve_opts *opts,
                             double *dl_dw, double *dl_dlambda, gslam_solve_stats *stats)
{
    if (!gh || !dl_dx || !opts) return GSLAM_ERR_ARG;
    DeviceGraph &g = gh->g;
    const gslam::HostGraph &h = g.h;
    if (h.n_obs_e != 0 || h.n_mot_e != 0) {
        set_error("gslam_adjoint: only graphs of between edges are supported");
        return GSLAM_ERR_ARG;
    }
    if (!g.anchor_prior) {
        set_error("gslam_adjoint: needs desc.anchor_prior = 1 (well-defined converged poses)");
        return GSLAM_ERR_STATE;
    }
    if (!(opts->tol >= 0.0) || opts->max_iter <= 0) {
        set_error("gslam_adjoint: tol must be >= 0 and max_iter > 0");
        return GSLAM_ERR_ARG;
    }
    const int64_t dim = 3 * static_cast<int64_t>(h.n), nb = h.n_bet_e;
    GSLAM_TRY(g.vecs_reserve());
    double *b = g.vecs, *v = g.vecs + dim, *gw = g.vecs + 4 * dim;
    gslam_solve_opts o = *opts;
    o.solver = GSLAM_SOLVER_PCG_FP64;
    gslam_solve_stats st{};
    g.assembled = false;
    g.solved = false;
    g.factors_valid = false;
    int rc = gslam::launch_assembly(g, true);
    if (rc == GSLAM_OK) rc = g.upload(b, dl_dx, dim * 8);
    if (rc == GSLAM_OK) rc = gslam::pcg_solve(g, o, b, v, st);
    const bool capped = rc == GSLAM_OK && !st.converged && !(st.flags & 1) && st.iters >= o.max_iter;
    if (rc == GSLAM_OK) rc = gslam::launch_adjoint_grad(g, v, gw);
    std::vector<double> host(static_cast<size_t>(nb));
    if (rc == GSLAM_OK && nb > 0) rc = g.download(host.data(), gw, nb * 8);
    if (stats) *stats = st;
    if (rc != GSLAM_OK) return rc;
    if (dl_dw && nb > 0) std::copy(host.begin(), host.end(), dl_dw);
    if (dl_dlambda) {
        double sum = 0.0;
        for (double x : host) sum += x;
        *dl_dlambda = sum / g.lambda;
    }
    return capped ? GSLAM_ERR_NOT_CONVERGED : GSLAM_OK;
}
