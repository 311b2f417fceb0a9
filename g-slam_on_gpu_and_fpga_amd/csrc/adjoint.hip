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

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
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

// ---- reference edge model: dL/dlambda by a tangent solve ----
//
// Landmark-pair and velocity-model motion edges (src/cpu_pcg_fp64/ObsEdge.c,
// MotionEdge.c) give xi(x) = -sum_e J_e^T W_e(x) e_e with information that
// depends on the poses, so xi is not the gradient of a cost and its Jacobian
// is not symmetric: the GN fixed point x* (xi(x*, lambda) = 0, with the
// anchor prior) moves with lambda as
//   K u = d xi / d lambda,   K = -d xi / dx at x*,   u = dx*/dlambda,
// and dL/dlambda = dL/dx* . u.  xi is linear in lambda (lambda scales the
// motion edges only, MotionEdge.c:113-175; ObsEdge.c:209 leaves landmark
// pairs unscaled), so d xi / d lambda = (xi(lambda) - xi(0)) / lambda from
// two assemblies.  K is applied matrix-free by central differences of the
// assembled xi, K v = -(xi(x* + h v) - xi(x* - h v)) / 2h, and the system is
// solved by GMRES preconditioned on the right by Omega(x*)^-1 (the FP64 PCG):
// Omega is K without the second-order terms, which vanish with the
// residuals, so the preconditioned operator is close to the identity.
namespace {

double dot(const std::vector<double> &a, const std::vector<double> &b)
{
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

}  // namespace

// xi at host poses x (3N) and the given lambda (device assembly), into xi_h.
static int xi_at(DeviceGraph &g, const std::vector<double> &x, double lambda,
                 std::vector<double> &xi_h)
{
    GSLAM_TRY(g.upload(g.poses, x.data(), g.dim() * 8));
    g.lambda = lambda;
    GSLAM_TRY(launch_assembly(g));
    return g.download(xi_h.data(), g.xi, g.dim() * 8);
}

int lambda_tangent(DeviceGraph &g, const double *dl_dx, const gslam_solve_opts &opts,
                   double *dl_dlambda, gslam_solve_stats &st)
{
    const int64_t dim = g.dim();
    const size_t nd = static_cast<size_t>(dim);
    const double lam = g.lambda;
    if (!(lam > 0.0)) {
        set_error("gslam_adjoint: lambda must be > 0");
        return GSLAM_ERR_ARG;
    }
    GSLAM_TRY(g.vecs_reserve());
    std::vector<double> xs(nd), xi1(nd), xi0(nd), b(nd), t(nd), tp(nd), tm(nd);
    GSLAM_TRY(g.download(xs.data(), g.poses, dim * 8));
    GSLAM_TRY(xi_at(g, xs, lam, xi1));
    GSLAM_TRY(xi_at(g, xs, 0.0, xi0));
    g.lambda = lam;
    for (size_t i = 0; i < nd; ++i) b[i] = (xi1[i] - xi0[i]) / lam;
    double xscale = 1.0;
    for (double v : xs) xscale = std::max(xscale, std::fabs(v));

    gslam_solve_opts po = opts;
    po.solver = GSLAM_SOLVER_PCG_FP64;
    po.mixed_handover = 0.0;
    int64_t pcg_iters = 0;
    // z = Omega(x*)^-1 r
    auto precond = [&](const std::vector<double> &r, std::vector<double> &z) -> int {
        GSLAM_TRY(g.upload(g.poses, xs.data(), dim * 8));
        g.lambda = lam;
        GSLAM_TRY(launch_assembly(g));
        GSLAM_TRY(g.upload(g.vecs, r.data(), dim * 8));
        gslam_solve_stats ps{};
        GSLAM_TRY(pcg_solve(g, po, g.vecs, g.vecs + dim, ps));
        pcg_iters += ps.iters;
        return g.download(z.data(), g.vecs + dim, dim * 8);
    };
    // w = K z by central differences of xi
    auto apply_k = [&](const std::vector<double> &z, std::vector<double> &w) -> int {
        double zmax = 0.0;
        for (double v : z) zmax = std::max(zmax, std::fabs(v));
        if (zmax == 0.0) {
            std::fill(w.begin(), w.end(), 0.0);
            return GSLAM_OK;
        }
        const double h = 1e-6 * xscale / zmax;
        for (size_t i = 0; i < nd; ++i) t[i] = xs[i] + h * z[i];
        GSLAM_TRY(xi_at(g, t, lam, tp));
        for (size_t i = 0; i < nd; ++i) t[i] = xs[i] - h * z[i];
        GSLAM_TRY(xi_at(g, t, lam, tm));
        for (size_t i = 0; i < nd; ++i) w[i] = -(tp[i] - tm[i]) / (2.0 * h);
        return GSLAM_OK;
    };

    // right-preconditioned restarted GMRES on K u = b
    const int m = 40;
    const int max_it = std::max(1, std::min(opts.max_iter, 400));
    const double tol = opts.tol > 0.0 ? opts.tol : 1e-10;
    const double bn = std::sqrt(dot(b, b));
    std::vector<double> u(nd, 0.0), r(nd), w(nd), z(nd);
    std::vector<std::vector<double>> V(m + 1, std::vector<double>(nd)), Z(m, std::vector<double>(nd));
    std::vector<double> H((m + 1) * m), cs(m), sn(m), gam(m + 1);
    int it = 0;
    double res = bn;
    bool conv = bn == 0.0;
    const bool trace = std::getenv("GSLAM_TANGENT_TRACE") != nullptr;
    // the relative accuracy of the central-difference Jacobian products
    constexpr double kFdFloor = 1e-4;
    int cycles_done = 0;
    while (!conv && it < max_it) {
        GSLAM_TRY(apply_k(u, w));
        for (size_t i = 0; i < nd; ++i) r[i] = b[i] - w[i];
        const double beta = std::sqrt(dot(r, r));
        res = beta;
        if (beta <= tol * bn) {
            conv = true;
            break;
        }
        for (size_t i = 0; i < nd; ++i) V[0][i] = r[i] / beta;
        std::fill(gam.begin(), gam.end(), 0.0);
        gam[0] = beta;
        int j = 0;
        for (; j < m && it < max_it; ++j, ++it) {
            GSLAM_TRY(precond(V[j], Z[j]));
            GSLAM_TRY(apply_k(Z[j], w));
            for (int i = 0; i <= j; ++i) {  // modified Gram-Schmidt
                const double hij = dot(w, V[i]);
                H[i * m + j] = hij;
                for (size_t k = 0; k < nd; ++k) w[k] -= hij * V[i][k];
            }
            const double hn = std::sqrt(dot(w, w));
            H[(j + 1) * m + j] = hn;
            if (hn > 0.0)
                for (size_t k = 0; k < nd; ++k) V[j + 1][k] = w[k] / hn;
            for (int i = 0; i < j; ++i) {  // apply the earlier rotations
                const double a = H[i * m + j], c = H[(i + 1) * m + j];
                H[i * m + j] = cs[i] * a + sn[i] * c;
                H[(i + 1) * m + j] = -sn[i] * a + cs[i] * c;
            }
            const double a = H[j * m + j], c = H[(j + 1) * m + j];
            const double rr = std::hypot(a, c);
            cs[j] = rr > 0.0 ? a / rr : 1.0;
            sn[j] = rr > 0.0 ? c / rr : 0.0;
            H[j * m + j] = rr;
            H[(j + 1) * m + j] = 0.0;
            gam[j + 1] = -sn[j] * gam[j];
            gam[j] = cs[j] * gam[j];
            res = std::fabs(gam[j + 1]);
            if (trace) std::fprintf(stderr, "tangent: GMRES it %d rel res %.3e\n", it + 1, res / bn);
            if (res <= tol * bn || hn == 0.0) {
                ++j;
                ++it;
                break;
            }
        }
        // y = H^-1 gam (upper triangular j x j), u += Z y
        std::vector<double> y(j);
        for (int i = j - 1; i >= 0; --i) {
            double s = gam[i];
            for (int k = i + 1; k < j; ++k) s -= H[i * m + k] * y[k];
            y[i] = s / H[i * m + i];
        }
        for (int i = 0; i < j; ++i)
            for (size_t k = 0; k < nd; ++k) u[k] += y[i] * Z[i][k];
        if (res <= tol * bn) {
            // the residual of u through a fresh central-difference product:
            // its noise (~1e-5 of |b| here) bounds what a restart can gain,
            // so one restart at most, then accept at that level
            GSLAM_TRY(apply_k(u, w));
            for (size_t i = 0; i < nd; ++i) r[i] = b[i] - w[i];
            res = std::sqrt(dot(r, r));
            if (trace) std::fprintf(stderr, "tangent: true rel res %.3e\n", res / bn);
            ++cycles_done;
            if (res <= 10.0 * tol * bn || cycles_done >= 2) {
                conv = res <= kFdFloor * bn;
                break;
            }
        }
    }
    // restore the converged poses; the handle is left unassembled
    GSLAM_TRY(g.upload(g.poses, xs.data(), dim * 8));
    g.lambda = lam;
    double d = 0.0;
    for (size_t i = 0; i < nd; ++i) d += dl_dx[i] * u[i];
    *dl_dlambda = d;
    st.iters = it;
    st.iters_fp64 = static_cast<int32_t>(std::min<int64_t>(pcg_iters, INT32_MAX));
    st.converged = conv ? 1 : 0;
    st.rel_res = bn > 0.0 ? res / bn : 0.0;
    st.true_rel_res = st.rel_res;
    return conv ? GSLAM_OK : GSLAM_ERR_NOT_CONVERGED;
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
    if ((h.n_obs_e != 0 || h.n_mot_e != 0) && h.n_bet_e != 0) {
        set_error("gslam_adjoint: graphs that mix between edges with the reference's "
                  "landmark-pair / motion edges are not supported");
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
    if (h.n_obs_e != 0 || h.n_mot_e != 0) {
        // reference edge model: no per-edge weights, dL/dlambda by the tangent solve
        gslam_solve_stats st{};
        g.assembled = false;
        g.solved = false;
        g.factors_valid = false;
        double dlam = 0.0;
        const int rc = gslam::lambda_tangent(g, dl_dx, *opts, &dlam, st);
        if (dl_dlambda) *dl_dlambda = dlam;
        if (stats) *stats = st;
        return rc;
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
