// The cost term of the reference Gauss-Jordan drivers' use_diff = 0 stopping
// test (src/gpu_gj_fp64/GSLAM.cu:245-257, the same lines in
// src/cpu_gj_mixed/GSLAM.c:74-86), evaluated on the host in the reference's
// operation order so that the printed "cost:" values match to the last bit.
//
// Per edge the drivers form e^T M e * lambda_e with e the edge's stored
// residual read as a 3-vector and M its stored 3x3 "Omega":
//   * landmark-pair edges store the 2-vector residual hat_e (ObsEdge.c:76-82)
//     and the 2x2 information inv(Sigma) flattened into Omega[0..3], zeros in
//     Omega[4..8] (ObsEdge.c:126-140, 249-256), lambda = 1;
//   * motion edges store hat_e = 0 (MotionEdge.c:234) and contribute 0.
// The third component of e is read past the end of hat_e[2], i.e. from the
// first bytes of the next edge record (two small unsigned indices, a
// subnormal double); its products are far below half an ulp of the running
// sum and are taken as zero here.  The running sum is never reset between
// rounds in the reference (GSLAM.cu:193, 257), so the caller passes it in.
#include <cmath>
#include <vector>

#include "graph_build.h"

namespace gslam {
namespace {

// result = a (ay x ax) . b (ax x bx); each entry a left-to-right sum from
// 0.0 (MatrixOps.c dot / vec_dot)
void ref_dot(double *res, const double *a, int ay, int ax, const double *b, int bx)
{
    for (int i = 0; i < ay * bx; ++i) {
        double s = 0.0;
        for (int j = 0; j < ax; ++j) s += a[(i / bx) * ax + j] * b[j * bx + i % bx];
        res[i] = s;
    }
}

void ref_tra2(double *res, const double *m)
{
    res[0] = m[0];
    res[1] = m[2];
    res[2] = m[1];
    res[3] = m[3];
}

// unpivoted Gauss-Jordan inverse of a 2x2 (MatrixOps.c inv, dim = 2); a zero
// pivot gives the zero matrix
void ref_inv2(double *res, const double *m)
{
    double t[4] = {m[0], m[1], m[2], m[3]};
    res[0] = 1.0;
    res[1] = 0.0;
    res[2] = 0.0;
    res[3] = 1.0;
    for (int i = 0; i < 2; ++i) {
        if (t[i * 2 + i] == 0.0) {
            res[0] = res[1] = res[2] = res[3] = 0.0;
            return;
        }
        for (int j = 0; j < 2; ++j) {
            if (i == j) continue;
            const double ratio = t[j * 2 + i] / t[i * 2 + i];
            for (int k = 0; k < 2; ++k) {
                t[j * 2 + k] = t[j * 2 + k] - ratio * t[i * 2 + k];
                res[j * 2 + k] = res[j * 2 + k] - ratio * res[i * 2 + k];
            }
        }
    }
    for (int i = 0; i < 4; ++i) res[i] = res[i] / t[(i / 2) * 2 + i / 2];
}

}  // namespace

double obs_cost_terms(const HostGraph &h, const double *poses, const double snr[2], double acc)
{
    for (int64_t e = 0; e < h.n_obs_e; ++e) {
        const double *x1 = poses + 3 * static_cast<int64_t>(h.t1[e]);
        const double *x2 = poses + 3 * static_cast<int64_t>(h.t2[e]);
        const double r1 = h.obs_param[4 * e], p1 = h.obs_param[4 * e + 1];
        const double r2 = h.obs_param[4 * e + 2], p2 = h.obs_param[4 * e + 3];
        const double s1 = std::sin(x1[2] + p1), c1 = std::cos(x1[2] + p1);
        const double s2 = std::sin(x2[2] + p2), c2 = std::cos(x2[2] + p2);
        const double he[2] = {x2[0] - x1[0] + r2 * c2 - r1 * c1, x2[1] - x1[1] + r2 * s2 - r1 * s1};
        const double q1[4] = {std::pow(r1 * snr[0], 2.0), 0.0, 0.0, std::pow(snr[1], 2.0)};
        const double R1[4] = {-c1, r1 * s1, -s1, -r1 * c1};
        const double q2[4] = {std::pow(r2 * snr[0], 2.0), 0.0, 0.0, std::pow(snr[1], 2.0)};
        const double R2[4] = {c2, -r2 * s2, s2, r2 * c2};
        double tR1[4], tR2[4], a1[4], b1[4], a2[4], b2[4], sig[4], om[4];
        ref_tra2(tR1, R1);
        ref_dot(a1, R1, 2, 2, q1, 2);
        ref_dot(b1, a1, 2, 2, tR1, 2);
        ref_tra2(tR2, R2);
        ref_dot(a2, R2, 2, 2, q2, 2);
        ref_dot(b2, a2, 2, 2, tR2, 2);
        for (int k = 0; k < 4; ++k) sig[k] = b1[k] + b2[k];
        ref_inv2(om, sig);
        // e = (he0, he1, 0), M = [om0 om1 om2; om3 0 0; 0 0 0] (3x3 view)
        const double ev[3] = {he[0], he[1], 0.0};
        const double M[9] = {om[0], om[1], om[2], om[3], 0.0, 0.0, 0.0, 0.0, 0.0};
        double v[3], s;
        ref_dot(v, ev, 1, 3, M, 3);
        ref_dot(&s, v, 1, 3, ev, 1);
        acc += s * 1.0;
    }
    return acc;
}

}  // namespace gslam
