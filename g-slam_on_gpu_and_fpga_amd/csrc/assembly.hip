// Edge evaluation and deterministic assembly of Omega (BSR 3x3) and xi.
//
// Every entry is summed in edge order, starting from the anchor -- the order
// of build_csr_from_edges (src/cpu_pcg_fp64/MatrixOps.c:223-272) -- with no
// atomics: Omega is bitwise repeatable run to run, unlike the reference's
// atomicAdd scatter (src/gpu_pcg_fp64/device_graph.cu:437-456, 805-826), and
// no per-round binary searches are needed.  Two schedules give the same
// values:
//   pose walk (graphs where no pose has more than kWalkMaxIncidence edge
//     ends): one thread per pose evaluates each incident edge in edge order
//     and adds its diagonal block, its off-diagonal block (a precomputed
//     incidence -> block map; the first contribution of a block stores) and
//     its xi half -- one pass, each edge evaluated once per end;
//   staged: one thread per edge writes its 42 values (4 blocks + 2 xi
//     halves) to a 48-double stage record, then one thread per Omega value
//     (9 per block) and per xi entry sums its contribution list.
#include "device_common.h"
#include "gslam_internal.h"

namespace gslam {

// 8 FP32 factors of one edge, stored at the slots of both of its incidence
// entries (the mixed solver's factorized operator, pcg_mixed.hip): factor c
// of slot 256 h + l at frec[(8 h + c) 256 + l].
__device__ inline void put_factor_slot(float *__restrict__ frec, int64_t s, float4 a, float4 b)
{
    const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    float *f = frec + (s / kTileBlocks) * 8 * kTileBlocks + s % kTileBlocks;
#pragma unroll
    for (int c = 0; c < 8; ++c) f[c * kTileBlocks] = v[c];
}

__device__ inline void put_factors(float *__restrict__ frec, const int32_t *__restrict__ mt_pos,
                                   int64_t e, float4 a, float4 b)
{
    if (!frec) return;
#pragma unroll
    for (int side = 0; side < 2; ++side) put_factor_slot(frec, mt_pos[2 * e + side], a, b);
}

// One edge evaluated at the current poses: its four 3x3 blocks UL (t1,t1),
// UR (t1,t2), BL (t2,t1), BR (t2,t2) row-major, then the xi halves of t1 and
// t2, and the 8 FP32 factors of the mixed operator.
struct EdgeRec {
    double s[42];
    float4 fa, fb;
};

__device__ inline void store_record(double *__restrict__ st, const EdgeRec &R)
{
#pragma unroll
    for (int k = 0; k < 42; ++k) st[k] = R.s[k];
}

// Velocity-model motion edge (i, i+1).  Model: src/cpu_pcg_fp64/MotionEdge.c
// :9-193 with state_transition (GSLAM.c:373-401): Omega_m = (A M A^T +
// 1e-4 I)^-1, blocks lambda*[F^T Om F, -F^T Om, -Om F, Om], xi1 = lambda F^T
// Om e, xi2 = -lambda Om e, e = x2 - x_pred (theta wrapped only in wrap
// mode).  The 3x3 inverse is the symmetric cofactor form.
__device__ inline void motion_record(const double *__restrict__ poses, const double *__restrict__ mp,
                                     int64_t i, double delta, double delta_t, double lambda,
                                     double m0, double m1, double m2, double m3, int wrap,
                                     EdgeRec &R)
{
    const double x10 = poses[3 * i], x11 = poses[3 * i + 1], th = poses[3 * i + 2];
    const double x20 = poses[3 * i + 3], x21 = poses[3 * i + 4], x22 = poses[3 * i + 5];
    const double nu = mp[2 * i];
    double om = mp[2 * i + 1];
    if (fabs(om) < 1e-5) om = 1e-5;  // MotionEdge.c:21-22 (sign dropped)

    const double M00 = m0 * m0 * fabs(nu) / delta + m1 * m1 * fabs(om) / delta;
    const double M11 = m2 * m2 * fabs(nu) / delta + m3 * m3 * fabs(om) / delta;
    double st, ct, stw, ctw;
    sincos(th, &st, &ct);
    sincos(th + om * delta, &stw, &ctw);
    const double nuom = nu / om;
    const double dS = stw - st, dC = ct - ctw;
    const double A00 = dS / om, A10 = dC / om;
    const double A01 = nuom * delta * ctw - nuom / om * dS;
    const double A11 = nuom * delta * stw - nuom / om * dC;
    const double F02 = -nuom * dC, F12 = nuom * dS;

    const double S00 = A00 * A00 * M00 + A01 * A01 * M11 + 1e-4;
    const double S01 = A00 * A10 * M00 + A01 * A11 * M11;
    const double S02 = A01 * delta * M11;
    const double S11 = A10 * A10 * M00 + A11 * A11 * M11 + 1e-4;
    const double S12 = A11 * delta * M11;
    const double S22 = delta * delta * M11 + 1e-4;
    const double c00 = S11 * S22 - S12 * S12, c01 = S02 * S12 - S01 * S22;
    const double c02 = S01 * S12 - S02 * S11, c11 = S00 * S22 - S02 * S02;
    const double c12 = S01 * S02 - S00 * S12, c22 = S00 * S11 - S01 * S01;
    double det = S00 * c00 + S01 * c01 + S02 * c02;
    if (det == 0.0) det = 1e-30;
    const double Om[9] = {c00 / det, c01 / det, c02 / det, c01 / det, c11 / det,
                          c12 / det, c02 / det, c12 / det, c22 / det};
    double T[9];  // F^T Om
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        T[j] = Om[j];
        T[3 + j] = Om[3 + j];
        T[6 + j] = F02 * Om[j] + F12 * Om[3 + j] + Om[6 + j];
    }
    double *s = R.s;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        s[0 + 3 * r + 0] = lambda * T[3 * r];
        s[0 + 3 * r + 1] = lambda * T[3 * r + 1];
        s[0 + 3 * r + 2] = lambda * (T[3 * r] * F02 + T[3 * r + 1] * F12 + T[3 * r + 2]);
        s[9 + 3 * r + 0] = -lambda * T[3 * r];
        s[9 + 3 * r + 1] = -lambda * T[3 * r + 1];
        s[9 + 3 * r + 2] = -lambda * T[3 * r + 2];
        s[18 + 3 * r + 0] = -lambda * Om[3 * r];
        s[18 + 3 * r + 1] = -lambda * Om[3 * r + 1];
        s[18 + 3 * r + 2] = -lambda * (Om[3 * r] * F02 + Om[3 * r + 1] * F12 + Om[3 * r + 2]);
        s[27 + 3 * r + 0] = lambda * Om[3 * r];
        s[27 + 3 * r + 1] = lambda * Om[3 * r + 1];
        s[27 + 3 * r + 2] = lambda * Om[3 * r + 2];
    }
    // prediction: the reference passes delta as unsigned int (GSLAM.h), so
    // the prediction uses the truncated step delta_t
    double sw = stw, cw = ctw;
    if (delta_t != delta) sincos(th + om * delta_t, &sw, &cw);
    const double e0 = x20 - (x10 + nuom * (sw - st));
    const double e1 = x21 - (x11 + nuom * (ct - cw));
    double e2 = x22 - (th + om * delta_t);
    if (wrap) e2 = wrap_pi(e2);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        s[36 + r] = lambda * (T[3 * r] * e0 + T[3 * r + 1] * e1 + T[3 * r + 2] * e2);
        s[39 + r] = -lambda * (Om[3 * r] * e0 + Om[3 * r + 1] * e1 + Om[3 * r + 2] * e2);
    }
    // factors: J1 = -F, J2 = I, W = lambda Om
    R.fa = make_float4(static_cast<float>(F02), static_cast<float>(F12),
                       static_cast<float>(lambda * Om[0]), static_cast<float>(lambda * Om[1]));
    R.fb = make_float4(static_cast<float>(lambda * Om[2]), static_cast<float>(lambda * Om[4]),
                       static_cast<float>(lambda * Om[5]), static_cast<float>(lambda * Om[8]));
}

__global__ void __launch_bounds__(kThreads)
motion_edges_kernel(const double *__restrict__ poses, const double *__restrict__ mp,
                    int64_t n_mot, int64_t e_off, double delta, double delta_t,
                    double lambda, double m0, double m1, double m2, double m3,
                    int wrap, double *__restrict__ stage, float *__restrict__ frec,
                    const int32_t *__restrict__ mt_pos)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n_mot) return;
    EdgeRec R;
    motion_record(poses, mp, i, delta, delta_t, lambda, m0, m1, m2, m3, wrap, R);
    store_record(stage + (e_off + i) * kStage, R);
    put_factors(frec, mt_pos, e_off + i, R.fa, R.fb);
}

// Landmark-pair observation edge.  Model: src/cpu_pcg_fp64/ObsEdge.c:10-212:
// e = x2 - x1 + r2 (c2, s2) - r1 (c1, s1), Sigma = R1 Q1 R1^T + R2 Q2 R2^T,
// Om = Sigma^-1, B1 = [-I | b1], B2 = [I | b2], blocks Bi^T Om Bj,
// xi = -B^T Om e.  No lambda (ObsEdge.c:209).
__device__ inline void obs_record(const double *__restrict__ poses, int64_t a, int64_t b,
                                  const double *__restrict__ op, int64_t e, double snr0,
                                  double snr1, EdgeRec &R)
{
    const double r1 = op[4 * e], p1 = op[4 * e + 1], r2 = op[4 * e + 2], p2 = op[4 * e + 3];
    double s1, c1, s2, c2;
    sincos(poses[3 * a + 2] + p1, &s1, &c1);
    sincos(poses[3 * b + 2] + p2, &s2, &c2);
    const double ex = poses[3 * b] - poses[3 * a] + r2 * c2 - r1 * c1;
    const double ey = poses[3 * b + 1] - poses[3 * a + 1] + r2 * s2 - r1 * s1;
    const double q1 = (r1 * snr0) * (r1 * snr0), q2 = (r2 * snr0) * (r2 * snr0);
    const double qb = snr1 * snr1;
    const double S00 = c1 * c1 * q1 + r1 * r1 * s1 * s1 * qb + c2 * c2 * q2 + r2 * r2 * s2 * s2 * qb;
    const double S01 = c1 * s1 * q1 - r1 * r1 * s1 * c1 * qb + c2 * s2 * q2 - r2 * r2 * s2 * c2 * qb;
    const double S11 = s1 * s1 * q1 + r1 * r1 * c1 * c1 * qb + s2 * s2 * q2 + r2 * r2 * c2 * c2 * qb;
    double det = S00 * S11 - S01 * S01;
    if (det == 0.0) det = 1e-30;
    const double o00 = S11 / det, o01 = -S01 / det, o11 = S00 / det;
    // third columns of B1, B2
    const double b1x = r1 * s1, b1y = -r1 * c1, b2x = -r2 * s2, b2y = r2 * c2;
    const double Ob1x = o00 * b1x + o01 * b1y, Ob1y = o01 * b1x + o11 * b1y;  // Om b1
    const double Ob2x = o00 * b2x + o01 * b2y, Ob2y = o01 * b2x + o11 * b2y;  // Om b2
    const double b1Ob1 = b1x * Ob1x + b1y * Ob1y, b1Ob2 = b1x * Ob2x + b1y * Ob2y;
    const double b2Ob1 = b2x * Ob1x + b2y * Ob1y, b2Ob2 = b2x * Ob2x + b2y * Ob2y;
    double *s = R.s;
    // UL = B1^T Om B1 = [[Om, -Om b1], [-b1^T Om, b1^T Om b1]]
    s[0] = o00;   s[1] = o01;   s[2] = -Ob1x;
    s[3] = o01;   s[4] = o11;   s[5] = -Ob1y;
    s[6] = -Ob1x; s[7] = -Ob1y; s[8] = b1Ob1;
    // UR = B1^T Om B2 = [[-Om, -Om b2], [b1^T Om, b1^T Om b2]]
    s[9] = -o00;  s[10] = -o01; s[11] = -Ob2x;
    s[12] = -o01; s[13] = -o11; s[14] = -Ob2y;
    s[15] = Ob1x; s[16] = Ob1y; s[17] = b1Ob2;
    // BL = B2^T Om B1 = [[-Om, Om b1], [-b2^T Om, b2^T Om b1]]
    s[18] = -o00;  s[19] = -o01;  s[20] = Ob1x;
    s[21] = -o01;  s[22] = -o11;  s[23] = Ob1y;
    s[24] = -Ob2x; s[25] = -Ob2y; s[26] = b2Ob1;
    // BR = B2^T Om B2 = [[Om, Om b2], [b2^T Om, b2^T Om b2]]
    s[27] = o00;  s[28] = o01;  s[29] = Ob2x;
    s[30] = o01;  s[31] = o11;  s[32] = Ob2y;
    s[33] = Ob2x; s[34] = Ob2y; s[35] = b2Ob2;
    const double Oex = o00 * ex + o01 * ey, Oey = o01 * ex + o11 * ey;
    s[36] = Oex;
    s[37] = Oey;
    s[38] = -(b1x * Oex + b1y * Oey);
    s[39] = -Oex;
    s[40] = -Oey;
    s[41] = -(b2x * Oex + b2y * Oey);
    // factors: B1 = [-I | b1], B2 = [I | b2], W = Om
    R.fa = make_float4(static_cast<float>(b1x), static_cast<float>(b1y), static_cast<float>(b2x),
                       static_cast<float>(b2y));
    R.fb = make_float4(static_cast<float>(o00), static_cast<float>(o01), static_cast<float>(o11), 0.f);
}

__global__ void __launch_bounds__(kThreads)
obs_edges_kernel(const double *__restrict__ poses, const int32_t *__restrict__ t1,
                 const int32_t *__restrict__ t2, const double *__restrict__ op,
                 int64_t n_obs_e, double snr0, double snr1, double *__restrict__ stage,
                 float *__restrict__ frec, const int32_t *__restrict__ mt_pos)
{
    const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (e >= n_obs_e) return;
    EdgeRec R;
    obs_record(poses, t1[e], t2[e], op, e, snr0, snr1, R);
    store_record(stage + e * kStage, R);
    put_factors(frec, mt_pos, e, R.fa, R.fb);
}

// SE(2) between edge (no reference counterpart):
// e = [R(th_i)^T (t_j - t_i) - z_t ; th_j - th_i - z_th], L = w diag(info),
// blocks Ji^T L Jj..., xi = -J^T L e.
__device__ inline void between_record(const double *__restrict__ poses, int64_t a, int64_t b,
                                      const double *__restrict__ bp, const double *__restrict__ bw,
                                      int64_t k, double lambda, int wrap, int exact, EdgeRec &R)
{
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
    // Ji = [[-c,-s,ga],[s,-c,gb],[0,0,-1]], Jj = [[c,s,0],[-s,c,0],[0,0,1]]
    const double Ji[9] = {-c, -s, ga, s, -c, gb, 0.0, 0.0, -1.0};
    const double Jj[9] = {c, s, 0.0, -s, c, 0.0, 0.0, 0.0, 1.0};
    const double L[3] = {L0, L1, L2};
    const double ev[3] = {e0, e1, e2};
    double *st = R.s;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double *JA = (q < 2) ? Ji : Jj;
        const double *JB = (q % 2 == 0) ? Ji : Jj;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
                st[9 * q + 3 * r + cc] = JA[r] * L[0] * JB[cc] + JA[3 + r] * L[1] * JB[3 + cc] +
                                         JA[6 + r] * L[2] * JB[6 + cc];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        st[36 + r] = -(Ji[r] * L[0] * ev[0] + Ji[3 + r] * L[1] * ev[1] + Ji[6 + r] * L[2] * ev[2]);
        st[39 + r] = -(Jj[r] * L[0] * ev[0] + Jj[3 + r] * L[1] * ev[1] + Jj[6 + r] * L[2] * ev[2]);
    }
    if (exact) {
        // second-order terms sum_k (L e)_k d2 e_k of the exact Hessian: the
        // translation residual R(th_i)^T (t_j - t_i) is linear in t and
        // quadratic only through th_i, so they touch the th_i row and column:
        // (th_i, th_i) c_tt, (th_i, t_i) (cx, cy), (th_i, t_j) -(cx, cy)
        const double g0 = L0 * e0, g1 = L1 * e1;
        const double c_tt = -g0 * (c * dx + s * dy) + g1 * (s * dx - c * dy);
        const double cx = g0 * s + g1 * c, cy = -g0 * c + g1 * s;
        st[2] += cx;
        st[5] += cy;
        st[6] += cx;
        st[7] += cy;
        st[8] += c_tt;
        st[9 + 6] -= cx;
        st[9 + 7] -= cy;
        st[18 + 2] -= cx;
        st[18 + 5] -= cy;
    }
    // factors: rotation (c, s), theta column (ga, gb), W = diag(L)
    R.fa = make_float4(static_cast<float>(c), static_cast<float>(s), static_cast<float>(ga),
                       static_cast<float>(gb));
    R.fb = make_float4(static_cast<float>(L0), static_cast<float>(L1), static_cast<float>(L2), 0.f);
}

__global__ void __launch_bounds__(kThreads)
between_edges_kernel(const double *__restrict__ poses, const int32_t *__restrict__ t1,
                     const int32_t *__restrict__ t2, const double *__restrict__ bp,
                     const double *__restrict__ bw, int64_t n_bet, int64_t e_off,
                     double lambda, int wrap, int exact, double *__restrict__ stage,
                     float *__restrict__ frec, const int32_t *__restrict__ mt_pos)
{
    const int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (k >= n_bet) return;
    const int64_t e = e_off + k;
    EdgeRec R;
    between_record(poses, t1[e], t2[e], bp, bw, k, lambda, wrap, exact, R);
    store_record(stage + e * kStage, R);
    put_factors(frec, mt_pos, e, R.fa, R.fb);
}

// Pose walk: one thread per pose i.  Its incidence entries (e*2+side, in
// edge order) give, per edge, the diagonal contribution (UL for side 0, BR
// for side 1) -- summed in registers after the anchor (MatrixOps.c:223-231)
// --, the off-diagonal one (UR into block (t1,t2), BL into (t2,t1): block
// row i, which only this thread writes; xblk holds -(b+1) for a block's first
// contribution, which stores 0.0 + v, as the staged sum starts from 0.0) and
// the xi half; the edge's FP32 factors go to this end's operator slot.
__global__ void __launch_bounds__(kThreads)
assemble_walk_kernel(const double *__restrict__ poses, const int32_t *__restrict__ t1,
                     const int32_t *__restrict__ t2, const double *__restrict__ op,
                     const double *__restrict__ mp, const double *__restrict__ bp,
                     const double *__restrict__ bw, int64_t n_obs_e, int64_t n_mot_e,
                     double snr0, double snr1, double delta, double delta_t, double lambda,
                     double m0, double m1, double m2, double m3, int wrap, int exact,
                     const int32_t *__restrict__ xptr, const int32_t *__restrict__ xlist,
                     const int32_t *__restrict__ xblk, const int32_t *__restrict__ bdiag,
                     int32_t n, double anchor, int prior, double ref0, double ref1, double ref2,
                     double *__restrict__ bval, double *__restrict__ xi, float *__restrict__ frec,
                     const int32_t *__restrict__ mt_pos)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double dg[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) dg[k] = (i == 0 && (k == 0 || k == 4 || k == 8)) ? anchor : 0.0;
    double xv[3] = {0.0, 0.0, 0.0};
    const int32_t c1 = xptr[i + 1];
    for (int32_t c = xptr[i]; c < c1; ++c) {
        const int32_t ent = xlist[c];
        const int64_t e = ent >> 1;
        const bool side = ent & 1;
        EdgeRec R;
        if (e < n_obs_e)
            obs_record(poses, t1[e], t2[e], op, e, snr0, snr1, R);
        else if (e < n_obs_e + n_mot_e)
            motion_record(poses, mp, e - n_obs_e, delta, delta_t, lambda, m0, m1, m2, m3, wrap, R);
        else
            between_record(poses, t1[e], t2[e], bp, bw, e - n_obs_e - n_mot_e, lambda, wrap, exact,
                           R);
#pragma unroll
        for (int k = 0; k < 9; ++k) dg[k] += side ? R.s[27 + k] : R.s[k];
        const int32_t xb = xblk[ent];
        double *bv = bval + 9 * static_cast<int64_t>(xb < 0 ? -(xb + 1) : xb);
#pragma unroll
        for (int k = 0; k < 9; ++k) bv[k] = (xb < 0 ? 0.0 : bv[k]) + (side ? R.s[18 + k] : R.s[9 + k]);
#pragma unroll
        for (int r = 0; r < 3; ++r) xv[r] += side ? R.s[39 + r] : R.s[36 + r];
        if (frec) put_factor_slot(frec, mt_pos[ent], R.fa, R.fb);
    }
    double *bd = bval + 9 * static_cast<int64_t>(bdiag[i]);
#pragma unroll
    for (int k = 0; k < 9; ++k) bd[k] = dg[k];
    const int64_t o = 3 * static_cast<int64_t>(i);
    const double ref[3] = {ref0, ref1, ref2};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        double v = xv[r];
        if (prior && i == 0) v += anchor * (ref[r] - poses[r]);
        xi[o + r] = v;
    }
}

// Omega values: one thread per (block, element); contributions in edge order
// after the anchor (MatrixOps.c:223-231).
__global__ void __launch_bounds__(kThreads)
gather_blocks_kernel(const int32_t *__restrict__ cptr, const int32_t *__restrict__ clist,
                     const double *__restrict__ stage, int64_t n_vals, int64_t diag0,
                     double anchor, double *__restrict__ bval)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n_vals) return;
    const int64_t b = t / 9;
    const int k = static_cast<int>(t - 9 * b);
    double v = (b == diag0 && (k == 0 || k == 4 || k == 8)) ? anchor : 0.0;
    const int32_t c1 = cptr[b + 1];
    for (int32_t c = cptr[b]; c < c1; ++c) {
        const int32_t ent = clist[c];
        v += stage[static_cast<int64_t>(ent >> 2) * kStage + (ent & 3) * 9 + k];
    }
    bval[t] = v;
}

// xi entries: contributions in edge order; with the anchor prior, pose 0
// also gets anchor * (ref - pose0) (the reference has no xi anchor term).
__global__ void __launch_bounds__(kThreads)
gather_xi_kernel(const int32_t *__restrict__ xptr, const int32_t *__restrict__ xlist,
                 const double *__restrict__ stage, int64_t dim, double *__restrict__ xi,
                 const double *__restrict__ poses, int prior, double anchor, double ref0,
                 double ref1, double ref2)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= dim) return;
    const int64_t i = t / 3;
    const int r = static_cast<int>(t - 3 * i);
    double v = 0.0;
    const int32_t c1 = xptr[i + 1];
    for (int32_t c = xptr[i]; c < c1; ++c) {
        const int32_t ent = xlist[c];
        v += stage[static_cast<int64_t>(ent >> 1) * kStage + 36 + (ent & 1) * 3 + r];
    }
    if (prior && i == 0) v += anchor * ((r == 0 ? ref0 : r == 1 ? ref1 : ref2) - poses[r]);
    xi[t] = v;
}

__global__ void __launch_bounds__(kThreads)
bsr_to_csr_kernel(const int32_t *__restrict__ pos, const double *__restrict__ bval,
                  int64_t n_vals, double *__restrict__ csr)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n_vals) return;
    const int32_t p = pos[t];
    if (p >= 0) csr[p] = bval[t];
}

__global__ void __launch_bounds__(kThreads)
csr_to_bsr_kernel(const int32_t *__restrict__ pos, const double *__restrict__ csr,
                  int64_t n_vals, double *__restrict__ bval)
{
    const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (t >= n_vals) return;
    const int32_t p = pos[t];
    bval[t] = p >= 0 ? csr[p] : 0.0;
}

static inline unsigned int nblk(int64_t n) { return static_cast<unsigned int>((n + kThreads - 1) / kThreads); }

int launch_assembly(DeviceGraph &g, bool exact_hessian)
{
    const HostGraph &h = g.h;
    const double delta_t = static_cast<double>(static_cast<unsigned int>(g.delta));
    if (g.walk) {
        hipLaunchKernelGGL(assemble_walk_kernel, dim3(nblk(h.n)), dim3(kThreads), 0, g.stream,
                           g.poses, g.e_t1, g.e_t2, g.obs_param, g.mot_param, g.bet_param, g.bet_w,
                           h.n_obs_e, h.n_mot_e, g.snr[0], g.snr[1], g.delta, delta_t, g.lambda,
                           g.mns[0], g.mns[1], g.mns[2], g.mns[3], g.wrap, exact_hessian ? 1 : 0,
                           g.xptr, g.xlist, g.xblk, g.bdiag, h.n, g.anchor, g.anchor_prior,
                           g.pose0_ref[0], g.pose0_ref[1], g.pose0_ref[2], g.bval, g.xi, g.frec,
                           g.mt_pos);
        GSLAM_HIP(hipGetLastError());
        return GSLAM_OK;
    }
    if (h.n_obs_e > 0)
        hipLaunchKernelGGL(obs_edges_kernel, dim3(nblk(h.n_obs_e)), dim3(kThreads), 0, g.stream,
                           g.poses, g.e_t1, g.e_t2, g.obs_param, h.n_obs_e, g.snr[0], g.snr[1],
                           g.stage, g.frec, g.mt_pos);
    if (h.n_mot_e > 0) {
        hipLaunchKernelGGL(motion_edges_kernel, dim3(nblk(h.n_mot_e)), dim3(kThreads), 0,
                           g.stream, g.poses, g.mot_param, h.n_mot_e, h.n_obs_e, g.delta,
                           delta_t, g.lambda, g.mns[0], g.mns[1], g.mns[2], g.mns[3], g.wrap,
                           g.stage, g.frec, g.mt_pos);
    }
    if (h.n_bet_e > 0)
        hipLaunchKernelGGL(between_edges_kernel, dim3(nblk(h.n_bet_e)), dim3(kThreads), 0,
                           g.stream, g.poses, g.e_t1, g.e_t2, g.bet_param, g.bet_w, h.n_bet_e,
                           h.n_obs_e + h.n_mot_e, g.lambda, g.wrap, exact_hessian ? 1 : 0, g.stage,
                           g.frec, g.mt_pos);
    const int64_t nv = 9 * h.n_blocks;
    hipLaunchKernelGGL(gather_blocks_kernel, dim3(nblk(nv)), dim3(kThreads), 0, g.stream,
                       g.cptr, g.clist, g.stage, nv, static_cast<int64_t>(h.bdiag[0]), g.anchor,
                       g.bval);
    const int64_t dim = 3 * static_cast<int64_t>(h.n);
    hipLaunchKernelGGL(gather_xi_kernel, dim3(nblk(dim)), dim3(kThreads), 0, g.stream, g.xptr,
                       g.xlist, g.stage, dim, g.xi, g.poses, g.anchor_prior, g.anchor,
                       g.pose0_ref[0], g.pose0_ref[1], g.pose0_ref[2]);
    GSLAM_HIP(hipGetLastError());
    return GSLAM_OK;
}

int launch_bsr_to_csr(DeviceGraph &g, const double *bval, double *csr_vals)
{
    const int64_t nv = 9 * g.h.n_blocks;
    hipLaunchKernelGGL(bsr_to_csr_kernel, dim3(nblk(nv)), dim3(kThreads), 0, g.stream,
                       g.csr_pos, bval, nv, csr_vals);
    GSLAM_HIP(hipGetLastError());
    return GSLAM_OK;
}

int launch_csr_to_bsr(DeviceGraph &g, const double *csr_vals, double *bval)
{
    const int64_t nv = 9 * g.h.n_blocks;
    hipLaunchKernelGGL(csr_to_bsr_kernel, dim3(nblk(nv)), dim3(kThreads), 0, g.stream,
                       g.csr_pos, csr_vals, nv, bval);
    GSLAM_HIP(hipGetLastError());
    return GSLAM_OK;
}

}  // namespace gslam
