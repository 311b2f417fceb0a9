/*
 * gslam_oracle.c -- CPU FP64 restatement of the reference Graph-SLAM
 * Gauss-Newton step.  TEST INFRASTRUCTURE ONLY (the checker): used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
 * product library.
 *
 * The arithmetic follows the reference C operation by operation (same
 * generic small-matrix products, same summation order, same libm calls) so
 * that, compiled with gcc -O3 on x86-64, it reproduces the reference
 * src/cpu_pcg_fp64 build bit for bit.  tests/test_oracle_vs_ref.py checks
 * exactly that against a build of the reference sources, and
 * tests/test_oracle_golden.py against committed fixtures.
 *
 * Parity status per function (see the tests named above):
 *   orc_log_count/read, orc_obs_pairs, orc_motion_edge, orc_obs_edge,
 *   orc_csr_pattern/values, orc_pcg, orc_block_jacobi, orc_spmv,
 *   orc_gj_inv, orc_gj_inv_f32, orc_dense_values, orc_dense_matvec,
 *   orc_log_gn_run, orc_diff, orc_pose_update, orc_landmark_means: pinned
 *   bitwise to the reference build and fixtures.
 *   orc_between_edge / orc_between_edges / orc_between_gn_step: no reference
 *   counterpart; pinned by central finite differences of the residual
 *   (J), J^T L J structure, and a direct sparse solve (tests).
 *
 * The OpenMP pragmas in the PCG loops take effect only in liboracle_omp.so
 * (the same file built with -fopenmp), a multi-core CPU timing baseline whose
 * dot products are summed in another order: within rounding of liboracle.so
 * (tests/test_oracle_omp.py), not bitwise, and the same bits on every run
 * for a given team size.  liboracle.so ignores them.
 */
#include "gslam_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef _OPENMP
/* Sum over i in [0, n) of the terms that BODY adds to s, [0, n) split into
 * one contiguous range per thread and the per-thread sums added in thread
 * order: the same bits on every run for a given team size (a reduction
 * clause adds them in the order in which the threads finish, which moved the
 * converged PCG iteration count on the victoria-park-100 system over
 * 380-394). */
#define ORC_MAX_THREADS 256
#define ORC_OMP_SUM(out, n, BODY)                                             \
    do {                                                                      \
        double part_[ORC_MAX_THREADS];                                        \
        int nt_ = 1;                                                          \
        int req_ = omp_get_max_threads();                                     \
        if (req_ > ORC_MAX_THREADS) req_ = ORC_MAX_THREADS;                   \
        _Pragma("omp parallel num_threads(req_)")                             \
        {                                                                     \
            const int t_ = omp_get_thread_num(), T_ = omp_get_num_threads();  \
            const int64_t lo_ = (n) * t_ / T_, hi_ = (n) * (t_ + 1) / T_;     \
            double s = 0.0;                                                   \
            for (int64_t i = lo_; i < hi_; i++) { BODY; }                     \
            part_[t_] = s;                                                    \
            if (t_ == 0) nt_ = T_;                                            \
        }                                                                     \
        for (int t_ = 0; t_ < nt_; t_++) (out) += part_[t_];                  \
    } while (0)
#endif

/* Threads the PCG loops use: 1 in liboracle.so, the OpenMP team size in
 * liboracle_omp.so. */
int32_t orc_threads(void)
{
#ifdef _OPENMP
    return (int32_t)omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Generic small dense helpers with the reference's loop/summation order.
 * transpose: MatrixOps.c:18-27; product: MatrixOps.c:123-159 (each output is
 * a sequential dot starting at 0.0); scale: :93-99; add/sub: :101-121.      */
static void m_tra(double *res, const double *mat, int my, int mx)
{
    for (int m = 0; m < mx; m++)
        for (int n = 0; n < my; n++)
            res[m * my + n] = mat[n * mx + m];
}

static void m_dot(double *res, const double *a, int ay, int ax,
                  const double *b, int bx)
{
    for (int i = 0; i < ay * bx; i++) {
        double s = 0.0;
        for (int j = 0; j < ax; j++)
            s += a[(i / bx) * ax + j] * b[j * bx + i % bx];
        res[i] = s;
    }
}

static void m_mul(double *res, const double *mat, int n, double k)
{
    for (int i = 0; i < n; i++) res[i] = mat[i] * k;
}

static void m_add(double *res, const double *a, const double *b, int n)
{
    for (int i = 0; i < n; i++) res[i] = a[i] + b[i];
}

static void m_sub(double *res, const double *a, const double *b, int n)
{
    for (int i = 0; i < n; i++) res[i] = a[i] - b[i];
}

/* Unpivoted augmented Gauss-Jordan, src/cpu_pcg_fp64/MatrixOps.c:29-91. */
static int gj_inv_core(int32_t n, double *t, double *r)
{
    for (int64_t m = 0; m < n; m++)
        for (int64_t k = 0; k < n; k++)
            r[m * n + k] = (m == k) ? 1.0 : 0.0;
    for (int64_t i = 0; i < n; i++) {
        if (t[i * n + i] == 0.0) {
            memset(r, 0, sizeof(double) * (size_t)n * (size_t)n);
            return 1;
        }
        for (int64_t j = 0; j < n; j++) {
            if (i == j) continue;
            double ratio = t[j * n + i] / t[i * n + i];
            for (int64_t k = 0; k < n; k++) {
                t[j * n + k] = t[j * n + k] - ratio * t[i * n + k];
                r[j * n + k] = r[j * n + k] - ratio * r[i * n + k];
            }
        }
    }
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++)
            r[i * n + j] = r[i * n + j] / t[i * n + i];
    return 0;
}

int orc_gj_inv(int32_t n, const double *a, double *a_inv)
{
    double *t = (double *)malloc(sizeof(double) * (size_t)n * (size_t)n);
    if (!t) return -1;
    memcpy(t, a, sizeof(double) * (size_t)n * (size_t)n);
    int rc = gj_inv_core(n, t, a_inv);
    free(t);
    return rc;
}

/* src/cpu_gj_mixed/MatrixOps.c:174-212 */
int orc_gj_inv_f32(int32_t n, const float *a, float *r)
{
    float *t = (float *)malloc(sizeof(float) * (size_t)n * (size_t)n);
    if (!t) return -1;
    for (int64_t m = 0; m < n; m++)
        for (int64_t k = 0; k < n; k++) {
            t[m * n + k] = a[m * n + k];
            r[m * n + k] = (m == k) ? 1.0f : 0.0f;
        }
    for (int64_t i = 0; i < n; i++) {
        if (t[i * n + i] == 0.0f) {
            memset(r, 0, sizeof(float) * (size_t)n * (size_t)n);
            free(t);
            return 1;
        }
        for (int64_t j = 0; j < n; j++) {
            if (i == j) continue;
            float ratio = t[j * n + i] / t[i * n + i];
            for (int64_t k = 0; k < n; k++) {
                t[j * n + k] -= ratio * t[i * n + k];
                r[j * n + k] -= ratio * r[i * n + k];
            }
        }
    }
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++)
            r[i * n + j] /= t[i * n + i];
    free(t);
    return 0;
}

void orc_dense_matvec(int32_t n, const double *a, const double *x, double *y)
{
    for (int64_t i = 0; i < n; i++) {
        double s = 0.0;
        for (int64_t j = 0; j < n; j++) s += a[i * n + j] * x[j];
        y[i] = s;
    }
}

/* ------------------------------------------------------------------------ */
/* .log.txt reader: src/cpu_pcg_fp64/GSLAM.c:415-493.  Tokens split on ' '
 * (strtok), integers via atoi, reals via atof; any first token other than
 * u/x/z sets delta from the second token (GSLAM.c:480-483).                 */
static int log_scan(const char *path, double *delta, int32_t *n_u,
                    int32_t *n_x, int32_t *n_z, int32_t *u_step, double *u_nu,
                    double *u_omega, int32_t *x_step, double *x_pose,
                    int32_t *z_step, int32_t *z_lm, double *z_range,
                    double *z_bearing)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char line[1024];
    int32_t nu = 0, nx = 0, nz = 0;
    int rc = 0;
    while (fgets(line, sizeof(line), f) != NULL) {
        char *tok = strtok(line, " ");
        if (!tok) { rc = -2; break; }
        char *t[4];
        int need = 0;
        if (strcmp(tok, "u") == 0) need = 3;
        else if (strcmp(tok, "x") == 0) need = 4;
        else if (strcmp(tok, "z") == 0) need = 4;
        else need = 1;
        for (int i = 0; i < need; i++) {
            t[i] = strtok(NULL, " ");
            if (!t[i]) { rc = -2; break; }
        }
        if (rc) break;
        if (strcmp(tok, "u") == 0) {
            if (u_step) {
                u_step[nu] = (int32_t)(unsigned int)atoi(t[0]);
                u_nu[nu] = atof(t[1]);
                u_omega[nu] = atof(t[2]);
            }
            nu++;
        } else if (strcmp(tok, "x") == 0) {
            if (x_step) {
                x_step[nx] = (int32_t)(unsigned int)atoi(t[0]);
                x_pose[3 * nx + 0] = atof(t[1]);
                x_pose[3 * nx + 1] = atof(t[2]);
                x_pose[3 * nx + 2] = atof(t[3]);
            }
            nx++;
        } else if (strcmp(tok, "z") == 0) {
            if (z_step) {
                z_step[nz] = (int32_t)(unsigned int)atoi(t[0]);
                z_lm[nz] = (int32_t)(unsigned int)atoi(t[1]);
                z_range[nz] = atof(t[2]);
                z_bearing[nz] = atof(t[3]);
            }
            nz++;
        } else {
            if (delta) *delta = atof(t[0]);
        }
    }
    fclose(f);
    if (n_u) *n_u = nu;
    if (n_x) *n_x = nx;
    if (n_z) *n_z = nz;
    return rc;
}

int orc_log_count(const char *path, int32_t *n_u, int32_t *n_x, int32_t *n_z)
{
    return log_scan(path, NULL, n_u, n_x, n_z, NULL, NULL, NULL, NULL, NULL,
                    NULL, NULL, NULL, NULL);
}

int orc_log_read(const char *path, double *delta, int32_t *u_step,
                 double *u_nu, double *u_omega, int32_t *x_step,
                 double *x_pose, int32_t *z_step, int32_t *z_lm,
                 double *z_range, double *z_bearing)
{
    return log_scan(path, delta, NULL, NULL, NULL, u_step, u_nu, u_omega,
                    x_step, x_pose, z_step, z_lm, z_range, z_bearing);
}

/* ------------------------------------------------------------------------ */
/* Landmark buckets and pairs j<k.  Order 0 restates the first-appearance id
 * list of GSLAM.c:246-261 and the bucket fill of :267-276; order 1 the
 * ascending-id buckets of device_graph.cu:505-551.  Within a bucket the
 * observations keep file order in both.                                      */
static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

int64_t orc_obs_pairs(int32_t n_obs, const int32_t *z_lm, int32_t order,
                      int32_t *pair_a, int32_t *pair_b)
{
    if (n_obs <= 0) return 0;
    /* distinct landmark ids in the requested order */
    int32_t *ids = (int32_t *)malloc(sizeof(int32_t) * (size_t)n_obs);
    int32_t n_ids = 0;
    if (order == 1) {
        memcpy(ids, z_lm, sizeof(int32_t) * (size_t)n_obs);
        qsort(ids, (size_t)n_obs, sizeof(int32_t), cmp_i32);
        for (int32_t i = 0; i < n_obs; i++)
            if (i == 0 || ids[i] != ids[n_ids - 1]) ids[n_ids++] = ids[i];
    } else {
        /* first appearance: sort (id, first index) pairs */
        int32_t *sorted = (int32_t *)malloc(sizeof(int32_t) * (size_t)n_obs);
        memcpy(sorted, z_lm, sizeof(int32_t) * (size_t)n_obs);
        qsort(sorted, (size_t)n_obs, sizeof(int32_t), cmp_i32);
        int32_t nd = 0;
        for (int32_t i = 0; i < n_obs; i++)
            if (i == 0 || sorted[i] != sorted[nd - 1]) sorted[nd++] = sorted[i];
        char *seen = (char *)calloc((size_t)nd, 1);
        for (int32_t i = 0; i < n_obs; i++) {
            int32_t *p = (int32_t *)bsearch(&z_lm[i], sorted, (size_t)nd,
                                            sizeof(int32_t), cmp_i32);
            size_t pos = (size_t)(p - sorted);
            if (!seen[pos]) { seen[pos] = 1; ids[n_ids++] = z_lm[i]; }
        }
        free(seen);
        free(sorted);
    }
    /* bucket members in file order, buckets in id order */
    int32_t *sorted_ids = (int32_t *)malloc(sizeof(int32_t) * (size_t)n_ids);
    memcpy(sorted_ids, ids, sizeof(int32_t) * (size_t)n_ids);
    qsort(sorted_ids, (size_t)n_ids, sizeof(int32_t), cmp_i32);
    int32_t *cnt = (int32_t *)calloc((size_t)n_ids + 1, sizeof(int32_t));
    int32_t *slot = (int32_t *)malloc(sizeof(int32_t) * (size_t)n_obs);
    for (int32_t i = 0; i < n_obs; i++) {
        int32_t *p = (int32_t *)bsearch(&z_lm[i], sorted_ids, (size_t)n_ids,
                                        sizeof(int32_t), cmp_i32);
        slot[i] = (int32_t)(p - sorted_ids);
        cnt[slot[i] + 1]++;
    }
    for (int32_t i = 0; i < n_ids; i++) cnt[i + 1] += cnt[i];
    int32_t *members = (int32_t *)malloc(sizeof(int32_t) * (size_t)n_obs);
    int32_t *fill = (int32_t *)calloc((size_t)n_ids, sizeof(int32_t));
    for (int32_t i = 0; i < n_obs; i++)
        members[cnt[slot[i]] + fill[slot[i]]++] = i;
    int64_t e = 0;
    for (int32_t b = 0; b < n_ids; b++) {
        int32_t *p = (int32_t *)bsearch(&ids[b], sorted_ids, (size_t)n_ids,
                                        sizeof(int32_t), cmp_i32);
        int32_t s = (int32_t)(p - sorted_ids);
        int32_t lo = cnt[s], n = cnt[s + 1] - cnt[s];
        for (int32_t j = 0; j < n; j++)
            for (int32_t k = j + 1; k < n; k++) {
                if (pair_a) {
                    pair_a[e] = members[lo + j];
                    pair_b[e] = members[lo + k];
                }
                e++;
            }
    }
    free(ids); free(sorted_ids); free(cnt); free(slot); free(members);
    free(fill);
    return e;
}

/* ------------------------------------------------------------------------ */
/* MotionEdge_create, src/cpu_pcg_fp64/MotionEdge.c:9-193.                    */
void orc_motion_edge(const double x1[3], const double x2[3], double nu,
                     double omega, double delta, double lambda,
                     const double mns[4], double out[ORC_EDGE_VALS],
                     double *omega_out)
{
    if (fabs(omega) < 1e-5) omega = 1e-5;                 /* :21-22 */
    double M[4] = {                                        /* :27-30 */
        mns[0] * mns[0] * fabs(nu) / delta + mns[1] * mns[1] * fabs(omega) / delta, 0.0,
        0.0, mns[2] * mns[2] * fabs(nu) / delta + mns[3] * mns[3] * fabs(omega) / delta};
    double th = x1[2];
    double st = sin(th), ct = cos(th);                     /* :37-41 */
    double stw = sin(th + omega * delta), ctw = cos(th + omega * delta);
    double A[6] = {                                        /* :54-58 */
        stw / omega - st / omega,
        nu / omega * delta * ctw - nu / omega / omega * stw + nu / omega / omega * st,
        ct / omega - ctw / omega,
        nu / omega * delta * stw - nu / omega / omega * ct + nu / omega / omega * ctw,
        0.0, delta};
    double F[9] = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0};
    F[2] = nu / omega * cos(th + omega * delta) - nu / omega * cos(th); /* :73 */
    F[5] = nu / omega * sin(th + omega * delta) - nu / omega * sin(th); /* :74 */
    const double I3[9] = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0};

    double tA[6], AM[6], AMAt[9], reg[9], S[9], Om[9];
    m_tra(tA, A, 3, 2);                                    /* :82-89 */
    m_dot(AM, A, 3, 2, M, 2);
    m_dot(AMAt, AM, 3, 2, tA, 3);
    m_mul(reg, I3, 9, 0.0001);                             /* :91-95 */
    m_add(S, AMAt, reg, 9);
    double tmp[9];
    memcpy(tmp, S, sizeof(tmp));
    gj_inv_core(3, tmp, Om);                               /* :97-102 */
    if (omega_out) memcpy(omega_out, Om, sizeof(Om));

    double tF[9], FtOm[9], FtOmF[9], OmF[9];
    m_tra(tF, F, 3, 3);                                    /* :104-118 */
    m_dot(FtOm, tF, 3, 3, Om, 3);
    m_dot(FtOmF, FtOm, 3, 3, F, 3);
    m_mul(out + 0, FtOmF, 9, lambda);                      /* UL */
    m_mul(out + 9, FtOm, 9, -lambda);                      /* UR :120-125 */
    m_dot(OmF, Om, 3, 3, F, 3);                            /* :127-135 */
    m_mul(out + 18, OmF, 9, -lambda);                      /* BL */
    m_mul(out + 27, Om, 9, lambda);                        /* BR :137-142 */

    /* state_transition, src/cpu_pcg_fp64/GSLAM.c:373-401: its time argument
     * is unsigned int (GSLAM.h), so delta is truncated here only. */
    unsigned int time = (unsigned int)delta;
    double pred[3];
    if (fabs(omega) < 1e-10) {
        double v[3] = {nu * cos(x1[2]), nu * sin(x1[2]), omega};
        double mv[3];
        m_mul(mv, v, 3, (double)time);
        m_add(pred, x1, mv, 3);
    } else {
        double v[3] = {
            nu / omega * sin(x1[2] + omega * time) - nu / omega * sin(x1[2]),
            nu / omega * cos(x1[2]) - nu / omega * cos(x1[2] + omega * time),
            omega * time};
        m_add(pred, x1, v, 3);
    }
    double e[3], FtOme[3], Ome[3];
    m_sub(e, x2, pred, 3);                                 /* :147-148 */
    m_dot(FtOme, FtOm, 3, 3, e, 1);                        /* :150-158 */
    m_mul(out + 36, FtOme, 3, lambda);
    m_dot(Ome, Om, 3, 3, e, 1);                            /* :160-175 */
    m_mul(out + 39, Ome, 3, -lambda);
}

/* ------------------------------------------------------------------------ */
/* ObsEdge_create, src/cpu_pcg_fp64/ObsEdge.c:10-212.                         */
void orc_obs_edge(const double x1[3], const double x2[3], double r1,
                  double phi1, double r2, double phi2, const double snr[2],
                  double out[ORC_EDGE_VALS], double *hat_e_out, double *om_out)
{
    double s1 = sin(x1[2] + phi1), c1 = cos(x1[2] + phi1);   /* :20-23 */
    double s2 = sin(x2[2] + phi2), c2 = cos(x2[2] + phi2);
    double e[2] = {x2[0] - x1[0] + r2 * c2 - r1 * c1,        /* :25-28 */
                   x2[1] - x1[1] + r2 * s2 - r1 * s1};
    if (hat_e_out) { hat_e_out[0] = e[0]; hat_e_out[1] = e[1]; }
    double q1 = r1 * snr[0], q2 = r2 * snr[0];
    double Q1[4] = {q1 * q1, 0.0, 0.0, snr[1] * snr[1]};      /* :34-37 */
    double R1[4] = {-c1, r1 * s1, -s1, -r1 * c1};             /* :39-42 */
    double Q2[4] = {q2 * q2, 0.0, 0.0, snr[1] * snr[1]};      /* :44-47 */
    double R2[4] = {c2, -r2 * s2, s2, r2 * c2};               /* :49-52 */
    double tR1[4], R1Q1[4], S1[4], tR2[4], R2Q2[4], S2[4], Sig[4], Om[4];
    m_tra(tR1, R1, 2, 2);                                     /* :54-61 */
    m_dot(R1Q1, R1, 2, 2, Q1, 2);
    m_dot(S1, R1Q1, 2, 2, tR1, 2);
    m_tra(tR2, R2, 2, 2);                                     /* :63-70 */
    m_dot(R2Q2, R2, 2, 2, Q2, 2);
    m_dot(S2, R2Q2, 2, 2, tR2, 2);
    m_add(Sig, S1, S2, 4);                                    /* :72-73 */
    gj_inv_core(2, Sig, Om);                                  /* :83 */
    if (om_out) memcpy(om_out, Om, sizeof(Om));
    double B1[6] = {-1.0, -0.0, r1 * s1, -0.0, -1.0, -r1 * c1}; /* :91-94 */
    double B2[6] = {1.0, 0.0, -r2 * s2, 0.0, 1.0, r2 * c2};    /* :102-105 */
    double tB1[6], tB2[6], tB1Om[6], tB2Om[6];
    m_tra(tB1, B1, 2, 3);                                     /* :113-123 */
    m_tra(tB2, B2, 2, 3);
    m_dot(tB1Om, tB1, 3, 2, Om, 2);
    m_dot(tB2Om, tB2, 3, 2, Om, 2);
    m_dot(out + 0, tB1Om, 3, 2, B1, 3);                       /* UL :125 */
    m_dot(out + 9, tB1Om, 3, 2, B2, 3);                       /* UR :137 */
    m_dot(out + 18, tB2Om, 3, 2, B1, 3);                      /* BL :149 */
    m_dot(out + 27, tB2Om, 3, 2, B2, 3);                      /* BR :161 */
    double v1[3], v2[3];
    m_dot(v1, tB1Om, 3, 2, e, 1);                             /* :173-189 */
    m_dot(v2, tB2Om, 3, 2, e, 1);
    for (int i = 0; i < 3; i++) {
        out[36 + i] = -v1[i];
        out[39 + i] = -v2[i];
    }
}

/* ------------------------------------------------------------------------ */
/* SE(2) between edge.  No reference counterpart.                             */
static double wrap_pi(double a)
{
    const double two_pi = 6.283185307179586476925286766559;
    return a - two_pi * floor((a + 3.1415926535897932384626433832795) / two_pi);
}

void orc_between_edge(const double xi[3], const double xj[3],
                      const double z[3], const double info[3], double weight,
                      int32_t wrap, double out[ORC_EDGE_VALS], double *e_out)
{
    double c = cos(xi[2]), s = sin(xi[2]);
    double dx = xj[0] - xi[0], dy = xj[1] - xi[1];
    double e[3];
    e[0] = c * dx + s * dy - z[0];
    e[1] = -s * dx + c * dy - z[1];
    e[2] = xj[2] - xi[2] - z[2];
    if (wrap) e[2] = wrap_pi(e[2]);
    if (e_out) { e_out[0] = e[0]; e_out[1] = e[1]; e_out[2] = e[2]; }
    double L[3] = {info[0] * weight, info[1] * weight, info[2] * weight};
    double Ji[9] = {-c, -s, -s * dx + c * dy,
                    s, -c, -c * dx - s * dy,
                    0.0, 0.0, -1.0};
    double Jj[9] = {c, s, 0.0, -s, c, 0.0, 0.0, 0.0, 1.0};
    const double *J[2] = {Ji, Jj};
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            double *blk = out + 9 * (2 * a + b);
            for (int r = 0; r < 3; r++)
                for (int cc = 0; cc < 3; cc++) {
                    double acc = 0.0;
                    for (int k = 0; k < 3; k++)
                        acc += J[a][k * 3 + r] * L[k] * J[b][k * 3 + cc];
                    blk[r * 3 + cc] = acc;
                }
        }
    for (int a = 0; a < 2; a++)
        for (int r = 0; r < 3; r++) {
            double acc = 0.0;
            for (int k = 0; k < 3; k++) acc += J[a][k * 3 + r] * L[k] * e[k];
            out[36 + 3 * a + r] = -acc;
        }
}

void orc_between_edges(int32_t n_poses, const double *poses, int64_t n_edges,
                       const int32_t *ei, const int32_t *ej,
                       const double *meas, const double *info,
                       const double *weight, double lambda, int32_t wrap,
                       double *edge_vals)
{
    (void)n_poses;
    for (int64_t e = 0; e < n_edges; e++) {
        double w = lambda * (weight ? weight[e] : 1.0);
        orc_between_edge(poses + 3 * (int64_t)ei[e], poses + 3 * (int64_t)ej[e],
                         meas + 3 * e, info + 3 * e, w, wrap,
                         edge_vals + ORC_EDGE_VALS * e, NULL);
    }
}

/* ------------------------------------------------------------------------ */
/* CSR pattern and values.  Same pattern and summation order as
 * build_csr_from_edges, src/cpu_pcg_fp64/MatrixOps.c:173-273, built from a
 * per-row column sort instead of the dim x dim marker (:177).                */
int64_t orc_csr_pattern(int32_t n_poses, int64_t n_edges, const int32_t *t1,
                        const int32_t *t2, int32_t *row_ptr, int32_t *col_idx)
{
    int64_t N = n_poses;
    int64_t *deg = (int64_t *)calloc((size_t)N + 1, sizeof(int64_t));
    for (int64_t e = 0; e < n_edges; e++) {
        deg[t1[e] + 1] += 2;
        deg[t2[e] + 1] += 2;
    }
    for (int64_t i = 0; i < N; i++) deg[i + 1] += deg[i];
    int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * (size_t)(deg[N] + 1));
    int64_t *fill = (int64_t *)calloc((size_t)N, sizeof(int64_t));
    for (int64_t e = 0; e < n_edges; e++) {
        int32_t a = t1[e], b = t2[e];
        nb[deg[a] + fill[a]++] = a;
        nb[deg[a] + fill[a]++] = b;
        nb[deg[b] + fill[b]++] = a;
        nb[deg[b] + fill[b]++] = b;
    }
    /* unique sorted block columns per block row */
    int64_t *bcnt = (int64_t *)calloc((size_t)N, sizeof(int64_t));
    for (int64_t i = 0; i < N; i++) {
        int32_t *p = nb + deg[i];
        int64_t n = deg[i + 1] - deg[i];
        qsort(p, (size_t)n, sizeof(int32_t), cmp_i32);
        int64_t u = 0;
        for (int64_t k = 0; k < n; k++)
            if (k == 0 || p[k] != p[u - 1]) p[u++] = p[k];
        bcnt[i] = u;
    }
    int anchor_only = (N > 0 && bcnt[0] == 0); /* pose 0 touched by no edge */
    int64_t nnz = 0;
    if (row_ptr) row_ptr[0] = 0;
    for (int64_t i = 0; i < N; i++)
        for (int a = 0; a < 3; a++) {
            int64_t row = 3 * i + a;
            int64_t n = 3 * bcnt[i];
            if (i == 0 && anchor_only) n = 1;
            if (col_idx) {
                int64_t pos = nnz;
                if (i == 0 && anchor_only) {
                    col_idx[pos] = a;
                } else {
                    for (int64_t k = 0; k < bcnt[i]; k++)
                        for (int c = 0; c < 3; c++)
                            col_idx[pos++] = 3 * nb[deg[i] + k] + c;
                }
            }
            nnz += n;
            if (row_ptr) row_ptr[row + 1] = (int32_t)nnz;
        }
    free(deg); free(nb); free(fill); free(bcnt);
    return nnz;
}

static int64_t find_slot(const int32_t *row_ptr, const int32_t *col_idx,
                         int64_t row, int64_t col)
{
    int64_t lo = row_ptr[row], hi = row_ptr[row + 1] - 1;
    while (lo <= hi) {
        int64_t mid = (lo + hi) >> 1;
        if (col_idx[mid] == col) return mid;
        if (col_idx[mid] < col) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

int orc_csr_values(int32_t n_poses, const int32_t *row_ptr,
                   const int32_t *col_idx, int64_t n_edges, const int32_t *t1,
                   const int32_t *t2, const double *edge_vals, double anchor,
                   double *values, double *xi)
{
    int64_t dim = 3 * (int64_t)n_poses;
    memset(values, 0, sizeof(double) * (size_t)row_ptr[dim]);
    memset(xi, 0, sizeof(double) * (size_t)dim);
    for (int d = 0; d < 3 && d < dim; d++) {               /* :223-231 */
        int64_t s = find_slot(row_ptr, col_idx, d, d);
        if (s >= 0) values[s] += anchor;
    }
    for (int64_t e = 0; e < n_edges; e++) {                /* :233-272 */
        int64_t f1 = 3 * (int64_t)t1[e], f2 = 3 * (int64_t)t2[e];
        const double *v = edge_vals + ORC_EDGE_VALS * e;
        for (int r = 0; r < 3; r++) {
            for (int c = 0; c < 3; c++) {
                int64_t s;
                s = find_slot(row_ptr, col_idx, f1 + r, f1 + c);
                if (s < 0) return -1;
                values[s] += v[0 + r * 3 + c];
                s = find_slot(row_ptr, col_idx, f1 + r, f2 + c);
                if (s < 0) return -1;
                values[s] += v[9 + r * 3 + c];
                s = find_slot(row_ptr, col_idx, f2 + r, f1 + c);
                if (s < 0) return -1;
                values[s] += v[18 + r * 3 + c];
                s = find_slot(row_ptr, col_idx, f2 + r, f2 + c);
                if (s < 0) return -1;
                values[s] += v[27 + r * 3 + c];
            }
            xi[f1 + r] += v[36 + r];
            xi[f2 + r] += v[39 + r];
        }
    }
    return 0;
}

/* Dense: anchor first (src/gpu_gj_fp64/GSLAM.cu:315-328), then add_edge,
 * src/cpu_pcg_fp64/GSLAM.c:320-371 (UL, UR, BL, BR, xi_upper, xi_bottom).   */
int orc_dense_values(int32_t n_poses, int64_t n_edges, const int32_t *t1,
                     const int32_t *t2, const double *edge_vals,
                     double anchor, double *omega, double *xi)
{
    int64_t dim = 3 * (int64_t)n_poses;
    memset(omega, 0, sizeof(double) * (size_t)(dim * dim));
    memset(xi, 0, sizeof(double) * (size_t)dim);
    for (int64_t d = 0; d < 3 && d < dim; d++) omega[d * dim + d] = anchor;
    for (int64_t e = 0; e < n_edges; e++) {
        int64_t f1 = 3 * (int64_t)t1[e], f2 = 3 * (int64_t)t2[e];
        const double *v = edge_vals + ORC_EDGE_VALS * e;
        for (int q = 0; q < 4; q++) {
            int64_t ro = (q < 2) ? f1 : f2, co = (q % 2 == 0) ? f1 : f2;
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++)
                    omega[(ro + r) * dim + co + c] += v[9 * q + r * 3 + c];
        }
        for (int r = 0; r < 3; r++) xi[f1 + r] += v[36 + r];
        for (int r = 0; r < 3; r++) xi[f2 + r] += v[39 + r];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Block-Jacobi PCG, src/cpu_pcg_fp64/MatrixOps.c:281-391.                    */
void orc_block_jacobi(int32_t dim, const int32_t *row_ptr,
                      const int32_t *col_idx, const double *values,
                      double *diag_inv)
{
    int64_t nb = dim / 3;
    for (int64_t blk = 0; blk < nb; blk++) {               /* :287-315 */
        int64_t base = blk * 3;
        double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int r = 0; r < 3; r++)
            for (int64_t p = row_ptr[base + r]; p < row_ptr[base + r + 1]; p++) {
                int64_t c = col_idx[p] - base;
                if (c >= 0 && c < 3) a[r * 3 + c] = values[p];
            }
        double det = a[0] * (a[4] * a[8] - a[5] * a[7])
                   - a[1] * (a[3] * a[8] - a[5] * a[6])
                   + a[2] * (a[3] * a[7] - a[4] * a[6]);
        if (det == 0.0) det = 1e-30;
        double *inv = diag_inv + blk * 9;
        inv[0] = (a[4] * a[8] - a[5] * a[7]) / det;
        inv[1] = (a[2] * a[7] - a[1] * a[8]) / det;
        inv[2] = (a[1] * a[5] - a[2] * a[4]) / det;
        inv[3] = (a[5] * a[6] - a[3] * a[8]) / det;
        inv[4] = (a[0] * a[8] - a[2] * a[6]) / det;
        inv[5] = (a[2] * a[3] - a[0] * a[5]) / det;
        inv[6] = (a[3] * a[7] - a[4] * a[6]) / det;
        inv[7] = (a[1] * a[6] - a[0] * a[7]) / det;
        inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
    }
}

void orc_spmv(int32_t dim, const int32_t *row_ptr, const int32_t *col_idx,
              const double *values, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < dim; i++) {
        y[i] = 0.0;
        for (int64_t j = row_ptr[i]; j < row_ptr[i + 1]; j++)
            y[i] += values[j] * x[col_idx[j]];
    }
}

static void apply_bj(int64_t dim, const double *diag_inv, const double *r,
                     double *z)
{
    int64_t nb = dim / 3;
#pragma omp parallel for schedule(static)
    for (int64_t blk = 0; blk < nb; blk++) {
        int64_t base = blk * 3;
        const double *inv = diag_inv + blk * 9;
        for (int rr = 0; rr < 3; rr++) {
            z[base + rr] = 0.0;
            for (int c = 0; c < 3; c++)
                z[base + rr] += inv[rr * 3 + c] * r[base + c];
        }
    }
}

int32_t orc_pcg(int32_t dim, const int32_t *row_ptr, const int32_t *col_idx,
                const double *values, const double *b, double tol_rel,
                int32_t max_iter, double *x, double *res_hist)
{
    int64_t n = dim;
    double *diag_inv = (double *)malloc(sizeof(double) * (size_t)(n / 3 * 9 + 9));
    double *r = (double *)calloc((size_t)n + 1, sizeof(double));
    double *z = (double *)calloc((size_t)n + 1, sizeof(double));
    double *p = (double *)calloc((size_t)n + 1, sizeof(double));
    double *q = (double *)calloc((size_t)n + 1, sizeof(double));
    orc_block_jacobi(dim, row_ptr, col_idx, values, diag_inv);
    for (int64_t i = 0; i < n; i++) { x[i] = 0.0; r[i] = b[i]; }   /* :323 */
    apply_bj(n, diag_inv, r, z);                                    /* :326 */
    for (int64_t i = 0; i < n; i++) p[i] = z[i];
    double rz_old = 0.0;
    for (int64_t i = 0; i < n; i++) rz_old += r[i] * z[i];
    double b_norm = 0.0;
    for (int64_t i = 0; i < n; i++) b_norm += b[i] * b[i];
    b_norm = sqrt(b_norm);
    double tol = tol_rel * b_norm;                                  /* :345 */
    int32_t it = 0;
    for (int32_t k = 0; k < max_iter; k++) {                        /* :347 */
        orc_spmv(dim, row_ptr, col_idx, values, p, q);
        double pq = 0.0;
#ifdef _OPENMP
        ORC_OMP_SUM(pq, n, s += p[i] * q[i]);
#else
        for (int64_t i = 0; i < n; i++) pq += p[i] * q[i];
#endif
        if (pq == 0.0) break;                                       /* :357 */
        double alpha = rz_old / pq;
        double r_norm = 0.0;
#ifdef _OPENMP
        ORC_OMP_SUM(r_norm, n, x[i] += alpha * p[i]; r[i] -= alpha * q[i]; s += r[i] * r[i]);
#else
        for (int64_t i = 0; i < n; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            r_norm += r[i] * r[i];
        }
#endif
        it = k + 1;
        if (res_hist) res_hist[k] = sqrt(r_norm);
        if (sqrt(r_norm) < tol) break;                              /* :366 */
        apply_bj(n, diag_inv, r, z);
        double rz_new = 0.0;
#ifdef _OPENMP
        ORC_OMP_SUM(rz_new, n, s += r[i] * z[i]);
#else
        for (int64_t i = 0; i < n; i++) rz_new += r[i] * z[i];
#endif
        double beta = rz_new / rz_old;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        rz_old = rz_new;
    }
    free(diag_inv); free(r); free(z); free(p); free(q);
    return it;
}

/* ------------------------------------------------------------------------ */
double orc_diff(int32_t dim, const double *dx)
{
    double s = 0.0;                                 /* MatrixOps.c:4-16 */
    for (int64_t i = 0; i < dim; i++) s += dx[i] * dx[i];
    return sqrt(s) / sqrt((double)dim);             /* GSLAM.c:104-105 */
}

void orc_pose_update(int32_t n_poses, double *poses, const double *dx,
                     int32_t wrap)
{
    for (int64_t i = 0; i < n_poses; i++) {         /* GSLAM.c:98-102 */
        poses[3 * i + 0] = poses[3 * i + 0] + dx[3 * i + 0];
        poses[3 * i + 1] = poses[3 * i + 1] + dx[3 * i + 1];
        poses[3 * i + 2] = poses[3 * i + 2] + dx[3 * i + 2];
        if (wrap) poses[3 * i + 2] = wrap_pi(poses[3 * i + 2]);
    }
}

int64_t orc_log_edges(int32_t n_poses, const double *poses,
                      const double *u_nu, const double *u_omega, double delta,
                      double lambda, const double mns[4], int32_t n_obs,
                      const int32_t *z_step, const int32_t *z_lm,
                      const double *z_range, const double *z_bearing,
                      int32_t order, int32_t *t1, int32_t *t2,
                      double *edge_vals)
{
    int64_t n_pairs = orc_obs_pairs(n_obs, z_lm, order, NULL, NULL);
    int64_t n_motion = n_poses > 0 ? n_poses - 1 : 0;
    if (!t1) return n_pairs + n_motion;
    int32_t *pa = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_pairs + 1));
    int32_t *pb = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_pairs + 1));
    orc_obs_pairs(n_obs, z_lm, order, pa, pb);
    const double snr[2] = {0.14, 0.05};                       /* GSLAM.c:301 */
    for (int64_t e = 0; e < n_pairs; e++) {
        int32_t a = pa[e], b = pb[e];
        t1[e] = z_step[a];
        t2[e] = z_step[b];
        orc_obs_edge(poses + 3 * (int64_t)z_step[a], poses + 3 * (int64_t)z_step[b],
                     z_range[a], z_bearing[a], z_range[b], z_bearing[b], snr,
                     edge_vals + ORC_EDGE_VALS * e, NULL, NULL);
    }
    for (int64_t i = 0; i < n_motion; i++) {                  /* GSLAM.c:68-70 */
        int64_t e = n_pairs + i;
        t1[e] = (int32_t)i;
        t2[e] = (int32_t)(i + 1);
        orc_motion_edge(poses + 3 * i, poses + 3 * (i + 1), u_nu[i + 1],
                        u_omega[i + 1], delta, lambda, mns,
                        edge_vals + ORC_EDGE_VALS * e, NULL);
    }
    free(pa); free(pb);
    return n_pairs + n_motion;
}

int32_t orc_log_gn_run(int32_t n_poses, double *poses, const double *u_nu,
                       const double *u_omega, double delta, double lambda,
                       int32_t n_obs, const int32_t *z_step,
                       const int32_t *z_lm, const double *z_range,
                       const double *z_bearing, int32_t num_rounds,
                       double tol_rel, int32_t max_iter, double *diffs)
{
    const double mns[4] = {0.19, 0.001, 0.13, 0.2};           /* GSLAM.c:66 */
    int64_t E = orc_log_edges(n_poses, poses, u_nu, u_omega, delta, lambda,
                              mns, n_obs, z_step, z_lm, z_range, z_bearing, 0,
                              NULL, NULL, NULL);
    int32_t *t1 = (int32_t *)malloc(sizeof(int32_t) * (size_t)(E + 1));
    int32_t *t2 = (int32_t *)malloc(sizeof(int32_t) * (size_t)(E + 1));
    double *ev = (double *)malloc(sizeof(double) * ORC_EDGE_VALS * (size_t)(E + 1));
    int64_t dim = 3 * (int64_t)n_poses;
    int32_t *row_ptr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(dim + 1));
    double *xi = (double *)malloc(sizeof(double) * (size_t)dim);
    double *dx = (double *)malloc(sizeof(double) * (size_t)dim);
    orc_log_edges(n_poses, poses, u_nu, u_omega, delta, lambda, mns, n_obs,
                  z_step, z_lm, z_range, z_bearing, 0, t1, t2, ev);
    int64_t nnz = orc_csr_pattern(n_poses, E, t1, t2, row_ptr, NULL);
    int32_t *col_idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz + 1));
    double *values = (double *)malloc(sizeof(double) * (size_t)(nnz + 1));
    orc_csr_pattern(n_poses, E, t1, t2, row_ptr, col_idx);
    double best = 1e30;
    int32_t no_improve = 0, n_exec = 0;
    for (int32_t n = 1; n <= num_rounds; n++) {               /* GSLAM.c:55 */
        if (n > 1)
            orc_log_edges(n_poses, poses, u_nu, u_omega, delta, lambda, mns,
                          n_obs, z_step, z_lm, z_range, z_bearing, 0, t1, t2,
                          ev);
        orc_csr_values(n_poses, row_ptr, col_idx, E, t1, t2, ev, 1000000.0,
                       values, xi);
        orc_pcg((int32_t)dim, row_ptr, col_idx, values, xi, tol_rel, max_iter,
                dx, NULL);
        orc_pose_update(n_poses, poses, dx, 0);
        double diff = orc_diff((int32_t)dim, dx);
        if (diffs) diffs[n - 1] = diff;
        n_exec = n;
        if (diff < 0.01) break;                               /* :119-130 */
        if (diff < best) { best = diff; no_improve = 0; }
        else if (++no_improve >= 20) break;
    }
    free(t1); free(t2); free(ev); free(row_ptr); free(xi); free(dx);
    free(col_idx); free(values);
    return n_exec;
}

int32_t orc_between_gn_step(int32_t n_poses, double *poses, int64_t n_edges,
                            const int32_t *ei, const int32_t *ej,
                            const double *meas, const double *info,
                            const double *weight, double lambda,
                            double anchor, int32_t wrap,
                            const int32_t *row_ptr, const int32_t *col_idx,
                            double tol_rel, int32_t max_iter, double *dx_out,
                            double *diff_out)
{
    int64_t dim = 3 * (int64_t)n_poses;
    int64_t nnz = row_ptr[dim];
    double *ev = (double *)malloc(sizeof(double) * ORC_EDGE_VALS * (size_t)(n_edges + 1));
    double *values = (double *)malloc(sizeof(double) * (size_t)(nnz + 1));
    double *xi = (double *)malloc(sizeof(double) * (size_t)dim);
    double *dx = dx_out ? dx_out : (double *)malloc(sizeof(double) * (size_t)dim);
    if (!ev || !values || !xi || !dx) return -1;
    orc_between_edges(n_poses, poses, n_edges, ei, ej, meas, info, weight,
                      lambda, wrap, ev);
    if (orc_csr_values(n_poses, row_ptr, col_idx, n_edges, ei, ej, ev, anchor,
                       values, xi) != 0) {
        free(ev); free(values); free(xi);
        if (!dx_out) free(dx);
        return -2;
    }
    int32_t it = orc_pcg((int32_t)dim, row_ptr, col_idx, values, xi, tol_rel,
                         max_iter, dx, NULL);
    orc_pose_update(n_poses, poses, dx, wrap);
    if (diff_out) *diff_out = orc_diff((int32_t)dim, dx);
    free(ev); free(values); free(xi);
    if (!dx_out) free(dx);
    return it;
}

/* ------------------------------------------------------------------------ */
/* Landmark estimates after the GN loop, src/cpu_pcg_fp64/GSLAM.c:141-196:
 * landmarks in first-appearance order (make_edges, GSLAM.c:222-285), each
 * observation gives m = (x + r cos(th + phi), y + r sin(th + phi))
 * (MapEdge_create, MapEdge.c:12-29: m1 + m2), the estimate is the mean of a
 * landmark's m in observation order (vec_mean, MatrixOps.c:161-171).  The
 * reference computes these and frees them without output; its sentinel and
 * 1024-entry id array limits are not reproduced.  Returns the number of
 * landmarks; ids / means (2 per landmark) may be NULL to count.            */
int32_t orc_landmark_means(int32_t n_obs, const double *poses, const int32_t *z_step,
                           const int32_t *z_lm, const double *z_range,
                           const double *z_bearing, int32_t *ids_out, double *means_out)
{
    if (n_obs <= 0) return 0;
    int32_t *sorted = (int32_t *)malloc(sizeof(int32_t) * (size_t)n_obs);
    memcpy(sorted, z_lm, sizeof(int32_t) * (size_t)n_obs);
    qsort(sorted, (size_t)n_obs, sizeof(int32_t), cmp_i32);
    int32_t nd = 0;
    for (int32_t i = 0; i < n_obs; i++)
        if (i == 0 || sorted[i] != sorted[nd - 1]) sorted[nd++] = sorted[i];
    int32_t *rank = (int32_t *)malloc(sizeof(int32_t) * (size_t)nd);  /* first-appearance rank */
    for (int32_t i = 0; i < nd; i++) rank[i] = -1;
    int32_t n_ids = 0;
    int32_t *slot = (int32_t *)malloc(sizeof(int32_t) * (size_t)n_obs);
    for (int32_t i = 0; i < n_obs; i++) {
        int32_t *p = (int32_t *)bsearch(&z_lm[i], sorted, (size_t)nd, sizeof(int32_t), cmp_i32);
        size_t pos = (size_t)(p - sorted);
        if (rank[pos] < 0) rank[pos] = n_ids++;
        slot[i] = rank[pos];
    }
    if (ids_out && means_out) {
        double *sum = (double *)calloc(2 * (size_t)n_ids, sizeof(double));
        int32_t *cnt = (int32_t *)calloc((size_t)n_ids, sizeof(int32_t));
        for (int32_t i = 0; i < n_obs; i++) {   /* file order within each landmark */
            const double *x = poses + 3 * (int64_t)z_step[i];
            const double m0 = x[0] + z_range[i] * cos(x[2] + z_bearing[i]);
            const double m1 = x[1] + z_range[i] * sin(x[2] + z_bearing[i]);
            sum[2 * slot[i]] += m0;
            sum[2 * slot[i] + 1] += m1;
            if (cnt[slot[i]]++ == 0) ids_out[slot[i]] = z_lm[i];
        }
        for (int32_t b = 0; b < n_ids; b++) {
            means_out[2 * b] = sum[2 * b] / cnt[b];
            means_out[2 * b + 1] = sum[2 * b + 1] / cnt[b];
        }
        free(sum);
        free(cnt);
    }
    free(sorted);
    free(rank);
    free(slot);
    return n_ids;
}
