/* Test infrastructure: drives the CPU oracle through every entry point on a
 * small .log.txt graph and a small between-edge graph, for the
 * AddressSanitizer / UndefinedBehaviorSanitizer build of `make sanitize`
 * (host code only).  Exit status 0 when every call returned its documented
 * success value; the sanitizers abort on the first memory or UB error.
 *
 *   oracle_sanitized <file.log.txt>                                          */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gslam_oracle.h"

#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                              \
        }                                                          \
    } while (0)

static int log_graph(const char *path)
{
    int32_t nu = 0, nx = 0, nz = 0;
    CHECK(orc_log_count(path, &nu, &nx, &nz) == 0 && nx > 1);
    double delta = 0.0;
    int32_t *u_step = malloc(sizeof(int32_t) * nu), *x_step = malloc(sizeof(int32_t) * nx);
    double *u_nu = malloc(sizeof(double) * nu), *u_om = malloc(sizeof(double) * nu);
    double *poses = malloc(sizeof(double) * 3 * nx);
    int32_t *z_step = malloc(sizeof(int32_t) * nz), *z_lm = malloc(sizeof(int32_t) * nz);
    double *z_r = malloc(sizeof(double) * nz), *z_b = malloc(sizeof(double) * nz);
    CHECK(orc_log_read(path, &delta, u_step, u_nu, u_om, x_step, poses, z_step, z_lm, z_r, z_b) == 0);

    /* one round of edges, CSR and dense assembly */
    const double mns[4] = {0.19, 0.001, 0.13, 0.2};
    const int64_t ne = orc_log_edges(nx, poses, u_nu, u_om, delta, 0.1, mns, nz, z_step, z_lm, z_r,
                                     z_b, 0, NULL, NULL, NULL);
    CHECK(ne > 0);
    int32_t *t1 = malloc(sizeof(int32_t) * ne), *t2 = malloc(sizeof(int32_t) * ne);
    double *ev = malloc(sizeof(double) * ORC_EDGE_VALS * ne);
    CHECK(orc_log_edges(nx, poses, u_nu, u_om, delta, 0.1, mns, nz, z_step, z_lm, z_r, z_b, 0, t1,
                        t2, ev) == ne);
    const int32_t dim = 3 * nx;
    int32_t *rp = malloc(sizeof(int32_t) * (dim + 1));
    const int64_t nnz = orc_csr_pattern(nx, ne, t1, t2, rp, NULL);
    CHECK(nnz > 0);
    int32_t *ci = malloc(sizeof(int32_t) * nnz);
    CHECK(orc_csr_pattern(nx, ne, t1, t2, rp, ci) == nnz);
    double *va = malloc(sizeof(double) * nnz), *xi = malloc(sizeof(double) * dim);
    CHECK(orc_csr_values(nx, rp, ci, ne, t1, t2, ev, 1e6, va, xi) == 0);
    double *dinv = malloc(sizeof(double) * 9 * nx), *x = calloc(dim, sizeof(double));
    double *y = malloc(sizeof(double) * dim), *hist = malloc(sizeof(double) * 500);
    orc_block_jacobi(dim, rp, ci, va, dinv);
    orc_spmv(dim, rp, ci, va, xi, y);
    const int32_t it = orc_pcg(dim, rp, ci, va, xi, 1e-6, 500, x, hist);
    CHECK(it > 0 && it <= 500);
    CHECK(isfinite(orc_diff(dim, x)));

    double *om = malloc(sizeof(double) * dim * dim), *inv = malloc(sizeof(double) * dim * dim);
    double *xi2 = malloc(sizeof(double) * dim);
    CHECK(orc_dense_values(nx, ne, t1, t2, ev, 1e6, om, xi2) == 0);
    CHECK(orc_gj_inv(dim, om, inv) == 0);
    orc_dense_matvec(dim, inv, xi2, y);
    float *omf = malloc(sizeof(float) * dim * dim), *invf = malloc(sizeof(float) * dim * dim);
    for (int64_t k = 0; k < (int64_t)dim * dim; ++k) omf[k] = (float)om[k];
    (void)orc_gj_inv_f32(dim, omf, invf);
    orc_pose_update(nx, poses, y, 1);

    /* the full reference loop, two rounds */
    double diffs[2];
    CHECK(orc_log_gn_run(nx, poses, u_nu, u_om, delta, 0.1, nz, z_step, z_lm, z_r, z_b, 2, 1e-6, 500,
                         diffs) >= 1);
    int32_t *ids = malloc(sizeof(int32_t) * (nz > 0 ? nz : 1));
    double *means = malloc(sizeof(double) * 2 * (nz > 0 ? nz : 1));
    CHECK(orc_landmark_means(nz, poses, z_step, z_lm, z_r, z_b, ids, means) >= 0);

    free(u_step); free(x_step); free(u_nu); free(u_om); free(poses); free(z_step); free(z_lm);
    free(z_r); free(z_b); free(t1); free(t2); free(ev); free(rp); free(ci); free(va); free(xi);
    free(dinv); free(x); free(y); free(hist); free(om); free(inv); free(xi2); free(omf);
    free(invf); free(ids); free(means);
    return 0;
}

/* A ring of n poses with odometry and a few loop closures. */
static int between_graph(void)
{
    const int32_t n = 60;
    const int64_t ne = n - 1 + 6;
    int32_t *ei = malloc(sizeof(int32_t) * ne), *ej = malloc(sizeof(int32_t) * ne);
    double *meas = malloc(sizeof(double) * 3 * ne), *info = malloc(sizeof(double) * 3 * ne);
    double *poses = malloc(sizeof(double) * 3 * n);
    for (int32_t i = 0; i < n; ++i) {
        const double a = 2.0 * M_PI * i / n;
        poses[3 * i] = 10.0 * cos(a) + 0.01 * (i % 7);
        poses[3 * i + 1] = 10.0 * sin(a) - 0.02 * (i % 5);
        poses[3 * i + 2] = a + M_PI / 2;
    }
    for (int64_t k = 0; k < ne; ++k) {
        ei[k] = (int32_t)(k < n - 1 ? k : (k - (n - 1)) * 9);
        ej[k] = (int32_t)(k < n - 1 ? k + 1 : ((k - (n - 1)) * 9 + 30) % n);
        meas[3 * k] = k < n - 1 ? 2.0 * M_PI * 10.0 / n : 0.0;
        meas[3 * k + 1] = 0.0;
        meas[3 * k + 2] = k < n - 1 ? 2.0 * M_PI / n : M_PI;
        info[3 * k] = info[3 * k + 1] = 400.0;
        info[3 * k + 2] = 1e4;
    }
    int32_t *rp = malloc(sizeof(int32_t) * (3 * n + 1));
    const int64_t nnz = orc_csr_pattern(n, ne, ei, ej, rp, NULL);
    int32_t *ci = malloc(sizeof(int32_t) * nnz);
    CHECK(orc_csr_pattern(n, ne, ei, ej, rp, ci) == nnz);
    double *dx = malloc(sizeof(double) * 3 * n), diff = 0.0;
    for (int r = 0; r < 3; ++r)
        CHECK(orc_between_gn_step(n, poses, ne, ei, ej, meas, info, NULL, 0.1, 1e6, 1, rp, ci,
                                  1e-10, 20000, dx, &diff) > 0);
    CHECK(isfinite(diff));
    free(ei); free(ej); free(meas); free(info); free(poses); free(rp); free(ci); free(dx);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: %s <file.log.txt>\n", argv[0]);
        return 2;
    }
    if (log_graph(argv[1])) return 1;
    if (between_graph()) return 1;
    printf("oracle_sanitized: ok\n");
    return 0;
}
