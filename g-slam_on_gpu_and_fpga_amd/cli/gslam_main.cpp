// GSLAM -- drop-in command-line driver over libgslam.
//
//   GSLAM <use_diff> <lambda> <num_rounds> <drawing_mode> <input_file>
//
// Same argv, stdout lines and pose dump as the reference executables
// (src/cpu_pcg_fp64/GSLAM.c:13-220, src/gpu_pcg_fp64/GSLAM.cu:28-335):
//   "Diff: %s Lambda: %s Rounds: %s Drawing: %s Input file: %s"
//   "delta: %f size_of_us: %d size_of_hat_xs: %d size_of_zlist: %d dim: %d"
//   "%d rounds executed: %.17g"                       (one per GN round)
// and the final poses as "%u %.17g %.17g %.17g" lines into POSE_DUMP_FILE.
//
// PCG solvers (pcg64, mixed) follow the reference GPU PCG driver
// (src/gpu_pcg_fp64/GSLAM.cu:150-200): diff = ||dx||/sqrt(dim) whatever
// use_diff says, stop at diff < 0.01 or after 20 rounds without a new best
// diff (src/cpu_pcg_fp64/GSLAM.c:119-130), pose dump when drawing_mode >= 1.
//
// Gauss-Jordan solvers (gj64, gj64b, gjmixed) follow the reference GPU GJ
// drivers (src/gpu_gj_fp64/GSLAM.cu:193-591, the same in gpu_gj_mixed):
//   use_diff 1: diff = ||dx||/sqrt(dim); stop at diff < 0.01, or once the
//     even rounds or the odd rounds have gone 10 of their own rounds without
//     a new best (:518-530, the 2-cycle plateau);
//   use_diff 0: diff = |cost - last_cost| / cost with the running cost of
//     gslam_obs_cost (never reset, :245-257, 454); stop at diff < 1e-4,
//     otherwise print "last_cost: %.17g" and "cost: %.17g" (:533-591);
// and the pose dump is written whenever POSE_DUMP_FILE is set (:757-770).
//
// Pose graphs: GSLAM_BETWEEN=<file> adds SE(2) between edges from a sidecar
// file (lines "e i j dx dy dth info_x info_y info_th", '#' comments; i, j
// index the x records in file order), kept out of the .log.txt because the
// reference reader takes any unknown first token as a delta record
// (src/cpu_pcg_fp64/GSLAM.c:480-483).  With a sidecar and no u records, no
// motion edges are formed.
//
// Extra knobs (environment, defaults reproduce the reference):
//   GSLAM_SOLVER   pcg64 | mixed | gj64 | gj64b | gjmixed  (pcg64)
//   GSLAM_TOL      PCG relative tolerance      (1e-6, MatrixOps.c:345)
//   GSLAM_REFINE   gjmixed refinement: ir | cg (ir)
//   GSLAM_SWEEPS   gjmixed refinement iterations (3)
//   GSLAM_MAX_ITER PCG iteration cap           (500,  MatrixOps.c:347)
//   GSLAM_HANDOVER mixed: true relative residual at which FP64 PCG takes over
//                  the correction (0: FP32 iterations all the way)
//   GSLAM_WRAP     1: wrap theta residuals and poses (0)
//   GSLAM_ORDER    landmark pair order 0 (CPU reference) | 1 (GPU reference)
//   GSLAM_STATS    1: one JSON line per round on stderr
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gslam.h"

namespace {

struct LogData {
    double delta = 0.0;
    std::vector<unsigned int> u_step, x_step, z_step, z_lm;
    std::vector<double> u_nu, u_omega, poses, z_range, z_bearing;
};

// read_data, src/cpu_pcg_fp64/GSLAM.c:415-493: ' '-separated tokens, atoi /
// atof conversions, any first token other than u/x/z sets delta from the
// second token.  A line too short for its record is an error here (the
// reference dereferences a NULL token).
int read_log(const char *path, LogData &d)
{
    FILE *f = std::fopen(path, "r");
    if (!f) {
        std::printf("Error! Cannot open the input file!");
        return -1;
    }
    char line[1024];
    int lineno = 0;
    while (std::fgets(line, sizeof(line), f)) {
        ++lineno;
        char *tok = std::strtok(line, " ");
        const int need = !tok ? 0 : (!std::strcmp(tok, "u") ? 3 : (!std::strcmp(tok, "x") || !std::strcmp(tok, "z")) ? 4 : 1);
        char *t[4] = {nullptr, nullptr, nullptr, nullptr};
        bool ok = tok != nullptr;
        for (int i = 0; ok && i < need; ++i) ok = (t[i] = std::strtok(nullptr, " ")) != nullptr;
        if (!ok) {
            std::fprintf(stderr, "GSLAM: malformed record at %s:%d\n", path, lineno);
            std::fclose(f);
            return -2;
        }
        if (!std::strcmp(tok, "u")) {
            d.u_step.push_back(static_cast<unsigned int>(std::atoi(t[0])));
            d.u_nu.push_back(std::atof(t[1]));
            d.u_omega.push_back(std::atof(t[2]));
        } else if (!std::strcmp(tok, "x")) {
            d.x_step.push_back(static_cast<unsigned int>(std::atoi(t[0])));
            d.poses.push_back(std::atof(t[1]));
            d.poses.push_back(std::atof(t[2]));
            d.poses.push_back(std::atof(t[3]));
        } else if (!std::strcmp(tok, "z")) {
            d.z_step.push_back(static_cast<unsigned int>(std::atoi(t[0])));
            d.z_lm.push_back(static_cast<unsigned int>(std::atoi(t[1])));
            d.z_range.push_back(std::atof(t[2]));
            d.z_bearing.push_back(std::atof(t[3]));
        } else {
            d.delta = std::atof(t[0]);
        }
    }
    std::fclose(f);
    return 0;
}

// Between-edge sidecar: "e i j dx dy dth info_x info_y info_th" per line.
int read_between(const char *path, std::vector<int32_t> &bi, std::vector<int32_t> &bj,
                 std::vector<double> &meas, std::vector<double> &info)
{
    FILE *f = std::fopen(path, "r");
    if (!f) {
        std::fprintf(stderr, "GSLAM: cannot open between-edge file %s\n", path);
        return -1;
    }
    char line[1024];
    int lineno = 0;
    while (std::fgets(line, sizeof(line), f)) {
        ++lineno;
        const char *c = line;
        while (*c == ' ' || *c == '\t') ++c;
        if (*c == '#' || *c == '\n' || *c == '\r' || *c == '\0') continue;
        int i = 0, j = 0;
        double v[6];
        if (std::sscanf(c, "e %d %d %lf %lf %lf %lf %lf %lf", &i, &j, &v[0], &v[1], &v[2], &v[3],
                        &v[4], &v[5]) != 8) {
            std::fprintf(stderr, "GSLAM: malformed between edge at %s:%d\n", path, lineno);
            std::fclose(f);
            return -2;
        }
        bi.push_back(i);
        bj.push_back(j);
        meas.insert(meas.end(), v, v + 3);
        info.insert(info.end(), v + 3, v + 6);
    }
    std::fclose(f);
    return 0;
}

double env_double(const char *name, double dflt)
{
    const char *v = std::getenv(name);
    return v ? std::atof(v) : dflt;
}

int env_int(const char *name, int dflt)
{
    const char *v = std::getenv(name);
    return v ? std::atoi(v) : dflt;
}

}  // namespace

// The poses as "%u %.17g %.17g %.17g" lines (x_step, x, y, theta), the
// POSE_DUMP_FILE format (GSLAM.c:198-213); written to a temporary file and
// renamed, so that a checkpoint is never seen half-written.
static void dump_poses(gslam_graph *g, const LogData &d, int n, const char *path)
{
    std::vector<double> poses(3 * static_cast<size_t>(n));
    if (gslam_get_poses(g, poses.data()) != GSLAM_OK) return;
    const std::string tmp = std::string(path) + ".tmp";
    FILE *df = std::fopen(tmp.c_str(), "w");
    if (!df) return;
    for (int i = 0; i < n; ++i)
        std::fprintf(df, "%u %.17g %.17g %.17g\n", d.x_step[i], poses[3 * i], poses[3 * i + 1],
                     poses[3 * i + 2]);
    if (std::fclose(df) == 0) std::rename(tmp.c_str(), path);
}

int main(int argc, char *argv[])
{
    if (argc == 6) {
        std::printf("Diff: %s ", argv[1]);
        std::printf("Lambda: %s ", argv[2]);
        std::printf("Rounds: %s ", argv[3]);
        std::printf("Drawing: %s ", argv[4]);
        std::printf("Input file: %s\n", argv[5]);
    } else {
        std::printf("Usage: ");
        std::printf("%s Diff_0|1 Lambda Number_of_rounds Drawing_0|1|2|3 path/to/input/file\n", argv[0]);
        return 1;
    }
    const double lambda = std::atof(argv[2]);
    const unsigned int rounds = static_cast<unsigned int>(std::atoi(argv[3]));
    const unsigned int draw = static_cast<unsigned int>(std::atoi(argv[4]));

    LogData d;
    if (read_log(argv[5], d) != 0) return 1;
    const unsigned int dim = static_cast<unsigned int>(d.u_nu.size()) * 3;
    std::printf("delta: %f ", d.delta);
    std::printf("size_of_us: %d ", static_cast<int>(d.u_nu.size()));
    std::printf("size_of_hat_xs: %d ", static_cast<int>(d.x_step.size()));
    std::printf("size_of_zlist: %d ", static_cast<int>(d.z_step.size()));
    std::printf("dim: %d\n", dim);
    std::fflush(stdout);

    const int n = static_cast<int>(d.x_step.size());
    std::vector<int32_t> be_i, be_j;
    std::vector<double> be_meas, be_info;
    const char *between = std::getenv("GSLAM_BETWEEN");
    if (between && read_between(between, be_i, be_j, be_meas, be_info) != 0) return 1;
    const bool motion = !(between && d.u_nu.empty());
    if (n == 0 || (motion && d.u_nu.size() != d.x_step.size())) {
        std::fprintf(stderr, "GSLAM: need one u record per x record (u=%zu, x=%zu)\n",
                     d.u_nu.size(), d.x_step.size());
        return 1;
    }
    std::vector<int32_t> zs(d.z_step.begin(), d.z_step.end()), zl(d.z_lm.begin(), d.z_lm.end());

    gslam_graph_desc desc;
    gslam_default_desc(&desc);
    desc.n_poses = n;
    desc.poses = d.poses.data();
    desc.motion = motion ? 1 : 0;
    desc.u_nu = motion ? d.u_nu.data() : nullptr;
    desc.u_omega = motion ? d.u_omega.data() : nullptr;
    desc.delta = d.delta;
    desc.lambda = lambda;
    desc.n_obs = static_cast<int32_t>(zs.size());
    desc.z_step = zs.data();
    desc.z_lm = zl.data();
    desc.z_range = d.z_range.data();
    desc.z_bearing = d.z_bearing.data();
    desc.pair_order = env_int("GSLAM_ORDER", 0);
    desc.wrap = env_int("GSLAM_WRAP", 0);
    desc.n_between = static_cast<int64_t>(be_i.size());
    desc.be_i = be_i.data();
    desc.be_j = be_j.data();
    desc.be_meas = be_meas.data();
    desc.be_info = be_info.data();

    gslam_graph *g = nullptr;
    int rc = gslam_create(&desc, &g);
    if (rc != GSLAM_OK) {
        std::fprintf(stderr, "GSLAM: gslam_create failed (%d): %s\n", rc, gslam_last_error());
        return 1;
    }
    gslam_solve_opts opts;
    gslam_default_opts(&opts);
    const char *solver = std::getenv("GSLAM_SOLVER");
    if (solver && !std::strcmp(solver, "mixed")) opts.solver = GSLAM_SOLVER_PCG_MIXED;
    if (solver && !std::strcmp(solver, "gj64")) opts.solver = GSLAM_SOLVER_GJ_FP64;
    if (solver && !std::strcmp(solver, "gjmixed")) opts.solver = GSLAM_SOLVER_GJ_MIXED;
    if (solver && !std::strcmp(solver, "gj64b")) opts.solver = GSLAM_SOLVER_GJ_BLOCKED64;
    const char *refine = std::getenv("GSLAM_REFINE");
    if (refine && !std::strcmp(refine, "cg")) opts.refine_mode = GSLAM_REFINE_CG;
    opts.refine_sweeps = env_int("GSLAM_SWEEPS", opts.refine_sweeps);
    opts.tol = env_double("GSLAM_TOL", opts.tol);
    opts.max_iter = env_int("GSLAM_MAX_ITER", opts.max_iter);
    opts.mixed_handover = env_double("GSLAM_HANDOVER", opts.mixed_handover);
    const bool stats_json = env_int("GSLAM_STATS", 0) != 0;
    // checkpoints: the poses after every k-th round into POSE_DUMP_FILE (the
    // final dump overwrites the last one); rewriting the x records of the
    // input with them resumes the run (gslam.resume_log)
    const int dump_every = env_int("GSLAM_DUMP_EVERY", 0);
    const char *dump_path = std::getenv("POSE_DUMP_FILE");

    const bool gj = opts.solver == GSLAM_SOLVER_GJ_FP64 || opts.solver == GSLAM_SOLVER_GJ_MIXED ||
                    opts.solver == GSLAM_SOLVER_GJ_BLOCKED64;
    const bool use_diff = std::atoi(argv[1]) != 0;
    double best = 1e30, best_even = 1e30, best_odd = 1e30;
    unsigned int no_improve = 0, no_improve_even = 0, no_improve_odd = 0;
    double cost = 0.0, last_cost = 0.0;
    for (unsigned int r = 1; r <= rounds; ++r) {
        // GJ drivers, use_diff 0: the cost is taken before the solve, at the
        // poses this round linearises at (GSLAM.cu:245-257)
        if (gj && !use_diff) {
            rc = gslam_obs_cost(g, &cost);
            if (rc != GSLAM_OK) {
                std::fprintf(stderr, "GSLAM: cost failed (%d): %s\n", rc, gslam_last_error());
                gslam_destroy(g);
                return 1;
            }
        }
        gslam_solve_stats st;
        double diff = 0.0;
        rc = gslam_gn_step(g, &opts, &st, &diff);
        if (rc != GSLAM_OK && rc != GSLAM_ERR_NOT_CONVERGED && rc != GSLAM_ERR_SINGULAR) {
            std::fprintf(stderr, "GSLAM: GN step failed (%d): %s\n", rc, gslam_last_error());
            gslam_destroy(g);
            return 1;
        }
        if (gj && !use_diff) diff = std::fabs(cost - last_cost) / cost;
        std::printf("%d rounds executed: %.17g\n", r, diff);
        std::fflush(stdout);
        if (stats_json)
            std::fprintf(stderr,
                         "{\"round\": %u, \"iters\": %d, \"converged\": %d, \"rel_res\": %.6e, "
                         "\"true_rel_res\": %.6e, \"solve_ms\": %.4f, \"assembly_ms\": %.4f, "
                         "\"setup_ms\": %.4f, \"update_ms\": %.4f}\n",
                         r, st.iters, st.converged, st.rel_res, st.true_rel_res, st.ms,
                         st.ms_assembly, st.ms_setup, st.ms_update);
        if (dump_every > 0 && dump_path && r % static_cast<unsigned int>(dump_every) == 0)
            dump_poses(g, d, n, dump_path);
        if (gj && use_diff) {
            if (diff < 0.01) break;
            // separate plateaus for even and odd rounds (a 2-cycle of the
            // iteration does not stop the run by itself)
            double &b = (r % 2 == 0) ? best_even : best_odd;
            unsigned int &k = (r % 2 == 0) ? no_improve_even : no_improve_odd;
            if (diff < b) {
                b = diff;
                k = 0;
            } else if (++k >= 10) {
                break;
            }
        } else if (gj) {
            if (diff < 1e-4) break;
            std::printf("last_cost: %.17g\n", last_cost);
            std::printf("cost: %.17g\n", cost);
            std::fflush(stdout);
            last_cost = cost;
        } else {
            if (diff < 0.01) break;
            if (diff < best) {
                best = diff;
                no_improve = 0;
            } else if (++no_improve >= 20) {
                break;
            }
        }
    }
    if (draw >= 1 || gj) {
        const char *dump = std::getenv("POSE_DUMP_FILE");
        if (dump) dump_poses(g, d, n, dump);
    }
    gslam_destroy(g);
    return 0;
}
