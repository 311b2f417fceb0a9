// Test infrastructure: the host-side graph structure of libgslam
// (g-slam_on_gpu_and_fpga_amd/csrc/graph_build.cpp, reference_cost.cpp; pure
// C++, no HIP) driven through its entry points with invariant checks, for an
// AddressSanitizer + UndefinedBehaviorSanitizer build (tests/test_host_sanitized.py).
//
// Graphs: a between-edge chain with loop closures, a star whose hub pose has
// more incidence entries than one 256-slot tile (the extra-chunk path of the
// mixed solver's tiles), a landmark / motion graph in both pair orders, and
// descriptions that must be refused.  Exit status 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "graph_build.h"

using namespace gslam;

static int failures = 0;
#define CHECK(c)                                                               \
    do {                                                                       \
        if (!(c)) {                                                            \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++failures;                                                        \
        }                                                                      \
    } while (0)

static gslam_graph_desc base_desc(int32_t n, const std::vector<double> &poses)
{
    gslam_graph_desc d;
    std::memset(&d, 0, sizeof(d));
    d.n_poses = n;
    d.poses = poses.data();
    d.delta = 1.0;
    d.lambda = 0.1;
    const double mns[4] = {0.19, 0.001, 0.13, 0.2};
    for (int k = 0; k < 4; ++k) d.mns[k] = mns[k];
    d.snr[0] = 0.14;
    d.snr[1] = 0.05;
    d.anchor = 1e6;
    return d;
}

// Structure invariants of the BSR pattern, the incidence lists and the
// mixed-operator tiles.
static void check_structure(const HostGraph &g)
{
    const int32_t N = g.n;
    CHECK(static_cast<int64_t>(g.brow_ptr.size()) == N + 1 && g.brow_ptr[0] == 0);
    for (int32_t i = 0; i < N; ++i) {
        CHECK(g.brow_ptr[i] < g.brow_ptr[i + 1]);
        CHECK(g.bcol[g.bdiag[i]] == i);
        for (int32_t b = g.brow_ptr[i] + 1; b < g.brow_ptr[i + 1]; ++b) CHECK(g.bcol[b - 1] < g.bcol[b]);
    }
    CHECK(g.xptr[N] == 2 * g.n_e);
    MixedTiles m;
    std::string err;
    CHECK(build_mixed_tiles(g, m, err) == GSLAM_OK);
    const int64_t slots = m.n_chunks * kTileBlocks;
    CHECK(static_cast<int64_t>(m.ent.size()) == slots && static_cast<int64_t>(m.pos.size()) == 2 * g.n_e);
    std::vector<int> seen(static_cast<size_t>(2 * g.n_e), 0);
    for (int64_t s = 0; s < slots; ++s) {
        const int32_t ent = m.ent[s];
        CHECK(m.own[s] >= 0 && m.own[s] < N && m.oth[s] >= 0 && m.oth[s] < N);
        if (ent < 0) continue;
        CHECK(ent < 2 * g.n_e);
        if (ent >= 2 * g.n_e) continue;
        ++seen[ent];
        CHECK(m.pos[ent] == s);
        const int64_t e = ent >> 1;
        const bool from = (ent & 1) == 0;
        CHECK(m.own[s] == (from ? g.t1[e] : g.t2[e]) && m.oth[s] == (from ? g.t2[e] : g.t1[e]));
    }
    for (int v : seen) CHECK(v == 1);
    int32_t next_row = 0;
    int64_t extra = 0;
    for (int32_t t = 0; t < m.n_tiles; ++t) {
        const int32_t r0 = m.info[4 * t], r1 = m.info[4 * t + 1], x0 = m.info[4 * t + 2],
                      x1 = m.info[4 * t + 3];
        CHECK(r0 == next_row && r1 > r0 && x0 == g.xptr[r0] && x1 == g.xptr[r1]);
        next_row = r1;
        if (x1 - x0 > kTileBlocks) {
            CHECK(r1 == r0 + 1 && m.ext[t] >= m.n_tiles);
            extra += (x1 - x0 - 1) / kTileBlocks;
        } else {
            CHECK(r1 - r0 <= kTileRows && m.ext[t] == -1);
        }
        // entry j of the tile sits in chunk t (j < 256) or in its extra chunks
        for (int32_t k = x0; k < x1; ++k) {
            const int64_t j = k - x0;
            const int64_t chunk = j < kTileBlocks ? t : m.ext[t] + j / kTileBlocks - 1;
            CHECK(m.ent[chunk * kTileBlocks + j % kTileBlocks] == g.xlist[k]);
        }
    }
    CHECK(next_row == N && m.n_chunks == m.n_tiles + extra);
}

static int between_graphs()
{
    std::mt19937 rng(7);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    // chain of 3000 with 400 loop closures
    {
        const int32_t n = 3000;
        std::vector<double> poses(3 * n);
        for (auto &v : poses) v = u(rng);
        std::vector<int32_t> bi, bj;
        for (int32_t i = 0; i + 1 < n; ++i) { bi.push_back(i); bj.push_back(i + 1); }
        std::uniform_int_distribution<int32_t> pick(0, n - 1);
        for (int k = 0; k < 400; ++k) {
            int32_t a = pick(rng), b = pick(rng);
            if (a != b) { bi.push_back(a); bj.push_back(b); }
        }
        std::vector<double> meas(3 * bi.size()), info(3 * bi.size(), 100.0);
        for (auto &v : meas) v = u(rng);
        gslam_graph_desc d = base_desc(n, poses);
        d.n_between = static_cast<int64_t>(bi.size());
        d.be_i = bi.data();
        d.be_j = bj.data();
        d.be_meas = meas.data();
        d.be_info = info.data();
        HostGraph g;
        std::string err;
        CHECK(build_host_graph(d, g, err) == GSLAM_OK);
        if (!failures) check_structure(g);
    }
    // star: pose 0 joined to 700 others, plus a chain (hub degree 701)
    {
        const int32_t n = 800;
        std::vector<double> poses(3 * n);
        for (auto &v : poses) v = u(rng);
        std::vector<int32_t> bi, bj;
        for (int32_t i = 1; i <= 700; ++i) { bi.push_back(0); bj.push_back(i); }
        for (int32_t i = 1; i + 1 < n; ++i) { bi.push_back(i); bj.push_back(i + 1); }
        std::vector<double> meas(3 * bi.size(), 0.5), info(3 * bi.size(), 10.0);
        gslam_graph_desc d = base_desc(n, poses);
        d.n_between = static_cast<int64_t>(bi.size());
        d.be_i = bi.data();
        d.be_j = bj.data();
        d.be_meas = meas.data();
        d.be_info = info.data();
        HostGraph g;
        std::string err;
        CHECK(build_host_graph(d, g, err) == GSLAM_OK);
        CHECK(g.xptr[1] - g.xptr[0] == 700);
        if (!failures) check_structure(g);
    }
    return failures;
}

static int landmark_graph()
{
    std::mt19937 rng(11);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    const int32_t n = 200, n_obs = 900;
    std::vector<double> poses(3 * n), nu(n), om(n), zr(n_obs), zb(n_obs);
    for (auto &v : poses) v = u(rng);
    for (int32_t i = 0; i < n; ++i) { nu[i] = 1.0 + 0.1 * u(rng); om[i] = 0.05 * u(rng); }
    std::vector<int32_t> zs(n_obs), zl(n_obs);
    std::uniform_int_distribution<int32_t> pose(0, n - 1), lm(0, 59);
    for (int32_t k = 0; k < n_obs; ++k) {
        zs[k] = pose(rng);
        zl[k] = lm(rng);
        zr[k] = 2.0 + u(rng);
        zb[k] = u(rng);
    }
    for (int order = 0; order < 2; ++order) {
        gslam_graph_desc d = base_desc(n, poses);
        d.motion = 1;
        d.u_nu = nu.data();
        d.u_omega = om.data();
        d.n_obs = n_obs;
        d.z_step = zs.data();
        d.z_lm = zl.data();
        d.z_range = zr.data();
        d.z_bearing = zb.data();
        d.pair_order = order;
        HostGraph g;
        std::string err;
        CHECK(build_host_graph(d, g, err) == GSLAM_OK);
        CHECK(g.n_mot_e == n - 1 && g.n_obs_e > 0);
        if (failures) return failures;
        check_structure(g);
        const double snr[2] = {0.14, 0.05};
        const double c = obs_cost_terms(g, poses.data(), snr, 0.0);
        CHECK(std::isfinite(c) && c > 0.0);
        CHECK(obs_cost_terms(g, poses.data(), snr, c) > c);
        std::vector<int32_t> a, b;
        landmark_pairs(n_obs, zl.data(), order, a, b);
        CHECK(static_cast<int64_t>(a.size()) == g.n_obs_e);
    }
    return failures;
}

static int refused()
{
    std::vector<double> poses = {0.0, 0.0, 0.0, 1.0, 0.0, 0.0};
    std::vector<int32_t> bi = {0}, bj = {1}, bad_j = {2};
    std::vector<double> meas = {1.0, 0.0, 0.0}, info = {1.0, 1.0, 1.0};
    auto desc = [&](const std::vector<int32_t> &j) {
        gslam_graph_desc d = base_desc(2, poses);
        d.n_between = 1;
        d.be_i = bi.data();
        d.be_j = j.data();
        d.be_meas = meas.data();
        d.be_info = info.data();
        return d;
    };
    HostGraph g;
    std::string err;
    CHECK(build_host_graph(desc(bad_j), g, err) == GSLAM_ERR_ARG);
    gslam_graph_desc d = desc(bj);
    d.lambda = std::nan("");
    CHECK(build_host_graph(d, g, err) == GSLAM_ERR_ARG && err.find("lambda") != std::string::npos);
    poses[4] = INFINITY;
    CHECK(build_host_graph(desc(bj), g, err) == GSLAM_ERR_ARG && err.find("pose") != std::string::npos);
    poses[4] = 0.0;
    d = desc(bj);
    d.n_poses = 0;
    CHECK(build_host_graph(d, g, err) == GSLAM_ERR_ARG);
    return failures;
}

int main()
{
    between_graphs();
    landmark_graph();
    refused();
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("graph_build_check: ok\n");
    return 0;
}
