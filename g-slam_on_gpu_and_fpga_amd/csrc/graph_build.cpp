// Host-side structure build for libgslam.  Replaces the reference's
// dim x dim byte marker (src/cpu_pcg_fp64/MatrixOps.c:177-221,
// src/gpu_pcg_fp64/device_graph.cu:553-615), which overflows 32-bit indices
// past N ~ 15k poses, by per-block-row neighbour sorting: O(E log deg)
// time, O(N + E) memory, 64-bit counts throughout.
#include "graph_build.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <unordered_map>

namespace gslam {

void landmark_pairs(int32_t n_obs, const int32_t *z_lm, int order,
                    std::vector<int32_t> &a, std::vector<int32_t> &b)
{
    a.clear();
    b.clear();
    if (n_obs <= 0) return;
    // rank of each landmark's bucket: first appearance or ascending id
    std::unordered_map<int32_t, int64_t> rank;
    rank.reserve(static_cast<size_t>(n_obs) * 2);
    for (int32_t i = 0; i < n_obs; ++i) rank.emplace(z_lm[i], static_cast<int64_t>(i));
    if (order == 1) {
        std::vector<int32_t> ids;
        ids.reserve(rank.size());
        for (auto &kv : rank) ids.push_back(kv.first);
        std::sort(ids.begin(), ids.end());
        for (size_t k = 0; k < ids.size(); ++k) rank[ids[k]] = static_cast<int64_t>(k);
    }
    std::vector<int32_t> idx(static_cast<size_t>(n_obs));
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) {
        return rank[z_lm[x]] < rank[z_lm[y]];
    });
    size_t lo = 0;
    while (lo < idx.size()) {
        size_t hi = lo;
        while (hi < idx.size() && z_lm[idx[hi]] == z_lm[idx[lo]]) ++hi;
        for (size_t j = lo; j < hi; ++j)
            for (size_t k = j + 1; k < hi; ++k) {
                a.push_back(idx[j]);
                b.push_back(idx[k]);
            }
        lo = hi;
    }
}

static int fail(std::string &err, int code, const std::string &msg)
{
    err = msg;
    return code;
}

int build_host_graph(const gslam_graph_desc &d, HostGraph &g, std::string &err)
{
    if (d.n_poses <= 0) return fail(err, GSLAM_ERR_ARG, "n_poses must be > 0");
    const int32_t N = d.n_poses;
    g.n = N;
    if (!d.poses) return fail(err, GSLAM_ERR_ARG, "poses is NULL");

    // ---- observation pairs (ObsEdges) ----
    std::vector<int32_t> pa, pb;
    if (d.n_obs > 0) {
        if (!d.z_step || !d.z_lm || !d.z_range || !d.z_bearing)
            return fail(err, GSLAM_ERR_ARG, "observation arrays missing");
        for (int32_t i = 0; i < d.n_obs; ++i)
            if (d.z_step[i] < 0 || d.z_step[i] >= N)
                return fail(err, GSLAM_ERR_ARG, "z_step[" + std::to_string(i) +
                            "] outside [0, n_poses)");
        landmark_pairs(d.n_obs, d.z_lm, d.pair_order, pa, pb);
    }
    g.n_obs_e = static_cast<int64_t>(pa.size());
    g.n_mot_e = (d.motion && N >= 2) ? N - 1 : 0;
    if (g.n_mot_e && (!d.u_nu || !d.u_omega))
        return fail(err, GSLAM_ERR_ARG, "motion edges need u_nu/u_omega");
    if (g.n_mot_e && !(d.delta > 0.0))
        return fail(err, GSLAM_ERR_ARG, "delta must be > 0 for motion edges");
    g.n_bet_e = d.n_between > 0 ? d.n_between : 0;
    if (g.n_bet_e && (!d.be_i || !d.be_j || !d.be_meas || !d.be_info))
        return fail(err, GSLAM_ERR_ARG, "between-edge arrays missing");
    g.n_e = g.n_obs_e + g.n_mot_e + g.n_bet_e;
    // contribution entries are int32 (e*4+q)
    if (g.n_e > (std::numeric_limits<int32_t>::max() / 4))
        return fail(err, GSLAM_ERR_TOO_LARGE, "more than 536,870,911 edges");

    g.t1.resize(g.n_e);
    g.t2.resize(g.n_e);
    g.obs_param.resize(4 * g.n_obs_e);
    for (int64_t e = 0; e < g.n_obs_e; ++e) {
        int32_t a = pa[e], b = pb[e];
        g.t1[e] = d.z_step[a];
        g.t2[e] = d.z_step[b];
        g.obs_param[4 * e + 0] = d.z_range[a];
        g.obs_param[4 * e + 1] = d.z_bearing[a];
        g.obs_param[4 * e + 2] = d.z_range[b];
        g.obs_param[4 * e + 3] = d.z_bearing[b];
    }
    g.mot_param.resize(2 * g.n_mot_e);
    for (int64_t i = 0; i < g.n_mot_e; ++i) {
        int64_t e = g.n_obs_e + i;
        g.t1[e] = static_cast<int32_t>(i);
        g.t2[e] = static_cast<int32_t>(i + 1);
        g.mot_param[2 * i + 0] = d.u_nu[i + 1];
        g.mot_param[2 * i + 1] = d.u_omega[i + 1];
    }
    g.bet_param.resize(6 * g.n_bet_e);
    for (int64_t k = 0; k < g.n_bet_e; ++k) {
        int64_t e = g.n_obs_e + g.n_mot_e + k;
        int32_t i = d.be_i[k], j = d.be_j[k];
        if (i < 0 || i >= N || j < 0 || j >= N)
            return fail(err, GSLAM_ERR_ARG, "between edge " + std::to_string(k) +
                        " references a pose outside [0, n_poses)");
        g.t1[e] = i;
        g.t2[e] = j;
        for (int c = 0; c < 3; ++c) {
            g.bet_param[6 * k + c] = d.be_meas[3 * k + c];
            g.bet_param[6 * k + 3 + c] = d.be_info[3 * k + c];
        }
    }

    // ---- every value the edge kernels read must be finite ----
    auto finite = [](const double *v, int64_t n) {
        for (int64_t k = 0; k < n; ++k)
            if (!std::isfinite(v[k])) return false;
        return true;
    };
    const double scal[9] = {d.lambda, d.anchor, d.mns[0], d.mns[1], d.mns[2], d.mns[3],
                            d.snr[0], d.snr[1], g.n_mot_e ? d.delta : 1.0};
    if (!finite(d.poses, 3 * static_cast<int64_t>(N)))
        return fail(err, GSLAM_ERR_ARG, "a pose value is NaN or Inf");
    if (!finite(scal, 9))
        return fail(err, GSLAM_ERR_ARG, "lambda, anchor, delta, mns or snr is NaN or Inf");
    if (!finite(g.obs_param.data(), 4 * g.n_obs_e) || !finite(g.mot_param.data(), 2 * g.n_mot_e) ||
        !finite(g.bet_param.data(), 6 * g.n_bet_e))
        return fail(err, GSLAM_ERR_ARG, "an observation, control or between-edge value is NaN or Inf");

    // ---- BSR pattern: per block row, sorted unique neighbour columns ----
    std::vector<int64_t> deg(static_cast<size_t>(N) + 1, 0);
    std::vector<uint8_t> touched(static_cast<size_t>(N), 0);
    for (int64_t e = 0; e < g.n_e; ++e) {
        deg[g.t1[e] + 1] += 2;
        deg[g.t2[e] + 1] += 2;
        touched[g.t1[e]] = touched[g.t2[e]] = 1;
    }
    for (int32_t i = 0; i < N; ++i) deg[i + 1] += 1;  // the diagonal block
    for (int32_t i = 0; i < N; ++i) deg[i + 1] += deg[i];
    std::vector<int32_t> nb(static_cast<size_t>(deg[N]));
    std::vector<int64_t> fill(deg.begin(), deg.end() - 1);
    for (int32_t i = 0; i < N; ++i) nb[fill[i]++] = i;
    for (int64_t e = 0; e < g.n_e; ++e) {
        int32_t a = g.t1[e], b = g.t2[e];
        nb[fill[a]++] = a;
        nb[fill[a]++] = b;
        nb[fill[b]++] = a;
        nb[fill[b]++] = b;
    }
    g.brow_ptr.assign(static_cast<size_t>(N) + 1, 0);
    g.bcol.clear();
    g.bcol.reserve(static_cast<size_t>(deg[N]));
    g.bdiag.resize(N);
    for (int32_t i = 0; i < N; ++i) {
        auto first = nb.begin() + deg[i], last = nb.begin() + deg[i + 1];
        std::sort(first, last);
        last = std::unique(first, last);
        for (auto it = first; it != last; ++it) {
            if (*it == i) g.bdiag[i] = static_cast<int32_t>(g.bcol.size());
            g.bcol.push_back(*it);
        }
        if (g.bcol.size() > static_cast<size_t>(std::numeric_limits<int32_t>::max() / 9))
            return fail(err, GSLAM_ERR_TOO_LARGE, "more than 238M 3x3 blocks");
        g.brow_ptr[i + 1] = static_cast<int32_t>(g.bcol.size());
    }
    g.n_blocks = static_cast<int64_t>(g.bcol.size());

    // SpMV tiles: greedy packing of whole block rows
    g.tile_row.clear();
    g.tile_row.push_back(0);
    {
        int32_t r0 = 0;
        while (r0 < N) {
            int32_t r1 = r0 + 1;
            while (r1 < N && r1 - r0 < kTileRows &&
                   g.brow_ptr[r1 + 1] - g.brow_ptr[r0] <= kTileBlocks)
                ++r1;
            g.tile_row.push_back(r1);
            r0 = r1;
        }
    }

    auto slot = [&](int32_t row, int32_t col) -> int32_t {
        auto first = g.bcol.begin() + g.brow_ptr[row];
        auto last = g.bcol.begin() + g.brow_ptr[row + 1];
        auto it = std::lower_bound(first, last, col);
        return static_cast<int32_t>(it - g.bcol.begin());
    };

    // ---- contribution lists in edge order (stable counting sort) ----
    const int64_t B = g.n_blocks;
    std::vector<int32_t> eslot(static_cast<size_t>(4 * g.n_e));
    g.cptr.assign(static_cast<size_t>(B) + 1, 0);
    for (int64_t e = 0; e < g.n_e; ++e) {
        int32_t a = g.t1[e], b = g.t2[e];
        int32_t s[4] = {g.bdiag[a], slot(a, b), slot(b, a), g.bdiag[b]};
        for (int q = 0; q < 4; ++q) {
            eslot[4 * e + q] = s[q];
            g.cptr[s[q] + 1]++;
        }
    }
    for (int64_t b = 0; b < B; ++b) g.cptr[b + 1] += g.cptr[b];
    g.clist.resize(static_cast<size_t>(4 * g.n_e));
    {
        std::vector<int32_t> f(g.cptr.begin(), g.cptr.end() - 1);
        for (int64_t e = 0; e < g.n_e; ++e)
            for (int q = 0; q < 4; ++q)
                g.clist[f[eslot[4 * e + q]]++] = static_cast<int32_t>(4 * e + q);
    }
    g.xblk.resize(static_cast<size_t>(2 * g.n_e));
    for (int64_t e = 0; e < g.n_e; ++e) g.self_loops = g.self_loops || g.t1[e] == g.t2[e];
    for (int64_t e = 0; e < g.n_e; ++e)
        for (int side = 0; side < 2; ++side) {
            const int32_t b = eslot[4 * e + 1 + side];
            const bool first = g.clist[g.cptr[b]] == static_cast<int32_t>(4 * e + 1 + side);
            g.xblk[2 * e + side] = first ? -(b + 1) : b;
        }
    g.xptr.assign(static_cast<size_t>(N) + 1, 0);
    for (int64_t e = 0; e < g.n_e; ++e) {
        g.xptr[g.t1[e] + 1]++;
        g.xptr[g.t2[e] + 1]++;
    }
    for (int32_t i = 0; i < N; ++i) g.max_incidence = std::max(g.max_incidence, g.xptr[i + 1]);
    for (int32_t i = 0; i < N; ++i) g.xptr[i + 1] += g.xptr[i];
    g.xlist.resize(static_cast<size_t>(2 * g.n_e));
    {
        std::vector<int32_t> f(g.xptr.begin(), g.xptr.end() - 1);
        for (int64_t e = 0; e < g.n_e; ++e) {
            g.xlist[f[g.t1[e]]++] = static_cast<int32_t>(2 * e);
            g.xlist[f[g.t2[e]]++] = static_cast<int32_t>(2 * e + 1);
        }
    }

    // ---- reference CSR layout: only blocks some edge touches, plus the
    // three anchor diagonal entries of pose 0 when no edge touches it.
    std::vector<uint8_t> btouched(static_cast<size_t>(B), 0);
    for (int64_t e = 0; e < 4 * g.n_e; ++e) btouched[eslot[e]] = 1;
    const int64_t dim = 3 * static_cast<int64_t>(N);
    g.csr_row_ptr.assign(static_cast<size_t>(dim) + 1, 0);
    g.csr_pos.assign(static_cast<size_t>(9 * B), -1);
    int64_t pos = 0;
    g.csr_col_idx.clear();
    for (int32_t i = 0; i < N; ++i) {
        int64_t cnt = 0;
        for (int32_t b = g.brow_ptr[i]; b < g.brow_ptr[i + 1]; ++b) cnt += btouched[b];
        for (int a = 0; a < 3; ++a) {
            if (i == 0 && !touched[0]) {
                // anchor only (MatrixOps.c:180-182): one diagonal entry per row
                g.csr_pos[9 * g.bdiag[0] + 4 * a] = static_cast<int32_t>(pos);
                g.csr_col_idx.push_back(a);
                ++pos;
            } else {
                for (int32_t b = g.brow_ptr[i]; b < g.brow_ptr[i + 1]; ++b) {
                    if (!btouched[b]) continue;
                    for (int c = 0; c < 3; ++c) {
                        g.csr_pos[9 * b + 3 * a + c] = static_cast<int32_t>(pos);
                        g.csr_col_idx.push_back(3 * g.bcol[b] + c);
                        ++pos;
                    }
                }
            }
            if (pos > std::numeric_limits<int32_t>::max())
                return fail(err, GSLAM_ERR_TOO_LARGE, "CSR nnz exceeds 2^31-1");
            g.csr_row_ptr[3 * i + a + 1] = static_cast<int32_t>(pos);
        }
        (void)cnt;
    }
    g.nnz = pos;
    return GSLAM_OK;
}

int build_mixed_tiles(const HostGraph &g, MixedTiles &m, std::string &err)
{
    const int32_t N = g.n;
    std::vector<int32_t> rows{0};
    for (int32_t r0 = 0; r0 < N;) {
        int32_t r1 = r0 + 1;
        while (r1 < N && r1 - r0 < kTileRows && g.xptr[r1 + 1] - g.xptr[r0] <= kTileBlocks) ++r1;
        rows.push_back(r1);
        r0 = r1;
    }
    const int64_t T = static_cast<int64_t>(rows.size()) - 1;
    m.ext.assign(static_cast<size_t>(T), -1);
    int64_t chunks = T;
    for (int64_t t = 0; t < T; ++t) {
        const int64_t len = g.xptr[rows[t + 1]] - g.xptr[rows[t]];
        if (len > kTileBlocks) {
            m.ext[t] = static_cast<int32_t>(chunks);
            chunks += (len - 1) / kTileBlocks;
        }
    }
    if (chunks * kTileBlocks > std::numeric_limits<int32_t>::max())
        return fail(err, GSLAM_ERR_TOO_LARGE, "mixed operator: more than 2^31 incidence slots");
    m.n_tiles = static_cast<int32_t>(T);
    m.n_chunks = chunks;
    const size_t slots = static_cast<size_t>(chunks) * kTileBlocks;
    m.info.resize(static_cast<size_t>(4 * T));
    m.ent.assign(slots, -1);
    m.own.assign(slots, 0);
    m.oth.assign(slots, 0);
    m.pos.assign(static_cast<size_t>(2 * g.n_e), -1);
    for (int64_t t = 0; t < T; ++t) {
        const int32_t r0 = rows[t], r1 = rows[t + 1], x0 = g.xptr[r0], x1 = g.xptr[r1];
        m.info[4 * t] = r0;
        m.info[4 * t + 1] = r1;
        m.info[4 * t + 2] = x0;
        m.info[4 * t + 3] = x1;
        const int64_t n_ch = x1 - x0 > kTileBlocks ? 1 + (x1 - x0 - 1) / kTileBlocks : 1;
        for (int64_t c = 0; c < n_ch; ++c) {
            const int64_t chunk = c == 0 ? t : m.ext[t] + c - 1;
            for (int l = 0; l < kTileBlocks; ++l) {
                const size_t slot = static_cast<size_t>(chunk) * kTileBlocks + l;
                m.own[slot] = r0;
                m.oth[slot] = r0;
            }
        }
        for (int32_t k = x0; k < x1; ++k) {
            const int64_t j = k - x0;
            const int64_t chunk = j < kTileBlocks ? t : m.ext[t] + j / kTileBlocks - 1;
            const size_t slot = static_cast<size_t>(chunk) * kTileBlocks + j % kTileBlocks;
            const int32_t ent = g.xlist[k];
            const int64_t e = ent >> 1;
            const bool from = (ent & 1) == 0;  // this pose is the edge's t1
            m.ent[slot] = ent;
            m.own[slot] = from ? g.t1[e] : g.t2[e];
            m.oth[slot] = from ? g.t2[e] : g.t1[e];
            m.pos[ent] = static_cast<int32_t>(slot);
        }
    }
    return GSLAM_OK;
}

}  // namespace gslam
