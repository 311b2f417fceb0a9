// C-ABI of libgslam (include/gslam.h): handle lifecycle, uploads, and the
// GN-step entry points over the assembly / PCG / GJ kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "gslam.h"
#include "gslam_internal.h"
#include "pcg_common.h"

namespace gslam {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

DeviceGraph::~DeviceGraph()
{
    for (void *p : allocs) (void)hipFree(p);
    if (status_host) (void)hipHostFree(status_host);
    if (status_ring) (void)hipHostFree(status_ring);
    if (ev_poll[0]) (void)hipEventDestroy(ev_poll[0]);
    if (ev_poll[1]) (void)hipEventDestroy(ev_poll[1]);
    if (scal_host) (void)hipHostFree(scal_host);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev2) (void)hipEventDestroy(ev2);
    for (hipEvent_t e : ev_phase)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
    dense.release();
    if (vecs) (void)hipFree(vecs);
    if (csr_tmp) (void)hipFree(csr_tmp);
    if (stream) (void)hipStreamDestroy(stream);
}

int DeviceGraph::vecs_reserve()
{
    if (vecs) return GSLAM_OK;
    const size_t count = static_cast<size_t>(4 * dim() + h.n_bet_e + 8);
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&vecs), count * 8);
    if (e != hipSuccess) {
        vecs = nullptr;
        set_error(std::string("hipMalloc(vector work space): ") + hipGetErrorString(e));
        return GSLAM_ERR_NOMEM;
    }
    bytes += static_cast<int64_t>(count * 8);
    return GSLAM_OK;
}

int DeviceGraph::csr_tmp_reserve()
{
    if (csr_tmp) return GSLAM_OK;
    const size_t count = static_cast<size_t>(h.nnz > 0 ? h.nnz : 1);
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&csr_tmp), count * 8);
    if (e != hipSuccess) {
        csr_tmp = nullptr;
        set_error(std::string("hipMalloc(CSR staging): ") + hipGetErrorString(e));
        return GSLAM_ERR_NOMEM;
    }
    bytes += static_cast<int64_t>(count * 8);
    return GSLAM_OK;
}

int DeviceGraph::tval_reserve()
{
    if (tval) return GSLAM_OK;
    return alloc(&tval, static_cast<int64_t>(9 * kTileBlocks) * n_tiles);
}

int DeviceGraph::fused_reserve()
{
    if (fpart) return GSLAM_OK;
    GSLAM_TRY(alloc(&r_alt, dim()));
    GSLAM_TRY(alloc(&q_alt, dim()));
    GSLAM_TRY(alloc(&zwp, 2 * 3 * dim()));
    double *f = nullptr;
    GSLAM_TRY(alloc(&f, static_cast<int64_t>(2 * kFusedParts) * kFusedMaxGroups));
    fpart = f;  // last: fpart != nullptr means every fused buffer exists
    return GSLAM_OK;
}

int DeviceGraph::mixed_factors()
{
    if (factors_valid) return GSLAM_OK;
    if (!from_edges) {
        set_error("mixed PCG needs the edge factors of gslam_assemble (system was set from CSR)");
        return GSLAM_ERR_STATE;
    }
    if (!frec) {
        const int64_t N = h.n, dim = 3 * N, E = h.n_e;
        MixedTiles m;
        std::string err;
        const int rc = build_mixed_tiles(h, m, err);
        if (rc != GSLAM_OK) {
            set_error(err);
            return rc;
        }
        const int64_t slots = m.n_chunks * kTileBlocks;
        GSLAM_TRY(alloc(&dinv32, 9 * N));
        GSLAM_TRY(alloc(&r32, dim));
        GSLAM_TRY(alloc(&z32, dim));
        GSLAM_TRY(alloc(&q32, dim));
        GSLAM_TRY(alloc(&p32[0], dim));
        GSLAM_TRY(alloc(&p32[1], dim));
        GSLAM_TRY(alloc(&mt_info, m.n_tiles));
        GSLAM_TRY(alloc(&mt_ext, m.n_tiles));
        GSLAM_TRY(alloc(&mt_ent, slots));
        GSLAM_TRY(alloc(&mt_own, slots));
        GSLAM_TRY(alloc(&mt_oth, slots));
        GSLAM_TRY(alloc(&mt_pos, 2 * E));
        GSLAM_TRY(upload(mt_info, m.info.data(), static_cast<int64_t>(m.info.size()) * 4));
        GSLAM_TRY(upload(mt_ext, m.ext.data(), static_cast<int64_t>(m.n_tiles) * 4));
        GSLAM_TRY(upload(mt_ent, m.ent.data(), slots * 4));
        GSLAM_TRY(upload(mt_own, m.own.data(), slots * 4));
        GSLAM_TRY(upload(mt_oth, m.oth.data(), slots * 4));
        GSLAM_TRY(upload(mt_pos, m.pos.data(), 2 * E * 4));
        float *f = nullptr;
        GSLAM_TRY(alloc(&f, 8 * slots));
        // padding slots keep zero factors
        GSLAM_HIP(hipMemsetAsync(f, 0, static_cast<size_t>(8 * slots) * 4, stream));
        mt_tiles = m.n_tiles;
        mt_slots = slots;
        frec = f;  // last: frec != nullptr means every mixed buffer exists
    }
    GSLAM_TRY(launch_assembly(*this));
    GSLAM_HIP(hipStreamSynchronize(stream));
    factors_valid = true;
    return GSLAM_OK;
}

int DeviceGraph::upload(void *dst, const void *src, int64_t nbytes)
{
    if (nbytes <= 0) return GSLAM_OK;
    GSLAM_HIP(hipMemcpyAsync(dst, src, static_cast<size_t>(nbytes), hipMemcpyHostToDevice, stream));
    GSLAM_HIP(hipStreamSynchronize(stream));
    return GSLAM_OK;
}

int DeviceGraph::download(void *dst, const void *src, int64_t nbytes)
{
    if (nbytes <= 0) return GSLAM_OK;
    GSLAM_HIP(hipMemcpyAsync(dst, src, static_cast<size_t>(nbytes), hipMemcpyDeviceToHost, stream));
    GSLAM_HIP(hipStreamSynchronize(stream));
    return GSLAM_OK;
}

static int create_device(const gslam_graph_desc &d, DeviceGraph &g)
{
    std::string err;
    int rc = build_host_graph(d, g.h, err);
    if (rc != GSLAM_OK) {
        set_error(err);
        return rc;
    }
    const HostGraph &h = g.h;
    g.lambda = d.lambda;
    g.anchor = d.anchor;
    g.delta = d.delta;
    for (int k = 0; k < 4; ++k) g.mns[k] = d.mns[k];
    g.snr[0] = d.snr[0];
    g.snr[1] = d.snr[1];
    g.wrap = d.wrap ? 1 : 0;
    g.anchor_prior = d.anchor_prior ? 1 : 0;
    if (d.poses && d.n_poses > 0)
        for (int k = 0; k < 3; ++k) g.pose0_ref[k] = d.poses[k];

    GSLAM_HIP(hipGetDevice(&g.dev));
    GSLAM_HIP(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    GSLAM_HIP(hipEventCreate(&g.ev0));
    GSLAM_HIP(hipEventCreate(&g.ev1));
    GSLAM_HIP(hipEventCreate(&g.ev2));
    GSLAM_HIP(hipEventCreate(&g.ev_phase[0]));
    GSLAM_HIP(hipEventCreate(&g.ev_phase[1]));
    GSLAM_HIP(hipEventCreateWithFlags(&g.ev_poll[0], hipEventDisableTiming));
    GSLAM_HIP(hipEventCreateWithFlags(&g.ev_poll[1], hipEventDisableTiming));
    GSLAM_HIP(hipHostMalloc(reinterpret_cast<void **>(&g.status_host), sizeof(PcgStatus)));
    GSLAM_HIP(hipHostMalloc(reinterpret_cast<void **>(&g.status_ring), 2 * sizeof(PcgStatus)));
    GSLAM_HIP(hipHostMalloc(reinterpret_cast<void **>(&g.scal_host), 16 * sizeof(double)));

    // assembly schedule: the pose walk when no pose has more than
    // kWalkMaxIncidence edge ends (its thread walks them one after another)
    // and no edge joins a pose to itself (its four blocks would be one)
    if (d.assembly < GSLAM_ASSEMBLY_AUTO || d.assembly > GSLAM_ASSEMBLY_STAGED) {
        set_error("assembly must be GSLAM_ASSEMBLY_AUTO, _POSE_WALK or _STAGED");
        return GSLAM_ERR_ARG;
    }
    if (d.assembly == GSLAM_ASSEMBLY_POSE_WALK && h.self_loops) {
        set_error("GSLAM_ASSEMBLY_POSE_WALK: an edge joins a pose to itself");
        return GSLAM_ERR_ARG;
    }
    g.walk = d.assembly == GSLAM_ASSEMBLY_POSE_WALK ||
             (d.assembly == GSLAM_ASSEMBLY_AUTO && !h.self_loops && h.max_incidence <= kWalkMaxIncidence);

    const int64_t N = h.n, dim = 3 * N, E = h.n_e, B = h.n_blocks;
    GSLAM_TRY(g.alloc(&g.poses, dim));
    GSLAM_TRY(g.alloc(&g.dx, dim));
    GSLAM_TRY(g.alloc(&g.e_t1, E));
    GSLAM_TRY(g.alloc(&g.e_t2, E));
    GSLAM_TRY(g.alloc(&g.obs_param, 4 * h.n_obs_e));
    GSLAM_TRY(g.alloc(&g.mot_param, 2 * h.n_mot_e));
    GSLAM_TRY(g.alloc(&g.bet_param, 6 * h.n_bet_e));
    GSLAM_TRY(g.alloc(&g.bet_w, h.n_bet_e));
    GSLAM_TRY(g.alloc(&g.brow_ptr, N + 1));
    GSLAM_TRY(g.alloc(&g.bcol, B));
    GSLAM_TRY(g.alloc(&g.bdiag, N));
    if (!g.walk) {
        GSLAM_TRY(g.alloc(&g.cptr, B + 1));
        GSLAM_TRY(g.alloc(&g.clist, 4 * E));
    } else {
        GSLAM_TRY(g.alloc(&g.xblk, 2 * E));
    }
    GSLAM_TRY(g.alloc(&g.xptr, N + 1));
    GSLAM_TRY(g.alloc(&g.xlist, 2 * E));
    GSLAM_TRY(g.alloc(&g.csr_row_ptr, dim + 1));
    GSLAM_TRY(g.alloc(&g.csr_col_idx, h.nnz));
    GSLAM_TRY(g.alloc(&g.csr_pos, 9 * B));
    g.n_tiles = static_cast<int32_t>(h.tile_row.size()) - 1;
    GSLAM_TRY(g.alloc(&g.tile_info, g.n_tiles));
    GSLAM_TRY(g.alloc(&g.tcol, static_cast<int64_t>(kTileBlocks) * g.n_tiles));
    if (!g.walk) GSLAM_TRY(g.alloc(&g.stage, static_cast<int64_t>(kStage) * E));
    GSLAM_TRY(g.alloc(&g.bval, 9 * B));
    GSLAM_TRY(g.alloc(&g.xi, dim));
    GSLAM_TRY(g.alloc(&g.x, dim));
    GSLAM_TRY(g.alloc(&g.r, dim));
    GSLAM_TRY(g.alloc(&g.z, dim));
    GSLAM_TRY(g.alloc(&g.q, dim));
    GSLAM_TRY(g.alloc(&g.p[0], dim));
    GSLAM_TRY(g.alloc(&g.p[1], dim));
    GSLAM_TRY(g.alloc(&g.dinv, 9 * N));
    GSLAM_TRY(g.alloc(&g.part, kPartTotal));
    GSLAM_TRY(g.alloc(&g.status, 1));
    GSLAM_TRY(g.alloc(&g.scal, 16));

    GSLAM_TRY(g.upload(g.poses, d.poses, dim * 8));
    GSLAM_TRY(g.upload(g.e_t1, h.t1.data(), E * 4));
    GSLAM_TRY(g.upload(g.e_t2, h.t2.data(), E * 4));
    GSLAM_TRY(g.upload(g.obs_param, h.obs_param.data(), h.obs_param.size() * 8));
    GSLAM_TRY(g.upload(g.mot_param, h.mot_param.data(), h.mot_param.size() * 8));
    GSLAM_TRY(g.upload(g.bet_param, h.bet_param.data(), h.bet_param.size() * 8));
    {
        std::vector<double> ones(static_cast<size_t>(h.n_bet_e), 1.0);
        GSLAM_TRY(g.upload(g.bet_w, ones.data(), h.n_bet_e * 8));
    }
    GSLAM_TRY(g.upload(g.brow_ptr, h.brow_ptr.data(), (N + 1) * 4));
    GSLAM_TRY(g.upload(g.bcol, h.bcol.data(), B * 4));
    GSLAM_TRY(g.upload(g.bdiag, h.bdiag.data(), N * 4));
    if (!g.walk) {
        GSLAM_TRY(g.upload(g.cptr, h.cptr.data(), (B + 1) * 4));
        GSLAM_TRY(g.upload(g.clist, h.clist.data(), 4 * E * 4));
    } else {
        GSLAM_TRY(g.upload(g.xblk, h.xblk.data(), 2 * E * 4));
    }
    GSLAM_TRY(g.upload(g.xptr, h.xptr.data(), (N + 1) * 4));
    GSLAM_TRY(g.upload(g.xlist, h.xlist.data(), 2 * E * 4));
    GSLAM_TRY(g.upload(g.csr_row_ptr, h.csr_row_ptr.data(), (dim + 1) * 4));
    GSLAM_TRY(g.upload(g.csr_col_idx, h.csr_col_idx.data(), h.nnz * 4));
    GSLAM_TRY(g.upload(g.csr_pos, h.csr_pos.data(), 9 * B * 4));
    {
        std::vector<int4> ti(static_cast<size_t>(g.n_tiles));
        for (int32_t t = 0; t < g.n_tiles; ++t) {
            const int32_t r0 = h.tile_row[t], r1 = h.tile_row[t + 1];
            ti[t] = make_int4(r0, r1, h.brow_ptr[r0], h.brow_ptr[r1]);
        }
        GSLAM_TRY(g.upload(g.tile_info, ti.data(), static_cast<int64_t>(ti.size()) * 16));
        std::vector<int32_t> tc(static_cast<size_t>(kTileBlocks) * g.n_tiles);
        for (int32_t t = 0; t < g.n_tiles; ++t) {
            const int32_t b0 = h.brow_ptr[h.tile_row[t]], b1 = h.brow_ptr[h.tile_row[t + 1]];
            for (int32_t l = 0; l < kTileBlocks; ++l)
                tc[static_cast<size_t>(t) * kTileBlocks + l] = h.bcol[b0 + l < b1 ? b0 + l : b0];
        }
        GSLAM_TRY(g.upload(g.tcol, tc.data(), static_cast<int64_t>(tc.size()) * 4));
    }
    return GSLAM_OK;
}

}  // namespace gslam

using gslam::DeviceGraph;
using gslam::set_error;

extern "C" {

void gslam_default_opts(gslam_solve_opts *o)
{
    if (!o) return;
    o->solver = GSLAM_SOLVER_PCG_FP64;
    o->tol = 1e-6;
    o->max_iter = 500;
    o->refine_sweeps = 3;
    o->refresh_every = 64;
    o->check_every = 0;
    o->refine_mode = GSLAM_REFINE_IR;
    o->profile_every = 0;
    o->mixed_handover = 0.0;
    o->pcg_schedule = GSLAM_PCG_AUTO;
}

void gslam_default_desc(gslam_graph_desc *d)
{
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->delta = 1.0;
    d->lambda = 0.1;
    d->mns[0] = 0.19;
    d->mns[1] = 0.001;
    d->mns[2] = 0.13;
    d->mns[3] = 0.2;
    d->snr[0] = 0.14;
    d->snr[1] = 0.05;
    d->anchor = 1e6;
    d->pair_order = 0;
}

const char *gslam_last_error(void) { return gslam::g_last_error.c_str(); }

int gslam_create(const gslam_graph_desc *desc, gslam_graph **out)
{
    if (!desc || !out) {
        set_error("gslam_create: NULL argument");
        return GSLAM_ERR_ARG;
    }
    *out = nullptr;
    gslam_graph *h = new (std::nothrow) gslam_graph();
    if (!h) return GSLAM_ERR_NOMEM;
    int rc = gslam::create_device(*desc, h->g);
    if (rc != GSLAM_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return GSLAM_OK;
}

int gslam_destroy(gslam_graph *g)
{
    delete g;
    return GSLAM_OK;
}

int gslam_get_info(const gslam_graph *gh, gslam_graph_info *info)
{
    if (!gh || !info) return GSLAM_ERR_ARG;
    const auto &h = gh->g.h;
    info->n_poses = h.n;
    info->n_edges = h.n_e;
    info->n_obs_edges = h.n_obs_e;
    info->n_motion_edges = h.n_mot_e;
    info->n_between_edges = h.n_bet_e;
    info->n_blocks = h.n_blocks;
    info->nnz = h.nnz;
    info->device_bytes = gh->g.bytes + gh->g.dense.bytes + gh->g.dense64.bytes;
    info->n_tiles = gh->g.n_tiles;
    return GSLAM_OK;
}

int gslam_set_poses(gslam_graph *gh, const double *poses)
{
    if (!gh || !poses) return GSLAM_ERR_ARG;
    DeviceGraph &g = gh->g;
    for (int64_t k = 0; k < g.dim(); ++k)
        if (!std::isfinite(poses[k])) {
            set_error("gslam_set_poses: pose value " + std::to_string(k) + " is not finite");
            return GSLAM_ERR_ARG;
        }
    g.assembled = g.solved = false;
    return g.upload(g.poses, poses, 3 * static_cast<int64_t>(g.h.n) * 8);
}

int gslam_get_poses(gslam_graph *gh, double *poses)
{
    if (!gh || !poses) return GSLAM_ERR_ARG;
    DeviceGraph &g = gh->g;
    return g.download(poses, g.poses, 3 * static_cast<int64_t>(g.h.n) * 8);
}

int gslam_obs_cost(gslam_graph *gh, double *cost)
{
    if (!gh || !cost) return GSLAM_ERR_ARG;
    DeviceGraph &g = gh->g;
    std::vector<double> poses(3 * static_cast<size_t>(g.h.n));
    GSLAM_TRY(g.download(poses.data(), g.poses, 3 * static_cast<int64_t>(g.h.n) * 8));
    *cost = obs_cost_terms(g.h, poses.data(), g.snr, *cost);
    return GSLAM_OK;
}

int gslam_set_lambda(gslam_graph *gh, double lambda)
{
    if (!gh) return GSLAM_ERR_ARG;
    if (!std::isfinite(lambda)) {
        set_error("gslam_set_lambda: lambda is not finite");
        return GSLAM_ERR_ARG;
    }
    gh->g.lambda = lambda;
    gh->g.assembled = gh->g.solved = false;
    return GSLAM_OK;
}

int gslam_set_log_weights(gslam_graph *gh, const double *log_w)
{
    if (!gh) return GSLAM_ERR_ARG;
    DeviceGraph &g = gh->g;
    std::vector<double> w(static_cast<size_t>(g.h.n_bet_e), 1.0);
    if (log_w)
        for (size_t k = 0; k < w.size(); ++k) {
            w[k] = std::exp(log_w[k]);
            if (!std::isfinite(w[k])) {
                set_error("gslam_set_log_weights: exp(log_w[" + std::to_string(k) + "]) is not finite");
                return GSLAM_ERR_ARG;
            }
        }
    g.assembled = g.solved = false;
    return g.upload(g.bet_w, w.data(), g.h.n_bet_e * 8);
}

int gslam_assemble(gslam_graph *gh)
{
    if (!gh) return GSLAM_ERR_ARG;
    DeviceGraph &g = gh->g;
    GSLAM_HIP(hipEventRecord(g.ev_phase[0], g.stream));
    GSLAM_TRY(gslam::launch_assembly(g));
    GSLAM_HIP(hipEventRecord(g.ev_phase[1], g.stream));
    GSLAM_HIP(hipStreamSynchronize(g.stream));
    float ms = 0.f;
    GSLAM_HIP(hipEventElapsedTime(&ms, g.ev_phase[0], g.ev_phase[1]));
    g.ms_assembly = ms;
    g.assembled = true;
    g.from_edges = true;
    g.factors_valid = g.frec != nullptr;
    g.solved = false;
    return GSLAM_OK;
}

int gslam_set_system_csr(gslam_graph *gh, const double *values, const double *xi)
{
    if (!gh || !values || !xi) return GSLAM_ERR_ARG;
    DeviceGraph &g = gh->g;
    const int64_t nnz = g.h.nnz;
    GSLAM_TRY(g.csr_tmp_reserve());
    int rc = g.upload(g.csr_tmp, values, nnz * 8);
    if (rc == GSLAM_OK) rc = gslam::launch_csr_to_bsr(g, g.csr_tmp, g.bval);
    if (rc == GSLAM_OK) rc = g.upload(g.xi, xi, 3 * static_cast<int64_t>(g.h.n) * 8);
    (void)hipStreamSynchronize(g.stream);
    if (rc != GSLAM_OK) return rc;
    g.assembled = true;
    g.from_edges = false;
    g.factors_valid = false;
    g.solved = false;
    return GSLAM_OK;
}

int gslam_time_kernel(gslam_graph *gh, int32_t which, int32_t reps, double *us)
{
    if (!gh || !us || reps <= 0) return GSLAM_ERR_ARG;
    DeviceGraph &g = gh->g;
    if (which == GSLAM_KERNEL_GJ_TRAILING) {
        if (!g.dense.a) {
            set_error("gslam_time_kernel: the GJ trailing update needs a GJ_MIXED solve first");
            return GSLAM_ERR_STATE;
        }
        return gslam::bgj_time_trailing(g.stream, g.dense, reps, g.ev0, g.ev1, us);
    }
    if (!g.assembled) {
        set_error("gslam_time_kernel: call gslam_assemble first");
        return GSLAM_ERR_STATE;
    }
    return gslam::time_kernel(g, which, reps, us);
}

int gslam_get_system_csr(gslam_graph *gh, int32_t *row_ptr, int32_t *col_idx, double *values,
                         double *xi)
{
    if (!gh) return GSLAM_ERR_ARG;
    DeviceGraph &g = gh->g;
    const auto &h = g.h;
    if ((values || xi) && !g.assembled) {
        set_error("gslam_get_system_csr: call gslam_assemble first");
        return GSLAM_ERR_STATE;
    }
    if (row_ptr) std::memcpy(row_ptr, h.csr_row_ptr.data(), h.csr_row_ptr.size() * 4);
    if (col_idx) std::memcpy(col_idx, h.csr_col_idx.data(), h.csr_col_idx.size() * 4);
    if (values) {
        GSLAM_TRY(g.csr_tmp_reserve());
        GSLAM_TRY(gslam::launch_bsr_to_csr(g, g.bval, g.csr_tmp));
        GSLAM_TRY(g.download(values, g.csr_tmp, h.nnz * 8));
    }
    if (xi) GSLAM_TRY(g.download(xi, g.xi, 3 * static_cast<int64_t>(h.n) * 8));
    return GSLAM_OK;
}

int gslam_get_system_bsr(gslam_graph *gh, int32_t *brow_ptr, int32_t *bcol, double *bval,
                         double *xi)
{
    if (!gh) return GSLAM_ERR_ARG;
    DeviceGraph &g = gh->g;
    const auto &h = g.h;
    if ((bval || xi) && !g.assembled) {
        set_error("gslam_get_system_bsr: call gslam_assemble first");
        return GSLAM_ERR_STATE;
    }
    if (brow_ptr) std::memcpy(brow_ptr, h.brow_ptr.data(), h.brow_ptr.size() * 4);
    if (bcol) std::memcpy(bcol, h.bcol.data(), h.bcol.size() * 4);
    if (bval) GSLAM_TRY(g.download(bval, g.bval, 9 * h.n_blocks * 8));
    if (xi) GSLAM_TRY(g.download(xi, g.xi, 3 * static_cast<int64_t>(h.n) * 8));
    return GSLAM_OK;
}

static int check_opts(const gslam_solve_opts *o)
{
    if (!o) return GSLAM_ERR_ARG;
    if (o->solver < GSLAM_SOLVER_PCG_FP64 || o->solver > GSLAM_SOLVER_GJ_BLOCKED64) {
        set_error("unknown solver id");
        return GSLAM_ERR_ARG;
    }
    if (!(o->tol >= 0.0) || o->max_iter < 0) {
        set_error("tol must be >= 0 and max_iter >= 0");
        return GSLAM_ERR_ARG;
    }
    if (!(o->mixed_handover >= 0.0)) {
        set_error("mixed_handover must be >= 0");
        return GSLAM_ERR_ARG;
    }
    if (o->pcg_schedule != GSLAM_PCG_AUTO && o->pcg_schedule != GSLAM_PCG_TWO_LAUNCH) {
        set_error("pcg_schedule must be GSLAM_PCG_AUTO or GSLAM_PCG_TWO_LAUNCH");
        return GSLAM_ERR_ARG;
    }
    if (o->solver == GSLAM_SOLVER_GJ_MIXED &&
        (o->refine_sweeps < 0 || (o->refine_mode != GSLAM_REFINE_IR && o->refine_mode != GSLAM_REFINE_CG))) {
        set_error("GJ_MIXED: refine_sweeps must be >= 0 and refine_mode IR or CG");
        return GSLAM_ERR_ARG;
    }
    return GSLAM_OK;
}

// A NaN or Inf met by a solver's recurrence (flags bit 1) or left in the
// solution (a non-finite true residual) fails the solve with
// GSLAM_ERR_NONFINITE; the stats stay valid.
static int nonfinite(gslam_solve_stats &s)
{
    if (!(s.flags & 2) && std::isfinite(s.true_rel_res)) return GSLAM_OK;
    s.flags |= 2;
    s.converged = 0;
    set_error("NaN or Inf in the system, the right-hand side or the solution");
    return GSLAM_ERR_NONFINITE;
}

static int solve_into(DeviceGraph &g, const gslam_solve_opts &o, const double *b, double *x,
                      gslam_solve_stats &s)
{
    std::memset(&s, 0, sizeof(s));
    s.ms_assembly = g.ms_assembly;
    if (o.solver == GSLAM_SOLVER_GJ_FP64 || o.solver == GSLAM_SOLVER_GJ_MIXED ||
        o.solver == GSLAM_SOLVER_GJ_BLOCKED64) {
        const int rc = gslam::dense_solve(g, o, b, x, s);
        s.ms_assembly = g.ms_assembly;
        s.ms_setup = s.ms_factor;
        return rc == GSLAM_OK ? nonfinite(s) : rc;
    }
    if (o.solver == GSLAM_SOLVER_PCG_MIXED) GSLAM_TRY(g.mixed_factors());
    GSLAM_TRY(gslam::pcg_solve(g, o, b, x, s));
    GSLAM_TRY(nonfinite(s));
    if (!s.converged && !(s.flags & 1) && o.max_iter > 0 && s.iters >= o.max_iter)
        return GSLAM_ERR_NOT_CONVERGED;
    return GSLAM_OK;
}

int gslam_solve(gslam_graph *gh, const gslam_solve_opts *opts, gslam_solve_stats *stats)
{
    if (!gh) return GSLAM_ERR_ARG;
    GSLAM_TRY(check_opts(opts));
    DeviceGraph &g = gh->g;
    if (!g.assembled) {
        set_error("gslam_solve: call gslam_assemble first");
        return GSLAM_ERR_STATE;
    }
    gslam_solve_stats s;
    int rc = solve_into(g, *opts, g.xi, g.dx, s);
    if (stats) *stats = s;
    if (rc == GSLAM_OK || rc == GSLAM_ERR_NOT_CONVERGED) g.solved = true;
    return rc;
}

int gslam_solve_rhs(gslam_graph *gh, const gslam_solve_opts *opts, const double *rhs, double *x,
                    gslam_solve_stats *stats)
{
    if (!gh || !rhs || !x) return GSLAM_ERR_ARG;
    GSLAM_TRY(check_opts(opts));
    DeviceGraph &g = gh->g;
    if (!g.assembled) {
        set_error("gslam_solve_rhs: call gslam_assemble first");
        return GSLAM_ERR_STATE;
    }
    const int64_t dim = g.dim();
    GSLAM_TRY(g.vecs_reserve());
    double *b = g.vecs, *xd = g.vecs + dim;
    gslam_solve_stats s;
    int rc = g.upload(b, rhs, dim * 8);
    if (rc == GSLAM_OK) rc = solve_into(g, *opts, b, xd, s);
    if (rc == GSLAM_OK || rc == GSLAM_ERR_NOT_CONVERGED) {
        int rc2 = g.download(x, xd, dim * 8);
        if (rc2 != GSLAM_OK) rc = rc2;
    }
    if (stats) *stats = s;
    return rc;
}

int gslam_get_dx(gslam_graph *gh, double *dx)
{
    if (!gh || !dx) return GSLAM_ERR_ARG;
    DeviceGraph &g = gh->g;
    if (!g.solved) {
        set_error("gslam_get_dx: no solution yet");
        return GSLAM_ERR_STATE;
    }
    return g.download(dx, g.dx, 3 * static_cast<int64_t>(g.h.n) * 8);
}

int gslam_apply_update(gslam_graph *gh, double *diff_out)
{
    if (!gh) return GSLAM_ERR_ARG;
    DeviceGraph &g = gh->g;
    if (!g.solved) {
        set_error("gslam_apply_update: no solution yet");
        return GSLAM_ERR_STATE;
    }
    GSLAM_TRY(gslam::pose_update(g, diff_out));
    g.assembled = g.solved = false;
    return GSLAM_OK;
}

int gslam_gn_step(gslam_graph *gh, const gslam_solve_opts *opts, gslam_solve_stats *stats,
                  double *diff_out)
{
    GSLAM_TRY(gslam_assemble(gh));
    int rc = gslam_solve(gh, opts, stats);
    if (rc != GSLAM_OK && rc != GSLAM_ERR_NOT_CONVERGED) return rc;
    DeviceGraph &g = gh->g;
    GSLAM_HIP(hipEventRecord(g.ev_phase[0], g.stream));
    int rc2 = gslam_apply_update(gh, diff_out);
    if (rc2 == GSLAM_OK && stats) {
        GSLAM_HIP(hipEventRecord(g.ev_phase[1], g.stream));
        GSLAM_HIP(hipEventSynchronize(g.ev_phase[1]));
        float ms = 0.f;
        GSLAM_HIP(hipEventElapsedTime(&ms, g.ev_phase[0], g.ev_phase[1]));
        stats->ms_update = ms;
    }
    return rc2 != GSLAM_OK ? rc2 : rc;
}

}  // extern "C"
