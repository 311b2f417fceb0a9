// Host-side graph structure of libgslam: the edge list in summation order,
// the 3x3 block-sparse (BSR) pattern of Omega, the per-block contribution
// lists that make assembly a deterministic gather, and the map to the
// reference's scalar CSR layout.  Pure C++ (no HIP), built once per graph.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "gslam.h"

namespace gslam {

// Quadrant ids of an edge's four 3x3 blocks: (t1,t1) (t1,t2) (t2,t1) (t2,t2).
enum { Q_UL = 0, Q_UR = 1, Q_BL = 2, Q_BR = 3 };
// Per-edge staged values: 4 blocks x 9 + 2 x 3 xi = 42 doubles, padded to 48
// (384 bytes = three 128-byte cache lines per edge record).
constexpr int kStage = 48;
// SpMV tile limits: 256 blocks (2304 values, one per thread of a 256-thread
// workgroup in 9 passes) and 85 block rows (3 outputs per row <= 256 threads).
constexpr int kTileBlocks = 256;
constexpr int kTileRows = 85;
// Assembly by pose walk (assembly.hip) when no pose has more edge ends.
constexpr int kWalkMaxIncidence = 64;

struct HostGraph {
    int32_t n = 0;                 // poses
    int64_t n_obs_e = 0, n_mot_e = 0, n_bet_e = 0, n_e = 0;
    std::vector<int32_t> t1, t2;   // n_e, summation order
    std::vector<double> obs_param; // 4 per obs edge: r1 phi1 r2 phi2
    std::vector<double> mot_param; // 2 per motion edge: nu omega (u[i+1])
    std::vector<double> bet_param; // 6 per between edge: meas[3] info[3]

    // BSR pattern: every pose owns its diagonal block; off-diagonal blocks
    // for each distinct neighbour; columns sorted within a block row.
    std::vector<int32_t> brow_ptr; // n+1
    std::vector<int32_t> bcol;     // n_blocks
    std::vector<int32_t> bdiag;    // n: index of block (i,i)
    int64_t n_blocks = 0;

    // Contributions: block b receives stage quadrants clist[cptr[b]..cptr[b+1])
    // (entries e*4+q in edge order); pose i receives xi halves
    // xlist[xptr[i]..xptr[i+1]) (entries e*2+side in edge order).
    std::vector<int32_t> cptr, clist, xptr, xlist;
    // Per incidence entry e*2+side: the off-diagonal block it adds to (UR of
    // (t1,t2) for side 0, BL of (t2,t1) for side 1), stored as -(b+1) when it
    // is that block's first contribution in edge order; and the largest
    // number of incidence entries of one pose.
    std::vector<int32_t> xblk;
    int32_t max_incidence = 0;
    bool self_loops = false;  // an edge with t1 == t2 (its four blocks are one)

    // SpMV tiles: tile t covers block rows [tile_row[t], tile_row[t+1]),
    // whole rows, at most kTileBlocks blocks and kTileRows rows unless a
    // single row is longer (then it is a tile of its own).
    std::vector<int32_t> tile_row;

    // Reference CSR (src/cpu_pcg_fp64/MatrixOps.c:173-221 pattern): row_ptr,
    // col_idx and, per BSR value slot (9*b+k), its CSR position or -1.
    std::vector<int32_t> csr_row_ptr, csr_col_idx, csr_pos;
    int64_t nnz = 0;
};

// Tiles of the mixed solver's factorized FP32 operator (pcg_mixed.hip): the
// incidence lists xlist cut into tiles of whole poses, greedy, at most
// kTileBlocks entries and kTileRows poses unless one pose has more entries
// (then it is a tile of its own).  Entry j of tile t sits in slot
// 256 t + j (j < 256, "chunk" t); a longer pose continues in extra chunks
// numbered from n_tiles on (ext[t] is its first, -1 for a one-chunk tile).
// Per slot: the entry e*2+side (-1 in padding), its own pose and the other
// pose of the edge (padding: the tile's first pose); pos[2e+side] = slot.
struct MixedTiles {
    int32_t n_tiles = 0;
    int64_t n_chunks = 0;
    std::vector<int32_t> info;  // 4 per tile: first pose, end pose, xptr of both
    std::vector<int32_t> ext;   // per tile
    std::vector<int32_t> ent, own, oth;  // per slot (256 per chunk)
    std::vector<int32_t> pos;   // 2 per edge
};
int build_mixed_tiles(const HostGraph &g, MixedTiles &m, std::string &err);

// Validates the description and builds everything above.  Returns GSLAM_OK
// or a negative status with a message in *err.
int build_host_graph(const gslam_graph_desc &d, HostGraph &g, std::string &err);

// Landmark-pair enumeration (observation index pairs j<k per landmark).
// order 0: landmarks by first appearance (src/cpu_pcg_fp64/GSLAM.c:246-276);
// order 1: ascending landmark id (src/gpu_pcg_fp64/device_graph.cu:505-551).
// Observations keep file order inside a landmark in both.
void landmark_pairs(int32_t n_obs, const int32_t *z_lm, int order,
                    std::vector<int32_t> &a, std::vector<int32_t> &b);

// acc + the landmark-pair terms of the reference Gauss-Jordan drivers' cost
// (src/gpu_gj_fp64/GSLAM.cu:245-257) at `poses` (3n doubles), summed onto
// acc in edge order (reference_cost.cpp).
double obs_cost_terms(const HostGraph &h, const double *poses, const double snr[2], double acc);

}  // namespace gslam
