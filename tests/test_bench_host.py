"""Host-side pieces of bench.py that need no GPU: the algorithmic byte counts
the roofline figures use, and the reverse Cuthill-McKee relabelling behind
`--order rcm`.  The relabelled graph must be the same graph (a permutation of
the pose labels carrying poses and edge ends together), with nearby labels
on every edge."""
import numpy as np

import bench
from gslam import generators as G


def test_byte_counts():
    # per block 9 FP64 values + an int32 column; per row the int32 pointer,
    # z, p_old and x read and p_new, q and x written (6 x 3 doubles)
    assert bench.spmv_bytes(1, 1) == 76 + 8 + 144
    assert bench.spmv_bytes(100000, 430854) == 47544908
    # per pose r, q and D^-1 read, r and z written
    assert bench.update_bytes(100000) == 16800000
    # one-launch iteration: the SpMV's blocks and row pointers; per pose the
    # record {z, w, p}, r/q/x and D^-1 read, r/q/x and the record written
    assert bench.fused_bytes(1, 1) == 76 + 8 + 5 * 72
    assert bench.fused_bytes(100000, 430854) == 69144908
    # per incidence entry its edge, other pose and 8 FP32 factors; per pose
    # z and p_old read and p_new and q written (4 x 3 floats)
    assert bench.mixed_spmv_bytes(1, 2) == 2 * 40 + 48


def test_rcm_relabel_is_a_permutation():
    gc = G.grid_city(3000, seed=2)
    out = bench.rcm_relabel(gc)
    n = len(gc["poses"]) // 3
    p0 = gc["poses"].reshape(-1, 3)
    p1 = out["poses"].reshape(-1, 3)
    # every pose appears once under its new label: recover the permutation
    # from the edges and check poses and edge ends move together
    new_of = np.full(n, -1)
    new_of[gc["i"]] = out["i"]
    new_of[gc["j"]] = out["j"]
    assert (new_of >= 0).all() and len(np.unique(new_of)) == n
    assert np.array_equal(p1[new_of], p0)
    assert np.array_equal(out["meas"], gc["meas"])
    # nearby labels: the mean label distance over the edges shrinks
    assert np.abs(out["i"] - out["j"]).mean() < 0.2 * np.abs(gc["i"] - gc["j"]).mean()
