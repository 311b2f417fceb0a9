"""Landmark estimates after the GN loop (GSLAM.c:141-196, MapEdge.c:7-69).

* the oracle restatement equals the reference's own make_edges +
  MapEdge_create + vec_mean (compiled in oracle/_ref) bit for bit, on the
  reference data at the poses after one reference round;
* gslam.landmark_means (NumPy, no oracle) equals the oracle to 1e-12;
* closed form: one observation straight ahead at range r from a pose at the
  origin with heading th places the landmark at r (cos th, sin th).
"""
import numpy as np
import pytest

from conftest import GOLDEN_SETS, data_path, golden

import gslam
from oracle import ref_available


@pytest.mark.parametrize("key", ["ct100", "vp100", "ct800"])
def test_oracle_against_reference_and_numpy(oracle, key):
    g = oracle.read_log(data_path(GOLDEN_SETS[key]))
    poses = golden(key)["poses_r1"][:, 1:].ravel()
    ids_o, m_o = oracle.landmark_means(g, poses)
    assert len(ids_o) == len(np.unique(g["z_lm"]))
    if ref_available():
        from oracle import RefPCG64
        ids_r, m_r = RefPCG64().landmark_means(g, poses)
        assert np.array_equal(ids_o, ids_r) and np.array_equal(m_o, m_r)
    ids, m = gslam.landmark_means(poses, g["z_step"], g["z_lm"], g["z_range"], g["z_bearing"])
    assert np.array_equal(ids, ids_o)
    assert np.abs(m - m_o).max() <= 1e-12 * max(1.0, np.abs(m_o).max())


def test_closed_form_and_order():
    poses = np.array([0.0, 0.0, 0.3, 1.0, 2.0, 0.0])
    ids, m = gslam.landmark_means(poses, [0, 1, 0], [7, 3, 7], [2.0, 1.0, 4.0], [0.0, 0.5, 0.0])
    assert list(ids) == [7, 3]                       # first appearance, not sorted
    assert np.allclose(m[0], 3.0 * np.array([np.cos(0.3), np.sin(0.3)]))   # mean of r = 2, 4
    assert np.allclose(m[1], [1.0 + np.cos(0.5), 2.0 + np.sin(0.5)])
    ids, m = gslam.landmark_means(poses, [], [], [], [])
    assert ids.size == 0 and m.shape == (0, 2)


def test_resume_log_rewrites_only_the_x_records(tmp_path):
    path = data_path("cityTrees100.log.txt")
    d = gslam.read_log(path)
    rng = np.random.default_rng(1)
    p = d["poses"].reshape(-1, 3) + rng.normal(size=(d["n_poses"], 3))
    dump = tmp_path / "dump.txt"
    with open(dump, "w") as f:
        for s, row in zip(d["x_step"], p):
            f.write("%d %.17g %.17g %.17g\n" % (s, *row))
    out = str(tmp_path / "out.log.txt")
    assert gslam.resume_log(path, str(dump), out) == d["n_poses"]
    e = gslam.read_log(out)
    assert np.array_equal(e["poses"], p.ravel())
    for k in ("u_nu", "u_omega", "z_step", "z_lm", "z_range", "z_bearing", "x_step"):
        assert np.array_equal(e[k], d[k]), k
    assert e["delta"] == d["delta"]
    with open(dump, "a") as f:
        f.write("999 0 0 0\n")
    with pytest.raises(ValueError):
        gslam.resume_log(path, str(dump), out)
