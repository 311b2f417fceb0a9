"""The CPU oracle against the golden fixtures written by the reference build
(tests/golden/make_golden.py).  Every comparison is bitwise: the oracle
restates the reference C operation by operation."""
import numpy as np
import pytest

from conftest import GOLDEN_SETS, data_path, golden


@pytest.mark.parametrize("key", list(GOLDEN_SETS))
def test_round1_edges_csr_pcg(oracle, key):
    fx = golden(key)
    g = oracle.read_log(data_path(GOLDEN_SETS[key]))
    n = g["n_poses"]
    t1, t2, ev = oracle.log_edges(g, lam=float(fx["lam"]))
    assert np.array_equal(t1, fx["t1"]) and np.array_equal(t2, fx["t2"])
    if "edge_vals" in fx.files:
        assert np.array_equal(ev, fx["edge_vals"])
    rp, ci, va, xi = oracle.assemble_csr(n, t1, t2, ev)
    assert np.array_equal(rp, fx["row_ptr"])
    assert np.array_equal(ci, fx["col_idx"])
    assert np.array_equal(va, fx["values"])
    assert np.array_equal(xi, fx["xi"])
    x, it = oracle.pcg(rp, ci, va, xi, 1e-6, 500)
    assert np.array_equal(x, fx["pcg_x"])


@pytest.mark.parametrize("key", ["ct100", "vp100"])
def test_dense_gj(oracle, key):
    fx = golden(key)
    g = oracle.read_log(data_path(GOLDEN_SETS[key]))
    n = g["n_poses"]
    t1, t2, ev = oracle.log_edges(g, lam=float(fx["lam"]))
    om, xi = oracle.dense(n, t1, t2, ev)
    inv, rc = oracle.gj_inv(om)
    assert rc == 0
    assert np.array_equal(np.diag(inv), fx["gj_inv_diag"])
    assert np.array_equal(inv[:3], fx["gj_inv_rows012"])
    assert np.array_equal(oracle.dense_matvec(inv, xi), fx["gj_dx"])
    inv32, rc = oracle.gj_inv_f32(om.astype(np.float32))
    assert rc == 0
    assert np.array_equal(np.diag(inv32), fx["gj32_inv_diag"])
    assert np.array_equal(oracle.dense_matvec(inv32.astype(np.float64), xi), fx["gj32_dx"])


def _stdout_diffs(text):
    out = []
    for line in str(text).splitlines():
        if "rounds executed:" in line:
            out.append(float(line.split(":")[1]))
    return np.array(out)


@pytest.mark.parametrize("key", list(GOLDEN_SETS))
@pytest.mark.parametrize("rounds", [1, 3])
def test_gn_rounds_match_reference_executable(oracle, key, rounds):
    """orc_log_gn_run == reference GSLAM executable (stdout diffs and the
    POSE_DUMP_FILE poses, printed with %.17g so the round trip is exact)."""
    fx = golden(key)
    g = oracle.read_log(data_path(GOLDEN_SETS[key]))
    poses, diffs = oracle.log_gn_run(g, float(fx["lam"]), rounds)
    ref_diffs = _stdout_diffs(fx[f"stdout_r{rounds}"])
    assert np.array_equal(diffs, ref_diffs)
    ref = fx[f"poses_r{rounds}"]
    assert np.array_equal(ref[:, 0].astype(int), g["x_step"])
    assert np.array_equal(poses.reshape(-1, 3), ref[:, 1:])


@pytest.mark.parametrize("key", list(GOLDEN_SETS))
def test_header_lines(oracle, key):
    """Record counts the reference prints (GSLAM.c:45-49) follow from the
    reader: dim = #u * 3."""
    fx = golden(key)
    g = oracle.read_log(data_path(GOLDEN_SETS[key]))
    line = str(fx["stdout_r1"]).splitlines()[1]
    expect = (f"delta: {g['delta']:f} size_of_us: {len(g['u_nu'])} size_of_hat_xs: "
              f"{g['n_poses']} size_of_zlist: {len(g['z_step'])} dim: {3 * len(g['u_nu'])}")
    assert line == expect
