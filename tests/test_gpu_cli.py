"""GSLAM, the drop-in command-line driver: the Gauss-Jordan drivers' stop
rules, cost lines and pose dump (src/gpu_gj_fp64/GSLAM.cu:193-591, 757-770),
and pose graphs given as a between-edge sidecar file."""
import os
import subprocess

import numpy as np
import pytest

from conftest import data_path

import gslam
from gslam import generators as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GSLAM_BIN = os.path.join(ROOT, "g-slam_on_gpu_and_fpga_amd", "bin", "GSLAM")


def _gslam(args, env=None, timeout=300):
    out = subprocess.run([GSLAM_BIN] + [str(a) for a in args], capture_output=True, text=True,
                         timeout=timeout, env=dict(os.environ, **(env or {})))
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout.splitlines()


def _diffs(lines):
    return [float(l.split(":")[1]) for l in lines if "rounds executed" in l]


@pytest.mark.parametrize("name", ["cityTrees100.log.txt", "victoria-park-100.log.txt"])
def test_obs_cost_matches_oracle_bitwise(oracle, name):
    path = data_path(name)
    d = oracle.read_log(path)
    rng = np.random.default_rng(0)
    with gslam.Graph.from_log(path, lam=0.1) as g:
        c = g.obs_cost()
        c_o = oracle.gj_cost(d, d["poses"])
        assert c == c_o
        for _ in range(3):
            p = d["poses"] + rng.normal(size=d["poses"].size) * 0.05
            g.set_poses(p)
            c = g.obs_cost(c)
            c_o = oracle.gj_cost(d, p, acc=c_o)
            assert c == c_o


def test_gj_use_diff0_prints_the_running_cost(oracle, tmp_path):
    path = data_path("cityTrees100.log.txt")
    d = oracle.read_log(path)
    dump1 = str(tmp_path / "p1.txt")
    # one round, drawing 0: the GJ drivers dump the poses all the same
    _gslam([0, 0.1, 1, 0, path], env=dict(GSLAM_SOLVER="gj64", POSE_DUMP_FILE=dump1))
    p1 = np.loadtxt(dump1)[:, 1:].ravel()
    lines = _gslam([0, 0.1, 3, 0, path], env=dict(GSLAM_SOLVER="gj64"))
    c1 = oracle.gj_cost(d, d["poses"])
    c2 = oracle.gj_cost(d, p1, acc=c1)
    assert lines[2] == "1 rounds executed: 1"
    assert lines[3] == "last_cost: 0"
    assert lines[4] == f"cost: {c1:.17g}"           # printf %.17g, as the reference prints it
    assert float(lines[5].split(":")[1]) == abs(c2 - c1) / c2
    assert float(lines[6].split(":")[1]) == c1 and lines[6].startswith("last_cost: ")
    assert float(lines[7].split(":")[1]) == c2 and lines[7].startswith("cost: ")
    assert len(lines) == 2 + 3 * 3


def _oracle_gj_diffs(oracle, d, rounds, lam=0.1):
    """FP64 Gauss-Jordan GN rounds on the CPU oracle (reference edges, dense
    Omega, inv(), Omega^-1 xi, pose update; the pieces pinned bitwise to
    the reference build)."""
    poses = d["poses"].copy()
    out = []
    for _ in range(rounds):
        t1, t2, ev = oracle.log_edges(d, poses, lam)
        om, xi = oracle.dense(d["n_poses"], t1, t2, ev)
        inv, _ = oracle.gj_inv(om)
        dx = oracle.dense_matvec(inv, xi)
        poses = oracle.pose_update(poses, dx)
        out.append(oracle.diff(dx))
    return out


def test_gj_use_diff1_plateau_stop(oracle, tmp_path):
    path = data_path("cityTrees100.log.txt")
    rounds = 80
    dump = str(tmp_path / "p.txt")
    lines = _gslam([1, 0.1, rounds, 0, path], env=dict(GSLAM_SOLVER="gj64", POSE_DUMP_FILE=dump))
    diffs = _diffs(lines)
    assert os.path.exists(dump)
    # the stop rule applied to the printed diffs: diff < 0.01, or 10 rounds of
    # one parity without a new best of that parity
    best = {0: 1e30, 1: 1e30}
    stale = {0: 0, 1: 0}
    expect = rounds
    for r, df in enumerate(diffs, 1):
        if df < 0.01:
            expect = r
            break
        k = r % 2
        if df < best[k]:
            best[k], stale[k] = df, 0
        else:
            stale[k] += 1
            if stale[k] >= 10:
                expect = r
                break
    assert len(diffs) == expect
    ref = _oracle_gj_diffs(oracle, oracle.read_log(path), 3)
    assert np.allclose(diffs[:3], ref, rtol=1e-6), (diffs[:3], ref)


@pytest.mark.parametrize("solver", ["pcg64", "gj64"])
def test_between_sidecar_pose_graph(oracle, tmp_path, solver):
    gc = G.grid_city(600, seed=3)
    log, bet = str(tmp_path / "g.log.txt"), str(tmp_path / "g.between")
    G.write_pose_graph(gc, log, bet)
    dump = str(tmp_path / "p.txt")
    env = dict(GSLAM_BETWEEN=bet, GSLAM_WRAP="1", GSLAM_SOLVER=solver, GSLAM_TOL="1e-13",
               GSLAM_MAX_ITER="200000", POSE_DUMP_FILE=dump)
    lines = _gslam([1, 0.1, 8, 1, log], env=env)
    assert lines[1].startswith("delta: 0.000000 size_of_us: 0 size_of_hat_xs: 600")
    diffs = _diffs(lines)
    got = np.loadtxt(dump)
    assert np.array_equal(got[:, 0], np.arange(600))
    # the same rounds through the oracle's between-edge GN step
    rp, ci = oracle.csr_pattern(600, gc["i"], gc["j"])
    p = gc["poses"].copy()
    for r, df in enumerate(diffs):
        _, _, diff_o = oracle.between_gn_step(p, gc["i"], gc["j"], gc["meas"], gc["info"], None,
                                              0.1, rp, ci, tol=1e-13, max_iter=200000)
        assert abs(diff_o - df) <= 1e-6 * max(df, 1e-3), (r, df, diff_o)
    d = (got[:, 1:] - p.reshape(-1, 3))
    d[:, 2] = (d[:, 2] + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(d).max() <= 1e-6
    assert diffs[-1] < 0.01 and all(x >= 0.01 for x in diffs[:-1])


def test_checkpoint_resume_continues_the_run(tmp_path):
    """GSLAM_DUMP_EVERY writes the poses after every k-th round; the x records
    rewritten with a checkpoint (gslam.resume_log) start a run that continues
    the original one: rounds 4-6 of one 6-round run are, bit for bit, rounds
    1-3 of a run resumed from the round-3 checkpoint."""
    path = data_path("cityTrees100.log.txt")
    full, ck, fin = (str(tmp_path / n) for n in ("full.txt", "ck.txt", "fin.txt"))
    a = _gslam([1, 0.1, 6, 1, path], env=dict(POSE_DUMP_FILE=full))
    b1 = _gslam([1, 0.1, 3, 1, path], env=dict(POSE_DUMP_FILE=ck, GSLAM_DUMP_EVERY="3"))
    assert _diffs(b1) == _diffs(a)[:3] and len(_diffs(a)) == 6
    resumed = str(tmp_path / "resumed.log.txt")
    assert gslam.resume_log(path, ck, resumed) == 100
    b2 = _gslam([1, 0.1, 3, 1, resumed], env=dict(POSE_DUMP_FILE=fin))
    assert _diffs(b2) == _diffs(a)[3:]
    assert np.array_equal(np.loadtxt(fin), np.loadtxt(full))
