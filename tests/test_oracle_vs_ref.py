"""The oracle against the reference C itself (oracle/_ref, built from the
reference sources by `make -C oracle ref`) on randomised inputs, bitwise.
Skipped where oracle/_ref was not built."""
import os
import subprocess

import numpy as np
import pytest

from oracle import RefGJMixed, RefPCG64, ref_available

pytestmark = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")


@pytest.fixture(scope="module")
def ref():
    return RefPCG64()


def _rand_pose(rng):
    return np.array([rng.uniform(-50, 50), rng.uniform(-50, 50), rng.uniform(-7, 7)])


def test_motion_edge_random(oracle, ref):
    rng = np.random.default_rng(1)
    for k in range(400):
        x1, x2 = _rand_pose(rng), _rand_pose(rng)
        nu = rng.uniform(-3, 3)
        # cover the |omega| < 1e-5 clamp (MotionEdge.c:21-22), both signs
        omega = rng.choice([rng.uniform(-1, 1), rng.uniform(-1e-5, 1e-5), 0.0, -3e-6])
        delta = rng.choice([1.0, 0.5, 2.0, 1.7])
        lam = rng.choice([0.1, 1.0, 0.06, 3.3])
        a, oa = oracle.motion_edge(x1, x2, nu, omega, delta, lam)
        b, ob = ref.motion_edge(x1, x2, nu, omega, delta, lam)
        assert np.array_equal(a, b), k
        assert np.array_equal(oa, ob), k


def test_obs_edge_random(oracle, ref):
    rng = np.random.default_rng(2)
    for k in range(400):
        x1, x2 = _rand_pose(rng), _rand_pose(rng)
        r1, r2 = rng.uniform(0.1, 30, 2)
        p1, p2 = rng.uniform(-3.2, 3.2, 2)
        a, ea, oa = oracle.obs_edge(x1, x2, r1, p1, r2, p2)
        b, eb, ob = ref.obs_edge(x1, x2, r1, p1, r2, p2)
        assert np.array_equal(a, b), k
        assert np.array_equal(ea, eb) and np.array_equal(oa, ob), k


def _random_log_graph(rng, n, n_lm, n_obs, self_pairs=False):
    poses = np.stack([rng.uniform(-20, 20, n), rng.uniform(-20, 20, n),
                      rng.uniform(-4, 4, n)], 1).ravel()
    z_step = rng.integers(0, n, n_obs).astype(np.int32)
    if self_pairs:  # the same pose observing a landmark twice -> t1 == t2
        z_step[1::7] = z_step[0::7][: len(z_step[1::7])]
    g = dict(poses=poses, n_poses=n, x_step=np.arange(n, dtype=np.int32),
             u_nu=rng.uniform(-2, 2, n), u_omega=rng.uniform(-0.5, 0.5, n), delta=1.0,
             z_step=z_step, z_lm=rng.integers(0, n_lm, n_obs).astype(np.int32) * 3 + 1,
             z_range=rng.uniform(0.5, 9, n_obs), z_bearing=rng.uniform(-3, 3, n_obs))
    g["z_lm"] = np.ascontiguousarray(g["z_lm"], np.int32)
    return g


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_make_edges_csr_pcg_random(oracle, ref, seed):
    rng = np.random.default_rng(seed)
    g = _random_log_graph(rng, int(rng.integers(5, 60)), int(rng.integers(2, 12)),
                          int(rng.integers(4, 60)), self_pairs=(seed == 5))
    n = g["n_poses"]
    lam = 0.3
    t1, t2, ev = oracle.log_edges(g, lam=lam)
    edges, ne = ref.round_edges(g, lam)
    rt1, rt2, rev = ref.edge_arrays(edges, ne)
    assert np.array_equal(t1, rt1) and np.array_equal(t2, rt2)
    assert np.array_equal(ev, rev)
    rp, ci, va, xi = oracle.assemble_csr(n, t1, t2, ev)
    csr, rrp, rci, rva, rxi = ref.csr(edges, ne, 3 * n)
    assert np.array_equal(rp, rrp) and np.array_equal(ci, rci)
    assert np.array_equal(va, rva) and np.array_equal(xi, rxi)
    x, _ = oracle.pcg(rp, ci, va, xi, 1e-6, 500)
    assert np.array_equal(x, ref.pcg(csr, rxi))
    ref.free_csr(csr)


def test_inv_random(oracle, ref):
    rng = np.random.default_rng(6)
    for n in (1, 2, 3, 5, 17, 64):
        a = rng.normal(size=(n, n))
        a = a @ a.T + n * np.eye(n)
        inv, rc = oracle.gj_inv(a)
        assert rc == 0
        assert np.array_equal(inv, ref.inv(a)), n
    z = np.eye(4)
    z[2, 2] = 0.0  # zero pivot -> zero matrix (MatrixOps.c:60-68)
    inv, rc = oracle.gj_inv(z)
    assert rc == 1 and not inv.any()
    assert not ref.inv(z).any()


def test_inv_float_random(oracle):
    G = RefGJMixed()
    rng = np.random.default_rng(7)
    for n in (3, 9, 40):
        a = rng.normal(size=(n, n))
        a = (a @ a.T + n * np.eye(n)).astype(np.float32)
        inv, rc = oracle.gj_inv_f32(a)
        assert rc == 0
        assert np.array_equal(inv, G.inv_float(a)), n


def test_executable_stdout_and_dump(oracle, tmp_path):
    """Full reference run (stdout + POSE_DUMP_FILE) on a random graph."""
    rng = np.random.default_rng(8)
    g = _random_log_graph(rng, 30, 6, 25)
    path = tmp_path / "g.log.txt"
    with open(path, "w") as f:
        f.write("delta 1.0\n")
        for i in range(g["n_poses"]):
            r = lambda v: repr(float(v))  # noqa: E731
            f.write(f"u {i} {r(g['u_nu'][i])} {r(g['u_omega'][i])}\n")
            p = g["poses"][3 * i: 3 * i + 3]
            f.write(f"x {i} {r(p[0])} {r(p[1])} {r(p[2])}\n")
            for k in np.nonzero(g["z_step"] == i)[0]:
                f.write(f"z {i} {g['z_lm'][k]} {r(g['z_range'][k])} {r(g['z_bearing'][k])}\n")
    g2 = oracle.read_log(str(path))
    for k in ("u_nu", "u_omega", "poses"):
        assert np.array_equal(g2[k], g[k])
    dump = tmp_path / "p.txt"
    out = subprocess.run([RefPCG64.exe(), "1", "0.5", "4", "1", str(path)], capture_output=True,
                         text=True, env=dict(os.environ, POSE_DUMP_FILE=str(dump)), check=True)
    diffs = [float(l.split(":")[1]) for l in out.stdout.splitlines() if "rounds executed" in l]
    poses, od = oracle.log_gn_run(g2, 0.5, 4)
    assert np.array_equal(od, np.array(diffs))
    assert np.array_equal(poses.reshape(-1, 3), np.loadtxt(dump)[:, 1:])
