"""GPU assembly (edge evaluation + Omega/xi gather) against the oracle.

Tolerances.  The sparsity pattern and the CSR layout are compared exactly.
Values are FP64 and the GPU evaluates the same formulas in a different
operation order (cofactor inverses, fused multiply-adds, device sin/cos), so
they are compared relative to the largest entry of their row:
  * SE(2) between edges and landmark-pair edges: <= 1e-12;
  * the reference velocity motion edge: <= 2e-5.  Its A matrix
    (src/cpu_pcg_fp64/MotionEdge.c:54-58) subtracts terms of size nu/omega^2
    (1e10 when omega is clamped to 1e-5) whose difference is O(1), so any two
    FP64 evaluations differ by ~1e10 * 1e-16; the bound below is that
    cancellation, and motion edges with |omega| >= 0.05 meet 1e-12.
"""
import numpy as np
import pytest

from conftest import GOLDEN_SETS, data_path, golden

import gslam
from gslam import generators as G

pytestmark = pytest.mark.gpu


def _row_rel(rp, va, ref):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    scale = np.zeros(n)
    np.maximum.at(scale, rows, np.abs(ref))
    return (np.abs(va - ref) / np.maximum(scale[rows], 1e-300)).max(initial=0.0)


@pytest.mark.parametrize("key", list(GOLDEN_SETS))
def test_log_model_against_reference_fixture(key):
    fx = golden(key)
    with gslam.Graph.from_log(data_path(GOLDEN_SETS[key]), lam=float(fx["lam"])) as g:
        g.assemble()
        rp, ci, va, xi = g.system_csr()
    assert np.array_equal(rp, fx["row_ptr"]) and np.array_equal(ci, fx["col_idx"])
    assert _row_rel(rp, va, fx["values"]) <= 2e-5
    assert np.linalg.norm(xi - fx["xi"]) <= 1e-8 * np.linalg.norm(fx["xi"])


@pytest.mark.parametrize("key", ["ct100", "vp100", "ct800"])
def test_obs_edges_only_tight(oracle, key):
    """Landmark-pair edges alone (no motion edges): well conditioned."""
    g = oracle.read_log(data_path(GOLDEN_SETS[key]))
    n = g["n_poses"]
    obs = dict(z_step=g["z_step"], z_lm=g["z_lm"], z_range=g["z_range"], z_bearing=g["z_bearing"])
    for order in (0, 1):
        a, b = oracle.obs_pairs(g["z_lm"], order)
        ev = np.array([oracle.obs_edge(g["poses"][3 * g["z_step"][i]:3 * g["z_step"][i] + 3],
                                       g["poses"][3 * g["z_step"][j]:3 * g["z_step"][j] + 3],
                                       g["z_range"][i], g["z_bearing"][i], g["z_range"][j],
                                       g["z_bearing"][j])[0] for i, j in zip(a, b)])
        t1, t2 = g["z_step"][a], g["z_step"][b]
        rp, ci, va, xi = oracle.assemble_csr(n, t1, t2, ev)
        with gslam.Graph(g["poses"], obs=obs, pair_order=order) as gg:
            assert gg.info["n_obs_edges"] == len(a)
            gg.assemble()
            grp, gci, gva, gxi = gg.system_csr()
        assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
        assert _row_rel(rp, gva, va) <= 1e-12
        assert np.abs(gxi - xi).max() <= 1e-12 * np.abs(xi).max()


def test_motion_edges_well_conditioned_tight(oracle):
    rng = np.random.default_rng(3)
    n = 50
    poses = np.stack([rng.uniform(-20, 20, n), rng.uniform(-20, 20, n), rng.uniform(-3, 3, n)], 1).ravel()
    nu = rng.uniform(0.5, 2.0, n)
    om = rng.choice([-1, 1], n) * rng.uniform(0.05, 0.6, n)
    t1 = np.arange(n - 1, dtype=np.int32)
    t2 = t1 + 1
    ev = np.array([oracle.motion_edge(poses[3 * i:3 * i + 3], poses[3 * i + 3:3 * i + 6], nu[i + 1],
                                      om[i + 1], 1.0, 0.3)[0] for i in range(n - 1)])
    rp, ci, va, xi = oracle.assemble_csr(n, t1, t2, ev)
    with gslam.Graph(poses, motion=True, u_nu=nu, u_omega=om, lam=0.3) as g:
        g.assemble()
        grp, gci, gva, gxi = g.system_csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    assert _row_rel(rp, gva, va) <= 1e-12
    assert np.abs(gxi - xi).max() <= 1e-11 * np.abs(xi).max()


@pytest.mark.parametrize("seed", [0, 1])
def test_between_edges_against_oracle(oracle, seed):
    gc = G.grid_city(1500, seed=seed)
    n = gc["n_poses"]
    w = np.exp(np.random.default_rng(seed).normal(0, 0.3, len(gc["i"])))
    ev = oracle.between_edges(gc["poses"], gc["i"], gc["j"], gc["meas"], gc["info"], w, 0.1, 1)
    rp, ci, va, xi = oracle.assemble_csr(n, gc["i"], gc["j"], ev)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
        g.set_log_weights(np.log(w))
        g.assemble()
        grp, gci, gva, gxi = g.system_csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    assert _row_rel(rp, gva, va) <= 1e-12
    assert np.abs(gxi - xi).max() <= 1e-11 * np.abs(xi).max()


def test_edge_cases_self_pairs_isolated_poses_and_no_edges(oracle):
    """t1 == t2 edges (a pose seeing a landmark twice), poses no edge touches
    (empty reference CSR rows), pose 0 untouched (anchor-only rows), and a
    graph without edges."""
    rng = np.random.default_rng(5)
    n = 12
    poses = rng.normal(size=3 * n)
    ei = np.array([3, 3, 5, 7, 9], np.int32)
    ej = np.array([3, 5, 7, 9, 3], np.int32)   # edge 0 is a self pair; poses 0,1,2,4,... isolated
    meas = rng.normal(size=3 * len(ei))
    info = np.full(3 * len(ei), 4.0)
    ev = oracle.between_edges(poses, ei, ej, meas, info, None, 0.5, 1)
    rp, ci, va, xi = oracle.assemble_csr(n, ei, ej, ev)
    with gslam.Graph(poses, lam=0.5, between=dict(i=ei, j=ej, meas=meas, info=info), wrap=1) as g:
        assert g.info["nnz"] == len(ci)
        g.assemble()
        grp, gci, gva, gxi = g.system_csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    assert _row_rel(rp, gva, va) <= 1e-12
    assert np.abs(gxi - xi).max() <= 1e-12 * max(np.abs(xi).max(), 1.0)
    with gslam.Graph(np.zeros(6)) as g:   # no edges: anchor only
        g.assemble()
        grp, gci, gva, gxi = g.system_csr()
    assert list(grp) == [0, 1, 2, 3, 3, 3, 3] and list(gci) == [0, 1, 2]
    assert list(gva) == [1e6, 1e6, 1e6] and not gxi.any()


def test_assembly_is_bitwise_repeatable():
    gc = G.grid_city(3000, seed=4)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
        g.assemble()
        a = g.system_bsr()
        for _ in range(3):
            g.assemble()
            b = g.system_bsr()
            assert all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("case", ["grid", "park_prior", "ct100", "vp100"])
def test_pose_walk_and_staged_assembly_agree_bitwise(case):
    """The two assembly schedules (one thread per pose walking its edges;
    one thread per edge staging its record, then one per entry summing) add
    the same edge records in the same edge order, so Omega and xi agree bit
    for bit -- and so do the mixed operator's FP32 factors (through forty
    mixed PCG iterations) and, with the anchor prior, the exact-Hessian
    assembly of the adjoint (through its gradients)."""
    if case == "grid":
        gc = G.grid_city(2000, seed=3)
        make = lambda a: gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1, wrap=1), assembly=a)
    elif case == "park_prior":
        gc = G.park(1500, seed=2, route="laps")
        make = lambda a: gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 1.0, wrap=1),
                                     anchor_prior=1, assembly=a)
    else:
        path = data_path(GOLDEN_SETS[case])
        make = lambda a: gslam.Graph.from_log(path, lam=0.1, assembly=a)
    with make("pose_walk") as gw, make("staged") as gs:
        for g in (gw, gs):
            g.assemble()
        assert all(np.array_equal(x, y) for x, y in zip(gw.system_bsr(), gs.system_bsr()))
        dx = []
        for g in (gw, gs):
            st = g.solve("mixed", 1e-30, 40)
            assert st["iters"] == 40
            dx.append(g.dx())
        assert np.array_equal(dx[0], dx[1])
        if case == "park_prior":
            rng = np.random.default_rng(1)
            dl = rng.normal(size=3 * gw.n)
            for g in (gw, gs):
                g.gn_step("pcg64", 1e-10, 100000)
            aw, a_s = gw.adjoint(dl, tol=1e-10), gs.adjoint(dl, tol=1e-10)
            assert np.array_equal(aw[0], a_s[0]) and aw[1] == a_s[1]


def test_assembly_schedule_choice():
    """'auto' walks poses unless a pose has more than 64 edge ends or an edge
    joins a pose to itself; forcing the walk on such a self pair is refused."""
    poses = np.zeros(3 * 4)
    be = dict(i=np.array([1, 1], np.int32), j=np.array([1, 2], np.int32),
              meas=np.zeros(6), info=np.ones(6))
    with pytest.raises(gslam.GSlamError) as ei:
        gslam.Graph(poses, between=be, assembly="pose_walk")
    assert ei.value.code == gslam.ERR_ARG
    with pytest.raises(gslam.GSlamError) as ei:
        gslam.Graph(poses, between=be, assembly=9)
    assert ei.value.code == gslam.ERR_ARG
    # a hub pose with 80 loop closures: 'auto' stages, and both schedules agree
    n = 90
    rng = np.random.default_rng(2)
    hub = dict(i=np.r_[np.arange(n - 1), np.zeros(80, np.int64)].astype(np.int32),
               j=np.r_[np.arange(1, n), np.arange(10, 90)].astype(np.int32),
               meas=rng.normal(size=3 * (n - 1 + 80)), info=np.full(3 * (n - 1 + 80), 4.0))
    p = rng.normal(size=3 * n)
    with gslam.Graph(p, between=hub, lam=0.5, wrap=1) as ga, \
            gslam.Graph(p, between=hub, lam=0.5, wrap=1, assembly="pose_walk") as gw:
        ga.assemble()
        gw.assemble()
        assert all(np.array_equal(x, y) for x, y in zip(ga.system_bsr(), gw.system_bsr()))
