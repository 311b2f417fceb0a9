"""GPU solvers against the oracle and direct solves.

* FP64 PCG on the reference's own Omega (the golden CSR loaded with
  gslam_set_system_csr) follows the reference iteration (1e-6, 500):
  iteration counts within three of the oracle's (the stopping test sits
  right at the 1e-6 threshold) and solutions within the relative bound
  stated per case (the recurrences are summed in a different order, and 500
  unconverged iterations on kappa ~ 1e9 amplify rounding).
* Converged solves (tol 1e-12) meet a relative dx error <= 1e-8 against a
  direct LU solve with one refinement step on graphs where the FP64
  conditioning allows it.
* The mixed solver always certifies its tolerance on the FP64 true residual.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import GOLDEN_SETS, data_path, golden
from refsolve import refined_direct

import gslam
from gslam import generators as G

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def direct(rp, ci, va, xi):
    n = len(rp) - 1
    A = sp.csr_matrix((va, ci, rp), shape=(n, n)).tocsc()
    lu = spla.splu(A)
    x = lu.solve(xi)
    return x + lu.solve(xi - A @ x)


# measured GPU-vs-oracle distance after the reference's 500 capped iterations:
# ct100 2.9e-4, vp100 (converges at 345-347) 2.9e-7, ct800 1.5e-8
@pytest.mark.parametrize("key,bound", [("ct100", 1e-3), ("vp100", 1e-5), ("ct800", 1e-5)])
def test_pcg64_reference_config_on_reference_system(oracle, key, bound):
    fx = golden(key)
    with gslam.Graph.from_log(data_path(GOLDEN_SETS[key]), lam=float(fx["lam"])) as g:
        g.set_system_csr(fx["values"], fx["xi"])
        st = g.solve("pcg64", 1e-6, 500)
        dx = g.dx()
    x_o, it_o = oracle.pcg(fx["row_ptr"], fx["col_idx"], fx["values"], fx["xi"], 1e-6, 500)
    assert np.array_equal(x_o, fx["pcg_x"])
    assert abs(st["iters"] - it_o) <= 3
    assert rel(dx, fx["pcg_x"]) <= bound
    assert st["true_rel_res"] <= 1.01 * max(st["rel_res"], 1e-6) or not st["converged"]


SCHEDULES = ["auto", "two_launch"]


@pytest.mark.parametrize("schedule", SCHEDULES)
@pytest.mark.parametrize("name", ["grid800", "grid5000", "park2000"])
def test_pcg64_converged_against_direct(name, schedule):
    """Both launch schedules (on these graphs 'auto' is one launch per
    iteration) meet 1e-8 against the direct solve."""
    gc = {"grid800": lambda: G.grid_city(800, seed=1), "grid5000": lambda: G.grid_city(5000, seed=2),
          "park2000": lambda: G.park(2000, seed=3)}[name]()
    lam = 1.0 if name.startswith("park") else 0.1
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, lam)) as g:
        g.assemble()
        xs = direct(*g.system_csr())
        st = g.solve("pcg64", 1e-12, 200000, pcg_schedule=schedule)
        # the stop test is on the recursively updated residual; the true
        # residual drifts from it at the rounding level (1.006e-12 seen)
        assert st["converged"] and st["true_rel_res"] < 2e-12
        assert rel(g.dx(), xs) <= 1e-8


@pytest.mark.parametrize("schedule", SCHEDULES)
def test_pcg64_block_row_longer_than_a_tile(schedule):
    """Pose 0 tied to every other pose by loop closures: its block row holds
    700 blocks, more than one 256-block SpMV tile, so the FP64 SpMV (and the
    one-launch iteration) walks it in three chunks (the first from the padded
    tile slots, the rest from the BSR arrays).  Converged FP64 PCG against
    the direct solve."""
    rng = np.random.default_rng(5)
    n = 700
    truth = np.zeros((n, 3))
    truth[:, 0] = np.arange(n, dtype=float)
    truth[:, 1] = rng.normal(size=n) * 0.5
    gc = G._finish(truth, np.zeros(n - 2, dtype=np.int64), np.arange(2, n), rng, 0.05, 0.01)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
        g.assemble()
        rp, ci, va, xi = g.system_csr()
        assert rp[1] - rp[0] == 3 * n       # scalar row 0 spans all 700 block columns
        xs = direct(rp, ci, va, xi)
        st = g.solve("pcg64", 1e-12, 200000, pcg_schedule=schedule)
        assert st["converged"] and st["true_rel_res"] < 2e-12
        assert rel(g.dx(), xs) <= 1e-8


def test_pcg64_several_tiles_per_workgroup(oracle):
    """Grid city N = 300,000: more SpMV tiles (about 5,100) than the 4,096
    workgroups of the launch, so some workgroups take a second tile.  Fifty
    FP64 PCG iterations against fifty of the CPU oracle on the same system
    (downloaded from the device); after so few iterations the two summation
    orders differ at the rounding level only."""
    gc = G.grid_city(300000, seed=7)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
        g.assemble()
        rp, ci, va, xi = g.system_csr()
        st = g.solve("pcg64", 1e-30, 50)
        assert st["iters"] == 50 and not st["converged"]
        x_o, it_o = oracle.pcg(rp, ci, va, xi, tol=1e-30, max_iter=50)
        assert it_o == 50
        assert rel(g.dx(), x_o) <= 1e-9


def test_pcg64_one_launch_schedule_against_oracle(oracle):
    """Grid city N = 20,000 (about 340 SpMV tiles: one launch per iteration
    under 'auto').  Fifty iterations against fifty of the CPU oracle (the
    reference's two-reduction recurrence) on the same system: the one-step
    beta identity and the different summation orders stay at the rounding
    level.  Then both schedules to tol 1e-12: iteration counts within 2% and
    both within 1e-8 of the direct solve."""
    gc = G.grid_city(20000, seed=8)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
        g.assemble()
        rp, ci, va, xi = g.system_csr()
        x_o, it_o = oracle.pcg(rp, ci, va, xi, tol=1e-30, max_iter=50)
        for schedule in SCHEDULES:
            st = g.solve("pcg64", 1e-30, 50, pcg_schedule=schedule)
            assert st["iters"] == 50 and it_o == 50 and not st["converged"]
            assert rel(g.dx(), x_o) <= 1e-9, schedule
        xs = direct(rp, ci, va, xi)
        its = {}
        for schedule in SCHEDULES:
            st = g.solve("pcg64", 1e-12, 400000, pcg_schedule=schedule)
            assert st["converged"] and st["true_rel_res"] < 2e-12, schedule
            assert rel(g.dx(), xs) <= 1e-8, schedule
            its[schedule] = st["iters"]
        assert abs(its["auto"] - its["two_launch"]) <= 0.02 * its["two_launch"], its


# one-launch iteration with several consecutive tiles per workgroup: two
# (512 < tiles <= 1024) or four (1024 < tiles <= 2048)
GROUPED = {"two": (2, 512, 1024, 40000), "four": (4, 1024, 2048, 80000)}


def _grouped_graph(gc, kind):
    """The graph and its tile count, which must fall in the kind's range."""
    _, lo, hi, _ = GROUPED[kind]
    g = gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1))
    nt = g.info["n_tiles"]
    assert lo < nt <= hi, (kind, nt)
    return g, nt


@pytest.mark.parametrize("kind", list(GROUPED))
def test_pcg64_tile_groups_against_oracle(oracle, kind):
    """Grid city N = 40,000 (about 680 SpMV tiles: two per workgroup) and
    80,000 (about 1,360: four per workgroup), sizes chosen so that the last
    workgroup holds fewer tiles than the others.  Fifty iterations of both
    schedules against fifty of the CPU oracle on the same system; both to tol
    1e-12 within 1e-8 of the direct solve and within 5% of each other's
    iteration count (31,021 and 31,840 measured at N = 40k with four tiles
    per group); iteration caps, the p.q == 0 break of a zero right-hand side,
    and the GSLAM_ERR_NONFINITE stop of a NaN in the right-hand side and of
    an Inf in Omega."""
    tpw, _, _, n0 = GROUPED[kind]
    for n in (n0, n0 + 13, n0 + 29, n0 + 47, n0 + 61, n0 + 83):
        gc = G.grid_city(n, seed=10)
        g, nt = _grouped_graph(gc, kind)
        if nt % tpw:
            break
        g.close()
    assert nt % tpw, nt
    with g:
        g.assemble()
        rp, ci, va, xi = g.system_csr()
        x_o, it_o = oracle.pcg(rp, ci, va, xi, tol=1e-30, max_iter=50)
        assert it_o == 50
        for schedule in SCHEDULES:
            st = g.solve("pcg64", 1e-30, 50, pcg_schedule=schedule)
            assert st["iters"] == 50 and not st["converged"]
            assert rel(g.dx(), x_o) <= 1e-9, schedule
        xs = direct(rp, ci, va, xi)
        its = {}
        for schedule in SCHEDULES:
            st = g.solve("pcg64", 1e-12, 600000, pcg_schedule=schedule)
            # the FP64 true residual of these systems bottoms out at 2-6e-12
            # (2.4e-12 measured at N = 80k, 5.8e-12 at 100k)
            assert st["converged"] and st["true_rel_res"] < 1e-11, schedule
            assert rel(g.dx(), xs) <= 1e-8, (schedule, rel(g.dx(), xs))
            its[schedule] = st["iters"]
        assert abs(its["auto"] - its["two_launch"]) <= 0.05 * its["two_launch"], its
        for cap in (1, 2, 7):
            st = g.solve("pcg64", 1e-14, cap)
            assert st["iters"] == cap and not st["converged"], cap
        x, st = g.solve_rhs(np.zeros(len(xi)), "pcg64", 1e-6, 100)
        assert not x.any() and st["flags"] & 1 and st["iters"] == 0
        bad = xi.copy()
        bad[3 * (len(xi) // 6) + 1] = np.nan
        with pytest.raises(gslam.GSlamError) as ei:
            g.solve_rhs(bad, "pcg64", 1e-10, 100000)
        assert ei.value.code == gslam.ERR_NONFINITE
        va_bad = va.copy()
        va_bad[rp[len(xi) - 5]] = np.inf
        g.set_system_csr(va_bad, xi)
        with pytest.raises(gslam.GSlamError) as ei:
            g.solve("pcg64", 1e-10, 100000)
        assert ei.value.code == gslam.ERR_NONFINITE


@pytest.mark.parametrize("kind", list(GROUPED))
def test_pcg64_tile_groups_long_block_row(oracle, kind):
    """Grid city N = 40,000 / 80,000 plus 700 loop closures from its middle
    pose to poses drawn across the walk: that block row (~710 blocks) takes
    three rounds of 256, in a workgroup whose other tiles take one, so every
    part of that workgroup runs the extra rounds' barriers.  Fifty iterations
    of both schedules against fifty of the CPU oracle; then both converged
    (tol 1e-12) against the extended-precision refined direct solve.  The
    hub's lever arms of up to a few hundred metres leave this Omega
    ill-conditioned: at N = 40k the CPU oracle's own converged PCG lands
    1.4e-7 from the refined solve (and the plain LU solve with one
    refinement step 1.6e-7), hence the bound of 5e-7."""
    _, _, _, n = GROUPED[kind]
    base = G.grid_city(n, seed=11)
    rng = np.random.default_rng(12)
    hub = n // 2
    others = np.sort(rng.choice(np.setdiff1d(np.arange(n), np.arange(hub - 30, hub + 31)),
                                700, replace=False))
    lc = np.flatnonzero(np.arange(len(base["i"])) >= base["n_odometry"])
    I = np.concatenate([base["i"][lc], np.full(700, hub)])
    J = np.concatenate([base["j"][lc], others])
    gc = G._finish(base["truth"], I, J, rng, 0.05, 0.01)
    g, nt = _grouped_graph(gc, kind)
    with g:
        g.assemble()
        rp, ci, va, xi = g.system_csr()
        assert rp[3 * hub + 1] - rp[3 * hub] >= 3 * 700
        x_o, it_o = oracle.pcg(rp, ci, va, xi, tol=1e-30, max_iter=50)
        assert it_o == 50
        for schedule in SCHEDULES:
            st = g.solve("pcg64", 1e-30, 50, pcg_schedule=schedule)
            assert st["iters"] == 50 and not st["converged"]
            assert rel(g.dx(), x_o) <= 1e-9, schedule
        xr = refined_direct(rp, ci, va, xi)
        for schedule in SCHEDULES:
            st = g.solve("pcg64", 1e-12, 600000, pcg_schedule=schedule)
            assert st["converged"] and st["true_rel_res"] < 1e-11, schedule
            assert rel(g.dx(), xr) <= 5e-7, (schedule, rel(g.dx(), xr))


def test_nontemporal_streams_against_oracle(oracle):
    """Grid city N = 2,000,000: the tile copy (~700 MB), the six FP64 solver
    vectors (288 MB) and the FP32 operator's slots all exceed their
    Infinity-Cache thresholds, so the SpMV, update and mixed kernels run
    their nontemporal variants (same arithmetic, other load and store
    instructions).  Fifty FP64 iterations against fifty of the CPU oracle,
    and a mixed solve certified on its FP64 true residual."""
    gc = G.grid_city(2000000, seed=9)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
        g.assemble()
        assert 76 * 256 * (g.info["n_blocks"] // 256) > 256 << 20
        assert 144 * g.info["n_poses"] > 256 << 20
        rp, ci, va, xi = g.system_csr()
        st = g.solve("pcg64", 1e-30, 50)
        assert st["iters"] == 50 and not st["converged"]
        x_o, it_o = oracle.pcg(rp, ci, va, xi, tol=1e-30, max_iter=50)
        assert it_o == 50
        assert rel(g.dx(), x_o) <= 1e-9
        st = g.solve("mixed", 1e-3, 60000)
        assert st["converged"] and st["true_rel_res"] < 1e-3


def test_pcg_schedule_is_validated():
    gc = G.grid_city(300, seed=4)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
        g.assemble()
        with pytest.raises(gslam.GSlamError) as ei:
            g.solve("pcg64", 1e-10, 1000, pcg_schedule=7)
        assert ei.value.code == gslam.ERR_ARG


@pytest.mark.parametrize("name", ["grid800", "park2000"])
def test_mixed_certifies_true_residual(name):
    gc = {"grid800": lambda: G.grid_city(800, seed=1), "park2000": lambda: G.park(2000, seed=3)}[name]()
    lam = 1.0 if name.startswith("park") else 0.1
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, lam)) as g:
        g.assemble()
        xs = direct(*g.system_csr())
        st = g.solve("mixed", 1e-12, 200000)
        assert st["converged"] and st["true_rel_res"] < 1e-12
        assert st["reliable_updates"] >= 1
        e = rel(g.dx(), xs)
        st2 = g.solve("mixed", 1e-12, 200000, mixed_handover=1e-4)
        e2 = rel(g.dx(), xs)
        print(f"{name}: mixed rel dx err {e:.2e} ({st['iters']} its), with the FP64 correction "
              f"below 1e-4 {e2:.2e} ({st2['iters']} its, {st2['iters_fp64']} FP64)")
        assert st2["converged"] and st2["true_rel_res"] < 1e-12
        # measured: grid800 4.7e-10 / 2.1e-11, park2000 1.8e-8 / 1.3e-10
        assert e <= {"grid800": 5e-9, "park2000": 2e-7}[name]
        assert e2 <= 2e-9


def test_mixed_pose_with_more_entries_than_a_tile():
    """A hub pose with 700 between edges: its row of the FP32 factorized
    operator spans three 256-entry chunks (a tile of its own, summed chunk by
    chunk); the solve still meets its FP64 true residual and the direct
    solution."""
    rng = np.random.default_rng(4)
    n = 800
    poses = np.column_stack([rng.uniform(-20, 20, n), rng.uniform(-20, 20, n),
                             rng.uniform(-np.pi, np.pi, n)]).ravel()
    bi = np.r_[np.zeros(700, np.int32), np.arange(1, n - 1, dtype=np.int32)]
    bj = np.r_[np.arange(1, 701, dtype=np.int32), np.arange(2, n, dtype=np.int32)]
    meas = rng.normal(size=3 * len(bi)) * 0.1
    info = np.tile([100.0, 100.0, 400.0], len(bi))
    with gslam.Graph(poses, between=dict(i=bi, j=bj, meas=meas, info=info), lam=1.0) as g:
        g.assemble()
        xs = refined_direct(*g.system_csr())
        err, its = {}, {}
        for name, solver, kw in (("pcg64", "pcg64", {}), ("mixed", "mixed", {}),
                                 ("handover", "mixed", dict(mixed_handover=1e-4))):
            st = g.solve(solver, 1e-11, 200000, **kw)
            assert st["converged"] and st["true_rel_res"] < 1e-11, (name, st)
            err[name], its[name] = rel(g.dx(), xs), st["iters"]
        print("hub graph: rel dx err", err, "iterations", its)
        # measured 8.6e-11 / 3.9e-8 / 1.5e-10 (1431 / 2944 / 2082 iterations); hubs of
        # 200-256 entries (one chunk) give 1.1-1.3e-8 for mixed at the same ratio of
        # iterations, 257-300 entries 2.3-4.9e-8
        assert err["pcg64"] <= 1e-9 and err["mixed"] <= 2e-7 and err["handover"] <= 1e-9, err
        assert its["mixed"] <= 3 * its["pcg64"], its


def test_mixed_many_tiles_per_workgroup():
    """Grid city N=400,000: 5,000+ operator tiles, more than the mixed SpMV's
    4,096 workgroups, so workgroups walk several tiles (round robin).  To a
    true FP64 residual of 1e-4 / 1e-5 the FP32 iterations follow the FP64
    ones almost step for step (measured 6,272 vs 6,226 and 23,680 vs 23,630
    iterations); a tile skipped or summed twice would leave the FP64 true
    residual stuck."""
    gc = G.grid_city(400000, seed=2)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1, wrap=1)) as g:
        g.assemble()
        for tol in (1e-4, 1e-5):
            a = g.solve("pcg64", tol, 100000)
            b = g.solve("mixed", tol, 100000)
            assert a["converged"] and b["converged"] and b["true_rel_res"] < tol, (tol, a, b)
            assert abs(b["iters"] - a["iters"]) <= 0.02 * a["iters"], (tol, a["iters"], b["iters"])


def test_mixed_on_reference_log_model():
    """The factorized FP32 operator covers landmark-pair and motion edges."""
    with gslam.Graph.from_log(data_path("victoria-park-100.log.txt"), lam=0.1) as g:
        g.assemble()
        xs = direct(*g.system_csr())
        st = g.solve("mixed", 1e-10, 100000)
        assert st["converged"] and st["true_rel_res"] < 1e-10
        assert rel(g.dx(), xs) <= 1e-5
        with pytest.raises(gslam.GSlamError):     # factors unavailable after a CSR load
            rp, ci, va, xi = g.system_csr()
            g.set_system_csr(va, xi)
            g.solve("mixed", 1e-10, 1000)


def test_solver_buffers_allocated_on_first_use():
    """The FP64 tile copy and the mixed solver's FP32 operator are allocated
    by the first solve that uses them.  The first mixed solve re-runs the
    assembly to write the edge factors: Omega and xi keep their bits, and the
    solution is bitwise the one of a handle that solves mixed at once."""
    gc = G.grid_city(1500, seed=2)
    kw = G.graph_kwargs(gc, 0.1)
    with gslam.Graph(gc["poses"], **kw) as a, gslam.Graph(gc["poses"], **kw) as b:
        a.assemble()
        b.assemble()
        m0 = a.device_bytes()
        assert m0 == a.info["device_bytes"]
        _, _, bv, xi = a.system_bsr()
        a.solve("pcg64", 1e-10, 100000)
        m1 = a.device_bytes()
        sa = a.solve("mixed", 1e-10, 200000)
        m2 = a.device_bytes()
        _, _, bv2, xi2 = a.system_bsr()
        assert np.array_equal(bv, bv2) and np.array_equal(xi, xi2)
        sb = b.solve("mixed", 1e-10, 200000)
        assert sa["iters"] == sb["iters"] and np.array_equal(a.dx(), b.dx())
        n, e = 1500, a.info["n_between_edges"]
        assert m1 - m0 >= 8 * 3 * n                              # FP64 tile copy
        assert m2 - m1 >= 4 * (9 * n + 5 * 3 * n) + 4 * 4 * e + 4 * 16 * e   # FP32 operator
        # later assemblies write the factors themselves, later solves allocate nothing
        a.assemble()
        sa2 = a.solve("mixed", 1e-10, 200000)
        assert a.device_bytes() == m2
        assert sa2["iters"] == sa["iters"] and np.array_equal(a.dx(), b.dx())


def test_non_finite_system_fails_loudly():
    """A NaN in Omega or in the right-hand side stops every solver within an
    iteration or two with GSLAM_ERR_NONFINITE (flags bit 1), instead of running
    to the iteration cap or returning NaN poses; non-finite poses are refused."""
    gc = G.grid_city(400, seed=5)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
        g.assemble()
        rp, ci, va, xi = g.system_csr()
        bad_rhs = xi.copy()
        bad_rhs[7] = np.nan
        for solver in ("pcg64", "mixed"):
            with pytest.raises(gslam.GSlamError) as ei:
                g.solve_rhs(bad_rhs, solver, 1e-10, 100000)
            assert ei.value.code == gslam.ERR_NONFINITE, solver
        # the handle stays usable
        st = g.solve("pcg64", 1e-10, 100000)
        assert st["converged"]
        va_bad = va.copy()
        va_bad[rp[30]] = np.inf
        g.set_system_csr(va_bad, xi)
        for solver in ("pcg64", "gj64", "gj64_blocked", "gj_mixed"):
            with pytest.raises(gslam.GSlamError) as ei:
                g.solve(solver, 1e-10, 100000)
            assert ei.value.code == gslam.ERR_NONFINITE, solver
        p = gc["poses"].copy()
        p[5] = np.nan
        with pytest.raises(gslam.GSlamError) as ei:
            g.set_poses(p)
        assert ei.value.code == gslam.ERR_ARG


@pytest.mark.parametrize("schedule", SCHEDULES)
def test_zero_rhs_breaks_on_pq_zero_and_cap(schedule):
    gc = G.grid_city(400, seed=5)
    kw = dict(pcg_schedule=schedule)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
        g.assemble()
        x, st = g.solve_rhs(np.zeros(3 * 400), "pcg64", 1e-6, 100, **kw)
        assert not x.any() and st["flags"] & 1 and st["iters"] == 0   # MatrixOps.c:357
        for cap in (1, 2, 7, 300):                                     # iteration cap
            st = g.solve("pcg64", 1e-14, cap, **kw)
            assert st["iters"] == cap and not st["converged"], cap
        st = g.solve("pcg64", 1e-6, 0, **kw)
        assert st["iters"] == 0 and not g.dx().any()


def test_solve_rhs_matches_solve():
    gc = G.grid_city(600, seed=6)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
        g.assemble()
        _, _, _, xi = g.system_csr()
        g.solve("pcg64", 1e-12, 100000)
        x, st = g.solve_rhs(xi, "pcg64", 1e-12, 100000)
        assert np.array_equal(x, g.dx())


@pytest.mark.parametrize("key", ["ct100", "vp100"])
def test_gj64_against_reference_fixture(oracle, key):
    fx = golden(key)
    g0 = oracle.read_log(data_path(GOLDEN_SETS[key]))
    n = g0["n_poses"]
    om, xi = oracle.dense(n, fx["t1"], fx["t2"], fx["edge_vals"])
    inv = gslam.gj_invert(om)
    inv_o, _ = oracle.gj_inv(om)
    assert np.array_equal(np.diag(inv_o), fx["gj_inv_diag"])
    # same algorithm, different summation order: compare the solves
    assert rel(inv @ xi, fx["gj_dx"]) <= 1e-8
    assert np.abs(inv @ om - np.eye(3 * n)).max() <= 1e-6
    with gslam.Graph.from_log(data_path(GOLDEN_SETS[key]), lam=float(fx["lam"])) as g:
        g.set_system_csr(fx["values"], fx["xi"])
        st = g.solve("gj64")
        assert rel(g.dx(), fx["gj_dx"]) <= 1e-8 and st["true_rel_res"] < 1e-9


def test_gj64_kats():
    for n in (1, 5, 8):
        H = 1.0 / (np.arange(n)[:, None] + np.arange(n)[None, :] + 1.0)
        assert np.allclose(gslam.gj_invert(H), np.linalg.inv(H), rtol=1e-6, atol=1e-6)
    rng = np.random.default_rng(0)
    for n in (3, 40, 129, 300):
        a = rng.normal(size=(n, n))
        a = a @ a.T + n * np.eye(n)
        assert np.allclose(gslam.gj_invert(a), np.linalg.inv(a), rtol=1e-10, atol=1e-12)
    z = np.eye(6)
    z[3, 3] = 0.0
    assert not gslam.gj_invert(z).any()   # zero pivot -> zero matrix (MatrixOps.c:60-68)


def test_kernel_timing_entry_point():
    gc = G.grid_city(2000, seed=7)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
        g.assemble()
        for k in ("spmv64", "update64", "spmv_mixed", "assembly"):
            us = g.time_kernel(k, 20)
            assert 0.5 < us < 5000.0, (k, us)
        st = g.solve("pcg64", 1e-10, 100000)   # state after timing still solvable
        assert st["converged"]
