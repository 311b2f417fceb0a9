"""The benchmark configurations at their full sizes, against independent
reference solutions.

Every solve is compared with `refined_direct` (tests/refsolve.py: SuperLU
with refinement on extended-precision residuals, accurate to ~1e-11 relative
on these systems), with numpy.linalg.solve beside it for the dense cases:

* C1  plain FP64 Gauss-Jordan, grid-city N=800 (3N = 2400), ten GN rounds
      with the diff of use_diff=1 -- every step's dx against the direct solve
      of the same system;
* C3  blocked Gauss-Jordan at N=5000 (3N = 15000): FP64 (GJ_BLOCKED64) and
      FP32-MFMA (GJ_MIXED) with conjugate-gradient refinement on the FP64
      true residual;
* C4a block-Jacobi PCG on grid-city N=100,000: FP64 and mixed;
* C4b the same on the park graph (laps route) of 7,120 poses, lambda 1.0.

The bounds are the measured errors with a margin (each test prints what it
measured).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from refsolve import refined_direct, rel

import gslam
from gslam import generators as G

pytestmark = pytest.mark.gpu


def _maxerr(a, b):
    d = np.abs(np.asarray(a) - np.asarray(b)).reshape(-1, 3)
    return float(d[:, :2].max()), float(d[:, 2].max())


def test_c1_gj64_grid800_ten_rounds():
    gc = G.grid_city(800, seed=0)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
        errs, diffs = [], []
        for _ in range(10):
            g.assemble()
            rp, ci, va, xi = g.system_csr()
            x_ref = refined_direct(rp, ci, va, xi)
            A = sp.csr_matrix((va, ci, rp), shape=(len(xi), len(xi))).toarray()
            x_np = np.linalg.solve(A, xi)
            st = g.solve("gj64")
            dx = g.dx()
            errs.append((rel(dx, x_ref), rel(x_np, x_ref)))
            assert st["true_rel_res"] < 1e-9
            diffs.append(g.apply_update())
            assert abs(diffs[-1] - np.linalg.norm(dx) / np.sqrt(len(dx))) <= 1e-12 * diffs[-1] + 1e-300
    print("C1 gj64 N=800: rel dx err per round (gj64, numpy.linalg.solve) vs refined direct:",
          ", ".join(f"({a:.1e}, {b:.1e})" for a, b in errs), "diffs", ", ".join(f"{d:.3g}" for d in diffs))
    assert max(e[0] for e in errs) <= 1e-10      # measured <= 1.4e-11
    # GN converges on this graph: the last steps are tiny
    assert diffs[-1] < 1e-6 * diffs[0]


@pytest.mark.parametrize("solver,kw,bound", [
    ("gj64_blocked", {}, 5e-8),                    # measured 1.6e-8
    # CG refinement to an FP64 true residual of 1e-12: measured 3.8e-9 in 116
    # iterations (at 2e-12 the error sits right at 1e-8: 8e-9 or 2e-8
    # depending on the summation order of the reductions)
    ("gj_mixed", dict(tol=1e-12, refine_sweeps=400, refine_mode="cg"), 1e-8),
])
def test_c3_blocked_gj_grid5000(solver, kw, bound):
    gc = G.grid_city(5000, seed=0)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
        g.assemble()
        rp, ci, va, xi = g.system_csr()
        x_ref = refined_direct(rp, ci, va, xi)
        st = g.solve(solver, **kw)
        e = rel(g.dx(), x_ref)
    print(f"C3 {solver} N=5000: rel dx err {e:.3e} vs refined direct, {st['iters']} refinement "
          f"its, true rel res {st['true_rel_res']:.2e}, {st['ms']:.1f} ms (factor {st['ms_factor']:.1f})")
    assert e <= bound, e


@pytest.fixture(scope="module")
def grid100k():
    gc = G.grid_city(100000, seed=0)
    g = gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1, wrap=1))
    g.assemble()
    rp, ci, va, xi = g.system_csr()
    x_ref = refined_direct(rp, ci, va, xi)
    yield g, x_ref
    g.close()


def test_c4a_pcg64_grid100k(grid100k):
    g, x_ref = grid100k
    st = g.solve("pcg64", 3e-13, 600000)
    dx = g.dx()
    e = rel(dx, x_ref)
    em, er = _maxerr(dx, x_ref)
    print(f"C4a pcg64 N=100k tol 3e-13: {st['iters']} its, rel dx err {e:.3e}, max pose err "
          f"{em:.2e} m {er:.2e} rad, true rel res {st['true_rel_res']:.2e}")
    assert st["converged"]
    assert e <= 1e-8, e                             # measured 7.99e-9


def test_c4a_mixed_grid100k(grid100k):
    """FP32 iterations with FP64 true-residual refreshes stall at ~1e-6 here
    (the block-Jacobi preconditioned operator has kappa ~ 1e13, far beyond
    1/eps32); with the FP64 correction (mixed_handover) the mixed solver meets
    the FP64 solver's accuracy.  The handover at a true residual of 1e-2 is
    the workload's: from 1e-2, 3e-3 and 1e-3 the correction ends at 7.8e-9,
    7.9e-9 and 7.9e-9, from 3e-2 and 1e-4 at 1.5e-8 and 1.2e-8 -- at this
    conditioning the attainable error is ~1e-8 within a factor of three, set
    by the rounding path."""
    g, x_ref = grid100k
    st = g.solve("mixed", 1e-12, 1500000)
    e0 = rel(g.dx(), x_ref)
    assert st["converged"] and st["true_rel_res"] < 1e-12 and st["iters_fp64"] == 0
    st = g.solve("mixed", 3e-13, 1500000, mixed_handover=1e-2)
    dx = g.dx()
    e = rel(dx, x_ref)
    em, er = _maxerr(dx, x_ref)
    print(f"C4a mixed N=100k: FP32 only (tol 1e-12) rel dx err {e0:.3e}; with the FP64 "
          f"correction below 1e-2 (tol 3e-13): {st['iters']} its ({st['iters_fp64']} FP64), "
          f"rel dx err {e:.3e}, max pose err {em:.2e} m {er:.2e} rad, true rel res "
          f"{st['true_rel_res']:.2e}")
    assert 1e-7 <= e0 <= 3e-6, e0                     # measured 3.6e-7 and 1.2e-6: the FP32 floor
    assert st["converged"] and 0 < st["iters_fp64"] < st["iters"]
    assert e <= 1e-8, e                               # measured 7.75e-9


@pytest.mark.parametrize("solver,tol,kw", [("pcg64", 1e-12, {}), ("mixed", 1e-12, {}),
                                           ("mixed", 1e-10, dict(mixed_handover=1e-2))])
def test_c4b_park7120(solver, tol, kw):
    gc = G.park(7120, seed=0, route="laps")
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 1.0)) as g:
        g.assemble()
        x_ref = refined_direct(*g.system_csr())
        st = g.solve(solver, tol, 400000, **kw)
        dx = g.dx()
    e = rel(dx, x_ref)
    em, er = _maxerr(dx, x_ref)
    print(f"C4b {solver} {kw} park N=7120: {st['iters']} its ({st['iters_fp64']} FP64), rel dx "
          f"err {e:.3e}, max pose err {em:.2e} m {er:.2e} rad, {st['ms']:.1f} ms")
    assert st["converged"]
    # measured 1.9e-13 (FP64), 2.3e-12 (mixed), 9.6e-12 (mixed, FP64 correction)
    assert e <= 1e-10, e
