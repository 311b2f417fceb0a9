"""Blocked Gauss-Jordan in FP64 (GJ_BLOCKED64).

The same unpivoted elimination as the reference's FP64 Gauss-Jordan
(src/cpu_pcg_fp64/MatrixOps.c:29-91), taken 128 pivots at a time: FP64
pivot blocks, FP64 MFMA row panel and trailing updates.  Checks:
* inverses against numpy.linalg.inv (LAPACK, FP64) and against the
  one-pivot-per-launch FP64 kernel (gj64) on symmetric and non-symmetric
  matrices below, at and across the 128-wide panel and with several tiles
  per block step -- a transposed MFMA operand or a misplaced accumulator row
  fails these by O(1);
* a diagonal spread over 6 orders (the path Jacobi-scales the matrix, as the
  reference's own Gauss-Jordan would need no scaling in exact arithmetic):
  D^-1 M^-1 D^-1 recovered to FP64 accuracy in the scaled basis -- the
  folded pivot-block update loses digits without the scaling;
* a zero pivot gives the zero matrix and GSLAM_ERR_SINGULAR (the reference
  rule, MatrixOps.c:60-68), also in a later panel;
* the reference's own systems (golden CSR of cityTrees100 and
  victoria-park-100): within 1e-8 of the reference FP64 Gauss-Jordan solve
  (golden gj_dx), like gj64;
* a GN step on grid-city matches FP64 PCG.
"""
import numpy as np
import pytest

from conftest import GOLDEN_SETS, data_path, golden

import gslam
from gslam import generators as G

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("n", [1, 50, 128, 200, 300, 513, 3000])
def test_inverse_against_numpy_and_gj64(n):
    rng = np.random.default_rng(n + 7)
    b = rng.normal(size=(n, n))
    spd = b @ b.T / n + np.eye(n)
    nonsym = rng.normal(size=(n, n)) + 2.0 * np.sqrt(n) * np.eye(n)
    for a in (spd, nonsym):
        ref = np.linalg.inv(a)
        inv = gslam.gj_invert(a, "gj64_blocked")
        kappa = np.linalg.cond(a)
        assert rel(inv, ref) <= 100 * kappa * np.finfo(np.float64).eps * max(1.0, np.sqrt(n)), n
        assert np.abs(inv @ a - np.eye(n)).max() <= 1e-10
        if n <= 513:
            assert rel(inv, gslam.gj_invert(a, "gj64")) <= 1e-12


def test_inverse_tail_tiles_of_the_pivot_workgroup():
    """n = 9200: 72 panels, 5111 tiles per trailing launch besides the next
    pivot block, 510 other tile workers; the pivot workgroup's two workers
    take the last 11 tiles after the pivot block.  a = I + 0.25 G / sqrt(n)
    (singular values in [0.5, 1.5]); one trailing tile update of one block
    step left out gives 9.6e-5 against numpy.linalg.inv (numpy emulation)."""
    n = 9200
    rng = np.random.default_rng(n)
    a = np.eye(n) + 0.25 * rng.normal(size=(n, n)) / np.sqrt(n)
    inv = gslam.gj_invert(a, "gj64_blocked")
    assert rel(inv, np.linalg.inv(a)) <= 1e-11, rel(inv, np.linalg.inv(a))
    assert np.abs(inv @ a - np.eye(n)).max() <= 1e-10


def test_diagonal_spread():
    n = 400
    rng = np.random.default_rng(3)
    b = rng.normal(size=(n, n))
    scale = 10.0 ** rng.uniform(-3, 3, n)
    m = b @ b.T / n + np.eye(n)          # kappa ~ 5
    a = m * scale[:, None] * scale[None, :]    # kappa ~ 1e12
    inv = gslam.gj_invert(a, "gj64_blocked")
    # a^-1 = D^-1 m^-1 D^-1 exactly: compare D a^-1 D with m^-1 (the CPU
    # reference Gauss-Jordan reaches 1.3e-15 here, numpy.linalg.inv 5.9e-14)
    assert rel(inv * scale[:, None] * scale[None, :], np.linalg.inv(m)) <= 1e-12


def test_zero_pivot_gives_zero_matrix():
    z = np.eye(6)
    z[3, 3] = 0.0
    assert not gslam.gj_invert(z, "gj64_blocked").any()
    z = np.eye(300)
    z[250, 250] = 0.0     # third panel
    assert not gslam.gj_invert(z, "gj64_blocked").any()


@pytest.mark.parametrize("key", ["ct100", "vp100"])
def test_reference_systems(key):
    fx = golden(key)
    with gslam.Graph.from_log(data_path(GOLDEN_SETS[key]), lam=float(fx["lam"])) as g:
        g.set_system_csr(fx["values"], fx["xi"])
        st = g.solve("gj64_blocked")
        assert rel(g.dx(), fx["gj_dx"]) <= 1e-8 and st["true_rel_res"] < 1e-9
        assert st["ms_factor"] > 0 and st["ms"] >= st["ms_factor"] and st["iters"] == 0


def test_gn_step_matches_fp64_pcg():
    gc = G.grid_city(600, seed=4)
    kw = G.graph_kwargs(gc, 0.1)
    with gslam.Graph(gc["poses"], **kw) as a, gslam.Graph(gc["poses"], **kw) as b:
        for _ in range(3):
            a.gn_step("pcg64", 1e-13, 100000)
            b.gn_step("gj64_blocked")
        pa, pb = a.get_poses(), b.get_poses()
    assert np.abs(pa - pb).max() <= 1e-6
