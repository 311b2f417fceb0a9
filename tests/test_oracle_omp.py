"""The OpenMP build of the CPU oracle (oracle/liboracle_omp.so), bench.py's
multi-core CPU baseline.  It is the same source as liboracle.so with the PCG
loops (SpMV rows, block-Jacobi apply, the three dot products and the vector
updates) shared among threads; only the dot products change, summed per
thread and then across threads.  Checks:
* liboracle.so runs one thread (its results stay bitwise the reference's:
  tests/test_oracle_golden.py), the OpenMP build the team size;
* a fixed number of PCG iterations on a well-conditioned grid-city system
  agrees with the serial oracle to rounding (measured 2.8e-13 after 1000
  iterations at N=20k) -- a race on a reduction or a vector gives O(1);
* converged to 1e-10 on the reference's victoria-park-100 system, the same
  solution within 1e-8 and an iteration count within 2% (rounding changes
  the CG trajectory: 391 vs 389 iterations with 8 threads, 2.4e-11 apart);
* the OpenMP sums are deterministic: two runs give the same bits (the
  per-thread sums are added in thread order; a reduction clause adds them as
  the threads finish, and moved that count over 380-394).
"""
import numpy as np
import pytest

from conftest import golden

from gslam import generators as G


@pytest.fixture(scope="module")
def omp():
    from oracle import LIB_OMP, Oracle
    return Oracle(LIB_OMP)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_thread_counts(oracle, omp):
    assert oracle.threads() == 1
    assert omp.threads() >= 1


def test_fixed_iterations_match_serial(oracle, omp):
    n = 5000
    gc = G.grid_city(n, seed=1)
    ev = oracle.between_edges(gc["poses"], gc["i"], gc["j"], gc["meas"], gc["info"], None, 0.1)
    rp, ci, va, xi = oracle.assemble_csr(n, gc["i"], gc["j"], ev)
    xs, its = oracle.pcg(rp, ci, va, xi, 1e-300, 300)
    xp, itp = omp.pcg(rp, ci, va, xi, 1e-300, 300)
    assert its == itp == 300
    assert rel(xp, xs) <= 1e-10, rel(xp, xs)


def test_converged_reference_system(oracle, omp):
    fx = golden("vp100")
    args = (fx["row_ptr"], fx["col_idx"], fx["values"], fx["xi"], 1e-10, 5000)
    xs, its = oracle.pcg(*args)
    xp, itp = omp.pcg(*args)
    assert abs(itp - its) <= 0.02 * its, (its, itp)
    assert rel(xp, xs) <= 1e-8, rel(xp, xs)


def test_runs_repeat_bitwise(omp):
    fx = golden("vp100")
    args = (fx["row_ptr"], fx["col_idx"], fx["values"], fx["xi"], 1e-10, 5000)
    runs = [omp.pcg(*args) for _ in range(20)]
    assert len({it for _, it in runs}) == 1
    assert all(np.array_equal(x, runs[0][0]) for x, _ in runs)
