"""The direct-solve reference of the solver tests (tests/refsolve.py),
pinned on systems with an exactly known solution."""
import numpy as np
import scipy.sparse as sp

from refsolve import refined_direct, rel


def _path_laplacian(n, anchor=1e6):
    # integer Laplacian of a path graph plus an anchor on node 0: right-hand
    # sides A x of integer x are exact in FP64, so the solution is known
    main = np.full(n, 2.0)
    main[0] = main[-1] = 1.0
    main[0] += anchor
    A = sp.diags([main, -np.ones(n - 1), -np.ones(n - 1)], [0, -1, 1], format="csr")
    return A


def test_recovers_exact_solution_of_ill_conditioned_path():
    n = 30000                                   # kappa ~ 4 n^2 / pi^2 * anchor-scale
    A = _path_laplacian(n)
    x_true = np.random.default_rng(0).integers(-1000, 1000, n).astype(np.float64)
    b = A @ x_true                              # exact (integers below 2^53)
    assert np.array_equal(b, np.round(b))
    A.sort_indices()
    hist = []
    x = refined_direct(A.indptr, A.indices, A.data, b, steps=4, history=hist)
    plain = sp.linalg.splu(A.tocsc()).solve(b)
    # floor: kappa * eps(longdouble) ~ 4e-11 worst case; measured 3.4e-13,
    # against 1.8e-11 for the unrefined FP64 LU solve
    assert rel(x, x_true) <= 2e-12, rel(x, x_true)
    assert rel(x, x_true) <= 0.1 * rel(plain, x_true)
    assert hist[-1][0] <= 1e-17


def test_block_system_against_dense_solve():
    rng = np.random.default_rng(1)
    n = 300
    M = rng.normal(size=(n, n))
    M = M @ M.T + n * np.eye(n)                 # well conditioned: FP64 dense solve is exact enough
    M[np.abs(M) < 2.0] = 0.0
    M = 0.5 * (M + M.T) + n * np.eye(n)
    A = sp.csr_matrix(M)
    b = rng.normal(size=n)
    x = refined_direct(A.indptr, A.indices, A.data, b)
    assert rel(x, np.linalg.solve(M, b)) <= 1e-14
