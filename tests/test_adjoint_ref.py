"""Pins the NumPy implicit-differentiation reference (tests/adjoint_ref.py) on
the CPU before the GPU adjoint is compared with it:
* central finite differences (step 1e-4, FP64) of the converged loss in
  w_e and lambda match the implicit gradient;
* the analytic gradient of the loss matches finite differences;
* the Gauss-Newton matrix alone would miss the second-order terms (the
  exact Hessian matters at noise-level residuals).
"""
import numpy as np

import adjoint_ref as R
from gslam import generators as G


def _setup(n=30, seed=0, lam=0.1):
    g = G.grid_city(n, seed=seed, lc_gap=5, lc_keep=0.6)
    w = np.zeros(len(g["i"]))
    ref = g["poses"][:3].copy()
    x = R.converge(g["poses"], g, w, lam, 1e6, ref)
    return g, w, ref, x


def test_loss_gradient_against_fd():
    g, w, ref, x = _setup()
    L, dL = R.loss_and_grad(x, g["truth"])
    h = 1e-6
    for k in (0, 7, 40, x.size - 1):
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        fd = (R.loss_and_grad(xp, g["truth"])[0] - R.loss_and_grad(xm, g["truth"])[0]) / (2 * h)
        assert abs(fd - dL[k]) <= 1e-6 * max(1.0, abs(fd))


def test_implicit_gradient_against_central_fd():
    lam = 0.1
    g, w, ref, x = _setup(lam=lam)
    assert np.abs(R.grad_f(x, g, w, lam, 1e6, ref)).max() < 1e-8   # converged
    _, dL = R.loss_and_grad(x, g["truth"])
    dw, dlam = R.implicit_grad(x, g, w, lam, 1e6, ref, dL)
    h = 1e-4
    rng = np.random.default_rng(0)
    ks = rng.choice(len(w), 8, replace=False)
    fd = []
    for k in ks:
        vals = []
        for sgn in (1, -1):
            wk = w.copy()
            wk[k] += sgn * h
            xk = R.converge(x, g, wk, lam, 1e6, ref)
            vals.append(R.loss_and_grad(xk, g["truth"])[0])
        fd.append((vals[0] - vals[1]) / (2 * h))
    fd = np.array(fd)
    assert np.linalg.norm(dw[ks] - fd) <= 1e-5 * np.linalg.norm(fd), (dw[ks], fd)
    vals = []
    for sgn in (1, -1):
        xk = R.converge(x, g, w, lam * (1 + sgn * h), 1e6, ref)
        vals.append(R.loss_and_grad(xk, g["truth"])[0])
    fd_lam = (vals[0] - vals[1]) / (2 * h * lam)
    assert abs(dlam - fd_lam) <= 1e-4 * max(abs(fd_lam), 1e-3 * np.abs(fd).max())
    # every edge scales with lambda: dL/dlam = sum_e dL/dw_e / lam
    assert abs(dlam - dw.sum() / lam) <= 1e-6 * np.abs(dw).sum() / lam


def test_gauss_newton_matrix_is_not_enough():
    lam = 0.1
    g, w, ref, x = _setup(lam=lam)
    _, dL = R.loss_and_grad(x, g["truth"])
    dw, _ = R.implicit_grad(x, g, w, lam, 1e6, ref, dL)
    dw_gn, _ = R.implicit_grad(x, g, w, lam, 1e6, ref, dL, H=R.gn_matrix(x, g, w, lam, 1e6))
    assert np.linalg.norm(dw_gn - dw) > 1e-3 * np.linalg.norm(dw)
