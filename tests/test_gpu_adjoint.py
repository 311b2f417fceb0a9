"""GPU adjoint (gslam_adjoint, gslam.autodiff.ConvergedPoses).

* against the FP64 NumPy implicit differentiation of tests/adjoint_ref.py
  (exact Hessian by finite differences of the analytic gradient, dense
  solve) at the same converged poses;
* against central finite differences (step 1e-4, FP64) of the converged
  loss on random edges and on lambda, each side re-converged on the GPU;
* the torch.autograd.Function returns the same gradients, and Adam on
  (w_e, log lambda) lowers the loss.
"""
import numpy as np
import pytest
import torch

import adjoint_ref as R
import gslam
from gslam import autodiff as AD
from gslam import generators as G

pytestmark = pytest.mark.gpu

CONV = dict(solver="pcg64", tol=1e-13, max_iter=400000, diff_tol=1e-13, max_rounds=60)


def _graph(gc, lam=0.1):
    return gslam.Graph(gc["poses"], anchor_prior=1, **G.graph_kwargs(gc, lam))


def test_converged_poses_match_numpy():
    gc = G.grid_city(200, seed=3)
    with _graph(gc) as g:
        rounds = g.converge(**CONV)
        x = g.get_poses()
    w = np.zeros(len(gc["i"]))
    xr = R.converge(gc["poses"], gc, w, 0.1, 1e6, gc["poses"][:3])
    assert rounds < CONV["max_rounds"]
    assert np.abs(x - xr).max() <= 1e-9


@pytest.mark.parametrize("seed", [3, 11])
def test_adjoint_matches_numpy_implicit_gradient(seed):
    gc = G.grid_city(200, seed=seed)
    lam = 0.1
    rng = np.random.default_rng(seed)
    w = rng.normal(0.0, 0.3, len(gc["i"]))
    with _graph(gc, lam) as g:
        g.set_log_weights(w)
        g.converge(**CONV)
        x = g.get_poses()
        _, dl_dx = R.loss_and_grad(x, gc["truth"])
        dw, dlam, st = g.adjoint(dl_dx, tol=1e-13)
    assert st["converged"]
    dw_r, dlam_r = R.implicit_grad(x, gc, w, lam, 1e6, gc["poses"][:3], dl_dx)
    assert np.linalg.norm(dw - dw_r) <= 1e-6 * np.linalg.norm(dw_r)
    assert abs(dlam - dlam_r) <= 1e-6 * np.abs(dw_r).sum() / lam


def test_adjoint_matches_central_differences():
    gc = G.grid_city(1500, seed=5)
    lam, h = 0.1, 1e-4
    with _graph(gc, lam) as g:
        g.converge(**CONV)
        x0 = g.get_poses()
        L0, dl_dx = R.loss_and_grad(x0, gc["truth"])
        dw, dlam, _ = g.adjoint(dl_dx, tol=1e-13)
        n_e = len(gc["i"])
        ks = np.random.default_rng(0).choice(n_e, 24, replace=False)
        fd = []
        for k in ks:
            v = []
            for sgn in (1.0, -1.0):
                w = np.zeros(n_e)
                w[k] = sgn * h
                g.set_log_weights(w)
                g.set_poses(x0)
                g.converge(**CONV)
                v.append(R.loss_and_grad(g.get_poses(), gc["truth"])[0])
            fd.append((v[0] - v[1]) / (2 * h))
        fd = np.array(fd)
        g.set_log_weights(None)
        v = []
        for sgn in (1.0, -1.0):
            g.set_lambda(lam * (1 + sgn * h))
            g.set_poses(x0)
            g.converge(**CONV)
            v.append(R.loss_and_grad(g.get_poses(), gc["truth"])[0])
        fd_lam = (v[0] - v[1]) / (2 * h * lam)
    err = np.linalg.norm(dw[ks] - fd) / np.linalg.norm(fd)
    print(f"adjoint vs central FD on {len(ks)} edges: rel {err:.2e}; dL/dlam {dlam:.3e} fd {fd_lam:.3e}")
    assert err <= 1e-4
    # every edge scales with lambda and the edges exert no net rigid-body
    # force on the anchored pose, so x* and L do not depend on lambda:
    # dL/dlambda = 0 up to rounding, and the central difference is at the
    # noise level of the re-converged loss (|dL| ~ 1e-12 relative)
    assert abs(dlam) <= 1e-6 * np.abs(dw).sum() / lam
    assert abs(fd_lam) <= 1e-3 * np.abs(fd).max()


def test_torch_function_and_adam():
    gc = G.grid_city(300, seed=8)
    with _graph(gc) as g:
        log_w = torch.zeros(len(gc["i"]), dtype=torch.float32, requires_grad=True)
        lam = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
        poses = AD.ConvergedPoses.apply(log_w, lam, g, gc["poses"], CONV)
        loss = AD.drift_and_truth_loss(poses, gc["truth"])
        loss.backward()
        x = poses.detach().numpy().ravel()
        Lr, dl_dx = R.loss_and_grad(x, gc["truth"])
        assert abs(float(loss.detach()) - Lr) <= 1e-12 * max(1.0, Lr)
        dw, dlam, _ = g.adjoint(dl_dx, tol=1e-13)
        assert log_w.grad.dtype == torch.float32
        assert np.allclose(log_w.grad.numpy(), dw.astype(np.float32), rtol=1e-5, atol=1e-9)
        assert abs(float(lam.grad) - dlam) <= 1e-9 + 1e-6 * abs(dlam)
        _, lam_out, hist = AD.tune(g, gc["poses"], gc["truth"], steps=8, lr=0.05, solve_kw=CONV)
    assert hist[-1] < hist[0]
    assert lam_out > 0


def test_backward_after_a_second_forward_on_the_same_graph():
    """Two forwards on one graph (other weights, other lambda), then backward
    through the first: the gradient is the first forward's, because backward
    restores that forward's weights, lambda and converged poses on the graph
    before the adjoint solve."""
    gc = G.grid_city(300, seed=9)
    rng = np.random.default_rng(10)
    with _graph(gc) as g:
        w1 = torch.zeros(len(gc["i"]), dtype=torch.float32, requires_grad=True)
        lam1 = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
        p1 = AD.ConvergedPoses.apply(w1, lam1, g, gc["poses"], CONV)
        w2 = torch.tensor(rng.normal(size=len(gc["i"])) * 0.3, dtype=torch.float32, requires_grad=True)
        lam2 = torch.tensor(0.25, dtype=torch.float64, requires_grad=True)
        p2 = AD.ConvergedPoses.apply(w2, lam2, g, gc["poses"], CONV)
        assert not torch.allclose(p1, p2)
        AD.drift_and_truth_loss(p1, gc["truth"]).backward()
        g1w, g1l = w1.grad.clone(), lam1.grad.clone()
    with _graph(gc) as h:
        w = torch.zeros(len(gc["i"]), dtype=torch.float32, requires_grad=True)
        lam = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
        p = AD.ConvergedPoses.apply(w, lam, h, gc["poses"], CONV)
        assert torch.allclose(p, p1, rtol=0, atol=1e-9)
        AD.drift_and_truth_loss(p, gc["truth"]).backward()
    assert np.allclose(g1w.numpy(), w.grad.numpy(), rtol=1e-5, atol=1e-9)
    assert abs(float(g1l) - float(lam.grad)) <= 1e-9 + 1e-5 * abs(float(lam.grad))
