"""Differentiable converged pose-graph solve (torch.autograd.Function).

Forward: set the per-edge log-weights w_e and lambda on a libgslam graph,
reset the poses to the initial guess, run Gauss-Newton to convergence on the
GPU and return the converged poses x* (N x 3, float64).

Backward: gslam_adjoint -- one FP64 PCG on the exact Hessian at x* and an
edge-parallel kernel -- gives dL/dw_e and dL/dlambda from dL/dx*.

The reference has no gradient path: lambda is tuned by a manual grid search
over whole runs (README.md:137-163); `tune` replaces that loop by Adam on
(w_e, log lambda).
"""
from __future__ import annotations

import numpy as np
import torch

from . import Graph, check_single_runtime

check_single_runtime()


class ConvergedPoses(torch.autograd.Function):
    """poses = ConvergedPoses.apply(log_w, lam, graph, init_poses, solve_kw)

    graph: a gslam.Graph of between edges created with anchor_prior=1.
    log_w: per-between-edge log-weights (any float dtype, any device).
    lam: 0-d tensor, the global information scale.
    solve_kw: keyword arguments of Graph.converge (tolerances, rounds) plus
      optional 'adjoint_tol' / 'adjoint_max_iter' for the backward PCG.
    """

    @staticmethod
    def forward(ctx, log_w, lam, graph: Graph, init_poses, solve_kw):
        kw = dict(solve_kw or {})
        adj_tol = kw.pop("adjoint_tol", 1e-12)
        adj_iter = kw.pop("adjoint_max_iter", 200000)
        w = log_w.detach().to("cpu", torch.float64).numpy().copy()
        lam_v = float(lam.detach().cpu())
        graph.set_log_weights(w)
        graph.set_lambda(lam_v)
        graph.set_poses(init_poses)
        graph.converge(**kw)
        x = graph.get_poses()
        # the graph is shared: backward restores this forward's weights, lambda
        # and converged poses, whatever later forwards left on it
        ctx.graph, ctx.x, ctx.w, ctx.lam = graph, x, w, lam_v
        ctx.adj = (adj_tol, adj_iter)
        ctx.w_meta = (log_w.dtype, log_w.device)
        ctx.lam_meta = (lam.dtype, lam.device)
        return torch.from_numpy(x.reshape(-1, 3).copy())

    @staticmethod
    def backward(ctx, grad_poses):
        g = grad_poses.detach().to("cpu", torch.float64).reshape(-1).numpy()
        ctx.graph.set_log_weights(ctx.w)
        ctx.graph.set_lambda(ctx.lam)
        ctx.graph.set_poses(ctx.x)
        dw, dlam, _ = ctx.graph.adjoint(g, tol=ctx.adj[0], max_iter=ctx.adj[1])
        gw = torch.from_numpy(dw).to(dtype=ctx.w_meta[0], device=ctx.w_meta[1])
        gl = torch.tensor(dlam, dtype=ctx.lam_meta[0], device=ctx.lam_meta[1])
        return gw, gl, None, None, None


def _wrap(a):
    return torch.atan2(torch.sin(a), torch.cos(a))


def drift_and_truth_loss(poses: torch.Tensor, truth) -> torch.Tensor:
    """Squared end-to-start drift error plus squared pose error against the
    ground truth (heading differences wrapped)."""
    t = torch.as_tensor(np.asarray(truth).reshape(-1, 3), dtype=poses.dtype)
    drift = (poses[-1, :2] - poses[0, :2]) - (t[-1, :2] - t[0, :2])
    return (drift.square().sum() + (poses[:, :2] - t[:, :2]).square().sum()
            + _wrap(poses[:, 2] - t[:, 2]).square().sum())


def tune(graph: Graph, init_poses, truth, steps: int = 50, lr: float = 0.05,
         lam0: float = 0.1, solve_kw: dict | None = None, log_w0=None):
    """Adam on per-edge log-weights (FP32 parameters) and log lambda, the loss
    drift_and_truth_loss of the converged poses.  Returns (log_w, lam, history)."""
    n_e = graph.info["n_between_edges"]
    log_w = torch.zeros(n_e, dtype=torch.float32) if log_w0 is None else \
        torch.as_tensor(log_w0, dtype=torch.float32).clone()
    log_w.requires_grad_(True)
    log_lam = torch.tensor(float(np.log(lam0)), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([log_w, log_lam], lr=lr)
    hist = []
    for _ in range(steps):
        opt.zero_grad()
        poses = ConvergedPoses.apply(log_w, log_lam.exp(), graph, init_poses, solve_kw)
        loss = drift_and_truth_loss(poses, truth)
        loss.backward()
        opt.step()
        hist.append(float(loss.detach()))
    return log_w.detach(), float(log_lam.detach().exp()), hist


def drift_loss(poses) -> tuple:
    """The README's lambda-search criterion (README.md:137-163): squared
    distance between the first and the last pose position, and its gradient
    with respect to the poses (3N)."""
    p = np.asarray(poses).reshape(-1, 3)
    d = p[-1, :2] - p[0, :2]
    g = np.zeros_like(p)
    g[-1, :2] = 2.0 * d
    g[0, :2] = -2.0 * d
    return float(d @ d), g.ravel()


def tune_lambda_drift(graph: Graph, init_poses, steps: int = 10, lr: float = 0.1,
                      lam0: float = 1.0, converge_kw: dict | None = None, adjoint_tol=1e-10):
    """Adam on log lambda of the converged poses' squared drift -- the
    gradient replacement of the README's coarse-then-fine lambda grid search
    (README.md:137-163).  Works on the reference's landmark-pair / motion edge
    model (graph created with anchor_prior=1): each step runs Gauss-Newton to
    convergence from init_poses, then gslam_adjoint's tangent solve for
    dL/dlambda.  Returns (lambda, history of (lambda, L, dL/dlambda))."""
    kw = dict(solver="pcg64", tol=1e-13, max_iter=400000, diff_tol=1e-12, max_rounds=3000,
              patience=30)
    kw.update(converge_kw or {})
    log_lam = float(np.log(lam0))
    m = v = 0.0
    b1, b2, eps = 0.9, 0.999, 1e-12
    hist = []
    for t in range(1, steps + 1):
        lam = float(np.exp(log_lam))
        graph.set_lambda(lam)
        graph.set_poses(init_poses)
        graph.converge(**kw)
        L, g = drift_loss(graph.get_poses())
        _, dlam, _ = graph.adjoint(g, tol=adjoint_tol)
        hist.append((lam, L, dlam))
        grad = dlam * lam                      # dL / d log lambda
        m = b1 * m + (1 - b1) * grad
        v = b2 * v + (1 - b2) * grad * grad
        log_lam -= lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return float(np.exp(log_lam)), hist
