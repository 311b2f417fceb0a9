"""FP64 NumPy implicit differentiation of converged between-edge pose graphs.

Written independently of the HIP kernels, for small graphs (dense
matrices):
* f(x) = 1/2 sum_e e_e^T L_e e_e + 1/2 a |x_0 - ref|^2 with
  e = [R(th_i)^T (t_j - t_i) - z_t, wrap(th_j - th_i - z_th)] and
  L_e = exp(w_e) lam diag(info_e);
* the analytic gradient grad f, Gauss-Newton to convergence;
* the exact Hessian as central differences of grad f (so the second-order
  terms are not taken from the kernels' closed form);
* dL/dw_e = -v^T dF/dw_e, H v = dL/dx, dF/dw_e = J_e^T L_e e_e;
  dL/dlam = -v^T dF/dlam with dF/dlam = (grad f - anchor gradient) / lam.
"""
import numpy as np


def wrap(a):
    return (a + np.pi) % (2 * np.pi) - np.pi


def residuals(x, g):
    p = x.reshape(-1, 3)
    I, J = g["i"], g["j"]
    z = g["meas"].reshape(-1, 3)
    c, s = np.cos(p[I, 2]), np.sin(p[I, 2])
    dx, dy = p[J, 0] - p[I, 0], p[J, 1] - p[I, 1]
    e = np.stack([c * dx + s * dy - z[:, 0], -s * dx + c * dy - z[:, 1],
                  wrap(p[J, 2] - p[I, 2] - z[:, 2])], 1)
    Ji = np.zeros((len(I), 3, 3))
    Ji[:, 0, 0], Ji[:, 0, 1], Ji[:, 0, 2] = -c, -s, -s * dx + c * dy
    Ji[:, 1, 0], Ji[:, 1, 1], Ji[:, 1, 2] = s, -c, -c * dx - s * dy
    Ji[:, 2, 2] = -1.0
    Jj = np.zeros((len(I), 3, 3))
    Jj[:, 0, 0], Jj[:, 0, 1] = c, s
    Jj[:, 1, 0], Jj[:, 1, 1] = -s, c
    Jj[:, 2, 2] = 1.0
    return e, Ji, Jj


def weights(g, log_w, lam):
    return np.exp(log_w)[:, None] * lam * g["info"].reshape(-1, 3)


def grad_f(x, g, log_w, lam, anchor, ref, edge_terms=False):
    e, Ji, Jj = residuals(x, g)
    Le = weights(g, log_w, lam) * e
    gi = np.einsum("kab,ka->kb", Ji, Le)
    gj = np.einsum("kab,ka->kb", Jj, Le)
    n = x.size // 3
    G = np.zeros((n, 3))
    np.add.at(G, g["i"], gi)
    np.add.at(G, g["j"], gj)
    G_anchor = np.zeros((n, 3))
    G_anchor[0] = anchor * (x[:3] - ref)
    if edge_terms:
        return (G + G_anchor).ravel(), gi, gj, G.ravel()
    return (G + G_anchor).ravel()


def gn_matrix(x, g, log_w, lam, anchor):
    e, Ji, Jj = residuals(x, g)
    L = weights(g, log_w, lam)
    n = x.size // 3
    H = np.zeros((3 * n, 3 * n))
    for k, (a, b) in enumerate(zip(g["i"], g["j"])):
        for (p, JA) in ((a, Ji[k]), (b, Jj[k])):
            for (q, JB) in ((a, Ji[k]), (b, Jj[k])):
                H[3 * p:3 * p + 3, 3 * q:3 * q + 3] += JA.T @ (L[k][:, None] * JB)
    H[:3, :3] += anchor * np.eye(3)
    return H


def converge(x0, g, log_w, lam, anchor, ref, tol=1e-14, rounds=60):
    x = x0.copy()
    for _ in range(rounds):
        dx = -np.linalg.solve(gn_matrix(x, g, log_w, lam, anchor), grad_f(x, g, log_w, lam, anchor, ref))
        x = x + dx
        x[2::3] = wrap(x[2::3])
        if np.linalg.norm(dx) < tol * max(1.0, np.linalg.norm(x)):
            break
    return x


def hessian_fd(x, g, log_w, lam, anchor, ref, h=1e-6):
    n = x.size
    H = np.zeros((n, n))
    for k in range(n):
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        H[:, k] = (grad_f(xp, g, log_w, lam, anchor, ref) - grad_f(xm, g, log_w, lam, anchor, ref)) / (2 * h)
    return 0.5 * (H + H.T)


def implicit_grad(x, g, log_w, lam, anchor, ref, dl_dx, H=None):
    """(dL/dw_e, dL/dlam) at converged x for the loss gradient dl_dx."""
    if H is None:
        H = hessian_fd(x, g, log_w, lam, anchor, ref)
    v = np.linalg.solve(H, dl_dx)
    _, gi, gj, G_edges = grad_f(x, g, log_w, lam, anchor, ref, edge_terms=True)
    V = v.reshape(-1, 3)
    dw = -(np.einsum("ka,ka->k", V[g["i"]], gi) + np.einsum("ka,ka->k", V[g["j"]], gj))
    dlam = -(v @ G_edges) / lam
    return dw, dlam


def loss_and_grad(x, truth):
    """Squared end-to-start drift error + squared pose error (wrapped heading),
    and its gradient in x."""
    p = x.reshape(-1, 3)
    t = np.asarray(truth).reshape(-1, 3)
    drift = (p[-1, :2] - p[0, :2]) - (t[-1, :2] - t[0, :2])
    d = p - t
    d[:, 2] = wrap(d[:, 2])
    L = drift @ drift + (d ** 2).sum()
    G = 2 * d
    G[-1, :2] += 2 * drift
    G[0, :2] -= 2 * drift
    return L, G.ravel()
