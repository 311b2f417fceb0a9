"""Mathematical pins of the oracle that do not rely on the reference text:
finite-difference Jacobians of each edge residual, closed forms, and library
solves.  A dropped term, a sign or an index slip in an edge model fails the
finite-difference checks; a wrong PCG recurrence fails the KATs."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _fd_jac(f, x, h=1e-6):
    x = np.asarray(x, float)
    f0 = f(x)
    J = np.zeros((len(f0), len(x)))
    for k in range(len(x)):
        d = np.zeros_like(x)
        d[k] = h
        J[:, k] = (f(x + d) - f(x - d)) / (2 * h)
    return J


def _blocks(out):
    return [out[9 * q: 9 * q + 9].reshape(3, 3) for q in range(4)], out[36:39], out[39:42]


def _wrap(a):
    return (a + np.pi) % (2 * np.pi) - np.pi


def test_between_edge_fd(oracle):
    rng = np.random.default_rng(11)
    for _ in range(50):
        xi = np.array([rng.uniform(-9, 9), rng.uniform(-9, 9), rng.uniform(-3, 3)])
        xj = xi + np.array([rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(-1, 1)])
        z = rng.normal(size=3) * 0.5
        info = 1.0 / rng.uniform(0.02, 0.3, 3) ** 2
        w = rng.uniform(0.05, 3.0)

        def res(v):
            a, b = v[:3], v[3:]
            c, s = np.cos(a[2]), np.sin(a[2])
            d = b[:2] - a[:2]
            return np.array([c * d[0] + s * d[1] - z[0], -s * d[0] + c * d[1] - z[1],
                             _wrap(b[2] - a[2] - z[2])])

        out, e = oracle.between_edge(xi, xj, z, info, w, 1)
        v = np.concatenate([xi, xj])
        assert np.allclose(e, res(v), rtol=0, atol=1e-12)
        J = _fd_jac(res, v)
        L = np.diag(info * w)
        H = J.T @ L @ J
        g = -J.T @ L @ res(v)
        (ul, ur, bl, br), x1, x2 = _blocks(out)
        Hq = np.block([[ul, ur], [bl, br]])
        scale = np.abs(H).max()
        assert np.abs(Hq - H).max() <= 1e-7 * scale
        assert np.abs(np.concatenate([x1, x2]) - g).max() <= 1e-7 * max(np.abs(g).max(), 1)


def test_obs_edge_fd(oracle):
    """ObsEdge (ObsEdge.c:10-212): B = de/dx, R = de/dz, Sigma = sum R Q R^T."""
    rng = np.random.default_rng(12)
    snr = (0.14, 0.05)
    for _ in range(50):
        x1 = np.array([rng.uniform(-9, 9), rng.uniform(-9, 9), rng.uniform(-3, 3)])
        x2 = np.array([rng.uniform(-9, 9), rng.uniform(-9, 9), rng.uniform(-3, 3)])
        r1, r2 = rng.uniform(0.5, 9, 2)
        p1, p2 = rng.uniform(-3, 3, 2)

        def res(v):
            a, b, zr1, zp1, zr2, zp2 = v[:3], v[3:6], v[6], v[7], v[8], v[9]
            return np.array([b[0] - a[0] + zr2 * np.cos(b[2] + zp2) - zr1 * np.cos(a[2] + zp1),
                             b[1] - a[1] + zr2 * np.sin(b[2] + zp2) - zr1 * np.sin(a[2] + zp1)])

        v = np.concatenate([x1, x2, [r1, p1, r2, p2]])
        J = _fd_jac(res, v)
        B, R1, R2 = J[:, :6], J[:, 6:8], J[:, 8:10]
        Q1 = np.diag([(r1 * snr[0]) ** 2, snr[1] ** 2])
        Q2 = np.diag([(r2 * snr[0]) ** 2, snr[1] ** 2])
        Om = np.linalg.inv(R1 @ Q1 @ R1.T + R2 @ Q2 @ R2.T)
        H = B.T @ Om @ B
        g = -B.T @ Om @ res(v)
        out, e, om = oracle.obs_edge(x1, x2, r1, p1, r2, p2)
        (ul, ur, bl, br), xa, xb = _blocks(out)
        assert np.allclose(e, res(v), atol=1e-12)
        assert np.abs(np.block([[ul, ur], [bl, br]]) - H).max() <= 1e-6 * np.abs(H).max()
        assert np.abs(np.concatenate([xa, xb]) - g).max() <= 1e-6 * max(np.abs(g).max(), 1e-3)


def test_motion_edge_fd(oracle):
    """MotionEdge (MotionEdge.c:9-193): F = dg/dx1, A = dg/du, e = x2 - g."""
    rng = np.random.default_rng(13)
    mns = (0.19, 0.001, 0.13, 0.2)
    for _ in range(50):
        x1 = np.array([rng.uniform(-9, 9), rng.uniform(-9, 9), rng.uniform(-3, 3)])
        x2 = x1 + rng.normal(size=3)
        nu, om_, lam = rng.uniform(0.2, 2), rng.choice([-1, 1]) * rng.uniform(0.05, 0.8), 0.7

        def g(v):
            th, n_, w_ = v[2], v[3], v[4]
            return np.array([v[0] + n_ / w_ * (np.sin(th + w_) - np.sin(th)),
                             v[1] + n_ / w_ * (np.cos(th) - np.cos(th + w_)), th + w_])

        v = np.concatenate([x1, [nu, om_]])
        J = _fd_jac(g, v)
        F, A = J[:, :3], J[:, 3:]
        M = np.diag([mns[0] ** 2 * abs(nu) + mns[1] ** 2 * abs(om_),
                     mns[2] ** 2 * abs(nu) + mns[3] ** 2 * abs(om_)])
        Om = np.linalg.inv(A @ M @ A.T + 1e-4 * np.eye(3))
        e = x2 - g(v)
        out, omo = oracle.motion_edge(x1, x2, nu, om_, 1.0, lam)
        (ul, ur, bl, br), xa, xb = _blocks(out)
        s = np.abs(Om).max()
        assert np.abs(omo.reshape(3, 3) - Om).max() <= 1e-6 * s
        assert np.abs(ul - lam * F.T @ Om @ F).max() <= 1e-6 * s * 10
        assert np.abs(ur + lam * F.T @ Om).max() <= 1e-6 * s * 10
        assert np.abs(bl + lam * Om @ F).max() <= 1e-6 * s * 10
        assert np.abs(br - lam * Om).max() <= 1e-6 * s
        assert np.abs(xa - lam * F.T @ Om @ e).max() <= 1e-6 * s * 10
        assert np.abs(xb + lam * Om @ e).max() <= 1e-6 * s * 10


def _csr_from_dense(a):
    m = sp.csr_matrix(a)
    m.sort_indices()
    return m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data.astype(np.float64)


def test_pcg_block_diagonal_one_step(oracle):
    """Block-Jacobi on a block-diagonal SPD matrix is exact: one iteration."""
    rng = np.random.default_rng(14)
    n = 20
    A = np.zeros((3 * n, 3 * n))
    for i in range(n):
        b = rng.normal(size=(3, 3))
        A[3 * i:3 * i + 3, 3 * i:3 * i + 3] = b @ b.T + 3 * np.eye(3)
    rhs = rng.normal(size=3 * n)
    x, it = oracle.pcg(*_csr_from_dense(A), rhs, 1e-12, 50)
    assert it == 1
    assert np.allclose(x, np.linalg.solve(A, rhs), rtol=1e-12, atol=1e-14)


def test_pcg_laplacian_finite_termination_and_spsolve(oracle):
    """CG on an SPD system of dimension d terminates in <= d steps (in exact
    arithmetic); at tight tolerance it agrees with a direct sparse solve."""
    rng = np.random.default_rng(15)
    n = 30
    d = 3 * n
    L = sp.diags([-np.ones(d - 1), 2.0 * np.ones(d), -np.ones(d - 1)], [-1, 0, 1]).toarray()
    L[0, 0] += 1.0
    rhs = rng.normal(size=d)
    x, it = oracle.pcg(*_csr_from_dense(L), rhs, 1e-13, 10 * d)
    assert it <= d + 5
    ref = spla.spsolve(sp.csc_matrix(L), rhs)
    assert np.linalg.norm(x - ref) <= 1e-9 * np.linalg.norm(ref)


def test_pcg_history_is_residual(oracle):
    """The recorded ||r|| equals ||b - A x_k|| (the recurrence is the true
    residual up to rounding)."""
    rng = np.random.default_rng(16)
    d = 60
    M = rng.normal(size=(d, d))
    A = M @ M.T + d * np.eye(d)
    b = rng.normal(size=d)
    rp, ci, va = _csr_from_dense(A)
    for k in (1, 3, 7):
        x, it, h = oracle.pcg(rp, ci, va, b, 1e-30, k, hist=True)
        assert it == k
        assert abs(h[-1] - np.linalg.norm(b - A @ x)) <= 1e-9 * np.linalg.norm(b)


@pytest.mark.parametrize("n", [1, 4, 8])
def test_gj_hilbert_and_numpy(oracle, n):
    H = 1.0 / (np.arange(n)[:, None] + np.arange(n)[None, :] + 1.0)
    inv, rc = oracle.gj_inv(H)
    assert rc == 0
    # exact integer inverse of the Hilbert matrix (closed form)
    from math import comb
    exact = np.array([[(-1) ** (i + j) * (i + j + 1) * comb(n + i, n - j - 1)
                       * comb(n + j, n - i - 1) * comb(i + j, i) ** 2 for j in range(n)]
                      for i in range(n)], float)
    assert np.abs(inv - exact).max() <= 1e-6 * np.abs(exact).max() * 10 ** (n / 2)
    rng = np.random.default_rng(n)
    A = rng.normal(size=(12, 12))
    A = A @ A.T + 12 * np.eye(12)
    inv, _ = oracle.gj_inv(A)
    assert np.allclose(inv, np.linalg.inv(A), rtol=1e-12, atol=1e-14)
    assert np.allclose(inv @ A, np.eye(12), atol=1e-12)


def test_csr_pattern_brute_force(oracle):
    """The sort-based pattern equals the dense marker of MatrixOps.c:177-221
    computed by brute force, including t1 == t2 edges and an isolated pose 0."""
    rng = np.random.default_rng(17)
    for n, E, iso in ((6, 9, False), (9, 14, True), (1, 0, True)):
        t1 = rng.integers(1 if iso else 0, n, E).astype(np.int32)
        t2 = rng.integers(1 if iso else 0, n, E).astype(np.int32)
        if E:
            t2[0] = t1[0]
        mark = np.zeros((3 * n, 3 * n), bool)
        mark[0, 0] = mark[1, 1] = mark[2, 2] = True
        for a, b in zip(t1, t2):
            for u in (a, b):
                for v in (a, b):
                    mark[3 * u:3 * u + 3, 3 * v:3 * v + 3] = True
        rp, ci = oracle.csr_pattern(n, t1, t2)
        r, c = np.nonzero(mark)
        assert np.array_equal(rp, np.concatenate([[0], np.cumsum(mark.sum(1))]))
        assert np.array_equal(ci, c)


def test_pose_update_and_diff(oracle):
    p = np.array([1.0, 2.0, 3.0, -1.0, 0.5, -3.0])
    d = np.array([0.5, -0.5, 0.3, 0.0, 0.0, -0.5])
    assert np.array_equal(oracle.pose_update(p, d, 0), p + d)
    w = oracle.pose_update(p, d, 1)
    assert np.allclose(w[[2, 5]], _wrap(p + d)[[2, 5]], atol=1e-15)
    assert -np.pi <= w[2] < np.pi and -np.pi <= w[5] < np.pi
    assert oracle.diff(d) == np.sqrt(np.sum(d * d)) / np.sqrt(6.0)
