"""Seeded synthetic pose graphs with ground truth (SE(2) between edges).

The reference ships only a landmark-world generator
(data/generate_citytrees.py:59-204) that writes no ground truth; these two
pose-graph generators produce odometry + loop-closure "between" edges, the
dead-reckoned initial guess and the true poses.

grid_city: random walk on a unit Manhattan grid (1 m steps, 90 degree turns
  with probability p_turn), odometry noise sigma_xy / sigma_th, loop closures
  between poses closer than lc_radius (true positions) with index gap >
  lc_gap, each kept with probability lc_keep.
park: smooth random trajectory inside a side x side square -- a free walk
  reflecting off the walls ("wander") or laps of a smooth random closed route
  with lateral wander ("laps") -- with loop closures within lc_radius and
  gap > lc_gap, subsampled to lc_per_pose closure edges per pose (fewer when
  fewer candidate pairs exist).

Measurements are the true relative poses plus Gaussian noise; information is
diag(1/sigma^2) per edge (lambda and per-edge weights are applied by the
solver).
"""
from __future__ import annotations

import numpy as np


def _wrap(a):
    return (a + np.pi) % (2.0 * np.pi) - np.pi


def relative_pose(pi: np.ndarray, pj: np.ndarray) -> np.ndarray:
    """z = [R(th_i)^T (t_j - t_i), th_j - th_i] for arrays of poses (k, 3)."""
    c, s = np.cos(pi[:, 2]), np.sin(pi[:, 2])
    dx, dy = pj[:, 0] - pi[:, 0], pj[:, 1] - pi[:, 1]
    return np.stack([c * dx + s * dy, -s * dx + c * dy, _wrap(pj[:, 2] - pi[:, 2])], 1)


def compose(p: np.ndarray, z: np.ndarray) -> np.ndarray:
    """p (+) z for single poses."""
    c, s = np.cos(p[2]), np.sin(p[2])
    return np.array([p[0] + c * z[0] - s * z[1], p[1] + s * z[0] + c * z[1], p[2] + z[2]])


def _dead_reckon(z_odo: np.ndarray, p0: np.ndarray) -> np.ndarray:
    n = z_odo.shape[0] + 1
    out = np.zeros((n, 3))
    out[0] = p0
    # vectorised composition: headings are a cumulative sum, translations a
    # cumulative sum of rotated steps
    th = p0[2] + np.concatenate([[0.0], np.cumsum(z_odo[:, 2])])
    c, s = np.cos(th[:-1]), np.sin(th[:-1])
    step = np.stack([c * z_odo[:, 0] - s * z_odo[:, 1], s * z_odo[:, 0] + c * z_odo[:, 1]], 1)
    out[1:, :2] = p0[:2] + np.cumsum(step, 0)
    out[:, 2] = th
    return out


def _loop_pairs(xy: np.ndarray, radius: float, gap: int, chunk_pairs: int = 1 << 24):
    """All (i, j), i < j - gap, with |xy_i - xy_j| < radius (cell hashing).

    Pairs come out grouped by neighbour-cell offset (9 offsets in a fixed
    order), then by i ascending, then in the cell's sorted order -- the order
    the seeded subsampling of the generators depends on.  Vectorised: each
    pose's candidate range in every neighbour cell comes from searchsorted
    over the sorted cell keys, expanded in chunks of at most chunk_pairs."""
    n = len(xy)
    cell = np.floor(xy / radius).astype(np.int64)
    key = (cell[:, 0] + (1 << 20)) * (1 << 21) + (cell[:, 1] + (1 << 20))
    order = np.argsort(key, kind="stable")
    ks = key[order]
    I, J = [], []
    idx = np.arange(n, dtype=np.int64)
    for dcx in (-1, 0, 1):
        for dcy in (-1, 0, 1):
            nk = key + dcx * (1 << 21) + dcy
            lo = np.searchsorted(ks, nk, side="left")
            cnt = np.searchsorted(ks, nk, side="right") - lo
            csum = np.concatenate([[0], np.cumsum(cnt)])
            a = 0
            while a < n:
                # poses [a, b) with at most chunk_pairs candidates (at least one pose)
                b = int(np.searchsorted(csum, csum[a] + chunk_pairs, side="right")) - 1
                b = min(max(b, a + 1), n)
                c = cnt[a:b]
                tot = int(c.sum())
                if tot:
                    ii = np.repeat(idx[a:b], c)
                    first = np.repeat(lo[a:b] - (csum[a:b] - csum[a]), c)
                    cand = order[first + np.arange(tot)]
                    m = cand > ii + gap
                    ii, cand = ii[m], cand[m]
                    d = np.hypot(xy[cand, 0] - xy[ii, 0], xy[cand, 1] - xy[ii, 1])
                    m = d < radius
                    I.append(ii[m])
                    J.append(cand[m])
                a = b
    if not I:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    return np.concatenate(I), np.concatenate(J)


def _finish(truth, I, J, rng, sxy, sth, initial_from_odometry=True):
    n = truth.shape[0]
    sig = np.array([sxy, sxy, sth])
    oi = np.arange(n - 1)
    z_odo = relative_pose(truth[oi], truth[oi + 1]) + rng.normal(size=(n - 1, 3)) * sig
    z_lc = relative_pose(truth[I], truth[J]) + rng.normal(size=(len(I), 3)) * sig
    init = _dead_reckon(z_odo, truth[0])
    init[:, 2] = _wrap(init[:, 2])
    ei = np.concatenate([oi, I]).astype(np.int32)
    ej = np.concatenate([oi + 1, J]).astype(np.int32)
    meas = np.concatenate([z_odo, z_lc])
    info = np.tile(1.0 / sig ** 2, (len(ei), 1))
    return dict(n_poses=n, truth=truth, poses=init.ravel().copy(), i=ei, j=ej,
                meas=meas.ravel().copy(), info=info.ravel().copy(), n_odometry=n - 1,
                n_loop=len(I))


def grid_city(n_poses: int, seed: int = 0, p_turn: float = 0.2, sigma_xy: float = 0.05,
              sigma_th: float = 0.01, lc_radius: float = 1.5, lc_gap: int = 20,
              lc_keep: float = 0.3, max_loop_per_pose: int | None = None) -> dict:
    """Grid-city pose graph (random walk on a unit Manhattan grid)."""
    rng = np.random.default_rng(seed)
    turns = rng.random(n_poses - 1) < p_turn
    sign = np.where(rng.random(n_poses - 1) < 0.5, -1.0, 1.0)
    dth = np.where(turns, sign * np.pi / 2, 0.0)
    th = np.concatenate([[0.0], np.cumsum(dth)])
    step = np.stack([np.round(np.cos(th[1:])), np.round(np.sin(th[1:]))], 1)
    xy = np.concatenate([[[0.0, 0.0]], np.cumsum(step, 0)])
    truth = np.concatenate([xy, _wrap(th)[:, None]], 1)
    I, J = _loop_pairs(xy, lc_radius, lc_gap)
    keep = rng.random(len(I)) < lc_keep
    I, J = I[keep], J[keep]
    if max_loop_per_pose is not None and len(I):
        perm = rng.permutation(len(I))
        cnt = np.zeros(n_poses, np.int64)
        sel = []
        for k in perm:
            if cnt[I[k]] < max_loop_per_pose and cnt[J[k]] < max_loop_per_pose:
                cnt[I[k]] += 1
                cnt[J[k]] += 1
                sel.append(k)
        sel = np.sort(np.array(sel, np.int64))
        I, J = I[sel], J[sel]
    return _finish(truth, I, J, rng, sigma_xy, sigma_th)


def _park_laps(n_poses, rng, side, speed, wander):
    """Laps of a smooth random closed route: r(phi) = R0 + sum_k a_k cos(k phi
    + phi_k), k = 2..5, around the centre of the square (R0 = side/3, the
    amplitudes keep the route inside the square), walked at `speed` metres
    per pose with a smooth lateral offset (an AR(1) process of standard
    deviation `wander` metres), so that every lap passes within a few metres
    of the earlier ones -- a jogger on a park path."""
    r0 = side / 3.0
    ks = np.arange(2, 6)
    amp = rng.uniform(0.03, 0.08, len(ks)) * r0
    ph = rng.uniform(0.0, 2.0 * np.pi, len(ks))
    fine = np.linspace(0.0, 2.0 * np.pi, 20001)
    rad = r0 + (amp[:, None] * np.cos(ks[:, None] * fine[None, :] + ph[:, None])).sum(0)
    pts = np.stack([side / 2 + rad * np.cos(fine), side / 2 + rad * np.sin(fine)], 1)
    arc = np.concatenate([[0.0], np.cumsum(np.hypot(*np.diff(pts, axis=0).T))])
    s = (speed * np.arange(n_poses)) % arc[-1]
    base = np.stack([np.interp(s, arc, pts[:, 0]), np.interp(s, arc, pts[:, 1])], 1)
    tang = np.stack([np.interp(s, arc, np.gradient(pts[:, 0])),
                     np.interp(s, arc, np.gradient(pts[:, 1]))], 1)
    tang /= np.linalg.norm(tang, axis=1, keepdims=True)
    normal = np.stack([-tang[:, 1], tang[:, 0]], 1)
    off = np.zeros(n_poses)
    a = 0.98
    e = rng.normal(0.0, wander * np.sqrt(1.0 - a * a), n_poses)
    for k in range(1, n_poses):
        off[k] = a * off[k - 1] + e[k]
    xy = np.clip(base + off[:, None] * normal, 0.0, side)
    d = np.diff(xy, axis=0)
    th = np.concatenate([np.arctan2(d[:, 1], d[:, 0]), [0.0]])
    th[-1] = th[-2]
    return xy, th


def park(n_poses: int = 7120, seed: int = 0, side: float = 300.0, speed: float = 1.0,
         sigma_xy: float = 0.1, sigma_th: float = 0.02, lc_radius: float = 3.0,
         lc_gap: int = 50, lc_per_pose: float = 4.0, route: str = "wander",
         wander: float = 1.0) -> dict:
    """Park-like graph: smooth random trajectory inside a side x side square.

    route "wander": a free random walk with smoothly varying turn rate that
    reflects off the walls (revisits are rare: about 1.6 closures per pose
    survive at 7,120 poses); route "laps": laps of a smooth random closed
    route with lateral wander (_park_laps), dense enough in revisits that the
    subsampling reaches lc_per_pose closures per pose."""
    rng = np.random.default_rng(seed)
    if route == "laps":
        xy, th = _park_laps(n_poses, rng, side, speed, wander)
    elif route == "wander":
        th = np.zeros(n_poses)
        xy = np.zeros((n_poses, 2))
        xy[0] = side / 2
        omega = 0.0
        for k in range(1, n_poses):
            omega = 0.9 * omega + rng.normal(0.0, 0.03)
            t = th[k - 1] + omega
            p = xy[k - 1] + speed * np.array([np.cos(t), np.sin(t)])
            # reflect off the walls
            if p[0] < 0 or p[0] > side:
                t = np.pi - t
            if p[1] < 0 or p[1] > side:
                t = -t
            p = np.clip(xy[k - 1] + speed * np.array([np.cos(t), np.sin(t)]), 0, side)
            th[k], xy[k] = t, p
    else:
        raise ValueError(f"unknown route {route!r}")
    truth = np.concatenate([xy, _wrap(th)[:, None]], 1)
    I, J = _loop_pairs(xy, lc_radius, lc_gap)
    if len(I):
        want = int(lc_per_pose * n_poses)
        keep = rng.permutation(len(I))[: min(want, len(I))]
        keep.sort()
        I, J = I[keep], J[keep]
    return _finish(truth, I, J, rng, sigma_xy, sigma_th)


def graph_kwargs(g: dict, lam: float, wrap: int = 1) -> dict:
    """Keyword arguments of gslam.Graph for a generated pose graph."""
    return dict(lam=lam, between=dict(i=g["i"], j=g["j"], meas=g["meas"], info=g["info"]),
                wrap=wrap)


def write_pose_graph(g: dict, log_path: str, between_path: str) -> None:
    """Writes a generated pose graph for the GSLAM command-line driver: the
    initial poses as .log.txt x records ("x <i> <x> <y> <theta>", the record
    src/cpu_pcg_fp64/GSLAM.c:450-466 reads) and the between edges as the
    sidecar GSLAM_BETWEEN reads ("e <i> <j> <dx> <dy> <dth> <info_x> <info_y>
    <info_th>").  Values as %.17g, so they read back to the same doubles."""
    p = np.asarray(g["poses"]).reshape(-1, 3)
    with open(log_path, "w") as f:
        for k, (x, y, t) in enumerate(p):
            f.write(f"x {k} {x:.17g} {y:.17g} {t:.17g}\n")
    m = np.asarray(g["meas"]).reshape(-1, 3)
    inf = np.asarray(g["info"]).reshape(-1, 3)
    with open(between_path, "w") as f:
        f.write("# SE(2) between edges: e i j dx dy dth info_x info_y info_th\n")
        for i, j, a, b in zip(g["i"], g["j"], m, inf):
            f.write(f"e {int(i)} {int(j)} {a[0]:.17g} {a[1]:.17g} {a[2]:.17g} "
                    f"{b[0]:.17g} {b[1]:.17g} {b[2]:.17g}\n")


def _velocity_step(x, y, th, nu, om, delta=1.0):
    """The velocity motion model the reference's motion edge linearises
    (state_transition, src/cpu_pcg_fp64/GSLAM.c:373-401)."""
    if abs(om) < 1e-10:
        return x + nu * np.cos(th) * delta, y + nu * np.sin(th) * delta, th
    return (x + nu / om * (np.sin(th + om * delta) - np.sin(th)),
            y + nu / om * (-np.cos(th + om * delta) + np.cos(th)), th + om * delta)


def landmark_world(n_poses: int, seed: int = 0, block: int = 25, odo_noise: float = 1.0,
                   obs_noise=(0.05, 0.01), sensor_range: float = 8.5, density: float = 3e-3,
                   max_obs_per_landmark: int = 10, strip: float = 20.0) -> dict:
    """A landmark world in the reference's .log.txt model (u / x / z records),
    after data/generate_citytrees.py:59-204: a robot drives laps of a closed
    block x block route of 1 m steps with 90 degree turns (so the last pose
    returns near the first, which makes the README's end-to-start drift a
    meaningful loss), trees are scattered along the route and observed by
    range and bearing within sensor_range (at most max_obs_per_landmark
    times each).  Unlike the reference generator, the controls written to the
    log are the noisy odometry (motion noise of the reference's motion edge,
    src/cpu_pcg_fp64/MotionEdge.c:23-26, scaled by odo_noise), the initial
    poses are dead-reckoned from them, and the observations carry Gaussian
    range / bearing noise obs_noise: with consistent data every residual
    vanishes at the solution and lambda has no effect on it.  Returns the
    fields of gslam.read_log plus the true poses ('truth')."""
    rng = np.random.default_rng(seed)
    us = [(0.0, 0.0)]
    k = 1
    while len(us) < n_poses:
        for _ in range(block - 1):
            us.append((1.0, 0.0))
        us.append((1.0, np.pi / 2))
        k += block
    us = us[:n_poses]
    gt = [(0.0, 0.0, 0.0)]
    for nu, om in us[1:]:
        gt.append(_velocity_step(*gt[-1], nu, om))
    gt = np.array(gt)
    odo = [(0.0, 0.0)]
    for nu, om in us[1:]:
        sn = np.sqrt(0.19 ** 2 * abs(nu) + 0.001 ** 2 * abs(om) + 1e-9) * odo_noise
        so = np.sqrt(0.13 ** 2 * abs(nu) + 0.2 ** 2 * abs(om) + 1e-9) * odo_noise
        odo.append((nu + rng.normal(0.0, sn), om + rng.normal(0.0, so)))
    init = [(0.0, 0.0, 0.0)]
    for nu, om in odo[1:]:
        init.append(_velocity_step(*init[-1], nu, om))
    init = np.array(init)
    init[:, 2] = _wrap(init[:, 2])
    lo, hi = gt[:, :2].min(0) - strip, gt[:, :2].max(0) + strip
    n_lm = max(10, int(np.prod(hi - lo) * density))
    lms = []
    while len(lms) < n_lm:
        m = rng.uniform(lo, hi)
        if (np.hypot(*(gt[:, :2] - m).T) <= sensor_range).any():
            lms.append(m)
    z, cnt = [], np.zeros(len(lms), np.int64)
    for s, (px, py, pt) in enumerate(gt):
        for li, (mx, my) in enumerate(lms):
            if cnt[li] >= max_obs_per_landmark:
                continue
            d = np.hypot(mx - px, my - py)
            if d > sensor_range:
                continue
            b = _wrap(np.arctan2(my - py, mx - px) - pt)
            z.append((s, li + 6, d + rng.normal(0.0, obs_noise[0]), b + rng.normal(0.0, obs_noise[1])))
            cnt[li] += 1
    z = np.array(z).reshape(-1, 4)
    return dict(delta=1.0, u_step=np.arange(n_poses, dtype=np.int32),
                u_nu=np.array([u[0] for u in odo]), u_omega=np.array([u[1] for u in odo]),
                x_step=np.arange(n_poses, dtype=np.int32), poses=init.ravel().copy(),
                z_step=z[:, 0].astype(np.int32), z_lm=z[:, 1].astype(np.int32),
                z_range=z[:, 2].copy(), z_bearing=z[:, 3].copy(), n_poses=n_poses, truth=gt)


def write_log(g: dict, path: str) -> None:
    """Writes a landmark world as a reference .log.txt ("delta", then per step
    its u, x and z records, the layout of data/generate_citytrees.py:178-196),
    values as %.17g."""
    zs = {}
    for s, l, r, b in zip(g["z_step"], g["z_lm"], g["z_range"], g["z_bearing"]):
        zs.setdefault(int(s), []).append((int(l), r, b))
    p = np.asarray(g["poses"]).reshape(-1, 3)
    with open(path, "w") as f:
        f.write(f"delta {g['delta']:.17g}\n")
        for s in range(g["n_poses"]):
            f.write(f"u {s} {g['u_nu'][s]:.17g} {g['u_omega'][s]:.17g}\n")
            f.write(f"x {s} {p[s, 0]:.17g} {p[s, 1]:.17g} {p[s, 2]:.17g}\n")
            for l, r, b in zs.get(s, []):
                f.write(f"z {s} {l} {r:.17g} {b:.17g}\n")
