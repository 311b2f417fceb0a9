"""SLAM++ ODOMETRY / LANDMARK files -> the reference's .log.txt records.

SLAM++ records (data/cityTrees10k.txt and the other cityTrees*.txt files):
    ODOMETRY from to dx dy dtheta  I_xx I_xy I_xt I_yy I_yt I_tt
    LANDMARK pose landmark dx dy   I_xx I_xy I_yy
Poses and landmarks share one id space; the .log.txt format numbers poses
0..N-1 in odometry order and keeps the landmark ids.

Two conventions:

* "data" (default) reproduces the reference's shipped conversions, e.g.
  data/cityTrees10k.log.txt from data/cityTrees10k.txt: poses are the SE(2)
  composition of the odometry with theta kept in [0, 2 pi); the control of
  step k is (nu, omega) with omega = dtheta and nu the speed of the
  constant-turn arc whose chord is the step, |d| |omega| / (2 |sin(omega/2)|);
  an observation is range = hypot(dx, dy), bearing = atan2(dx, dy) (the
  shipped files use this argument order).
* "script" restates data/convert_slampp.py:29-84 as written: pose ids are
  taken as consecutive (from_id -> from_id + 1, gaps repeat the pose with a
  zero control), poses are integrated with the velocity model of
  state_transition (GSLAM.c:373-401 restated in convert_slampp.py:17-27)
  with nu = |d|, omega = dtheta, and the LANDMARK values are copied as
  (range, bearing) unchanged.

    python -m gslam.slampp input.txt output.log.txt [delta] [--convention script]
"""
from __future__ import annotations

import math
import sys

import numpy as np


def read_slampp(path: str) -> dict:
    """ODOMETRY and LANDMARK records of a SLAM++ text file (other lines skipped)."""
    odo, lm = [], []
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if p[0] == "ODOMETRY":
                odo.append([float(v) for v in p[1:12]])
            elif p[0] == "LANDMARK":
                lm.append([float(v) for v in p[1:8]])
    odo = np.array(odo, dtype=np.float64).reshape(-1, 11)
    lm = np.array(lm, dtype=np.float64).reshape(-1, 7)
    return dict(o_from=odo[:, 0].astype(np.int64), o_to=odo[:, 1].astype(np.int64),
                o_meas=odo[:, 2:5].copy(), o_info=odo[:, 5:11].copy(),
                l_pose=lm[:, 0].astype(np.int64), l_id=lm[:, 1].astype(np.int64),
                l_meas=lm[:, 2:4].copy(), l_info=lm[:, 4:7].copy())


def _to_log_data(d: dict):
    n = len(d["o_from"]) + 1
    step = {0: 0} if n > 0 else {}
    if len(d["o_from"]):
        step = {int(d["o_from"][0]): 0}
    x = np.zeros((n, 3))
    u = np.zeros((n, 2))
    c_th = 0.0
    px = py = 0.0
    for k in range(n - 1):
        a, b = int(d["o_from"][k]), int(d["o_to"][k])
        if a not in step:
            raise ValueError(f"ODOMETRY {a} {b}: pose {a} not reached by the chain")
        step[b] = step[a] + 1
        dx, dy, dt = d["o_meas"][k]
        c, s = math.cos(c_th), math.sin(c_th)
        px, py, c_th = px + c * dx - s * dy, py + s * dx + c * dy, c_th + dt
        x[k + 1] = (px, py, c_th)
        dist = math.hypot(dx, dy)
        u[k + 1] = (dist * abs(dt) / (2.0 * abs(math.sin(dt / 2.0))) if dt != 0.0 else dist, dt)
    x[:, 2] = np.mod(x[:, 2], 2.0 * math.pi)
    z = np.zeros((len(d["l_pose"]), 4))
    for k, (pid, lid, (dx, dy)) in enumerate(zip(d["l_pose"], d["l_id"], d["l_meas"])):
        z[k] = (step[int(pid)], lid, math.hypot(dx, dy), math.atan2(dx, dy))
    return u, x, z


def _state_transition(x, y, th, nu, om, delta):
    # constant-turn motion model of convert_slampp.py:17-27
    if abs(om) < 1e-10:
        return x + nu * math.cos(th) * delta, y + nu * math.sin(th) * delta, th + om * delta
    return (x + nu / om * (math.sin(th + om * delta) - math.sin(th)),
            y + nu / om * (math.cos(th) - math.cos(th + om * delta)), th + om * delta)


def _to_log_script(d: dict, delta: float):
    odo = {int(a): tuple(m) for a, m in zip(d["o_from"], d["o_meas"])}
    n = max(odo) + 2
    x = np.zeros((n, 3))
    u = np.zeros((n, 2))
    for k in range(n - 1):
        if k not in odo:
            x[k + 1] = x[k]
            continue
        dx, dy, dt = odo[k]
        nu = math.sqrt(dx * dx + dy * dy)
        x[k + 1] = _state_transition(*x[k], nu, dt, delta)
        u[k + 1] = (nu, dt)
    z = np.column_stack([d["l_pose"], d["l_id"], d["l_meas"]]).astype(np.float64)
    return u, x, z


def to_log(d: dict, delta: float = 1.0, convention: str = "data"):
    """(u[N,2], x[N,3], z[M,4]) records of the .log.txt format."""
    if convention == "data":
        return _to_log_data(d)
    if convention == "script":
        return _to_log_script(d, delta)
    raise ValueError("convention must be 'data' or 'script'")


def write_log(path: str, u, x, z, delta: float = 1.0) -> None:
    """delta / u / x lines per step, then the z lines (GSLAM.c:415-493 reads them)."""
    with open(path, "w") as f:
        f.write(f"delta {delta!r}\n")
        for k in range(len(x)):
            f.write(f"u {k} {float(u[k, 0])!r} {float(u[k, 1])!r}\n")
            f.write(f"x {k} {float(x[k, 0])!r} {float(x[k, 1])!r} {float(x[k, 2])!r}\n")
        for r in z:
            f.write(f"z {int(r[0])} {int(r[1])} {float(r[2])!r} {float(r[3])!r}\n")


def convert(src: str, dst: str, delta: float = 1.0, convention: str = "data") -> tuple:
    u, x, z = to_log(read_slampp(src), delta, convention)
    write_log(dst, u, x, z, delta)
    return len(x), len(z)


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    conv = "data"
    if "--convention" in argv:
        i = argv.index("--convention")
        conv = argv[i + 1]
        del argv[i:i + 2]
    if len(argv) < 2:
        print("usage: python -m gslam.slampp input.txt output.log.txt [delta] "
              "[--convention data|script]", file=sys.stderr)
        return 1
    delta = float(argv[2]) if len(argv) > 2 else 1.0
    n, m = convert(argv[0], argv[1], delta, conv)
    print(f"{argv[0]} -> {argv[1]}: {n} poses, {m} observations, dim {3 * n}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
