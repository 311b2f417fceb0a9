"""ctypes front-end to the CPU oracle (liboracle.so) and to oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- the checker, never the product path.

`Oracle` wraps gslam_oracle.c (the restatement).  `RefPCG64` / `RefGJMixed`
wrap builds of the reference C (src/cpu_pcg_fp64, src/cpu_gj_mixed) made by
`make -C oracle ref`; they exist only where the reference sources were present
at build time and are used to pin the restatement bit for bit.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
LIB_OMP = os.path.join(HERE, "liboracle_omp.so")
REF_DIR = os.path.join(HERE, "_ref")
EDGE_VALS = 42

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


def _ptr(a):
    """Pointer or NULL for optional array arguments."""
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref when the reference sources exist)."""
    if force or not (os.path.exists(LIB) and os.path.exists(LIB_OMP)):
        subprocess.run(["make", "-s", "-C", HERE, "liboracle.so", "liboracle_omp.so"],
                       check=True)
    ref = os.environ.get("GSLAM_REFERENCE", "/root/reference")
    if os.path.isdir(os.path.join(ref, "src", "cpu_pcg_fp64")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={ref}"], check=True)


class Oracle:
    def __init__(self, path: str = LIB):
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        self.L = L
        vp = C.c_void_p
        L.orc_log_count.argtypes = [C.c_char_p, vp, vp, vp]
        L.orc_log_read.argtypes = [C.c_char_p] + [vp] * 10
        L.orc_obs_pairs.argtypes = [C.c_int32, vp, C.c_int32, vp, vp]
        L.orc_obs_pairs.restype = C.c_int64
        L.orc_landmark_means.argtypes = [C.c_int32] + [vp] * 7
        L.orc_landmark_means.restype = C.c_int32
        L.orc_motion_edge.argtypes = [vp, vp, C.c_double, C.c_double, C.c_double,
                                      C.c_double, vp, vp, vp]
        L.orc_obs_edge.argtypes = [vp, vp, C.c_double, C.c_double, C.c_double,
                                   C.c_double, vp, vp, vp, vp]
        L.orc_between_edge.argtypes = [vp, vp, vp, vp, C.c_double, C.c_int32, vp, vp]
        L.orc_between_edges.argtypes = [C.c_int32, vp, C.c_int64, vp, vp, vp, vp, vp,
                                        C.c_double, C.c_int32, vp]
        L.orc_csr_pattern.argtypes = [C.c_int32, C.c_int64, vp, vp, vp, vp]
        L.orc_csr_pattern.restype = C.c_int64
        L.orc_csr_values.argtypes = [C.c_int32, vp, vp, C.c_int64, vp, vp, vp,
                                     C.c_double, vp, vp]
        L.orc_dense_values.argtypes = [C.c_int32, C.c_int64, vp, vp, vp, C.c_double,
                                       vp, vp]
        L.orc_pcg.argtypes = [C.c_int32, vp, vp, vp, vp, C.c_double, C.c_int32, vp, vp]
        L.orc_pcg.restype = C.c_int32
        L.orc_threads.argtypes = []
        L.orc_threads.restype = C.c_int32
        L.orc_block_jacobi.argtypes = [C.c_int32, vp, vp, vp, vp]
        L.orc_spmv.argtypes = [C.c_int32, vp, vp, vp, vp, vp]
        L.orc_gj_inv.argtypes = [C.c_int32, vp, vp]
        L.orc_gj_inv_f32.argtypes = [C.c_int32, vp, vp]
        L.orc_dense_matvec.argtypes = [C.c_int32, vp, vp, vp]
        L.orc_log_edges.argtypes = [C.c_int32, vp, vp, vp, C.c_double, C.c_double, vp,
                                    C.c_int32, vp, vp, vp, vp, C.c_int32, vp, vp, vp]
        L.orc_log_edges.restype = C.c_int64
        L.orc_diff.argtypes = [C.c_int32, vp]
        L.orc_diff.restype = C.c_double
        L.orc_pose_update.argtypes = [C.c_int32, vp, vp, C.c_int32]
        L.orc_log_gn_run.argtypes = [C.c_int32, vp, vp, vp, C.c_double, C.c_double,
                                     C.c_int32, vp, vp, vp, vp, C.c_int32, C.c_double,
                                     C.c_int32, vp]
        L.orc_log_gn_run.restype = C.c_int32
        L.orc_between_gn_step.argtypes = [C.c_int32, vp, C.c_int64, vp, vp, vp, vp, vp,
                                          C.c_double, C.c_double, C.c_int32, vp, vp,
                                          C.c_double, C.c_int32, vp, vp]
        L.orc_between_gn_step.restype = C.c_int32

    # ---------------------------------------------------------------- I/O
    def read_log(self, path: str) -> dict:
        """read_data, src/cpu_pcg_fp64/GSLAM.c:415-493."""
        nu, nx, nz = C.c_int32(), C.c_int32(), C.c_int32()
        rc = self.L.orc_log_count(path.encode(), C.byref(nu), C.byref(nx), C.byref(nz))
        if rc != 0:
            raise IOError(f"orc_log_count({path}) = {rc}")
        nu, nx, nz = nu.value, nx.value, nz.value
        d = dict(
            u_step=np.zeros(nu, np.int32), u_nu=np.zeros(nu), u_omega=np.zeros(nu),
            x_step=np.zeros(nx, np.int32), poses=np.zeros(3 * nx),
            z_step=np.zeros(nz, np.int32), z_lm=np.zeros(nz, np.int32),
            z_range=np.zeros(nz), z_bearing=np.zeros(nz))
        delta = C.c_double(0.0)
        rc = self.L.orc_log_read(path.encode(), C.byref(delta),
                                 *[_ptr(d[k]) for k in ("u_step", "u_nu", "u_omega", "x_step",
                                                        "poses", "z_step", "z_lm", "z_range",
                                                        "z_bearing")])
        if rc != 0:
            raise IOError(f"orc_log_read({path}) = {rc}")
        d["delta"] = delta.value
        d["n_poses"] = nx
        return d

    # ---------------------------------------------------------------- edges
    def obs_pairs(self, z_lm: np.ndarray, order: int = 0):
        z_lm = np.ascontiguousarray(z_lm, np.int32)
        n = self.L.orc_obs_pairs(len(z_lm), _ptr(z_lm), order, None, None)
        a = np.zeros(n, np.int32)
        b = np.zeros(n, np.int32)
        self.L.orc_obs_pairs(len(z_lm), _ptr(z_lm), order, _ptr(a), _ptr(b))
        return a, b

    def motion_edge(self, x1, x2, nu, omega, delta, lam, mns=(0.19, 0.001, 0.13, 0.2)):
        x1 = np.ascontiguousarray(x1, np.float64)
        x2 = np.ascontiguousarray(x2, np.float64)
        m = np.ascontiguousarray(mns, np.float64)
        out = np.zeros(EDGE_VALS)
        om = np.zeros(9)
        self.L.orc_motion_edge(_ptr(x1), _ptr(x2), nu, omega, delta, lam, _ptr(m),
                               _ptr(out), _ptr(om))
        return out, om

    def obs_edge(self, x1, x2, r1, p1, r2, p2, snr=(0.14, 0.05)):
        x1 = np.ascontiguousarray(x1, np.float64)
        x2 = np.ascontiguousarray(x2, np.float64)
        s = np.ascontiguousarray(snr, np.float64)
        out = np.zeros(EDGE_VALS)
        e = np.zeros(2)
        om = np.zeros(4)
        self.L.orc_obs_edge(_ptr(x1), _ptr(x2), r1, p1, r2, p2, _ptr(s), _ptr(out),
                            _ptr(e), _ptr(om))
        return out, e, om

    def between_edge(self, xi, xj, z, info, weight=1.0, wrap=1):
        a = [np.ascontiguousarray(v, np.float64) for v in (xi, xj, z, info)]
        out = np.zeros(EDGE_VALS)
        e = np.zeros(3)
        self.L.orc_between_edge(*[_ptr(v) for v in a], weight, wrap, _ptr(out), _ptr(e))
        return out, e

    def between_edges(self, poses, ei, ej, meas, info, weight, lam, wrap=1):
        poses = np.ascontiguousarray(poses, np.float64)
        ei = np.ascontiguousarray(ei, np.int32)
        ej = np.ascontiguousarray(ej, np.int32)
        meas = np.ascontiguousarray(meas, np.float64)
        info = np.ascontiguousarray(info, np.float64)
        w = None if weight is None else np.ascontiguousarray(weight, np.float64)
        out = np.zeros((len(ei), EDGE_VALS))
        self.L.orc_between_edges(len(poses) // 3, _ptr(poses), len(ei), _ptr(ei), _ptr(ej),
                                 _ptr(meas), _ptr(info), _ptr(w), lam, wrap, _ptr(out))
        return out

    def log_edges(self, g: dict, poses=None, lam: float = 0.1, order: int = 0,
                  mns=(0.19, 0.001, 0.13, 0.2)):
        """ObsEdges then MotionEdges in reference order (GSLAM.c:62-72)."""
        poses = np.ascontiguousarray(g["poses"] if poses is None else poses, np.float64)
        n = len(poses) // 3
        m = np.ascontiguousarray(mns, np.float64)
        args = [n, _ptr(poses), _ptr(g["u_nu"]), _ptr(g["u_omega"]), g["delta"], lam, _ptr(m),
                len(g["z_step"]), _ptr(g["z_step"]), _ptr(g["z_lm"]), _ptr(g["z_range"]),
                _ptr(g["z_bearing"]), order]
        E = self.L.orc_log_edges(*args, None, None, None)
        t1 = np.zeros(E, np.int32)
        t2 = np.zeros(E, np.int32)
        ev = np.zeros((E, EDGE_VALS))
        self.L.orc_log_edges(*args, _ptr(t1), _ptr(t2), _ptr(ev))
        return t1, t2, ev

    # ---------------------------------------------------------------- assembly
    def csr_pattern(self, n_poses, t1, t2):
        t1 = np.ascontiguousarray(t1, np.int32)
        t2 = np.ascontiguousarray(t2, np.int32)
        rp = np.zeros(3 * n_poses + 1, np.int32)
        nnz = self.L.orc_csr_pattern(n_poses, len(t1), _ptr(t1), _ptr(t2), _ptr(rp), None)
        ci = np.zeros(nnz, np.int32)
        self.L.orc_csr_pattern(n_poses, len(t1), _ptr(t1), _ptr(t2), _ptr(rp), _ptr(ci))
        return rp, ci

    def csr_values(self, n_poses, rp, ci, t1, t2, ev, anchor=1e6):
        ev = np.ascontiguousarray(ev, np.float64)
        vals = np.zeros(len(ci))
        xi = np.zeros(3 * n_poses)
        rc = self.L.orc_csr_values(n_poses, _ptr(rp), _ptr(ci), len(t1), _ptr(t1), _ptr(t2),
                                   _ptr(ev), anchor, _ptr(vals), _ptr(xi))
        if rc != 0:
            raise ValueError("edge outside the CSR pattern")
        return vals, xi

    def assemble_csr(self, n_poses, t1, t2, ev, anchor=1e6):
        rp, ci = self.csr_pattern(n_poses, t1, t2)
        vals, xi = self.csr_values(n_poses, rp, ci, t1, t2, ev, anchor)
        return rp, ci, vals, xi

    def dense(self, n_poses, t1, t2, ev, anchor=1e6):
        ev = np.ascontiguousarray(ev, np.float64)
        dim = 3 * n_poses
        om = np.zeros((dim, dim))
        xi = np.zeros(dim)
        self.L.orc_dense_values(n_poses, len(t1), _ptr(np.ascontiguousarray(t1, np.int32)),
                                _ptr(np.ascontiguousarray(t2, np.int32)), _ptr(ev), anchor,
                                _ptr(om), _ptr(xi))
        return om, xi

    # ---------------------------------------------------------------- solve
    def threads(self) -> int:
        """Threads of the PCG loops (1 unless this is liboracle_omp.so)."""
        return int(self.L.orc_threads())

    def pcg(self, rp, ci, vals, b, tol=1e-6, max_iter=500, hist=False):
        dim = len(rp) - 1
        x = np.zeros(dim)
        h = np.zeros(max(max_iter, 1)) if hist else None
        it = self.L.orc_pcg(dim, _ptr(rp), _ptr(ci), _ptr(vals), _ptr(b), tol, max_iter,
                            _ptr(x), _ptr(h))
        if hist:
            return x, it, h[:it]
        return x, it

    def block_jacobi(self, rp, ci, vals):
        dim = len(rp) - 1
        out = np.zeros((dim // 3, 9))
        self.L.orc_block_jacobi(dim, _ptr(rp), _ptr(ci), _ptr(vals), _ptr(out))
        return out

    def spmv(self, rp, ci, vals, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(len(rp) - 1)
        self.L.orc_spmv(len(rp) - 1, _ptr(rp), _ptr(ci), _ptr(vals), _ptr(x), _ptr(y))
        return y

    def gj_inv(self, a):
        a = np.ascontiguousarray(a, np.float64)
        out = np.zeros_like(a)
        rc = self.L.orc_gj_inv(a.shape[0], _ptr(a), _ptr(out))
        return out, rc

    def gj_inv_f32(self, a):
        a = np.ascontiguousarray(a, np.float32)
        out = np.zeros_like(a)
        rc = self.L.orc_gj_inv_f32(a.shape[0], _ptr(a), _ptr(out))
        return out, rc

    def dense_matvec(self, a, x):
        a = np.ascontiguousarray(a, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(a.shape[0])
        self.L.orc_dense_matvec(a.shape[0], _ptr(a), _ptr(x), _ptr(y))
        return y

    def gj_cost(self, g: dict, poses=None, order: int = 0, acc: float = 0.0) -> float:
        """acc + the cost terms of the reference GJ drivers' use_diff=0 test
        (src/gpu_gj_fp64/GSLAM.cu:245-257, src/cpu_gj_mixed/GSLAM.c:74-86):
        per landmark-pair edge in make_edges order, the stored residual e
        (2 values) and information Om = inv(Sigma) (2x2, flattened into the
        first 4 of 9 slots, ObsEdge.c:76-82, 126-140, 200-207) read as a
        3-vector and a 3x3 matrix; e^T M e summed from 0.0 in the reference's
        dot order and added to the running cost.  Motion edges store e = 0
        and add 0.  The third component of e lies past the array (the next
        edge record's first bytes, a subnormal): taken as 0."""
        poses = np.ascontiguousarray(g["poses"] if poses is None else poses, np.float64)
        pa, pb = self.obs_pairs(g["z_lm"], order)
        zs, zr, zb = g["z_step"], g["z_range"], g["z_bearing"]
        for a, b in zip(pa, pb):
            _, e, om = self.obs_edge(poses[3 * zs[a]:3 * zs[a] + 3], poses[3 * zs[b]:3 * zs[b] + 3],
                                     float(zr[a]), float(zb[a]), float(zr[b]), float(zb[b]))
            ev = (float(e[0]), float(e[1]), 0.0)
            M = (float(om[0]), float(om[1]), float(om[2]), float(om[3]), 0.0, 0.0, 0.0, 0.0, 0.0)
            v = []
            for j in range(3):
                t = 0.0
                for k in range(3):
                    t += ev[k] * M[3 * k + j]
                v.append(t)
            s = 0.0
            for k in range(3):
                s += v[k] * ev[k]
            acc += s * 1.0
        return acc

    def landmark_means(self, g: dict, poses=None):
        """(ids, means[L, 2]) of the landmarks, GSLAM.c:141-196."""
        poses = np.ascontiguousarray(g["poses"] if poses is None else poses, np.float64)
        zs, zl = (np.ascontiguousarray(g[k], np.int32) for k in ("z_step", "z_lm"))
        zr, zb = (np.ascontiguousarray(g[k], np.float64) for k in ("z_range", "z_bearing"))
        n = self.L.orc_landmark_means(len(zs), _ptr(poses), _ptr(zs), _ptr(zl), _ptr(zr), _ptr(zb),
                                      None, None)
        ids = np.zeros(n, np.int32)
        means = np.zeros((n, 2))
        self.L.orc_landmark_means(len(zs), _ptr(poses), _ptr(zs), _ptr(zl), _ptr(zr), _ptr(zb),
                                  _ptr(ids), _ptr(means))
        return ids, means

    def diff(self, dx):
        dx = np.ascontiguousarray(dx, np.float64)
        return self.L.orc_diff(len(dx), _ptr(dx))

    def pose_update(self, poses, dx, wrap=0):
        poses = np.array(poses, np.float64)
        dx = np.ascontiguousarray(dx, np.float64)
        self.L.orc_pose_update(len(poses) // 3, _ptr(poses), _ptr(dx), wrap)
        return poses

    def log_gn_run(self, g: dict, lam: float, rounds: int, tol=1e-6, max_iter=500):
        poses = np.array(g["poses"], np.float64)
        diffs = np.zeros(max(rounds, 1))
        n = self.L.orc_log_gn_run(len(poses) // 3, _ptr(poses), _ptr(g["u_nu"]),
                                  _ptr(g["u_omega"]), g["delta"], lam, len(g["z_step"]),
                                  _ptr(g["z_step"]), _ptr(g["z_lm"]), _ptr(g["z_range"]),
                                  _ptr(g["z_bearing"]), rounds, tol, max_iter, _ptr(diffs))
        return poses, diffs[:n]

    def between_gn_step(self, poses, ei, ej, meas, info, weight, lam, rp, ci,
                        tol=1e-10, max_iter=100000, wrap=1, anchor=1e6):
        """One GN step on a between-edge pose graph; poses updated in place."""
        dim = len(poses)
        dx = np.zeros(dim)
        diff = C.c_double(0.0)
        w = None if weight is None else np.ascontiguousarray(weight, np.float64)
        it = self.L.orc_between_gn_step(dim // 3, _ptr(poses), len(ei), _ptr(ei), _ptr(ej),
                                        _ptr(meas), _ptr(info), _ptr(w), lam, anchor, wrap,
                                        _ptr(rp), _ptr(ci), tol, max_iter, _ptr(dx),
                                        C.byref(diff))
        if it < 0:
            raise RuntimeError(f"orc_between_gn_step = {it}")
        return dx, it, diff.value


# ------------------------------------------------------------------ reference
class _HAT_X(C.Structure):
    _fields_ = [("step", C.c_uint), ("hat_x", C.c_double * 3)]


class _U(C.Structure):
    _fields_ = [("step", C.c_uint), ("nu", C.c_double), ("omega", C.c_double)]


class _Z(C.Structure):
    _fields_ = [("step", C.c_uint), ("landmark_id", C.c_uint), ("z", C.c_double * 2)]


class _Edge(C.Structure):
    """struct Edge, src/cpu_pcg_fp64/Edge.h:8-15 (552 bytes)."""
    _fields_ = [("t1", C.c_uint), ("t2", C.c_uint), ("hat_x1", _HAT_X), ("hat_x2", _HAT_X),
                ("z1", _Z), ("z2", _Z), ("Omega", C.c_double * 9),
                ("ul", C.c_double * 9), ("ur", C.c_double * 9), ("bl", C.c_double * 9),
                ("br", C.c_double * 9), ("xiu", C.c_double * 3), ("xib", C.c_double * 3),
                ("lam", C.c_double), ("hat_e", C.c_double * 2)]


class _MapEdge(C.Structure):
    """struct MapEdge, src/cpu_pcg_fp64/MapEdge.h:7-11."""
    _fields_ = [("x", _HAT_X), ("z", _Z), ("m", C.c_double * 2), ("Omega", C.c_double * 4),
                ("xi", C.c_double * 2)]


class _LandmarkKeysZ(C.Structure):
    """struct LandmarkKeysZList, src/cpu_pcg_fp64/LandmarkKeysZList.h:4-7."""
    _fields_ = [("landmark_id", C.c_uint), ("z", _Z)]


class _CSR(C.Structure):
    """CSRMatrix, src/cpu_pcg_fp64/MatrixOps.h:21-27."""
    _fields_ = [("row_ptr", C.POINTER(C.c_int)), ("col_idx", C.POINTER(C.c_int)),
                ("values", C.POINTER(C.c_double)), ("dim", C.c_int), ("nnz", C.c_int)]


def ref_available() -> bool:
    return os.path.exists(os.path.join(REF_DIR, "libref_cpu_pcg_fp64.so"))


class RefPCG64:
    """Build of src/cpu_pcg_fp64 (reference C) -- used to pin the oracle."""

    def __init__(self):
        L = C.CDLL(os.path.join(REF_DIR, "libref_cpu_pcg_fp64.so"))
        self.L = L
        L.MotionEdge_create.argtypes = [C.POINTER(_Edge), C.c_uint, C.c_uint,
                                        C.POINTER(_HAT_X), C.POINTER(_U), C.c_double,
                                        C.c_double, C.POINTER(C.c_double)]
        L.MotionEdge_create.restype = C.c_void_p
        L.ObsEdge_create.argtypes = [C.POINTER(_Edge), _Z, _Z, C.POINTER(_HAT_X),
                                     C.POINTER(C.c_double)]
        L.ObsEdge_create.restype = C.c_void_p
        L.build_csr_from_edges.argtypes = [C.POINTER(_Edge), C.c_uint, C.c_uint,
                                           C.POINTER(_CSR), C.POINTER(C.c_double)]
        L.free_csr.argtypes = [C.POINTER(_CSR)]
        L.pcg_solve_sparse.argtypes = [C.POINTER(C.c_double), C.POINTER(_CSR),
                                       C.POINTER(C.c_double), C.c_uint]
        L.inv.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_uint]
        L.make_edges.argtypes = [C.POINTER(_HAT_X), C.c_uint, C.POINTER(_Z), C.c_uint,
                                 C.POINTER(_Edge), C.POINTER(C.c_uint), C.c_void_p,
                                 C.POINTER(C.c_uint)]
        L.MapEdge_create.argtypes = [C.POINTER(_MapEdge), _Z, _Z, C.POINTER(_HAT_X),
                                     C.POINTER(C.c_double)]
        L.MapEdge_create.restype = C.c_void_p
        L.vec_mean.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_uint, C.c_uint]

    @staticmethod
    def _poses(poses):
        n = len(poses) // 3
        arr = (_HAT_X * max(n, 1))()
        for i in range(n):
            arr[i].step = i
            for k in range(3):
                arr[i].hat_x[k] = poses[3 * i + k]
        return arr

    def motion_edge(self, x1, x2, nu, omega, delta, lam, mns=(0.19, 0.001, 0.13, 0.2)):
        hx = self._poses(np.concatenate([x1, x2]))
        us = (_U * 2)()
        us[1].nu = nu
        us[1].omega = omega
        m = (C.c_double * 4)(*mns)
        e = _Edge()
        self.L.MotionEdge_create(C.byref(e), 0, 1, hx, us, delta, lam, m)
        return self._edge_vals(e), np.array(e.Omega[:])

    def obs_edge(self, x1, x2, r1, p1, r2, p2, snr=(0.14, 0.05)):
        hx = self._poses(np.concatenate([x1, x2]))
        z1 = _Z(0, 7, (C.c_double * 2)(r1, p1))
        z2 = _Z(1, 7, (C.c_double * 2)(r2, p2))
        s = (C.c_double * 2)(*snr)
        e = _Edge()
        self.L.ObsEdge_create(C.byref(e), z1, z2, hx, s)
        return self._edge_vals(e), np.array(e.hat_e[:]), np.array(e.Omega[:4])

    @staticmethod
    def _edge_vals(e):
        return np.array(list(e.ul) + list(e.ur) + list(e.bl) + list(e.br) + list(e.xiu)
                        + list(e.xib))

    def make_edges(self, g: dict, poses=None):
        """make_edges, src/cpu_pcg_fp64/GSLAM.c:222-318: ObsEdges only."""
        poses = g["poses"] if poses is None else poses
        hx = self._poses(poses)
        nz = len(g["z_step"])
        zl = (_Z * max(nz, 1))()
        for i in range(nz):
            zl[i].step = int(g["z_step"][i])
            zl[i].landmark_id = int(g["z_lm"][i])
            zl[i].z[0] = g["z_range"][i]
            zl[i].z[1] = g["z_bearing"][i]
        n_pairs = int(len(self._pairs_count(g["z_lm"])))
        cap = n_pairs + len(poses) // 3 + 1
        edges = (_Edge * cap)()
        lkz = C.create_string_buffer(32 * max(nz, 1) + 64)
        ne = C.c_uint(0)
        nl = C.c_uint(0)
        self.L.make_edges(hx, len(poses) // 3, zl, nz, edges, C.byref(ne), lkz, C.byref(nl))
        return edges, ne.value, hx

    def gj_cost(self, g: dict, poses=None, order: int = 0, acc: float = 0.0) -> float:
        """acc + the cost terms of the reference GJ drivers' use_diff=0 test
        (src/gpu_gj_fp64/GSLAM.cu:245-257, src/cpu_gj_mixed/GSLAM.c:74-86):
        per landmark-pair edge in make_edges order, the stored residual e
        (2 values) and information Om = inv(Sigma) (2x2, flattened into the
        first 4 of 9 slots, ObsEdge.c:76-82, 126-140, 200-207) read as a
        3-vector and a 3x3 matrix; e^T M e summed from 0.0 in the reference's
        dot order and added to the running cost.  Motion edges store e = 0
        and add 0.  The third component of e lies past the array (the next
        edge record's first bytes, a subnormal): taken as 0."""
        poses = np.ascontiguousarray(g["poses"] if poses is None else poses, np.float64)
        pa, pb = self.obs_pairs(g["z_lm"], order)
        zs, zr, zb = g["z_step"], g["z_range"], g["z_bearing"]
        for a, b in zip(pa, pb):
            _, e, om = self.obs_edge(poses[3 * zs[a]:3 * zs[a] + 3], poses[3 * zs[b]:3 * zs[b] + 3],
                                     float(zr[a]), float(zb[a]), float(zr[b]), float(zb[b]))
            ev = (float(e[0]), float(e[1]), 0.0)
            M = (float(om[0]), float(om[1]), float(om[2]), float(om[3]), 0.0, 0.0, 0.0, 0.0, 0.0)
            v = []
            for j in range(3):
                t = 0.0
                for k in range(3):
                    t += ev[k] * M[3 * k + j]
                v.append(t)
            s = 0.0
            for k in range(3):
                s += v[k] * ev[k]
            acc += s * 1.0
        return acc

    def landmark_means(self, g: dict, poses=None):
        """The reference's post-loop (GSLAM.c:141-196) through its own
        make_edges, MapEdge_create and vec_mean: (ids, means[L, 2])."""
        poses = g["poses"] if poses is None else poses
        hx = self._poses(poses)
        nz = len(g["z_step"])
        zl = (_Z * max(nz, 1))()
        for i in range(nz):
            zl[i].step = int(g["z_step"][i])
            zl[i].landmark_id = int(g["z_lm"][i])
            zl[i].z[0] = g["z_range"][i]
            zl[i].z[1] = g["z_bearing"][i]
        cap = int(len(self._pairs_count(g["z_lm"]))) + 1
        edges = (_Edge * cap)()
        lkz = (_LandmarkKeysZ * max(nz, 1))()
        ne, nl = C.c_uint(0), C.c_uint(0)
        self.L.make_edges(hx, len(poses) // 3, zl, nz, edges, C.byref(ne), C.cast(lkz, C.c_void_p),
                          C.byref(nl))
        snr = (C.c_double * 2)(0.14, 0.05)
        ids, means, i = [], [], 0
        while i < nl.value:
            j = i
            while j < nl.value and lkz[j].landmark_id == lkz[i].landmark_id:
                j += 1
            m = (C.c_double * (2 * (j - i)))()
            for k in range(i, j):
                me = _MapEdge()
                self.L.MapEdge_create(C.byref(me), lkz[k].z, lkz[i].z, hx, snr)
                m[2 * (k - i)], m[2 * (k - i) + 1] = me.m[0], me.m[1]
            mean = (C.c_double * 2)(0.0, 0.0)
            self.L.vec_mean(mean, m, 2, j - i)
            ids.append(lkz[i].landmark_id)
            means.append((mean[0], mean[1]))
            i = j
        return np.array(ids, np.int32), np.array(means).reshape(-1, 2)

    @staticmethod
    def _pairs_count(z_lm):
        _, c = np.unique(np.asarray(z_lm), return_counts=True)
        return np.zeros(int((c * (c - 1) // 2).sum()))

    def round_edges(self, g: dict, lam: float, poses=None, mns=(0.19, 0.001, 0.13, 0.2)):
        """The edge array of one reference round (GSLAM.c:56-72)."""
        poses = g["poses"] if poses is None else poses
        edges, ne, hx = self.make_edges(g, poses)
        n = len(poses) // 3
        us = (_U * max(n, 1))()
        for i in range(n):
            us[i].nu = g["u_nu"][i]
            us[i].omega = g["u_omega"][i]
        m = (C.c_double * 4)(*mns)
        for i in range(n - 1):
            self.L.MotionEdge_create(C.byref(edges[ne + i]), i, i + 1, hx, us, g["delta"], lam, m)
        return edges, ne + n - 1

    def edge_arrays(self, edges, n):
        t1 = np.array([edges[i].t1 for i in range(n)], np.int32)
        t2 = np.array([edges[i].t2 for i in range(n)], np.int32)
        ev = np.array([self._edge_vals(edges[i]) for i in range(n)])
        return t1, t2, ev

    def csr(self, edges, n_edges, dim):
        csr = _CSR()
        xi = (C.c_double * dim)()
        self.L.build_csr_from_edges(edges, n_edges, dim, C.byref(csr), xi)
        nnz = csr.nnz
        rp = np.ctypeslib.as_array(csr.row_ptr, (dim + 1,)).copy()
        ci = np.ctypeslib.as_array(csr.col_idx, (nnz,)).copy() if nnz else np.zeros(0, np.int32)
        va = np.ctypeslib.as_array(csr.values, (nnz,)).copy() if nnz else np.zeros(0)
        return csr, rp.astype(np.int32), ci.astype(np.int32), va, np.array(xi[:])

    def pcg(self, csr, xi):
        dim = csr.dim
        x = (C.c_double * dim)()
        b = (C.c_double * dim)(*xi)
        self.L.pcg_solve_sparse(x, C.byref(csr), b, dim)
        return np.array(x[:])

    def free_csr(self, csr):
        self.L.free_csr(C.byref(csr))

    def inv(self, a):
        """inv(), src/cpu_pcg_fp64/MatrixOps.c:29-91 (stack VLA: keep n small)."""
        a = np.ascontiguousarray(a, np.float64)
        n = a.shape[0]
        out = np.zeros_like(a)
        self.L.inv(out.ctypes.data_as(C.POINTER(C.c_double)),
                   a.ctypes.data_as(C.POINTER(C.c_double)), n)
        return out

    @staticmethod
    def exe():
        return os.path.join(REF_DIR, "GSLAM_cpu_pcg_fp64")


class RefGJMixed:
    """Build of src/cpu_gj_mixed (C++): inv_float, MatrixOps.c:174-212, and
    the executable (its GN driver, GSLAM.c:28-250)."""

    @staticmethod
    def exe():
        return os.path.join(REF_DIR, "GSLAM_cpu_gj_mixed")

    def __init__(self):
        L = C.CDLL(os.path.join(REF_DIR, "libref_cpu_gj_mixed.so"))
        self.f = getattr(L, "_Z9inv_floatPfS_j")
        self.f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint]

    def inv_float(self, a):
        a = np.ascontiguousarray(a, np.float32)
        out = np.zeros_like(a)
        self.f(_ptr(out), _ptr(a), a.shape[0])
        return out
