"""Python binding of libgslam (ctypes over the C-ABI in include/gslam.h).

The binding is a thin mirror of the C entry points: a `Graph` owns one
`gslam_graph` handle; every call goes to the HIP library.  There is no CPU
fallback: if libgslam.so is missing or fails to load, `GSlamError` is raised.

    g = Graph.from_log("cityTrees800.txt", lam=0.1)     # reference .log.txt model
    for _ in range(10):
        stats, diff = g.gn_step(solver="pcg64", tol=1e-6, max_iter=500)
    poses = g.get_poses()
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .landmarks import landmark_means  # noqa: F401  (host post-processing, GSLAM.c:141-196)

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("GSLAM_LIB", os.path.join(PKG_DIR, "libgslam.so"))

OK = 0
ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_SINGULAR, ERR_STATE, ERR_TOO_LARGE, ERR_NOT_CONVERGED = (
    -1, -2, -3, -4, -5, -6, -7)
ERR_NONFINITE = -8
SOLVERS = {"pcg64": 0, "mixed": 1, "gj64": 2, "gj_mixed": 3, "gj64_blocked": 4}
KERNELS = {"spmv64": 0, "update64": 1, "spmv_mixed": 2, "assembly": 3, "gj_trailing": 4}
REFINE = {"ir": 0, "cg": 1}


class GSlamError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


class _Desc(C.Structure):
    _fields_ = [
        ("n_poses", C.c_int32), ("poses", C.c_void_p),
        ("motion", C.c_int32), ("u_nu", C.c_void_p), ("u_omega", C.c_void_p),
        ("delta", C.c_double), ("lambda_", C.c_double), ("mns", C.c_double * 4),
        ("n_obs", C.c_int32), ("z_step", C.c_void_p), ("z_lm", C.c_void_p),
        ("z_range", C.c_void_p), ("z_bearing", C.c_void_p), ("pair_order", C.c_int32),
        ("snr", C.c_double * 2),
        ("n_between", C.c_int64), ("be_i", C.c_void_p), ("be_j", C.c_void_p),
        ("be_meas", C.c_void_p), ("be_info", C.c_void_p),
        ("anchor", C.c_double), ("wrap", C.c_int32), ("anchor_prior", C.c_int32),
        ("assembly", C.c_int32)]


ASSEMBLY = {"auto": 0, "pose_walk": 1, "staged": 2}


class _Info(C.Structure):
    _fields_ = [("n_poses", C.c_int32), ("n_edges", C.c_int64), ("n_obs_edges", C.c_int64),
                ("n_motion_edges", C.c_int64), ("n_between_edges", C.c_int64),
                ("n_blocks", C.c_int64), ("nnz", C.c_int64), ("device_bytes", C.c_int64),
                ("n_tiles", C.c_int32)]


class _Opts(C.Structure):
    _fields_ = [("solver", C.c_int32), ("tol", C.c_double), ("max_iter", C.c_int32),
                ("refine_sweeps", C.c_int32), ("refresh_every", C.c_int32),
                ("check_every", C.c_int32), ("refine_mode", C.c_int32),
                ("profile_every", C.c_int32), ("mixed_handover", C.c_double),
                ("pcg_schedule", C.c_int32)]


class _Stats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("converged", C.c_int32), ("reliable_updates", C.c_int32),
                ("flags", C.c_int32), ("rel_res", C.c_double), ("true_rel_res", C.c_double),
                ("ms", C.c_double), ("ms_factor", C.c_double), ("spmv_us", C.c_double),
                ("update_us", C.c_double), ("iters_fp64", C.c_int32),
                ("ms_assembly", C.c_double), ("ms_setup", C.c_double), ("ms_update", C.c_double)]


_LIB = None

# every symbol of include/gslam.h and include/gslam_compat.h
EXPORTS = [
    "gslam_default_opts", "gslam_default_desc", "gslam_create", "gslam_destroy",
    "gslam_get_info", "gslam_set_poses", "gslam_get_poses", "gslam_set_lambda",
    "gslam_set_log_weights", "gslam_assemble", "gslam_get_system_csr", "gslam_get_system_bsr",
    "gslam_solve", "gslam_get_dx", "gslam_solve_rhs", "gslam_apply_update", "gslam_gn_step",
    "gslam_gj_invert", "gslam_last_error", "gslam_set_system_csr", "gslam_time_kernel",
    "gslam_adjoint", "gslam_obs_cost", "gslam_gj_invert_device",
    "device_graph_init", "device_graph_update", "device_graph_get_row_ptr",
    "device_graph_get_col_idx", "device_build_diag_inv", "device_pcg_solve",
    "device_graph_free", "inv_cuda",
]


def hip_runtimes() -> list:
    """Distinct files of the HIP runtime (libamdhip64) mapped in this process."""
    try:
        with open("/proc/self/maps") as f:
            paths = {ln.split(None, 5)[5].strip() for ln in f if "libamdhip64" in ln
                     and len(ln.split(None, 5)) == 6}
    except OSError:
        return []
    return sorted(os.path.realpath(p) for p in paths)


def check_single_runtime():
    """Raise GSlamError if two HIP runtimes are loaded in this process.

    libgslam.so needs libamdhip64.so.7; PyTorch's ROCm wheels ship their own
    copy, which their libraries request as libamdhip64.so.  Loaded first,
    torch's copy satisfies libgslam too (the dynamic linker matches the
    SONAME libamdhip64.so.7 it carries); loaded after libgslam, it is a second
    runtime with its own HSA instance, and GPU calls made through one of the
    two fail (seen as "No HIP GPUs are available" from torch after libgslam
    had allocated device memory)."""
    rt = hip_runtimes()
    if len(rt) > 1:
        raise GSlamError(ERR_HIP, "load",
                         "two HIP runtimes are loaded in this process (" + ", ".join(rt) +
                         "): import torch before loading libgslam.so, or use only one of them")


def lib():
    """Load libgslam.so (once).  Raises GSlamError if it is not built.

    If PyTorch is installed it is imported first, so that libgslam binds to
    the HIP runtime torch loads and the process holds one runtime (see
    check_single_runtime)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise GSlamError(ERR_STATE, "load", f"{LIB_PATH} not built (run __graft_entry__.build())")
    if os.environ.get("GSLAM_NO_TORCH", "0") != "1":
        try:
            import torch  # noqa: F401  (one HIP runtime: torch's, shared with libgslam)
        except ImportError:
            pass
    L = C.CDLL(LIB_PATH)
    check_single_runtime()
    vp = C.c_void_p
    L.gslam_default_opts.argtypes = [vp]
    L.gslam_default_desc.argtypes = [vp]
    L.gslam_create.argtypes = [vp, C.POINTER(vp)]
    L.gslam_destroy.argtypes = [vp]
    L.gslam_get_info.argtypes = [vp, vp]
    L.gslam_set_poses.argtypes = [vp, vp]
    L.gslam_get_poses.argtypes = [vp, vp]
    L.gslam_set_lambda.argtypes = [vp, C.c_double]
    L.gslam_set_log_weights.argtypes = [vp, vp]
    L.gslam_assemble.argtypes = [vp]
    L.gslam_get_system_csr.argtypes = [vp, vp, vp, vp, vp]
    L.gslam_get_system_bsr.argtypes = [vp, vp, vp, vp, vp]
    L.gslam_solve.argtypes = [vp, vp, vp]
    L.gslam_get_dx.argtypes = [vp, vp]
    L.gslam_solve_rhs.argtypes = [vp, vp, vp, vp, vp]
    L.gslam_apply_update.argtypes = [vp, vp]
    L.gslam_gn_step.argtypes = [vp, vp, vp, vp]
    L.gslam_gj_invert.argtypes = [vp, vp, C.c_int32, C.c_int32]
    L.gslam_last_error.restype = C.c_char_p
    L.gslam_set_system_csr.argtypes = [vp, vp, vp]
    L.gslam_time_kernel.argtypes = [vp, C.c_int32, C.c_int32, vp]
    L.gslam_obs_cost.argtypes = [vp, vp]
    _LIB = L
    return L


def _check(rc: int, where: str, allow=()):
    if rc != OK and rc not in allow:
        raise GSlamError(rc, where, lib().gslam_last_error().decode())
    return rc


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def read_log(path: str) -> dict:
    """read_data (src/cpu_pcg_fp64/GSLAM.c:415-493): records u/x/z, any other
    first token sets delta from its second token; tokens split on ' '."""
    d = {"delta": 0.0, "u": [], "x": [], "z": []}
    with open(path, "r") as f:
        for ln, line in enumerate(f, 1):
            tok = [t for t in line.split(" ") if t != ""]
            if not tok:
                raise ValueError(f"{path}:{ln}: empty record")
            need = {"u": 3, "x": 4, "z": 4}.get(tok[0], 1)
            if len(tok) < need + 1:
                raise ValueError(f"{path}:{ln}: truncated record")
            if tok[0] == "u":
                d["u"].append((int(float(tok[1])), float(tok[2]), float(tok[3])))
            elif tok[0] == "x":
                d["x"].append((int(float(tok[1])), float(tok[2]), float(tok[3]), float(tok[4])))
            elif tok[0] == "z":
                d["z"].append((int(float(tok[1])), int(float(tok[2])), float(tok[3]), float(tok[4])))
            else:
                d["delta"] = float(tok[1])
    u = np.array(d["u"], dtype=np.float64).reshape(-1, 3)
    x = np.array(d["x"], dtype=np.float64).reshape(-1, 4)
    z = np.array(d["z"], dtype=np.float64).reshape(-1, 4)
    return dict(delta=d["delta"], u_step=u[:, 0].astype(np.int32), u_nu=u[:, 1].copy(),
                u_omega=u[:, 2].copy(), x_step=x[:, 0].astype(np.int32),
                poses=x[:, 1:].ravel().copy(), z_step=z[:, 0].astype(np.int32),
                z_lm=z[:, 1].astype(np.int32), z_range=z[:, 2].copy(), z_bearing=z[:, 3].copy(),
                n_poses=len(x))


def resume_log(log_path: str, dump_path: str, out_path: str) -> int:
    """Checkpoint resume: writes `log_path` to `out_path` with the pose values
    of its x records replaced by those of a pose dump (POSE_DUMP_FILE lines
    `step x y theta`, matched to the x records by step), every other record
    unchanged.  The GN loop started on `out_path` continues from the dumped
    poses.  Returns the number of x records rewritten."""
    dumped = {}
    with open(dump_path) as f:
        for ln, line in enumerate(f, 1):
            tok = line.split()
            if len(tok) != 4:
                raise ValueError(f"{dump_path}:{ln}: expected 'step x y theta'")
            dumped[int(tok[0])] = tok[1:]
    n = 0
    with open(log_path) as f, open(out_path, "w") as g:
        for line in f:
            tok = [t for t in line.split(" ") if t != ""]
            if tok and tok[0] == "x":
                step = int(float(tok[1]))
                if step not in dumped:
                    raise ValueError(f"{dump_path}: no pose for x record {step}")
                line = "x %s %s\n" % (tok[1].strip(), " ".join(dumped[step]))
                n += 1
            g.write(line)
    if n != len(dumped):
        raise ValueError(f"{dump_path}: {len(dumped)} poses for {n} x records")
    return n


PCG_SCHEDULES = {"auto": 0, "two_launch": 1}


def make_opts(solver="pcg64", tol=1e-6, max_iter=500, refine_sweeps=3, refresh_every=64,
              check_every=0, refine_mode="ir", profile_every=0, mixed_handover=0.0,
              pcg_schedule="auto") -> _Opts:
    o = _Opts()
    lib().gslam_default_opts(C.byref(o))
    o.solver = SOLVERS[solver] if isinstance(solver, str) else int(solver)
    o.tol = tol
    o.max_iter = max_iter
    o.refine_sweeps = refine_sweeps
    o.refresh_every = refresh_every
    o.check_every = check_every
    o.refine_mode = REFINE[refine_mode] if isinstance(refine_mode, str) else int(refine_mode)
    o.profile_every = profile_every
    o.mixed_handover = mixed_handover
    o.pcg_schedule = (PCG_SCHEDULES[pcg_schedule] if isinstance(pcg_schedule, str)
                      else int(pcg_schedule))
    return o


def _stats(s: _Stats) -> dict:
    return dict(iters=s.iters, converged=bool(s.converged), reliable_updates=s.reliable_updates,
                flags=s.flags, rel_res=s.rel_res, true_rel_res=s.true_rel_res, ms=s.ms,
                ms_factor=s.ms_factor, spmv_us=s.spmv_us, update_us=s.update_us,
                iters_fp64=s.iters_fp64, ms_assembly=s.ms_assembly, ms_setup=s.ms_setup,
                ms_update=s.ms_update)


class Graph:
    """One libgslam graph handle (device-resident poses, Omega, xi, solver)."""

    def __init__(self, poses, *, motion=False, u_nu=None, u_omega=None, delta=1.0, lam=0.1,
                 mns=(0.19, 0.001, 0.13, 0.2), obs=None, pair_order=0, snr=(0.14, 0.05),
                 between=None, anchor=1e6, wrap=0, anchor_prior=0, assembly="auto"):
        L = lib()
        poses = _f64(poses)
        n = poses.size // 3
        self._keep = []
        d = _Desc()
        L.gslam_default_desc(C.byref(d))
        d.n_poses = n
        d.poses = _p(poses)
        self._keep.append(poses)
        if motion:
            u_nu, u_omega = _f64(u_nu, n), _f64(u_omega, n)
            self._keep += [u_nu, u_omega]
            d.motion = 1
            d.u_nu, d.u_omega = _p(u_nu), _p(u_omega)
        d.delta = delta
        d.lambda_ = lam
        for k in range(4):
            d.mns[k] = mns[k]
        if obs is not None:
            zs, zl = _i32(obs["z_step"]), _i32(obs["z_lm"])
            zr, zb = _f64(obs["z_range"]), _f64(obs["z_bearing"])
            self._keep += [zs, zl, zr, zb]
            d.n_obs = zs.size
            d.z_step, d.z_lm, d.z_range, d.z_bearing = _p(zs), _p(zl), _p(zr), _p(zb)
        d.pair_order = pair_order
        d.snr[0], d.snr[1] = snr
        if between is not None:
            bi, bj = _i32(between["i"]), _i32(between["j"])
            bm, bf = _f64(between["meas"], 3 * bi.size), _f64(between["info"], 3 * bi.size)
            self._keep += [bi, bj, bm, bf]
            d.n_between = bi.size
            d.be_i, d.be_j, d.be_meas, d.be_info = _p(bi), _p(bj), _p(bm), _p(bf)
        d.anchor = anchor
        d.wrap = wrap
        d.anchor_prior = anchor_prior
        d.assembly = ASSEMBLY[assembly] if isinstance(assembly, str) else int(assembly)
        h = C.c_void_p()
        _check(L.gslam_create(C.byref(d), C.byref(h)), "gslam_create")
        self._keep = []
        self.h = h
        self.n = n
        info = _Info()
        _check(L.gslam_get_info(self.h, C.byref(info)), "gslam_get_info")
        self.info = {k: getattr(info, k) for k, _ in _Info._fields_}

    def device_bytes(self) -> int:
        """Device memory the handle holds now: `info["device_bytes"]` is the
        figure at creation; solver work space is added on its first use."""
        info = _Info()
        _check(lib().gslam_get_info(self.h, C.byref(info)), "gslam_get_info")
        return int(info.device_bytes)

    @classmethod
    def from_log(cls, path: str | None = None, lam: float = 0.1, pair_order: int = 0,
                 wrap: int = 0, data: dict | None = None, anchor_prior: int = 0,
                 assembly="auto"):
        g = read_log(path) if data is None else data
        return cls(g["poses"], motion=True, u_nu=g["u_nu"], u_omega=g["u_omega"],
                   delta=g["delta"], lam=lam,
                   obs=dict(z_step=g["z_step"], z_lm=g["z_lm"], z_range=g["z_range"],
                            z_bearing=g["z_bearing"]),
                   pair_order=pair_order, wrap=wrap, anchor_prior=anchor_prior,
                   assembly=assembly)

    # -------------------------------------------------------------- lifecycle
    def close(self):
        if getattr(self, "h", None):
            lib().gslam_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -------------------------------------------------------------- state
    def set_poses(self, poses):
        p = _f64(poses, 3 * self.n)
        _check(lib().gslam_set_poses(self.h, _p(p)), "gslam_set_poses")

    def get_poses(self) -> np.ndarray:
        out = np.zeros(3 * self.n)
        _check(lib().gslam_get_poses(self.h, _p(out)), "gslam_get_poses")
        return out

    def set_lambda(self, lam: float):
        _check(lib().gslam_set_lambda(self.h, lam), "gslam_set_lambda")

    def set_log_weights(self, w):
        w = None if w is None else _f64(w, self.info["n_between_edges"])
        _check(lib().gslam_set_log_weights(self.h, _p(w)), "gslam_set_log_weights")

    # -------------------------------------------------------------- GN pieces
    def assemble(self):
        _check(lib().gslam_assemble(self.h), "gslam_assemble")

    def system_csr(self):
        nnz = self.info["nnz"]
        rp = np.zeros(3 * self.n + 1, np.int32)
        ci = np.zeros(nnz, np.int32)
        va = np.zeros(nnz)
        xi = np.zeros(3 * self.n)
        _check(lib().gslam_get_system_csr(self.h, _p(rp), _p(ci), _p(va), _p(xi)),
               "gslam_get_system_csr")
        return rp, ci, va, xi

    def system_bsr(self):
        nb = self.info["n_blocks"]
        rp = np.zeros(self.n + 1, np.int32)
        bc = np.zeros(nb, np.int32)
        bv = np.zeros((nb, 3, 3))
        xi = np.zeros(3 * self.n)
        _check(lib().gslam_get_system_bsr(self.h, _p(rp), _p(bc), _p(bv), _p(xi)),
               "gslam_get_system_bsr")
        return rp, bc, bv, xi

    def set_system_csr(self, values, xi):
        """Load Omega (reference CSR layout on this graph's pattern) and xi."""
        va = _f64(values, self.info["nnz"])
        x = _f64(xi, 3 * self.n)
        _check(lib().gslam_set_system_csr(self.h, _p(va), _p(x)), "gslam_set_system_csr")

    def time_kernel(self, kernel: str, reps: int = 100) -> float:
        """Mean per-launch microseconds of one solver kernel (HIP events)."""
        us = C.c_double()
        _check(lib().gslam_time_kernel(self.h, KERNELS[kernel], reps, C.byref(us)),
               "gslam_time_kernel")
        return us.value

    def solve(self, solver="pcg64", tol=1e-6, max_iter=500, **kw) -> dict:
        o = make_opts(solver, tol, max_iter, **kw)
        s = _Stats()
        _check(lib().gslam_solve(self.h, C.byref(o), C.byref(s)), "gslam_solve",
               allow=(ERR_NOT_CONVERGED, ERR_SINGULAR))
        return _stats(s)

    def adjoint(self, dl_dx, tol=1e-12, max_iter=200000):
        """dL/dw_e (between edges) and dL/dlambda at the current (converged)
        poses for the caller's dL/dx*.  Between-edge graphs: one FP64 PCG on
        the exact Hessian and an edge-parallel gradient kernel.  Graphs of the
        reference's landmark-pair and motion edges: dL/dlambda by a tangent
        solve (GMRES preconditioned by the FP64 PCG; tol is then its relative
        residual), no per-edge weights (dl_dw is empty)."""
        o = make_opts("pcg64", tol, max_iter)
        g = _f64(dl_dx, 3 * self.n)
        dw = np.zeros(self.info["n_between_edges"])
        dlam = C.c_double()
        s = _Stats()
        _check(lib().gslam_adjoint(self.h, _p(g), C.byref(o), _p(dw), C.byref(dlam), C.byref(s)),
               "gslam_adjoint")
        return dw, dlam.value, _stats(s)

    def converge(self, solver="pcg64", tol=1e-12, max_iter=200000, diff_tol=1e-12,
                 max_rounds=50, patience=3, **kw) -> int:
        """Gauss-Newton steps until ||dx||/sqrt(3N) < diff_tol, or until the
        diff has not improved for `patience` rounds (the rounding floor of the
        step, as the reference's 20-round no-improvement stop,
        src/cpu_pcg_fp64/GSLAM.c:119-130); returns the rounds run."""
        best, stale = float("inf"), 0
        for r in range(1, max_rounds + 1):
            _, diff = self.gn_step(solver, tol, max_iter, **kw)
            if diff < diff_tol:
                return r
            if diff < best:
                best, stale = diff, 0
            else:
                stale += 1
                if stale >= patience:
                    return r
        return max_rounds

    def dx(self) -> np.ndarray:
        out = np.zeros(3 * self.n)
        _check(lib().gslam_get_dx(self.h, _p(out)), "gslam_get_dx")
        return out

    def solve_rhs(self, rhs, solver="pcg64", tol=1e-10, max_iter=100000, **kw):
        o = make_opts(solver, tol, max_iter, **kw)
        b = _f64(rhs, 3 * self.n)
        x = np.zeros(3 * self.n)
        s = _Stats()
        _check(lib().gslam_solve_rhs(self.h, C.byref(o), _p(b), _p(x), C.byref(s)),
               "gslam_solve_rhs", allow=(ERR_NOT_CONVERGED, ERR_SINGULAR))
        return x, _stats(s)

    def obs_cost(self, acc: float = 0.0) -> float:
        """acc + the reference GJ drivers' use_diff=0 cost terms at the
        current poses (gslam_obs_cost)."""
        c = C.c_double(acc)
        _check(lib().gslam_obs_cost(self.h, C.byref(c)), "gslam_obs_cost")
        return c.value

    def apply_update(self) -> float:
        diff = C.c_double()
        _check(lib().gslam_apply_update(self.h, C.byref(diff)), "gslam_apply_update")
        return diff.value

    def gn_step(self, solver="pcg64", tol=1e-6, max_iter=500, **kw):
        o = make_opts(solver, tol, max_iter, **kw)
        s = _Stats()
        diff = C.c_double()
        _check(lib().gslam_gn_step(self.h, C.byref(o), C.byref(s), C.byref(diff)),
               "gslam_gn_step", allow=(ERR_NOT_CONVERGED, ERR_SINGULAR))
        return _stats(s), diff.value


def gj_invert(a, solver="gj64") -> np.ndarray:
    """Dense unpivoted Gauss-Jordan inverse on the GPU (host in, host out).

    "gj64": FP64, one pass per pivot; "gj64_blocked": FP64, 128 pivots per
    block step with FP64 MFMA trailing updates; "gj_mixed": blocked, FP32
    MFMA trailing updates on the Jacobi-scaled matrix (FP32 accuracy)."""
    a = np.ascontiguousarray(a, np.float64)
    n = a.shape[0]
    out = np.zeros_like(a)
    _check(lib().gslam_gj_invert(_p(a), _p(out), n, SOLVERS[solver]), "gslam_gj_invert",
           allow=(ERR_SINGULAR,))
    return out
