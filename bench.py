#!/usr/bin/env python3
"""bench.py -- ms per Gauss-Newton step, mixed vs FP64 PCG, grid-city N=100k.

One step = one Gauss-Newton step of the hot path on the GPU: edge evaluation
and Omega/xi assembly, block-Jacobi PCG to the tolerance, pose update.  Every
timed step starts from the same dead-reckoned initial guess (uploaded before
each step's timed region, so the step starts with its inputs in HBM), so every
step solves the same system and the accuracy check applies to each of them.

Headline workload: synthetic grid-city pose graph (gslam.generators.grid_city),
N = 100,000 poses, seed 0, lambda 0.1, angle wrap on.  Accuracy: relative dx
error against an independent direct solve of the same system (SuperLU,
refined on residuals in x87 extended precision: ~1e-11 relative; the FP64
LU solve alone is 1.3e-7 away), with the FP64 PCG solve at tolerance 1e-14
reported beside it.  The FP64 solver runs at tolerance 3e-13 (3.3e-9 rel dx
err on this graph; 1e-12 gives ~1.3e-8); the mixed solver runs FP32
iterations until the FP64 true residual is below 1e-2 (as on the park
graph) and then FP64 PCG on the correction to the same 3e-13 (7.8e-9;
hand-overs at 3e-3 and 1e-3 also meet 1e-8, 3e-2 and 1e-4 end at 1.2-1.5e-8
-- at kappa ~ 1e13 the attainable error is ~1e-8 within a factor of three,
set by the rounding path -- and FP32 iterations alone stall at ~1e-6
error, kappa far beyond 1/eps32).

Also reported:
* roofline: the dominant FP64 PCG kernel's mean per-launch duration inside
  the timed solves (HIP events around sampled launches on the solver
  stream) -- at N=100k the one-launch iteration pcg_fused_kernel, else the
  SpMV of the two-launch schedule --, its algorithmic bytes and the HBM
  fraction; HBM traffic per launch from the committed rocprofv3 counter
  summary when present;
* cpu_baseline: the CPU oracle (FP64, one thread) on a bounded sample of the
  same GN step, extrapolated to the FP64 solve's iteration count; the same
  with the PCG loops on all OpenMP threads (omp); and, as
  whole processes on the reference's full victoria-park and cityTrees10k
  logs, the reference's own CPU executable beside the drop-in GSLAM;
* gj: blocked Gauss-Jordan with FP32 MFMA trailing updates (GJ_MIXED) and in
  FP64 (GJ_BLOCKED64) against the plain FP64 Gauss-Jordan and
  numpy.linalg.solve (FP64), grid-city N=800, 1600 and 5000 (3N = 2400,
  4800, 15000; the plain FP64 Gauss-Jordan up to N=2000);
* park: the park graph (laps route) of 7,120 poses, lambda 1.0 -- FP64 and
  mixed PCG at the same tolerance against the same kind of direct solve;
* grid10k: the FP64 PCG GN run on grid-city N=10,000 (tol 1e-10, the
  reference stop rule within 20 rounds), with the device time per phase;
* large: FP64 and mixed PCG on grid-city N=1,200,000 (HBM-bound) to a true
  relative residual of 1e-5, with the FP64 SpMV's in-solve HBM fraction
  (`large.roofline_hbm`);
* adjoint: converged-pose adjoint at grid-city N=10,000 (forward, backward,
  central finite differences on sampled edges, optional Adam steps).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
import argparse
import gzip
import json
import os
import re
import shutil
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "g-slam_on_gpu_and_fpga_amd"))

METRIC = "ms per Gauss-Newton step, mixed vs FP64 PCG at N=100k, rel Δx err ≤1e-8"
HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK = 157.3   # TFLOP/s, MI355X dense FP32 matrix
FP64_PEAK = 78.6         # TFLOP/s, MI355X FP64 (vector and matrix)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pcg_spmv_grid100k_traffic.md")
FUSED_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pcg_fused_grid100k_traffic.md")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n_poses, n_blocks):
    """Algorithmic bytes of one FP64 PCG SpMV launch: the BSR values and
    column indices (9*8 + 4 B per block), block-row pointers (4 B per row),
    z, p_old and x read once (3 x 24 B per row) and p_new, q and x (the
    previous iteration's x += alpha p_old) written once (3 x 24 B per row)."""
    return n_blocks * (9 * 8 + 4) + 4 * (n_poses + 1) + 6 * 24 * n_poses


def fused_bytes(n_poses, n_blocks):
    """Algorithmic bytes of one launch of the one-launch FP64 PCG iteration
    (pcg_fused_kernel): the BSR values and column indices (9*8 + 4 B per
    block) and block-row pointers (4 B per row); per pose its record
    {z, w = D^-1 q, p} (72 B, the gathers of neighbouring rows hitting the
    caches), r, q and x (72 B) and its 3x3 block of D^-1 (72 B) read once,
    and r, q, x (72 B) and the new record (72 B) written once."""
    return n_blocks * (9 * 8 + 4) + 4 * (n_poses + 1) + 5 * 72 * n_poses


def update_bytes(n_poses):
    """Algorithmic bytes of one FP64 PCG update launch: r, q and the 3x3 block
    of D^-1 read (2 x 24 + 72 B per pose), r and z written (2 x 24 B per
    pose)."""
    return n_poses * (2 * 24 + 72 + 2 * 24)


def mixed_spmv_bytes(n_poses, n_inc):
    """Algorithmic bytes of one mixed (FP32 factorized) SpMV launch: per
    incidence entry (two per edge) its edge entry, the other pose's index and
    8 FP32 factors (40 B); per pose z and p_old read and p_new and q written
    (FP32: 4 x 12 B).  The kernel's tile-padded layout also stores each
    entry's own pose (4 B) and pads the tiles (about 1%); those bytes and the
    row pointers are not counted."""
    return n_inc * 40 + n_poses * 48


def assembly_bytes(n_poses, n_edges, n_blocks):
    """Algorithmic bytes of one assembly of a between-edge graph, counted as
    the minimal traffic: each edge's inputs once (its two poses, 48 B, and
    its parameters: i, j, measurement, information and weight, 64 B), each
    Omega block written once (72 B) and xi written once (24 B per pose).  The
    pose walk reads each edge at both of its ends, and the staged schedule
    writes and re-reads a 384-byte record per edge; neither is counted."""
    return n_edges * (48 + 64) + n_blocks * 72 + n_poses * 24


def measured_traffic(fused=False):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration)."""
    try:
        with open(FUSED_TRAFFIC_FILE if fused else TRAFFIC_FILE) as f:
            m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def refined_direct(rp, ci, va, xi, steps=4):
    """Direct solve of the CSR system: SuperLU (FP64) refined with residuals
    formed in x87 extended precision (numpy.longdouble) -- accurate to ~1e-11
    relative on the N=100k grid-city system."""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    n = len(rp) - 1
    lu = spla.splu(sp.csr_matrix((va, ci, rp), shape=(n, n)).tocsc())
    Al = sp.csr_matrix((va.astype(np.longdouble), ci, rp), shape=(n, n))
    x = lu.solve(xi).astype(np.longdouble)
    bl = xi.astype(np.longdouble)
    for _ in range(steps):
        x += lu.solve((bl - Al @ x).astype(np.float64))
    return x.astype(np.float64)


def accuracy(np, dx, dx_ref):
    err = float(np.linalg.norm(dx - dx_ref) / np.linalg.norm(dx_ref))
    d3 = (dx - dx_ref).reshape(-1, 3)
    return dict(rel_dx_err=err, meets_1e8=err <= 1e-8, max_abs_dx_err=float(np.abs(d3).max()),
                max_pose_err_m=float(np.abs(d3[:, :2]).max()),
                max_pose_err_rad=float(np.abs(d3[:, 2]).max()))


def hip_sync():
    import ctypes
    ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()


def pcg_section(args, gslam, G, np):
    t0 = time.time()
    gc = G.grid_city(args.n, seed=args.seed)
    log(f"graph N={args.n}: {len(gc['i'])} edges ({time.time() - t0:.1f}s)")
    init = gc["poses"].copy()
    g = gslam.Graph(init, **G.graph_kwargs(gc, 0.1, wrap=1))
    info = dict(g.info)
    log(f"handle: {info}")
    dx_ref = dx_pcg = None
    if not args.no_accuracy:
        g.set_poses(init)
        g.assemble()
        t = time.time()
        dx_ref = refined_direct(*g.system_csr())
        log(f"reference: refined direct solve on the host ({time.time() - t:.1f} s)")
        st = g.solve("pcg64", args.ref_tol, args.max_iter)
        dx_pcg = g.dx()
        log(f"FP64 PCG at tol {args.ref_tol:g}: {st['iters']} its, true rel res "
            f"{st['true_rel_res']:.2e}, rel dx err vs the direct solve "
            f"{np.linalg.norm(dx_pcg - dx_ref) / np.linalg.norm(dx_ref):.2e}")
    tols = {"pcg64": args.tol, "mixed": args.mixed_tol}
    kws = {"pcg64": {}, "mixed": dict(mixed_handover=args.handover)}
    results, samples = {}, []
    for solver in [s for s in args.solvers.split(",") if s]:
        tol = tols[solver]
        acc = {}
        if dx_ref is not None:
            g.set_poses(init)
            g.assemble()
            st = g.solve(solver, tol, args.max_iter, **kws[solver])
            dx = g.dx()
            acc = accuracy(np, dx, dx_ref)
            acc.update(rel_dx_err_vs_pcg_1e14=float(np.linalg.norm(dx - dx_pcg) / np.linalg.norm(dx_pcg)),
                       true_rel_res=st["true_rel_res"], converged=st["converged"],
                       refreshes=st["reliable_updates"], fp64_correction_iters=st["iters_fp64"])
            if solver == "mixed":
                acc["handover"] = args.handover
            log(f"{solver}: accuracy {acc}")
        for _ in range(args.warmup):
            g.set_poses(init)
            g.gn_step(solver, tol, args.max_iter, **kws[solver])
        hip_sync()
        iters = []
        t_total = 0.0
        for _ in range(args.steps):
            # the initial poses are uploaded before the timed region: a step
            # starts with its inputs resident in HBM (gslam_gn_step returns
            # once the updated poses and the diff are resident)
            g.set_poses(init)
            hip_sync()
            t0 = time.perf_counter()
            st, _ = g.gn_step(solver, tol, args.max_iter,
                              profile_every=args.profile_every if solver == "pcg64" else 0,
                              **kws[solver])
            hip_sync()
            t_total += time.perf_counter() - t0
            iters.append(st["iters"])
            if solver == "pcg64" and st["spmv_us"] > 0:
                # (HIP events of the solve: its launch loop per launch -- the
                # iterations plus the first launch --, and the sampled launches)
                samples.append((st["spmv_us"], st["update_us"],
                                1000.0 * (st["ms"] - st["ms_setup"]) / (st["iters"] + 1)))
        ms = 1000.0 * t_total / args.steps
        results[solver] = dict(ms_per_step=ms, tol=tol, pcg_iters=int(np.mean(iters)),
                               us_per_iter=1000.0 * ms / max(np.mean(iters), 1),
                               handle_device_mb=g.device_bytes() / 1e6, **acc)
        log(f"{solver}: {ms:.1f} ms per GN step, {np.mean(iters):.0f} PCG iterations")

    # dominant kernel, mean duration inside the timed solves: the one-launch
    # iteration (graphs of at most 2048 SpMV tiles; the sampled update time is
    # then zero) or the FP64 SpMV of the two-launch schedule
    fused = bool(samples) and all(s[1] == 0.0 for s in samples)
    if fused:
        byts = fused_bytes(info["n_poses"], info["n_blocks"])
        kname = ("pcg_fused_kernel<4> (one FP64 PCG iteration: SpMV, r/x/z/w update and "
                 "partials; four SpMV tiles per 1024-thread workgroup)")
    else:
        byts = spmv_bytes(info["n_poses"], info["n_blocks"])
        kname = "pcg_spmv_kernel (FP64 BSR SpMV + p.q partials)"
    roofline = None
    if samples:
        # one launch per iteration: the solve's launch loop is this kernel
        # (99.8% of its GPU time by rocprofv3), so its per-launch time is the
        # loop's HIP-event time over the launches; a launch bracketed by its own
        # events (the two-launch kernels' measure) reads ~1.3 us more
        us = float(np.mean([s[2] if fused else s[0] for s in samples]))
        achieved = byts / (us * 1e-6) / 1e9
        roofline = dict(bound="hbm", kernel=kname,
                        regime=("latency-bound: the 31 MB matrix and the vectors stay in the "
                                "256 MB Infinity Cache at N=100k; the fraction is algorithmic "
                                "bytes over time against the 8 TB/s HBM peak (counted HBM "
                                "traffic in 'traffic')"),
                        us_per_launch=us, bytes_per_launch=byts, achieved=achieved,
                        peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                        traffic=measured_traffic(fused),
                        how=(f"HIP events around the launch loop of the {args.steps} timed FP64 "
                             f"solves, per launch (events around every {args.profile_every}th "
                             f"launch: us_per_launch_sampled)" if fused else
                             f"HIP events around every {args.profile_every}th launch of "
                             f"the {args.steps} timed FP64 solves"))
        if fused:
            roofline["us_per_launch_sampled"] = float(np.mean([s[0] for s in samples]))
        log(f"{'fused iteration' if fused else 'spmv64'} in solve: {us:.2f} us/launch, "
            f"{achieved:.0f} GB/s algorithmic")
    g.set_poses(init)
    g.assemble()
    kernels = {k: g.time_kernel(k, 200) for k in ("spmv64", "update64", "spmv_mixed", "assembly")}
    if fused:
        kernels["fused64_in_solve"] = float(np.mean([s[2] for s in samples]))
    elif samples:
        kernels["spmv64_in_solve"] = float(np.mean([s[0] for s in samples]))
        kernels["update64_in_solve"] = float(np.mean([s[1] for s in samples]))
    g.close()
    return gc, init, info, results, roofline, kernels


def _cpu_gn_ms(O, args, gc, init, rp, ci, full_iters):
    """One CPU oracle GN step: edges + CSR assembly + pose update timed in full
    (0 PCG iterations), PCG timed for args.cpu_iters iterations and
    extrapolated to full_iters."""
    times = {}
    for it in (0, args.cpu_iters):
        p = init.copy()
        t = time.perf_counter()
        O.between_gn_step(p, gc["i"], gc["j"], gc["meas"], gc["info"], None, 0.1, rp, ci,
                          tol=1e-300, max_iter=it, wrap=1)
        times[it] = time.perf_counter() - t
    t_iter = (times[args.cpu_iters] - times[0]) / args.cpu_iters
    return 1000.0 * (times[0] + t_iter * full_iters), times[0], t_iter


def cpu_section(args, gc, init, info, results):
    if args.cpu_iters <= 0 or "pcg64" not in results:
        return None
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import LIB_OMP, Oracle
    O = Oracle()
    rp, ci = O.csr_pattern(info["n_poses"], gc["i"], gc["j"])
    full_iters = results["pcg64"]["pcg_iters"]
    cpu_ms, t0, t_iter = _cpu_gn_ms(O, args, gc, init, rp, ci, full_iters)
    log(f"cpu oracle: {cpu_ms:.0f} ms per GN step (extrapolated)")
    # the same step with the PCG loops on all host threads (liboracle_omp.so)
    P = Oracle(LIB_OMP)
    omp_ms, o0, o_iter = _cpu_gn_ms(P, args, gc, init, rp, ci, full_iters)
    log(f"cpu oracle, OpenMP {P.threads()} threads: {omp_ms:.0f} ms per GN step "
        f"({1000 * o_iter:.3f} ms/iteration)")
    omp = dict(value=omp_ms, unit="ms per GN step", cores=P.threads(), kind="port",
               sample=(f"oracle/liboracle_omp.so: the same step with the PCG loops on "
                       f"{P.threads()} OpenMP threads (edges, assembly and pose update on one), "
                       f"{1000 * o0:.0f} ms fixed + {args.cpu_iters} iterations at "
                       f"{1000 * o_iter:.3f} ms/iteration, extrapolated to {full_iters}"))
    return dict(value=cpu_ms, unit="ms per GN step", cores=1, kind="port",
                reference_cli=reference_cli(args), omp=omp,
                sample=(f"CPU oracle (oracle/gslam_oracle.c, FP64, 1 thread) GN step on the same "
                        f"N={info['n_poses']} graph: edges + CSR assembly + pose update timed in "
                        f"full ({1000 * t0:.0f} ms), block-Jacobi PCG timed for "
                        f"{args.cpu_iters} iterations ({1000 * t_iter:.2f} ms/iteration) and "
                        f"extrapolated to the GPU FP64 solve's {full_iters} iterations"))


def _dataset(name):
    """Path of a reference data set shipped (gzip) under tests/golden/data,
    inflated once into _build/data."""
    src = os.path.join(ROOT, "tests", "golden", "data", name)
    if os.path.exists(src):
        return src
    out = os.path.join(ROOT, "_build", "data", name)
    if not os.path.exists(out):
        os.makedirs(os.path.dirname(out), exist_ok=True)
        with gzip.open(src + ".gz", "rb") as fi, open(out + ".tmp", "wb") as fo:
            shutil.copyfileobj(fi, fo)
        os.replace(out + ".tmp", out)
    return out


def reference_cli(args):
    """Whole processes on the reference's own full data sets: the reference's
    CPU executable (oracle/_ref/GSLAM_cpu_pcg_fp64, its cpu_pcg_fp64 sources
    compiled by oracle/Makefile, one thread) against the drop-in GSLAM on the
    GPU, `<exe> 1 0.1 <rounds> 0 <file>` each (file reading, edges, PCG at the
    reference's 1e-6 / 500 iterations, pose updates), with the per-round diff
    lines both print."""
    ref = os.path.join(ROOT, "oracle", "_ref", "GSLAM_cpu_pcg_fp64")
    ours = os.path.join(ROOT, "g-slam_on_gpu_and_fpga_amd", "bin", "GSLAM")
    if args.ref_cli_rounds <= 0 or not (os.path.exists(ref) and os.path.exists(ours)):
        return None
    out = dict(rounds=args.ref_cli_rounds, cores=1)
    for name in ("victoria-park.log.txt", "cityTrees10k.log.txt"):
        path = _dataset(name)
        row = {}
        for tag, exe in (("reference_cpu", ref), ("gslam_gpu", ours)):
            t = time.perf_counter()
            r = subprocess.run([exe, "1", "0.1", str(args.ref_cli_rounds), "0", path],
                               capture_output=True, text=True, timeout=900)
            row[f"{tag}_s"] = time.perf_counter() - t
            row[f"{tag}_diffs"] = [float(l.split(":")[1]) for l in r.stdout.splitlines()
                                   if "rounds executed" in l]
        out[name] = row
        log(f"{name}, {args.ref_cli_rounds} GN rounds: reference CPU executable "
            f"{row['reference_cpu_s']:.2f} s, GSLAM on the GPU {row['gslam_gpu_s']:.2f} s "
            f"(diffs {row['reference_cpu_diffs']} vs {row['gslam_gpu_diffs']})")
    return out


def gj_section(args, gslam, G, np):
    import scipy.sparse as sp
    out = {}
    for N in [int(v) for v in args.gj_n.split(",") if v]:
        gc = G.grid_city(N, seed=0)
        n = 3 * N
        with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
            g.assemble()
            rp, ci, va, xi = g.system_csr()
            A = sp.csr_matrix((va, ci, rp), shape=(n, n)).toarray()
            t = time.perf_counter()
            x_np = np.linalg.solve(A, xi)
            t_np = time.perf_counter() - t
            del A
            nx = np.linalg.norm(x_np)
            row = dict(n=n, numpy_solve_s=t_np)
            for name, kw in (("ir3", dict(refine_sweeps=3, refine_mode="ir", tol=0.0)),
                             ("cg", dict(refine_sweeps=400, refine_mode="cg", tol=1e-12))):
                g.solve("gj_mixed", **kw)   # warm-up (work space allocation)
                st = g.solve("gj_mixed", **kw)
                e = float(np.linalg.norm(g.dx() - x_np) / nx)
                row[f"gj_mixed_{name}"] = dict(ms=st["ms"], factor_ms=st["ms_factor"],
                                               refine_iters=st["iters"], rel_dx_err_vs_numpy=e,
                                               factor_tflops=2.0 * n ** 3 / st["ms_factor"] / 1e9)
                log(f"gj_mixed {name} N={N}: {st['ms']:.2f} ms (factor {st['ms_factor']:.2f}), "
                    f"{st['iters']} refinement its, err vs numpy {e:.2e}")
            g.solve("gj64_blocked")   # warm-up (work space allocation)
            st = g.solve("gj64_blocked")
            e = float(np.linalg.norm(g.dx() - x_np) / nx)
            row["gj64_blocked"] = dict(ms=st["ms"], factor_ms=st["ms_factor"], rel_dx_err_vs_numpy=e,
                                       tflops=2.0 * n ** 3 / st["ms_factor"] / 1e9,
                                       frac_fp64_peak=2.0 * n ** 3 / st["ms_factor"] / 1e9 / FP64_PEAK)
            log(f"gj64_blocked N={N}: {st['ms']:.2f} ms (factor {st['ms_factor']:.2f}), "
                f"err vs numpy {e:.2e}")
            us = g.time_kernel("gj_trailing", 10)
            npd = -(-n // 128) * 128
            fl = 2.0 * (npd - 128) * npd * 128
            row["trailing_kernel"] = dict(us=us, tflops=fl / us / 1e6, bound="mfma",
                                          peak=FP32_MFMA_PEAK, frac=fl / us / 1e6 / FP32_MFMA_PEAK)
            if N <= args.gj64_max_n:
                st = g.solve("gj64")
                e = float(np.linalg.norm(g.dx() - x_np) / nx)
                row["gj64"] = dict(ms=st["ms"], rel_dx_err_vs_numpy=e,
                                   tflops=2.0 * n ** 3 / st["ms"] / 1e9,
                                   frac_fp64_peak=2.0 * n ** 3 / st["ms"] / 1e9 / FP64_PEAK)
                log(f"gj64 N={N}: {st['ms']:.1f} ms, err vs numpy {e:.2e}")
        out[f"N{N}"] = row
    return out


def adjoint_section(args, gslam, G, np):
    from gslam import autodiff as AD
    import torch
    N = args.adj_n
    gc = G.grid_city(N, seed=0)
    conv = dict(solver="pcg64", tol=1e-13, max_iter=400000, diff_tol=1e-13, max_rounds=60)
    out = dict(n_poses=N, n_edges=int(len(gc["i"])))
    truth = gc["truth"].reshape(-1, 3)

    def loss_grad(x):
        p = x.reshape(-1, 3)
        drift = (p[-1, :2] - p[0, :2]) - (truth[-1, :2] - truth[0, :2])
        d = p - truth
        d[:, 2] = (d[:, 2] + np.pi) % (2 * np.pi) - np.pi
        G_ = 2 * d
        G_[-1, :2] += 2 * drift
        G_[0, :2] -= 2 * drift
        return float(drift @ drift + (d ** 2).sum()), G_.ravel()

    with gslam.Graph(gc["poses"], anchor_prior=1, **G.graph_kwargs(gc, 0.1)) as g:
        hip_sync()
        t = time.perf_counter()
        rounds = g.converge(**conv)
        hip_sync()
        out["forward_ms"] = 1000.0 * (time.perf_counter() - t)
        out["forward_gn_rounds"] = rounds
        x0 = g.get_poses()
        L0, dl = loss_grad(x0)
        t = time.perf_counter()
        dw, dlam, st = g.adjoint(dl, tol=1e-13)
        out["adjoint_ms"] = 1000.0 * (time.perf_counter() - t)
        out["adjoint_pcg_iters"] = st["iters"]
        out["dl_dlambda"] = dlam
        log(f"adjoint N={N}: forward {out['forward_ms']:.0f} ms ({rounds} GN rounds), "
            f"backward {out['adjoint_ms']:.1f} ms ({st['iters']} PCG its)")
        if args.adj_fd > 0:
            h = 1e-4
            ks = np.random.default_rng(0).choice(len(dw), args.adj_fd, replace=False)
            fd = []
            for k in ks:
                v = []
                for sgn in (1.0, -1.0):
                    w = np.zeros(len(dw))
                    w[k] = sgn * h
                    g.set_log_weights(w)
                    g.set_poses(x0)
                    g.converge(**conv)
                    v.append(loss_grad(g.get_poses())[0])
                fd.append((v[0] - v[1]) / (2 * h))
            fd = np.array(fd)
            g.set_log_weights(None)
            out["fd_edges"] = int(len(ks))
            out["fd_rel_err"] = float(np.linalg.norm(dw[ks] - fd) / np.linalg.norm(fd))
            out["fd_max_rel_err"] = float(np.max(np.abs(dw[ks] - fd) / np.abs(fd).max()))
            log(f"adjoint vs central FD on {len(ks)} edges: rel {out['fd_rel_err']:.2e}")
        if args.adam > 0:
            t = time.perf_counter()
            _, lam_out, hist = AD.tune(g, gc["poses"], gc["truth"], steps=args.adam, lr=0.05,
                                       solve_kw=conv)
            out["adam"] = dict(steps=args.adam, loss_first=hist[0], loss_last=hist[-1],
                               lambda_last=lam_out, s_per_step=(time.perf_counter() - t) / args.adam)
            log(f"adam: loss {hist[0]:.4f} -> {hist[-1]:.4f} in {args.adam} steps")
    return out


def park_section(args, gslam, G, np):
    """C4b: FP64 and mixed PCG on the park graph at one tolerance, each with
    its error against the refined direct solve; ms of the solve alone (HIP
    events) and of a whole GN step."""
    gc = G.park(args.park_n, seed=0, route="laps")
    out = dict(n_poses=args.park_n, n_loop_closures=int(gc["n_loop"]), lambda_=1.0,
               tol=args.park_tol)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 1.0, wrap=1)) as g:
        g.assemble()
        dx_ref = refined_direct(*g.system_csr())
        for name, kw in (("pcg64", {}), ("mixed", dict(mixed_handover=args.park_handover))):
            g.set_poses(gc["poses"])
            g.assemble()
            g.solve(name, args.park_tol, 400000, **kw)
            st = g.solve(name, args.park_tol, 400000, **kw)
            acc = accuracy(np, g.dx(), dx_ref)
            t_total = 0.0
            for _ in range(args.steps):
                g.set_poses(gc["poses"])     # before the timed region, as above
                hip_sync()
                t = time.perf_counter()
                g.gn_step(name, args.park_tol, 400000, **kw)
                hip_sync()
                t_total += time.perf_counter() - t
            ms = 1000.0 * t_total / args.steps
            out[name] = dict(ms_per_step=ms, solve_ms=st["ms"], pcg_iters=st["iters"],
                             fp64_correction_iters=st["iters_fp64"], **acc, **kw)
            log(f"park {name}: {ms:.1f} ms per GN step (solve {st['ms']:.1f} ms, "
                f"{st['iters']} its), rel dx err {acc['rel_dx_err']:.2e}")
        # the FP64 solve with two launches per iteration beside the default
        # one-launch schedule of graphs this size
        g.set_poses(gc["poses"])
        g.assemble()
        st = g.solve("pcg64", args.park_tol, 400000, pcg_schedule="two_launch")
        st = g.solve("pcg64", args.park_tol, 400000, pcg_schedule="two_launch")
        out["pcg64_two_launch"] = dict(solve_ms=st["ms"], pcg_iters=st["iters"],
                                       us_per_iter=1000.0 * st["ms"] / max(st["iters"], 1),
                                       **accuracy(np, g.dx(), dx_ref))
        out["pcg64"]["us_per_iter"] = 1000.0 * out["pcg64"]["solve_ms"] / max(out["pcg64"]["pcg_iters"], 1)
        log(f"park pcg64 two launches per iteration: solve {st['ms']:.1f} ms, {st['iters']} its "
            f"(one launch: {out['pcg64']['solve_ms']:.1f} ms, {out['pcg64']['pcg_iters']} its)")
    out["mixed_speedup_vs_fp64"] = out["pcg64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    return out


def large_section(args, gslam, G, np):
    """Mixed vs FP64 PCG where both stream from HBM: grid-city N = args.large_n
    (1.2M: 1.4 GB of BSR blocks), one solve to a true relative residual of
    args.large_tol from the dead-reckoned guess's system.  There the FP32
    iterations are about half the cost of FP64 ones and, to this tolerance,
    just as many."""
    gc = G.grid_city(args.large_n, seed=2)
    out = dict(n_poses=args.large_n, n_edges=len(gc["i"]), lambda_=0.1, tol=args.large_tol)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1, wrap=1)) as g:
        g.assemble()
        for name in ("pcg64", "mixed"):
            kw = dict(profile_every=args.profile_every) if name == "pcg64" else {}
            st = g.solve(name, args.large_tol, 200000, **kw)
            out[name] = dict(solve_ms=st["ms"], pcg_iters=st["iters"], converged=st["converged"],
                             true_rel_res=st["true_rel_res"],
                             us_per_iter=1000.0 * st["ms"] / max(st["iters"], 1))
            log(f"large N={args.large_n} {name}: {st['ms']:.0f} ms, {st['iters']} its, true rel res "
                f"{st['true_rel_res']:.2e}")
            if name == "pcg64" and st["spmv_us"] > 0:
                # the dominant kernel where the matrix streams from HBM: HIP
                # events around sampled SpMV launches of this solve
                byts = spmv_bytes(g.info["n_poses"], g.info["n_blocks"])
                gbs = byts / (st["spmv_us"] * 1e-6) / 1e9
                out["roofline_hbm"] = dict(
                    bound="hbm", kernel="pcg_spmv_kernel (FP64 BSR SpMV + p.q partials)",
                    regime=f"HBM-bound: {byts / 1e6:.0f} MB per launch, beyond the 256 MB "
                           f"Infinity Cache", us_per_launch=st["spmv_us"],
                    update_us_per_launch=st["update_us"], bytes_per_launch=byts,
                    achieved=gbs, peak=HBM_PEAK_GBS, unit="GB/s", frac=gbs / HBM_PEAK_GBS,
                    how=f"HIP events around every {args.profile_every}th SpMV launch of the solve")
                log(f"large N={args.large_n} spmv64 in solve: {st['spmv_us']:.1f} us/launch, "
                    f"{gbs:.0f} GB/s = {100 * gbs / HBM_PEAK_GBS:.1f}% of HBM peak")
    out["mixed_speedup_vs_fp64"] = out["pcg64"]["solve_ms"] / out["mixed"]["solve_ms"]
    return out


def grid10k_section(args, gslam, G, np):
    """C2: FP64 PCG GN run on the N=10,000 grid city (lambda 0.1, PCG tol
    1e-10, the reference stop rule diff < 0.01 within 20 rounds); the poses
    of the same run are checked against the CPU oracle in
    tests/test_gpu_gn.py.  Whole-run and per-step milliseconds with the
    device time per phase."""
    gc = G.grid_city(args.grid10k_n, seed=0)
    out = dict(n_poses=args.grid10k_n, lambda_=0.1, tol=1e-10)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1, wrap=1)) as g:
        for schedule in ("two_launch", "auto"):     # the default ('auto') last
            g.set_poses(gc["poses"])
            g.gn_step("pcg64", 1e-10, 200000, pcg_schedule=schedule)   # warm-up round
            g.set_poses(gc["poses"])
            hip_sync()
            t = time.perf_counter()
            its, diffs, phases = [], [], []
            for _ in range(20):
                st, diff = g.gn_step("pcg64", 1e-10, 200000, pcg_schedule=schedule)
                its.append(st["iters"])
                diffs.append(diff)
                phases.append((st["ms_assembly"], st["ms_setup"], st["ms"], st["ms_update"]))
                if diff < 0.01:
                    break
            hip_sync()
            ms = 1000.0 * (time.perf_counter() - t)
            ph = np.mean(np.array(phases), axis=0)
            res = dict(rounds=len(its), ms_run=ms, ms_per_step=ms / len(its), pcg_iters=int(sum(its)),
                       us_per_iter=1000.0 * float(np.sum(np.array(phases)[:, 2])) / max(sum(its), 1),
                       last_diff=diffs[-1], phase_ms_mean=dict(assembly=ph[0], setup=ph[1],
                                                               solve=ph[2], update=ph[3]))
            log(f"grid10k pcg64 ({schedule}): {len(its)} GN rounds in {ms:.1f} ms "
                f"({ms / len(its):.1f} ms per step, {sum(its)} PCG its, "
                f"{res['us_per_iter']:.2f} us per iteration, last diff {diffs[-1]:.2e})")
            if schedule == "auto":
                out.update(res)
            else:
                out["two_launch"] = res
    return out


def gj_cpu_baseline(args, np):
    """The reference's FP64 Gauss-Jordan inv() (src/cpu_pcg_fp64/MatrixOps.c:29-91,
    oracle/_ref) on the host, one thread, at n = args.gj_cpu_n on the N=800
    grid-city Omega's leading block (its work is 4 n^3 flops as written), and
    scaled by (2400 / n)^3 to the N=800 system (n = 2400)."""
    if args.gj_cpu_n <= 0:
        return None
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        from oracle import Oracle, RefPCG64, ref_available
    except ImportError:
        return None
    if not ref_available():
        return None
    import threading
    import scipy.sparse as sp
    from gslam import generators as G
    gc = G.grid_city(800, seed=0)
    O = Oracle()
    ev = O.between_edges(gc["poses"], gc["i"], gc["j"], gc["meas"], gc["info"], None, 0.1, 1)
    rp, ci, va, _ = O.assemble_csr(800, gc["i"], gc["j"], ev)
    n = args.gj_cpu_n
    A = sp.csr_matrix((va, ci, rp), shape=(2400, 2400)).toarray()[:n, :n].copy()
    ref = RefPCG64()
    box = {}

    def run():
        t = time.perf_counter()
        box["inv"] = ref.inv(A)
        box["s"] = time.perf_counter() - t

    # inv() keeps an n x n VLA on the stack: run it on a thread with room for it
    old = threading.stack_size(max(64 << 20, 8 * n * n + (16 << 20)))
    try:
        th = threading.Thread(target=run)
        th.start()
        th.join()
    finally:
        threading.stack_size(old)
    s = box["s"]
    err = float(np.abs(box["inv"] @ A - np.eye(n)).max())
    log(f"reference inv() n={n}: {s:.2f} s on one host thread (|A inv - I| {err:.1e}); "
        f"x{(2400 / n) ** 3:.0f} for n=2400")
    return dict(value=1000.0 * s * (2400.0 / n) ** 3, unit="ms per GJ inversion (n=2400)", cores=1,
                kind="reference",
                sample=f"reference inv() (oracle/_ref build of src/cpu_pcg_fp64/MatrixOps.c) on "
                       f"the leading {n} x {n} block of the N=800 grid-city Omega: {s:.2f} s, "
                       f"scaled by (2400/{n})^3", measured_s=s, n=n)


def rcm_relabel(gc):
    """The grid-city graph with its poses relabelled in reverse Cuthill-McKee
    order of the edge graph (scipy), so that edges join nearby labels."""
    import numpy as np
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    n = len(gc["poses"]) // 3
    i, j = gc["i"], gc["j"]
    a = sp.coo_matrix((np.ones(len(i)), (i, j)), shape=(n, n)).tocsr()
    perm = reverse_cuthill_mckee(a + a.T, symmetric_mode=True)   # new -> old
    new = np.empty(n, np.int64)
    new[perm] = np.arange(n)
    out = dict(gc)
    out["poses"] = gc["poses"].reshape(-1, 3)[perm].ravel().copy()
    out["i"], out["j"] = new[i].astype(i.dtype), new[j].astype(j.dtype)
    log(f"rcm: mean |i - j| over the edges {np.abs(i - j).mean():.0f} -> "
        f"{np.abs(out['i'] - out['j']).mean():.0f}")
    return out


def profile_kernel(args, gslam, G):
    """A few launches of one kernel for counter collection."""
    if args.profile_kernel == "copy":
        import torch
        a = torch.ones(16 << 20, dtype=torch.float32, device="cuda")   # 64 MiB
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copy: 5 launches, 64 MiB read + 64 MiB written each")
        return
    if args.profile_kernel == "stream":
        import torch
        a = torch.ones(1 << 29, dtype=torch.float32, device="cuda")    # 2 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        us = 1000.0 * e0.elapsed_time(e1) / 10
        gbs = 2 * (1 << 31) / (us * 1e-6) / 1e9
        log(f"stream: 2 GiB copy {us:.0f} us, {gbs:.0f} GB/s read+write = "
            f"{100 * gbs / HBM_PEAK_GBS:.1f}% of the 8 TB/s HBM peak")
        del a, b
        # reads and copies of working sets the size of the PCG iteration's (69
        # MB at N = 100k), re-read by every launch as the iteration re-reads
        # its matrix and vectors: what plain streaming kernels achieve there
        probe = os.path.join(ROOT, "g-slam_on_gpu_and_fpga_amd", "bin", "gslam_stream_probe")
        if not os.path.exists(probe):
            raise SystemExit(f"{probe} is not built (python -c 'import __graft_entry__ as e; e.build()')")
        torch.cuda.synchronize()
        out = subprocess.run([probe, "31", "69", "1024"], check=True, capture_output=True, text=True)
        for line in out.stdout.splitlines():
            log(line)
        return
    if args.profile_kernel == "spmv_mixed":
        gc = G.grid_city(args.n, seed=args.seed)
        with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
            g.assemble()
            us = g.time_kernel("spmv_mixed", 20)
            byts = mixed_spmv_bytes(g.info["n_poses"], 2 * g.info["n_edges"])
            gbs = byts / (us * 1e-6) / 1e9
            log(f"spmv_mixed N={args.n}: {us:.2f} us/launch (23 launches), {byts / 1e6:.1f} MB "
                f"algorithmic, {gbs:.0f} GB/s = {100 * gbs / HBM_PEAK_GBS:.1f}% of HBM peak")
        return
    if args.profile_kernel == "assembly":
        gc = G.grid_city(args.n, seed=args.seed)
        for sched in ("pose_walk", "staged"):
            with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1, wrap=1), assembly=sched) as g:
                g.assemble()
                us = g.time_kernel("assembly", 20)
                byts = assembly_bytes(g.info["n_poses"], g.info["n_edges"], g.info["n_blocks"])
                log(f"assembly ({sched}) N={args.n}: {us:.2f} us per assembly (23 launches), "
                    f"{byts / 1e6:.1f} MB algorithmic, {byts / (us * 1e-6) / 1e9:.0f} GB/s = "
                    f"{100 * byts / (us * 1e-6) / 1e9 / HBM_PEAK_GBS:.1f}% of HBM peak; "
                    f"handle {g.device_bytes() / 1e6:.0f} MB")
        return
    if args.profile_kernel == "pcg_solve":
        # whole FP64 PCG solves (tol 1e-10) with the chosen launch schedule
        gc = G.grid_city(args.n, seed=args.seed) if args.graph == "grid" else \
            G.park(args.n, seed=0, route="laps")
        lam = 0.1 if args.graph == "grid" else 1.0
        with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, lam)) as g:
            g.assemble()
            for _ in range(3):
                st = g.solve("pcg64", 1e-10, args.pcg_iters, pcg_schedule=args.pcg_schedule)
                log(f"pcg_solve {args.graph} N={args.n} ({args.pcg_schedule}): {st['ms']:.2f} ms, "
                    f"{st['iters']} its, {1000 * st['ms'] / max(st['iters'], 1):.2f} us per iteration")
        return
    if args.profile_kernel in ("spmv64", "update64"):
        k = args.profile_kernel
        gc = G.grid_city(args.n, seed=args.seed)
        if args.order == "rcm":
            gc = rcm_relabel(gc)
        with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
            g.assemble()
            us = g.time_kernel(k, 20)
            byts = (spmv_bytes(g.info["n_poses"], g.info["n_blocks"]) if k == "spmv64"
                    else update_bytes(g.info["n_poses"]))
            gbs = byts / (us * 1e-6) / 1e9
            log(f"{k} N={args.n}: {us:.2f} us/launch (23 launches), {byts / 1e6:.1f} MB "
                f"algorithmic, {gbs:.0f} GB/s = {100 * gbs / HBM_PEAK_GBS:.1f}% of HBM peak")
        return
    N = int(args.gj_n.split(",")[0] or 1600)
    gc = G.grid_city(N, seed=0)
    with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
        g.assemble()
        g.solve("gj_mixed", tol=0.0, refine_sweeps=0)
        log(f"gj_trailing N={N}: {g.time_kernel('gj_trailing', 5):.1f} us/launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--order", choices=["gen", "rcm"], default="gen",
                    help="pose labels for --profile-kernel spmv64/update64: generator order "
                         "or reverse Cuthill-McKee")
    ap.add_argument("--tol", type=float, default=3e-13, help="FP64 PCG tolerance")
    ap.add_argument("--mixed-tol", type=float, default=3e-13)
    ap.add_argument("--handover", type=float, default=1e-2,
                    help="mixed: FP64 true relative residual at which FP64 PCG takes over the "
                         "correction (0: FP32 iterations all the way)")
    ap.add_argument("--park-n", type=int, default=7120, help="park section size (0 skips)")
    ap.add_argument("--large-n", type=int, default=1200000,
                    help="size of the HBM-regime mixed vs FP64 section (0 skips)")
    ap.add_argument("--large-tol", type=float, default=1e-5)
    ap.add_argument("--grid10k-n", type=int, default=10000,
                    help="size of the FP64 PCG GN-run section (0 skips)")
    ap.add_argument("--park-tol", type=float, default=1e-9)
    ap.add_argument("--park-handover", type=float, default=1e-2)
    ap.add_argument("--gj-cpu-n", type=int, default=1200,
                    help="dimension of the reference inv() CPU sample (0 skips)")
    ap.add_argument("--ref-tol", type=float, default=1e-14)
    ap.add_argument("--max-iter", type=int, default=600000)
    ap.add_argument("--solvers", default="mixed,pcg64")
    ap.add_argument("--profile-every", type=int, default=97,
                    help="sample every k-th SpMV/update launch of the timed FP64 solves")
    ap.add_argument("--cpu-iters", type=int, default=5000,
                    help="PCG iterations of the bounded CPU oracle sample (0: skip)")
    ap.add_argument("--ref-cli-rounds", type=int, default=3,
                    help="GN rounds of the reference executable vs GSLAM comparison (0: skip)")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--gj-n", default="800,1600,5000",
                    help="grid-city sizes of the GJ section ('' skips; --profile-kernel "
                         "gj_trailing uses the first)")
    ap.add_argument("--gj64-max-n", type=int, default=2000)
    ap.add_argument("--adj-n", type=int, default=10000, help="adjoint section size (0 skips)")
    ap.add_argument("--adj-fd", type=int, default=8, help="edges checked by central differences")
    ap.add_argument("--adam", type=int, default=0, help="Adam steps on (w_e, log lambda)")
    ap.add_argument("--graph", choices=["grid", "park"], default="grid",
                    help="--profile-kernel pcg_solve: grid city or the park graph (laps route)")
    ap.add_argument("--pcg-schedule", choices=["auto", "two_launch"], default="auto",
                    help="--profile-kernel pcg_solve: FP64 PCG launch schedule")
    ap.add_argument("--pcg-iters", type=int, default=400000,
                    help="--profile-kernel pcg_solve: iteration cap (a short run for counters)")
    ap.add_argument("--profile-kernel", choices=["spmv64", "update64", "spmv_mixed", "pcg_solve", "gj_trailing",
                                                 "assembly", "copy", "stream"],
                    help="only launch this kernel a few times and exit (rocprofv3 counter "
                         "passes; 'copy' is a 64 MiB device copy for calibrating the counters, "
                         "'stream' times a 2 GiB device copy: the achievable HBM bandwidth)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import gslam
    from gslam import generators as G

    if args.profile_kernel:
        profile_kernel(args, gslam, G)
        return
    gc, init, info, results, roofline, kernels = pcg_section(args, gslam, G, np)
    cpu = cpu_section(args, gc, init, info, results)
    gj = gj_section(args, gslam, G, np) if args.gj_n else None
    park = park_section(args, gslam, G, np) if args.park_n > 0 else None
    grid10k = grid10k_section(args, gslam, G, np) if args.grid10k_n > 0 else None
    large = large_section(args, gslam, G, np) if args.large_n > 0 else None
    gj_cpu = gj_cpu_baseline(args, np)
    if cpu is not None and gj_cpu is not None:
        cpu["gj_inv"] = gj_cpu
    adj = adjoint_section(args, gslam, G, np) if args.adj_n > 0 else None

    fp = results.get("pcg64", {})
    mx = results.get("mixed", {})
    head = fp if fp else mx
    out = {
        "metric": METRIC,
        "value": head.get("ms_per_step"),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head.get("ms_per_step"),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if fp else "f32",
        "data": "synthetic",
        "config": {"workload": f"grid-city pose graph, N={info['n_poses']} poses, seed {args.seed}, "
                               f"lambda 0.1, one GN step from the dead-reckoned initial guess, "
                               f"block-Jacobi PCG to rel tol {args.tol:g} (FP64); mixed: FP32 "
                               f"iterations to {args.handover:g}, then FP64 to {args.mixed_tol:g}",
                   "n_poses": info["n_poses"], "n_edges": info["n_edges"],
                   "n_blocks": info["n_blocks"], "nnz": info["nnz"],
                   "headline_solver": "pcg64" if fp else "mixed"},
        "fp64": fp,
        "mixed": mx,
        "mixed_speedup_vs_fp64": (fp["ms_per_step"] / mx["ms_per_step"]) if fp and mx else None,
        "kernels_us": kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gj": gj,
        "park": park,
        "grid10k": grid10k,
        "large": large,
        "adjoint": adj,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
