"""Writes the golden fixtures under tests/golden/ from oracle/_ref, the build
of the reference C sources (src/cpu_pcg_fp64, src/cpu_gj_mixed).

Run in a container that has the reference sources:
    make -C oracle ref && python tests/golden/make_golden.py

Every number in the fixtures is produced by the reference's own code:
  * round-1 edge blocks: MotionEdge_create / ObsEdge_create via make_edges
    (src/cpu_pcg_fp64/GSLAM.c:56-72)
  * round-1 CSR and xi: build_csr_from_edges (MatrixOps.c:173-273)
  * PCG solution at (1e-6, 500): pcg_solve_sparse (MatrixOps.c:281-391)
  * dense GJ solve: inv (MatrixOps.c:29-91) then dot (MatrixOps.c:123-147)
  * FP32 GJ: inv_float (src/cpu_gj_mixed/MatrixOps.c:174-212)
  * end-to-end: the reference GSLAM executable's stdout and POSE_DUMP_FILE
    after 1 and 3 rounds (GSLAM.c:13-220)
The inputs are the reference's own data files, copied to tests/golden/data/.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from oracle import Oracle, RefGJMixed, RefPCG64, ref_available  # noqa: E402

DATA = os.path.join(ROOT, "tests", "golden", "data")
OUT = os.path.join(ROOT, "tests", "golden")
SETS = {
    "ct100": "cityTrees100.log.txt",
    "vp100": "victoria-park-100.log.txt",
    "ct800": "cityTrees800.txt",
}
LAM = 0.1


def main():
    if not ref_available():
        sys.exit("oracle/_ref is not built: run `make -C oracle ref` first")
    O = Oracle()
    R = RefPCG64()
    G = RefGJMixed()
    for key, fname in SETS.items():
        path = os.path.join(DATA, fname)
        g = O.read_log(path)  # parse only; every value below comes from _ref
        n = g["n_poses"]
        edges, ne = R.round_edges(g, LAM)
        t1, t2, ev = R.edge_arrays(edges, ne)
        csr, rp, ci, va, xi = R.csr(edges, ne, 3 * n)
        x = R.pcg(csr, xi)
        R.free_csr(csr)
        fx = dict(lam=LAM, t1=t1, t2=t2, row_ptr=rp, col_idx=ci, values=va, xi=xi,
                  pcg_x=x)
        if key != "ct800":
            fx["edge_vals"] = ev
        if n <= 100:
            # dense GJ path (the reference inv() keeps a dim*dim VLA on the stack)
            om, xi_d = O.dense(n, t1, t2, ev)
            inv = R.inv(om)
            fx["gj_dx"] = O.dense_matvec(inv, xi_d)  # dot(): MatrixOps.c:123-147
            fx["gj_inv_diag"] = np.diag(inv).copy()
            fx["gj_inv_rows012"] = inv[:3].copy()
            inv32 = G.inv_float(om.astype(np.float32))
            fx["gj32_inv_diag"] = np.diag(inv32).copy()
            fx["gj32_dx"] = O.dense_matvec(inv32.astype(np.float64), xi_d)
        for rounds in (1, 3):
            with tempfile.TemporaryDirectory() as td:
                dump = os.path.join(td, "poses.txt")
                env = dict(os.environ, POSE_DUMP_FILE=dump)
                out = subprocess.run([R.exe(), "1", str(LAM), str(rounds), "1", path],
                                     capture_output=True, text=True, env=env, check=True,
                                     cwd=td)
                fx[f"stdout_r{rounds}"] = np.array(out.stdout.replace(path, "<input>"))
                fx[f"poses_r{rounds}"] = np.loadtxt(dump)
        np.savez_compressed(os.path.join(OUT, f"{key}_ref.npz"), **fx)
        print(key, n, "edges", ne, "nnz", len(ci))


if __name__ == "__main__":
    main()
