"""One HIP runtime per process: libgslam.so binds to the runtime PyTorch
loads (gslam.lib() imports torch first when it is installed), and a second
runtime -- libgslam loaded before torch by some other route -- is detected
and reported instead of failing later inside a GPU call."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "g-slam_on_gpu_and_fpga_amd")


def _run(code, **env):
    e = dict(os.environ, PYTHONPATH=PKG, **env)
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=e,
                          timeout=300)


def test_one_runtime_when_torch_comes_after_gslam():
    r = _run("import gslam; gslam.lib(); import torch; rt = gslam.hip_runtimes(); "
             "gslam.check_single_runtime(); print(len(rt))")
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "1"


def test_second_runtime_is_reported():
    # GSLAM_NO_TORCH=1 loads libgslam with its own runtime first; the later
    # torch import maps a second copy, which check_single_runtime names
    r = _run("import gslam; gslam.lib(); import torch; gslam.check_single_runtime()",
             GSLAM_NO_TORCH="1")
    assert r.returncode != 0
    assert "two HIP runtimes" in r.stderr


@pytest.mark.gpu
def test_torch_after_graph_in_fresh_process():
    code = """
import numpy as np, gslam
from gslam import generators as G
gc = G.grid_city(3000, seed=1)
with gslam.Graph(gc["poses"], **G.graph_kwargs(gc, 0.1)) as g:
    g.assemble()
    st = g.solve("pcg64", 1e-10, 100000)
    assert st["converged"], st
    import torch
    a = torch.ones(1 << 26, device="cuda")          # 256 MiB after libgslam's allocations
    assert float(a.sum()) == float(1 << 26)
    st2, _ = g.gn_step("pcg64", 1e-10, 100000)
    assert st2["converged"], st2
gslam.check_single_runtime()
print("ok", len(gslam.hip_runtimes()))
"""
    r = _run(code)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok 1"
