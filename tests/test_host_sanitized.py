"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

* The CPU oracle (oracle/gslam_oracle.c, `make -C oracle sanitize`) driven
  through every entry point by oracle/sanitize_main.c on reference logs.
* libgslam's host-side graph structure (csrc/graph_build.cpp,
  csrc/reference_cost.cpp: BSR pattern, incidence lists, the mixed
  operator's tiles including a pose with more than 256 incidence entries,
  landmark pairs in both orders, the Gauss-Jordan drivers' cost, refused
  descriptions), built with g++ and the sanitizers beside a checking driver
  (tests/host_checks/graph_build_check.cpp).

Device code is not built with sanitizers (not available for gfx950 here)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT, data_path

PKG = os.path.join(ROOT, "g-slam_on_gpu_and_fpga_amd")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def _run(cmd, **kw):
    return subprocess.run(cmd, capture_output=True, text=True, timeout=600, **kw)


@pytest.fixture(scope="module")
def oracle_sanitized():
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    out = _run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize"])
    if out.returncode != 0:
        pytest.skip("sanitizer runtime not available: " + out.stderr[-300:])
    return os.path.join(ROOT, "oracle", "_san", "oracle_sanitized")


@pytest.mark.parametrize("name", ["cityTrees100.log.txt", "victoria-park-100.log.txt",
                                  "cityTrees10k_head.log.txt"])
def test_oracle_clean_under_asan_ubsan(oracle_sanitized, name):
    out = _run([oracle_sanitized, data_path(name)])
    assert out.returncode == 0, out.stdout + out.stderr
    assert "oracle_sanitized: ok" in out.stdout
    assert "runtime error" not in out.stderr and "AddressSanitizer" not in out.stderr


def test_graph_build_clean_under_asan_ubsan(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "graph_build_check")
    src = [os.path.join(ROOT, "tests", "host_checks", "graph_build_check.cpp"),
           os.path.join(PKG, "csrc", "graph_build.cpp"), os.path.join(PKG, "csrc", "reference_cost.cpp")]
    out = _run(["g++", "-O1", "-g", "-std=c++17", *SAN, "-I" + os.path.join(ROOT, "include"),
                "-I" + os.path.join(PKG, "csrc"), *src, "-o", exe])
    if out.returncode != 0 and ("asan" in out.stderr or "ubsan" in out.stderr):
        pytest.skip("sanitizer runtime not available: " + out.stderr[-300:])
    assert out.returncode == 0, out.stderr
    out = _run([exe])
    assert out.returncode == 0, out.stdout + out.stderr
    assert "graph_build_check: ok" in out.stdout
