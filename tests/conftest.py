import gzip
import os
import shutil
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "g-slam_on_gpu_and_fpga_amd")
for p in (os.path.join(ROOT, "oracle"), PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
DATA = os.path.join(GOLDEN, "data")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct MI355X (gfx950)")
    config.addinivalue_line("markers", "slow: longer CPU case")
    # One HIP runtime per process: gslam.lib() imports torch before it loads
    # libgslam.so (tests/test_runtime.py); importing it here as well keeps
    # test files that load the library through other routes (the oracle's
    # ctypes handles, subprocess launches) in the same state.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle()


def golden(key: str):
    return np.load(os.path.join(GOLDEN, f"{key}_ref.npz"), allow_pickle=False)


GOLDEN_SETS = {"ct100": "cityTrees100.log.txt", "vp100": "victoria-park-100.log.txt",
               "ct800": "cityTrees800.txt"}


def data_path(name: str, tmp_dir: str | None = None) -> str:
    """Path of a dataset under tests/golden/data; .gz files are inflated once."""
    p = os.path.join(DATA, name)
    if os.path.exists(p):
        return p
    gz = p + ".gz"
    if not os.path.exists(gz):
        raise FileNotFoundError(p)
    out_dir = tmp_dir or os.path.join(ROOT, "_build", "data")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, name)
    if not os.path.exists(out):
        with gzip.open(gz, "rb") as fi, open(out + ".tmp", "wb") as fo:
            shutil.copyfileobj(fi, fo)
        os.replace(out + ".tmp", out)
    return out
