"""gslam_gj_invert_device: the allocation-free variant of gslam_gj_invert
(FP64 Gauss-Jordan on device buffers, scratch in a static device array),
called through the C-ABI on device memory, against the allocating entry
point on the same matrices: two calls in a row, the second smaller, in place,
a zero pivot, and a size beyond the static scratch."""
import ctypes as C

import numpy as np
import pytest

import gslam

pytestmark = pytest.mark.gpu


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n))
    return a @ a.T + n * np.eye(n)


def _lib():
    L = gslam.lib()
    L.gslam_gj_invert_device.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
    L.gslam_gj_invert_device.restype = C.c_int
    return L


def test_device_inverse_matches_allocating_entry_point():
    import torch
    L = _lib()
    for n, seed in ((300, 0), (120, 1)):          # the second call smaller
        a = _spd(n, seed)
        ref = gslam.gj_invert(a, "gj64")          # gslam_gj_invert, GJ_FP64
        ta = torch.from_numpy(a).to("cuda")
        tinv = torch.full_like(ta, float("nan"))
        torch.cuda.synchronize()
        assert L.gslam_gj_invert_device(ta.data_ptr(), tinv.data_ptr(), n) == gslam.OK
        torch.cuda.synchronize()
        assert np.array_equal(tinv.cpu().numpy(), ref)
        assert np.array_equal(ta.cpu().numpy(), a)            # input untouched
        # in place
        assert L.gslam_gj_invert_device(ta.data_ptr(), ta.data_ptr(), n) == gslam.OK
        torch.cuda.synchronize()
        assert np.array_equal(ta.cpu().numpy(), ref)


def test_device_inverse_zero_pivot_and_size_limit():
    import torch
    L = _lib()
    z = np.eye(40)
    z[7, 7] = 0.0
    tz = torch.from_numpy(z).to("cuda")
    out = torch.ones_like(tz)
    torch.cuda.synchronize()
    assert L.gslam_gj_invert_device(tz.data_ptr(), out.data_ptr(), 40) == gslam.ERR_SINGULAR
    torch.cuda.synchronize()
    assert not out.cpu().numpy().any()                        # zero matrix, as the reference
    # after a zero pivot the flag starts from zero again
    a = _spd(50, 2)
    ta = torch.from_numpy(a).to("cuda")
    torch.cuda.synchronize()
    assert L.gslam_gj_invert_device(ta.data_ptr(), ta.data_ptr(), 50) == gslam.OK
    torch.cuda.synchronize()
    assert np.array_equal(ta.cpu().numpy(), gslam.gj_invert(a, "gj64"))
    # beyond the static scratch: rejected before any device work
    assert L.gslam_gj_invert_device(ta.data_ptr(), ta.data_ptr(), 16385) == gslam.ERR_TOO_LARGE
    assert b"16384" in L.gslam_last_error()
