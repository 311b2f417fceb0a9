"""CPU-side checks of the C-ABI boundary: libgslam.so loads, exports every
function include/*.h declares, and fails cleanly (an error code, not a crash or
a CPU fallback) when no GPU is present.  No GPU compute is called here."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT

import gslam

HEADERS = [os.path.join(ROOT, "include", h) for h in ("gslam.h", "gslam_compat.h")]


def declared_functions():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?(?:int|void|char)\s*\**\s*(\w+)\s*\(", text, re.M):
            names.append(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 25, names
    L = gslam.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the binding's list covers the same set
    assert set(names) == set(gslam.EXPORTS), set(names) ^ set(gslam.EXPORTS)


def test_defaults_reproduce_reference_constants():
    o = gslam._Opts()
    gslam.lib().gslam_default_opts(C.byref(o))
    assert (o.solver, o.tol, o.max_iter) == (0, 1e-6, 500)  # MatrixOps.c:345-347
    assert (o.refine_sweeps, o.refine_mode, o.profile_every) == (3, 0, 0)
    assert (o.refresh_every, o.check_every) == (64, 0)
    d = gslam._Desc()
    gslam.lib().gslam_default_desc(C.byref(d))
    assert list(d.mns) == [0.19, 0.001, 0.13, 0.2]           # GSLAM.c:66
    assert list(d.snr) == [0.14, 0.05]                        # GSLAM.c:301
    assert d.anchor == 1e6                                    # MatrixOps.c:227
    assert d.wrap == 0 and d.pair_order == 0 and d.anchor_prior == 0


def _gpu_present():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_gpu_present(), reason="checks the no-GPU error path")
def test_create_without_gpu_is_an_error_not_a_fallback():
    with pytest.raises(gslam.GSlamError) as ei:
        gslam.Graph([0.0, 0.0, 0.0, 1.0, 0.0, 0.0],
                    between=dict(i=[0], j=[1], meas=[1.0, 0, 0], info=[1.0, 1, 1]))
    assert ei.value.code == gslam.ERR_HIP


def test_argument_validation_before_any_device_call():
    """Invalid descriptions are rejected by the host-side structure build."""
    with pytest.raises(gslam.GSlamError) as ei:
        gslam.Graph([0.0, 0.0, 0.0], between=dict(i=[0], j=[5], meas=[1.0, 0, 0],
                                                  info=[1.0, 1, 1]))
    assert ei.value.code == gslam.ERR_ARG
    with pytest.raises(gslam.GSlamError) as ei:
        gslam.Graph([0.0, 0.0, 0.0, 1, 0, 0], obs=dict(z_step=[0, 7], z_lm=[1, 1],
                                                       z_range=[1.0, 1.0], z_bearing=[0.0, 0.0]))
    assert ei.value.code == gslam.ERR_ARG


def test_non_finite_inputs_rejected_before_any_device_call():
    """NaN / Inf in poses, edge values or model constants: GSLAM_ERR_ARG from
    the host-side structure build, with the offending input named."""
    nan, inf = float("nan"), float("inf")
    bet = dict(i=[0], j=[1], meas=[1.0, 0, 0], info=[1.0, 1, 1])
    cases = [
        (dict(poses=[0.0, 0.0, nan, 1, 0, 0], between=bet), "pose"),
        (dict(poses=[0.0, 0, 0, 1, 0, 0], between=dict(bet, meas=[inf, 0, 0])), "between"),
        (dict(poses=[0.0, 0, 0, 1, 0, 0], between=dict(bet, info=[1.0, nan, 1])), "between"),
        (dict(poses=[0.0, 0, 0, 1, 0, 0], between=bet, lam=nan), "lambda"),
        (dict(poses=[0.0, 0, 0, 1, 0, 0], obs=dict(z_step=[0, 1], z_lm=[3, 3],
                                                    z_range=[1.0, nan], z_bearing=[0.0, 0.0])),
         "observation"),
    ]
    for kw, what in cases:
        poses = kw.pop("poses")
        with pytest.raises(gslam.GSlamError) as ei:
            gslam.Graph(poses, **kw)
        assert ei.value.code == gslam.ERR_ARG and what in str(ei.value), (what, str(ei.value))


def test_read_log_matches_oracle_reader(oracle):
    from conftest import data_path
    import numpy as np
    for f in ("cityTrees100.log.txt", "victoria-park-100.log.txt", "cityTrees800.txt"):
        a = gslam.read_log(data_path(f))
        b = oracle.read_log(data_path(f))
        for k in ("u_nu", "u_omega", "poses", "z_range", "z_bearing", "z_step", "z_lm"):
            assert np.array_equal(a[k], b[k]), (f, k)
        assert a["delta"] == b["delta"]


def test_ctypes_structs_match_the_header(tmp_path):
    """Field offsets and sizes of the binding's ctypes mirrors equal the C
    header's (compiled with gcc), so a field added on one side only fails."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    pairs = [("gslam_solve_opts", gslam._Opts), ("gslam_solve_stats", gslam._Stats),
             ("gslam_graph_desc", gslam._Desc), ("gslam_graph_info", gslam._Info)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gslam.h"', "int main(void) {"]
    for cname, cls in pairs:
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            cf = "lambda" if fname == "lambda_" else fname
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {cf}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    for cname, cls in pairs:
        assert int(out[f"{cname} size"]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(out[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)
