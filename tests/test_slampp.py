"""SLAM++ -> .log.txt conversion (gslam.slampp).

* "data" convention against the reference's shipped conversion: the first
  300 odometry steps of data/cityTrees10k.txt and the matching lines of
  data/cityTrees10k.log.txt (fixture pair from tests/golden/
  make_slampp_fixture.py);
* "script" convention (data/convert_slampp.py:29-84 as written) against
  closed forms: straight steps advance by nu along the heading, turning steps
  end on the chord of length 2 nu / |omega| |sin(omega delta / 2)|, gaps
  repeat the pose with a zero control, landmark values pass unchanged;
* the written file is read back by the .log.txt reader of the oracle.
"""
import math

import numpy as np
import pytest

from conftest import DATA

from gslam import slampp


def _read_log(path):
    u, x, z = {}, {}, []
    for line in open(path):
        p = line.split()
        if p[0] == "u":
            u[int(p[1])] = (float(p[2]), float(p[3]))
        elif p[0] == "x":
            x[int(p[1])] = (float(p[2]), float(p[3]), float(p[4]))
        elif p[0] == "z":
            z.append((int(p[1]), int(p[2]), float(p[3]), float(p[4])))
    n = len(x)
    return (np.array([u[k] for k in range(n)]), np.array([x[k] for k in range(n)]),
            np.array(z, dtype=np.float64))


def test_data_convention_matches_reference_conversion():
    d = slampp.read_slampp(f"{DATA}/cityTrees10k_head.txt")
    u, x, z = slampp.to_log(d)
    ur, xr, zr = _read_log(f"{DATA}/cityTrees10k_head.log.txt")
    assert x.shape == xr.shape == (301, 3) and len(z) == len(zr) == 91
    assert np.abs(x - xr).max() <= 1e-9          # printed to ~16 digits, 300 compositions
    assert np.abs(u - ur).max() <= 1e-9
    assert np.array_equal(z[:, :2], zr[:, :2])
    assert np.abs(z[:, 2:] - zr[:, 2:]).max() <= 1e-12


def test_script_convention_closed_forms(tmp_path):
    src = tmp_path / "in.txt"
    src.write_text(
        "ODOMETRY 0 1 1.0 0.0 0.0 1 0 0 1 0 1\n"
        "ODOMETRY 1 2 0.6 0.8 0.5 1 0 0 1 0 1\n"
        "ODOMETRY 3 4 2.0 0.0 0.0 1 0 0 1 0 1\n"      # pose 2 -> 3 missing: a gap
        "LANDMARK 1 9 3.5 -0.25 1 0 1\n")
    u, x, z = slampp.to_log(slampp.read_slampp(str(src)), delta=1.0, convention="script")
    assert x.shape == (5, 3)
    assert np.allclose(x[1], [1.0, 0.0, 0.0]) and np.allclose(u[1], [1.0, 0.0])
    nu, om = 1.0, 0.5                                 # |(0.6, 0.8)| = 1
    assert np.allclose(u[2], [nu, om])
    chord = math.hypot(*(x[2, :2] - x[1, :2]))
    assert abs(chord - 2 * nu / om * abs(math.sin(om / 2))) <= 1e-12
    assert abs(x[2, 2] - 0.5) <= 1e-15
    assert np.array_equal(x[3], x[2]) and np.array_equal(u[3], [0.0, 0.0])
    assert np.allclose(x[4, :2], x[3, :2] + 2.0 * np.array([math.cos(0.5), math.sin(0.5)]))
    assert np.array_equal(z, [[1, 9, 3.5, -0.25]])


def test_written_file_reads_back(tmp_path, oracle):
    out = tmp_path / "head.log.txt"
    n, m = slampp.convert(f"{DATA}/cityTrees10k_head.txt", str(out))
    g = oracle.read_log(str(out))
    assert g["n_poses"] == n == 301 and len(g["z_step"]) == m == 91
    ur, xr, zr = _read_log(f"{DATA}/cityTrees10k_head.log.txt")
    assert np.abs(g["poses"].reshape(-1, 3) - xr).max() <= 1e-9
    assert np.abs(g["z_range"] - zr[:, 2]).max() <= 1e-12


def test_cli(tmp_path, capsys):
    out = tmp_path / "o.log.txt"
    assert slampp.main([f"{DATA}/cityTrees10k_head.txt", str(out)]) == 0
    assert "301 poses, 91 observations" in capsys.readouterr().out
    assert slampp.main([]) == 1
    with pytest.raises(ValueError):
        slampp.to_log(slampp.read_slampp(f"{DATA}/cityTrees10k_head.txt"), convention="other")
