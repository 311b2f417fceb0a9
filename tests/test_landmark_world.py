"""gslam.generators.landmark_world: a seeded landmark world in the
reference's .log.txt model (after data/generate_citytrees.py)."""
import numpy as np

from conftest import ROOT  # noqa: F401  (sys.path)
from gslam import generators as G
import gslam


def test_landmark_world_shape_and_log_round_trip(tmp_path):
    a = G.landmark_world(300, seed=4)
    b = G.landmark_world(300, seed=4)
    assert np.array_equal(a["z_range"], b["z_range"]) and np.array_equal(a["poses"], b["poses"])
    t = a["truth"]
    # closed block route: every 4 turns the robot is back near the start
    assert np.hypot(*(t[200, :2] - t[0, :2])) < 2.0
    # observations lie within sensor range (noise aside) and use landmark ids from 6
    assert a["z_range"].max() < 8.5 + 0.5 and a["z_lm"].min() >= 6
    # consecutive initial poses follow the logged (noisy) controls exactly
    p = a["poses"].reshape(-1, 3)
    x, y, th = G._velocity_step(*p[10], a["u_nu"][11], a["u_omega"][11])
    assert np.allclose([x, y], p[11, :2], atol=1e-9)
    path = str(tmp_path / "w.log.txt")
    G.write_log(a, path)
    r = gslam.read_log(path)
    assert np.array_equal(r["poses"], a["poses"]) and np.array_equal(r["z_bearing"], a["z_bearing"])
    assert np.array_equal(r["u_omega"], a["u_omega"]) and r["delta"] == a["delta"]
