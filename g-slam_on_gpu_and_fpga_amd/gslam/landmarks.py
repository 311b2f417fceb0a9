"""Landmark position estimates after the Gauss-Newton loop.

The reference computes, after its last round, one estimate per landmark
(src/cpu_pcg_fp64/GSLAM.c:141-196 with MapEdge_create, MapEdge.c:12-29): each
observation (pose x, y, th; range r, bearing phi) places the landmark at
(x + r cos(th + phi), y + r sin(th + phi)), and the estimate is the mean over
the landmark's observations, landmarks in first-appearance order.  The
reference frees the result without output; here it is returned (for drawing
and evaluation).
"""
from __future__ import annotations

import numpy as np


def landmark_means(poses, z_step, z_lm, z_range, z_bearing):
    """(ids[L], means[L, 2]) in first-appearance order of the landmark ids."""
    p = np.asarray(poses, np.float64).reshape(-1, 3)
    zs = np.asarray(z_step, np.int64)
    zl = np.asarray(z_lm, np.int64)
    r = np.asarray(z_range, np.float64)
    phi = np.asarray(z_bearing, np.float64)
    if zs.size == 0:
        return np.zeros(0, np.int32), np.zeros((0, 2))
    uniq, first, inv = np.unique(zl, return_index=True, return_inverse=True)
    rank = np.empty(len(uniq), np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(len(uniq))
    slot = rank[inv]
    a = p[zs, 2] + phi
    m = np.stack([p[zs, 0] + r * np.cos(a), p[zs, 1] + r * np.sin(a)], 1)
    s = np.zeros((len(uniq), 2))
    np.add.at(s, slot, m)                      # observation order within a landmark
    cnt = np.bincount(slot, minlength=len(uniq))
    ids = np.empty(len(uniq), np.int64)
    ids[rank] = uniq
    return ids.astype(np.int32), s / cnt[:, None]
