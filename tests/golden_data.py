"""Loaders of the committed fixtures in tests/golden/ (written by tests/golden/make_fixtures.py)
and the reference's dense accuracy measure (benchmarks/README.md, "dense" row: a reference's
median relative depth error at its grid sites and its accurate coverage, the share of sampled
grid pixels with a point within 2% of the true depth)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# The reference's gate for the dense stage (benchmarks/README.md:60).
MAX_MEDIAN_RELATIVE_ERROR = 0.01
MIN_ACCURATE_COVERAGE = 0.4


def help_pan():
    """(luma frames, cameras as (frame_index, R, t), landmarks, focal, fixture arrays)."""
    z = dict(np.load(os.path.join(GOLDEN, "help_pan.npz")))
    cams = [(i, z["rotation"][i], z["translation"][i]) for i in range(len(z["luma"]))]
    return list(z["luma"]), cams, z["landmarks"], float(z["focal"]), z


def accuracy(points, grid_sites, sampled_pixels, ref, z):
    """(median relative depth error, accurate coverage) of one reference's points."""
    ix = {int(v): k for k, v in enumerate(z["grid_x"])}
    iy = {int(v): k for k, v in enumerate(z["grid_y"])}
    R, t = z["rotation"][ref], z["translation"][ref]
    P = np.stack([points["x"], points["y"], points["z"]], 1).astype(np.float64)
    depth = (P @ R.T + t)[:, 2]
    truth = np.array([z["depth_grid"][ref][iy[int(s["y"])], ix[int(s["x"])]]
                      for s in grid_sites])
    rel = np.abs(depth - truth) / truth
    return float(np.median(rel)), float((rel <= 0.02).sum()) / max(1, int(sampled_pixels))


def patch_points(m, p):
    """A multi-reference result's points of one reference patch (primary, then completed)."""
    idx = (list(range(p["primary_start"], p["primary_start"] + p["primary_points"]))
           + list(range(p["completion_start"], p["completion_start"] + p["completed_points"])))
    return m.combined.points[idx], m.combined.grid_sites[idx]
