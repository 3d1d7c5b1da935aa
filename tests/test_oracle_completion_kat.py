"""verified_surface_completion_adds_expected_plane_geometry
(crates/video-to-3d-core/src/surface_completion_tests.rs:54-110) on the oracle: the
three-camera textured plane at centres 0, 0.25 and -0.22, through estimate_depth_points."""
import numpy as np

from vt3d_dense import scenes

COUNTERS = ("surface_completed_points", "surface_completion_rejected_texture",
            "surface_completion_rejected_cross_view", "surface_completion_rejected_reciprocal",
            "surface_completion_rejected_fusion", "surface_completion_rejected_footprint")


def completion_fixture():
    W, H, f, depth = 68, 48, 60.0, 4.0
    centers = (0.0, 0.25, -0.22)
    frames = [scenes.plane_frame(W, H, f, c, depth) for c in centers]
    cams = [scenes.plane_camera(c) for c in centers]
    return frames, cams, scenes.plane_sparse_points(W, H, f, depth), f, depth


def check_completion_kat(analysis, depth):
    """The assertions of surface_completion_tests.rs:73-109."""
    s = analysis.stats
    assert s["attempted"]
    assert s["surface_completion_proposals"] > 0
    assert s["surface_completed_points"] > 0
    assert s["surface_completion_proposals"] == sum(s[k] for k in COUNTERS)
    pts, sites = analysis.points, analysis.grid_sites
    primary = len(pts) - s["surface_completed_points"]
    assert primary >= 0
    done, done_sites = pts[primary:], sites[primary:]
    assert len(done) == len(done_sites) == s["surface_completed_points"]
    assert np.all(np.abs(done["z"].astype(np.float64) - depth) < 0.55), done["z"]
    assert np.all((done["confidence"] >= 0.05) & (done["confidence"] <= 0.9)), done["confidence"]
    return s["surface_completed_points"]


def test_surface_completion_adds_expected_plane_geometry(oracle):
    frames, cams, sparse, f, depth = completion_fixture()
    m = oracle.estimate(frames, cams, sparse, f)
    assert check_completion_kat(m.combined, depth) > 0
    # a wrong completion depth (planted in a copy of the result) fails the KAT's depth test
    bad = m.combined
    bad.points = bad.points.copy()
    bad.points["z"][-1] += 0.6
    try:
        check_completion_kat(bad, depth)
    except AssertionError:
        pass
    else:
        raise AssertionError("the depth bound did not catch a 0.6 m error")
