"""Known-answer tests of the oracle's mesh consumer: the reference's own tests of
mesh_legacy.rs:560-909 and mesh.rs:278-381, plus the wrapper's skip paths and the poll
counts. No GPU."""
import numpy as np
import pytest

from vt3d_dense import abi

W, H, F = 68, 48, 60.0
IDENTITY = (0, np.eye(3), np.zeros(3))


def point_at_pixel(x, y, depth=4.0, conf=0.8):
    """point_at_pixel, mesh_legacy.rs:570-588 (the f64 values cast to f32)."""
    p = np.zeros(1, abi.POINT_DTYPE)
    p["x"] = np.float32((x - W * 0.5) / F * depth)
    p["y"] = np.float32((y - H * 0.5) / F * depth)
    p["z"] = np.float32(depth)
    p["confidence"] = np.float32(conf)
    p["r"], p["g"], p["b"] = 120, 140, 160
    return p[0]


def sites(*xy):
    return np.array(list(xy), abi.SITE_DTYPE)


QUAD = ((3, 3), (7, 3), (3, 7), (7, 7))
DENSE = {"attempted": True, "reference_frame": 0, "grid_stride": 4, "grid_border": 3}


def run(oracle, pts, st, dense=DENSE, patches=None, control=None, cams=(IDENTITY,)):
    return oracle.mesh(np.array(pts, abi.POINT_DTYPE), st, dense, list(cams), W, H, F,
                       patches=patches, control=control)


def sorted_tris(m):
    return {tuple(sorted((int(t["a"]), int(t["b"]), int(t["c"])))) for t in m.triangles}


def test_preserves_admission_stats_when_too_few_grid_vertices_survive(oracle):
    pts = [point_at_pixel(3, 3, conf=0.8), point_at_pixel(7, 3, conf=0.7),
           point_at_pixel(6, 7, conf=0.9)]
    m = run(oracle, pts, sites((3, 3), (7, 3), (3, 7)))
    s = m.stats
    assert s["attempted"] and s["skip_reason"] is not None
    assert s["grid_vertices"] == 2 and s["rejected_grid_vertices"] == 1
    assert len(m.triangles) == 0


def test_triangulates_a_smooth_reference_grid_quad(oracle):
    pts = [point_at_pixel(3, 3, conf=0.8), point_at_pixel(7, 3, conf=0.7),
           point_at_pixel(3, 7, conf=0.9), point_at_pixel(7, 7, conf=0.6)]
    m = run(oracle, pts, sites(*QUAD))
    s = m.stats
    assert s["attempted"] and s["grid_vertices"] == 4 and s["candidate_cells"] == 1
    assert s["candidate_triangles"] == 2 and s["accepted_triangles"] == 2
    assert s["rejected_discontinuities"] == 0 and s["rejected_degenerate"] == 0
    assert len(m.triangles) == 2
    assert all(max(t["a"], t["b"], t["c"]) < 4 for t in m.triangles)
    assert all(t["confidence"] <= np.float32(0.8) for t in m.triangles)
    # by construction: tl-br and tr-bl diagonals are equal, so <= takes [tl, tr, br], [tl, br,
    # bl], each with the minimum confidence of its corners
    assert [(int(t["a"]), int(t["b"]), int(t["c"])) for t in m.triangles] == [(0, 1, 3), (0, 3, 2)]
    assert m.triangles["confidence"].tolist() == [np.float32(0.6)] * 2


def test_bridges_one_missing_grid_sample_on_a_smooth_surface(oracle):
    st = sites((3, 3), (11, 3), (3, 7), (11, 7), (7, 7))
    pts = [point_at_pixel(x, y) for x, y in st.tolist()]
    m = run(oracle, pts, st)
    assert m.stats["attempted"] and m.stats["grid_vertices"] == 5
    assert m.stats["accepted_triangles"] == 3
    assert (0, 1, 4) in sorted_tris(m)


def test_does_not_bridge_a_missing_sample_across_a_depth_discontinuity(oracle):
    st = sites((3, 3), (11, 3), (3, 7), (11, 7), (7, 7))
    pts = [point_at_pixel(3, 3, 4.0), point_at_pixel(11, 3, 5.0), point_at_pixel(3, 7, 4.0),
           point_at_pixel(11, 7, 5.0), point_at_pixel(7, 7, 4.0)]
    m = run(oracle, pts, st)
    assert m.stats["attempted"] and m.stats["rejected_discontinuities"] > 0
    assert (0, 1, 4) not in sorted_tris(m)


def test_rejects_triangles_that_bridge_a_depth_discontinuity(oracle):
    pts = [point_at_pixel(3, 3, 4.0), point_at_pixel(7, 3, 4.0), point_at_pixel(3, 7, 4.0),
           point_at_pixel(7, 7, 7.0)]
    s = run(oracle, pts, sites(*QUAD)).stats
    assert s["attempted"] and s["candidate_triangles"] == 2
    assert 0 < s["accepted_triangles"] < s["candidate_triangles"]
    assert s["rejected_discontinuities"] > 0


def test_preserves_original_grid_site_after_fusion_shifts_reprojection(oracle):
    pts = [point_at_pixel(4.8, 3), point_at_pixel(7, 3), point_at_pixel(3, 7)]
    s = run(oracle, pts, sites((3, 3), (7, 3), (3, 7))).stats
    assert s["attempted"] and s["grid_vertices"] == 3 and s["accepted_triangles"] == 1


def test_rejects_fused_vertex_that_leaves_original_grid_footprint(oracle):
    pts = [point_at_pixel(5.4, 3), point_at_pixel(7, 3), point_at_pixel(3, 7),
           point_at_pixel(7, 7)]
    m = run(oracle, pts, sites(*QUAD))
    assert m.stats["attempted"] and m.stats["grid_vertices"] == 3
    assert m.stats["accepted_triangles"] == 1
    assert all(0 not in (t["a"], t["b"], t["c"]) for t in m.triangles)


def test_rejects_triangle_that_flips_reference_projection_winding(oracle):
    pts = [point_at_pixel(4.9, 1.1), point_at_pixel(5.1, 4.9), point_at_pixel(8.9, 5.1)]
    s = run(oracle, pts, sites((3, 3), (7, 3), (7, 7))).stats
    assert s["attempted"] and s["grid_vertices"] == 3 and s["candidate_triangles"] == 1
    assert s["accepted_triangles"] == 0 and s["rejected_degenerate"] == 1


def test_does_not_attempt_mesh_without_enough_dense_points(oracle):
    pts = [point_at_pixel(3, 3), point_at_pixel(7, 3)]
    m = run(oracle, pts, sites(*QUAD))
    assert not m.stats["attempted"] and len(m.triangles) == 0
    assert "three accepted fused dense points" in m.stats["skip_reason"]


def test_combines_reference_meshes_using_global_dense_point_indices(oracle):
    """mesh.rs:310-380: two patches of the same quad, indices mapped to the global points."""
    pts = [point_at_pixel(x, y) for x, y in QUAD] * 2
    st = sites(*(QUAD * 2))
    base = {"source_frames": [], "completed_points": 0, "sampled_pixels": 0,
            "reciprocal_consistent_points": 0, "grid_stride": 4, "grid_border": 3,
            "accepted_points": 4, "search_min_depth": None}
    patches = [dict(base, reference_frame=0, primary_start=0, primary_points=4,
                    completion_start=4),
               dict(base, reference_frame=1, primary_start=4, primary_points=4,
                    completion_start=8)]
    dense = dict(DENSE, reference_patches=patches)
    m = run(oracle, pts, st, dense=dense, patches=patches,
            cams=(IDENTITY, (1, np.eye(3), np.zeros(3))))
    s = m.stats
    assert s["attempted"] and s["reference_frames"] == [0, 1]
    assert len(s["reference_patches"]) == 2
    assert s["accepted_triangles"] == 4 and len(m.triangles) == 4
    assert any(min(t["a"], t["b"], t["c"]) >= 4 for t in m.triangles)
    # the second patch's triangles are the first's, shifted by 4
    t = m.triangles
    assert (t["a"][2:] - t["a"][:2]).tolist() == [4, 4]


@pytest.mark.parametrize("case,reason", [
    ("not-attempted", "dense reconstruction was not attempted"),
    ("stride-0", "dense reconstruction did not expose a valid reference-grid stride"),
    ("no-reference", "dense reconstruction did not select a reference frame"),
    ("no-camera", "the dense reference frame has no accepted registered camera"),
    ("counts", "dense point and reference-grid evidence counts disagree"),
])
def test_skip_reasons(oracle, case, reason):
    pts = [point_at_pixel(x, y) for x, y in QUAD]
    st, dense, cams = sites(*QUAD), dict(DENSE), (IDENTITY,)
    if case == "not-attempted":
        dense["attempted"] = False
    elif case == "stride-0":
        dense["grid_stride"] = 0
    elif case == "no-reference":
        dense["reference_frame"] = None
    elif case == "no-camera":
        cams = ((5, np.eye(3), np.zeros(3)),)
    elif case == "counts":
        st = st[:3]
    m = run(oracle, pts, st, dense=dense, cams=cams)
    assert m.stats["skip_reason"] == reason and len(m.triangles) == 0
    assert m.stats["attempted"] is False


def test_bad_patch_ranges_are_reported(oracle):
    pts = [point_at_pixel(x, y) for x, y in QUAD]
    p = {"reference_frame": 0, "primary_start": 2, "primary_points": 4, "completion_start": 6,
         "completed_points": 0, "accepted_points": 4, "source_frames": [], "sampled_pixels": 0,
         "reciprocal_consistent_points": 0, "grid_stride": 4, "grid_border": 3,
         "search_min_depth": None}
    m = run(oracle, pts, sites(*QUAD), patches=[p])
    ps = m.stats["reference_patches"][0]
    assert ps["skip_reason"] == "dense patch ranges exceed the accepted point evidence"
    assert not ps["attempted"] and ps["point_count"] == 4
    assert m.stats["skip_reason"] == "no reference patch produced an accepted surface triangle"


def test_polls_and_cancellation(oracle):
    """One poll per sample, one per cell origin, one per bridge column (mesh_legacy.rs:221,
    285, 310, 331); a cancel anywhere gives the canceled skip."""
    st = sites((3, 3), (11, 3), (3, 7), (11, 7), (7, 7))
    pts = [point_at_pixel(x, y) for x, y in st.tolist()]
    polls = []
    run(oracle, pts, st, control=lambda: polls.append(1) and False)
    # 5 samples; grid x in {0, 1, 2}, y in {0, 1}: cell origins are the 2x2 neighbourhoods of
    # the 5 sites (x -1..2, y -1..1 minus (-1..0, ...) gaps: 11 distinct), 1 horizontal
    # bridge column (x range 2), no vertical (y range 1)
    origins = {(gx + dx, gy + dy) for gx, gy in ((0, 0), (2, 0), (0, 1), (2, 1), (1, 1))
               for dx in (-1, 0) for dy in (-1, 0)}
    assert len(polls) == 5 + len(origins) + 1
    for at in (1, 5, 6, len(polls)):
        n = []
        m = run(oracle, pts, st, control=lambda: n.append(1) or len(n) >= at)
        assert len(n) == at
        assert m.stats["skip_reason"] == "mesh reconstruction canceled"
        assert not m.stats["attempted"] and len(m.triangles) == 0
