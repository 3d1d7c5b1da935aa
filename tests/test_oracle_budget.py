"""The working-set budget (crates/video-to-3d-core/src/dense.rs:354-605).

estimate_working_set charges Rust size_of values; without rustc the byte totals are stated for
the x86_64 layout and their exact values are parity unpinned. What is pinned: the library's
estimate equals the oracle's restatement on every input here, and both satisfy the
reference's own tests (dense.rs:1269-1490): the skip below the one-view estimate, the
identities of the estimate's parts, the retained-candidate and feature-option inequalities,
and the plan the budget picks (a tight budget keeps the leading patch of the full run bit for
bit; a two-view budget reports its limit only when it left a candidate out)."""
import itertools

import numpy as np
import pytest

from vt3d_dense import scenes


def _plane(centers):
    W, H, f, depth = 68, 48, 60.0, 4.0
    frames = [scenes.plane_frame(W, H, f, c, depth) for c in centers]
    cams = [scenes.plane_camera(c) for c in centers]
    return frames, cams, scenes.plane_sparse_points(W, H, f, depth), f


PARTS = ("frame_bytes", "retry_frame_bytes", "remapped_frame_bytes", "luminance_bytes",
         "sparse_pipeline_bytes", "dense_sample_bytes", "topology_bytes", "mesh_builder_bytes",
         "surface_fusion_bytes", "retained_candidate_bytes", "packed_output_bytes")


@pytest.mark.parametrize("sizes", [[(68, 48)] * 3, [(640, 360)] * 9, [(1920, 1080)] * 17,
                                   [(360, 640), (480, 270), (640, 360)], [(7, 5), (2, 2)], []])
def test_library_estimate_matches_the_oracle(product, oracle, sizes):
    for feats, radius, views, fine in itertools.product((0, 320, 1200), (0, 3, 7), (1, 2, 3),
                                                        (False, True)):
        a = product.estimate_working_set(sizes, feats, radius, views, fine)
        b = oracle.estimate_working_set(sizes, feats, radius, views, fine)
        assert a == b, (sizes, feats, radius, views, fine)


def test_estimate_identities(oracle):
    """budget_above_estimate_preserves_dense_output, dense.rs:1387-1427 (the estimate part)."""
    frames, _, _, _ = _plane((0.0, 0.24, -0.22))
    w = oracle.estimate_working_set(frames, 320, 3, 3, False)
    assert w["total_bytes"] == sum(w[k] for k in PARTS)
    assert w["retry_frame_bytes"] == w["frame_bytes"] == w["remapped_frame_bytes"]
    for k in ("retained_candidate_bytes", "mesh_builder_bytes", "surface_fusion_bytes",
              "sparse_pipeline_bytes"):
        assert w[k] > 0, k


def test_retained_candidate_is_charged_at_the_largest_plan(oracle):
    """dense.rs:1429-1450."""
    frames, _, _, _ = _plane((0.0, 0.24, -0.22))
    ordinary = oracle.estimate_working_set(frames, 320, 3, 3, False)
    retry = oracle.estimate_working_set(frames, 1200, 3, 1, False)
    assert retry["sparse_pipeline_bytes"] > ordinary["sparse_pipeline_bytes"]
    assert retry["dense_sample_bytes"] < ordinary["dense_sample_bytes"]
    assert retry["retained_candidate_bytes"] == ordinary["retained_candidate_bytes"]
    assert ordinary["retained_candidate_bytes"] >= ordinary["topology_bytes"]


def test_feature_options_increase_the_sparse_estimate(oracle):
    """dense.rs:1474-1486."""
    frames, _, _, _ = _plane((0.0, 0.24, -0.22))
    ordinary = oracle.estimate_working_set(frames, 320, 3, 3, False)
    heavy = oracle.estimate_working_set(frames, 640, 7, 3, False)
    assert heavy["sparse_pipeline_bytes"] > ordinary["sparse_pipeline_bytes"]
    assert heavy["total_bytes"] > ordinary["total_bytes"]


def test_budget_below_estimate_suppresses_dense_geometry(oracle):
    """dense.rs:1269-1295."""
    frames, cams, sparse, f = _plane((0.0, 0.24))
    m = oracle.estimate_with_budget(frames, cams, sparse, f, 1, 320, 3)
    s = m.stats
    assert not s["attempted"] and len(m.combined.points) == 0
    assert len(m.combined.grid_sites) == 0
    assert s["working_set_estimate"]["total_bytes"] > 1
    assert s["working_set_budget_bytes"] == 1
    assert "exceeds the configured" in s["skip_reason"]
    one = oracle.estimate_working_set(frames, 320, 3, 1, False)["total_bytes"]
    assert s["skip_reason"] == (f"estimated dense working set of {one} bytes exceeds the "
                                "configured 1 byte budget even for one reference view")


def test_tight_budget_keeps_the_leading_patch_of_the_full_run(oracle):
    """dense.rs:1297-1362."""
    frames, cams, sparse, f = _plane((0.0, 0.24, -0.22))
    full = oracle.estimate(frames, cams, sparse, f)
    assert len(full.stats["reference_patches"]) >= 2
    one = oracle.estimate_working_set(frames, 320, 3, 1, False)["total_bytes"]
    alls = oracle.estimate_working_set(frames, 320, 3, 3, False)["total_bytes"]
    assert one < alls
    lim = oracle.estimate_with_budget(frames, cams, sparse, f, one, 320, 3)
    s = lim.stats
    assert s["attempted"] and s["reference_view_limit"] == 1
    assert s["working_set_estimate"]["total_bytes"] == one
    assert s["full_working_set_bytes"] == alls
    assert len(s["reference_patches"]) == 1
    kept, lead = s["reference_patches"][0], full.stats["reference_patches"][0]
    assert kept["reference_frame"] == lead["reference_frame"]
    fp = full.combined.points
    expected = np.concatenate([fp[lead["primary_start"]:][:lead["primary_points"]],
                               fp[lead["completion_start"]:][:lead["completed_points"]]])
    assert lim.combined.points.tobytes() == expected.tobytes()
    two = oracle.estimate_working_set(frames, 320, 3, 2, False)["total_bytes"]
    wider = oracle.estimate_with_budget(frames, cams, sparse, f, two, 320, 3)
    assert wider.stats["reference_view_limit"] == 2
    assert len(wider.stats["reference_patches"]) == min(2, len(full.stats["reference_patches"]))
    unlimited = oracle.estimate_with_budget(frames, cams, sparse, f, alls, 320, 3)
    assert unlimited.stats["reference_view_limit"] is None


def test_generous_budget_preserves_dense_output(oracle):
    """dense.rs:1364-1427."""
    frames, cams, sparse, f = _plane((0.0, 0.24, -0.22))
    base = oracle.estimate(frames, cams, sparse, f)
    b = oracle.estimate_with_budget(frames, cams, sparse, f, 2 ** 64 - 1, 320, 3)
    for k in ("points", "grid_sites", "point_attributes"):
        assert getattr(b.combined, k).tobytes() == getattr(base.combined, k).tobytes(), k


def test_limit_that_omits_no_reference_is_not_reported(oracle):
    """dense.rs:1452-1472."""
    frames, cams, sparse, f = _plane((0.0, 0.24))
    two = oracle.estimate_working_set(frames, 320, 3, 2, False)["total_bytes"]
    lim = oracle.estimate_with_budget(frames, cams, sparse, f, two, 320, 3)
    assert lim.stats["full_working_set_bytes"] > two
    assert lim.stats["reference_view_limit"] is None
