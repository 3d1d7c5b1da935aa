"""Known-answer tests of the oracle's multi-reference estimate (dense.rs:161-330, 620-1082).

The cases restate the reference's own tests in dense.rs (multi_reference_tests,
dense.rs:1084-1562) where they concern this path, plus orderings fixed by construction: the
candidate order of a fixture whose landmark visibility is chosen per camera, and the
reference's statement that a limited run keeps the leading patches of the unlimited one
(dense.rs:244-250), checked bit for bit. No GPU.
"""
import ctypes as C

import numpy as np
import pytest

from vt3d_dense import abi, scenes


def _plane3(centers=(0.0, 0.24, -0.22)):
    W, H, f, depth = 68, 48, 60.0, 4.0
    frames = [scenes.plane_frame(W, H, f, c, depth) for c in centers]
    cams = [scenes.plane_camera(c) for c in centers]
    return frames, cams, scenes.plane_sparse_points(W, H, f, depth), f


def _flat240():
    W = H = 240
    f = 200.0
    flat = np.full((H, W), 128, np.uint8)
    world = lambda x, y: ((x - 120.0) / f * 4.0, (y - 120.0) / f * 4.0, 4.0)  # noqa: E731
    distributed = []
    for y in (30.0, 90.0, 150.0, 210.0):
        for x in (30.0, 90.0, 150.0, 210.0):
            distributed.append(world(x, y))
            distributed.append(world(x + 2.0, y))
    concentrated = [world(90.0 + i, 90.0) for i in range(len(distributed))]
    cams = [scenes.plane_camera(0.0), scenes.plane_camera(0.2)]
    return [flat, flat], cams, np.array(distributed), np.array(concentrated), f


# ---- has_distributed_landmarks ---------------------------------------------------------------

def test_clustered_landmark_scan_can_be_canceled_within_one_camera(oracle):
    """dense.rs:1089-1111: 4096 copies of one landmark, cancel on the third poll."""
    frame = np.zeros((240, 240), np.uint8)
    points = np.tile(np.array([[0.0, 0.0, 4.0]]), (4096, 1))
    polls = []

    def cancel():
        polls.append(1)
        return len(polls) == 3

    assert not oracle.distributed_landmarks([frame], [scenes.plane_camera(0.0)], points, 200.0,
                                            control=cancel)
    assert len(polls) == 3
    # without a cancel the scan polls once per chunk of 256 landmarks: 16 chunks
    polls.clear()
    assert not oracle.distributed_landmarks([frame], [scenes.plane_camera(0.0)], points, 200.0,
                                            control=lambda: polls.append(1) and False)
    assert len(polls) == 16


def test_distributed_and_concentrated_landmarks(oracle):
    frames, cams, distributed, concentrated, f = _flat240()
    assert oracle.distributed_landmarks(frames, cams, distributed, f)
    assert not oracle.distributed_landmarks(frames, cams, concentrated, f)
    # a majority of views must be covered: with a third camera that sees none of them,
    # two of three still suffice, one of two does not
    far = (0, np.eye(3), np.array([0.0, 0.0, -100.0]))  # landmarks behind it
    assert oracle.distributed_landmarks(frames, list(cams) + [(1, far[1], far[2])], distributed, f)
    assert not oracle.distributed_landmarks(frames, [cams[0], (1, far[1], far[2])], distributed, f)


def test_finer_dense_grid_requires_distributed_landmarks(oracle):
    """dense.rs:1113-1157 with an unlimited budget: stride 4 for distributed landmarks, 5 for
    concentrated ones, and no points on flat frames either way."""
    frames, cams, distributed, concentrated, f = _flat240()
    for sparse, stride, fine in ((distributed, 4, True), (concentrated, 5, False)):
        m = oracle.estimate(frames, cams, sparse, f)
        assert m.fine_grid == fine
        assert m.stats["grid_stride"] == stride
        assert len(m.combined.points) == 0


def test_fine_grid_is_not_checked_where_it_is_not_finer(oracle):
    """dense.rs:366-372: at 68x48 both spacings give stride 4, so the landmark scan (and its
    polling) is skipped and the coarse grid runs."""
    frames, cams, sparse, f = _plane3()
    lib = oracle.library()
    assert lib.vt3d_oracle_grid_stride(68, 48, 1) == lib.vt3d_oracle_grid_stride(68, 48, 0)
    polls = []
    m = oracle.estimate(frames, cams, sparse, f, control=lambda: polls.append(1) and False)
    assert not m.fine_grid
    # the first poll is the candidate loop's, not the scan's: cancel there gives the default
    first = []
    m1 = oracle.estimate(frames, cams, sparse, f, control=lambda: first.append(1) or True)
    assert len(first) == 1 and not m1.stats["attempted"] and len(m1.combined.points) == 0


# ---- reference_candidate_order -------------------------------------------------------------

def _pixel_landmarks(xs, ys, W=68, H=48, f=60.0, depth=4.0):
    return np.array([((x - W * 0.5) / f * depth, (y - H * 0.5) / f * depth, depth)
                     for y in ys for x in xs])


def test_candidate_order_ties_and_distance(oracle):
    """All three see the 20 landmarks: frame 0 leads (lowest index), the rest farthest first."""
    frames, cams, sparse, f = _plane3()
    assert oracle.reference_order(frames, cams, sparse, f) == [0, 1, 2]
    # permuted frame indices: the tie goes to the lowest frame index, not camera index
    perm = [(2, np.eye(3), np.array([0.0, 0.0, 0.0])), (0, np.eye(3), np.array([-0.24, 0, 0])),
            (1, np.eye(3), np.array([0.22, 0.0, 0.0]))]
    order = oracle.reference_order(frames, perm, sparse, f)
    # camera 1 (frame 0, x = 0.24) leads; from it camera 2 (x = -0.22) is farther than camera 0
    assert order == [1, 2, 0]


def test_candidate_order_follows_visibility(oracle):
    """Six landmarks at pixel x = -2 of camera 0 fall inside camera 2 only (it sees them
    3.3 px further right), so camera 2 sees 26 and leads; from x = -0.22 camera 1 (0.46 away)
    is farther than camera 0 (0.22 away)."""
    frames, cams, sparse, f = _plane3()
    extra = _pixel_landmarks([-2.0], [8, 14, 20, 26, 32, 38])
    assert oracle.reference_order(frames, cams, np.vstack([sparse, extra]), f) == [2, 1, 0]
    # fewer than 8 visible landmarks drops a camera from the candidates
    few = sparse[:7]
    assert oracle.reference_order(frames, cams, few, f) == []


# ---- attempt statistics and the fallback merge ----------------------------------------------

def test_rejected_reference_attempt_keeps_its_diagnostics(oracle):
    """dense.rs:1489-1520."""
    r = abi.DenseResult()
    r.stats.attempted = 1
    r.stats.sampled_pixels = 41
    r.stats.reciprocal_checked_points = 2
    r.stats.reciprocal_rejected_points = 1
    r.stats.reciprocal_consistent_points = 1
    r.stats.surface_completion_proposals = 3
    r.stats.surface_completed_points = 1
    r.stats.surface_completion_rejected_cross_view = 2
    r.n_points = 2
    a = abi.ReferenceAttempt()
    oracle.library().vt3d_oracle_reference_attempt(7, C.byref(r), 0, C.byref(a))
    d = abi._attempt(a)
    assert d["attempted"] and not d["accepted"] and d["reference_frame"] == 7
    assert (d["sampled_pixels"], d["accepted_points"], d["reciprocal_checked_points"],
            d["reciprocal_rejected_points"], d["reciprocal_consistent_points"],
            d["surface_completion_proposals"], d["surface_completed_points"],
            d["surface_completion_rejected_cross_view"]) == (41, 2, 2, 1, 1, 3, 1, 2)
    assert "at least 3" in d["skip_reason"]
    # not attempted, and a legacy skip reason taking precedence
    r.stats.attempted = 0
    oracle.library().vt3d_oracle_reference_attempt(7, C.byref(r), 0, C.byref(a))
    assert a.skip_reason.decode() == "dense estimation did not run for this reference"
    r.skip_code = 9
    r.skip_reason = b"dense reconstruction canceled"
    oracle.library().vt3d_oracle_reference_attempt(7, C.byref(r), 0, C.byref(a))
    assert a.skip_reason.decode() == "dense reconstruction canceled"


def test_successful_fallback_evidence_replaces_rejected_attempt_for_same_reference(oracle):
    """dense.rs:1522-1561."""
    def att(frame, accepted, points, reason=b""):
        a = abi.ReferenceAttempt()
        a.reference_frame, a.attempted, a.accepted = frame, 1, int(accepted)
        a.accepted_points, a.skip_reason = points, reason
        return a

    tried = (abi.ReferenceAttempt * 2)(att(3, False, 1, b"remapped reference rejected"),
                                       att(5, False, 0, b"another reference rejected"))
    fb = (abi.ReferenceAttempt * 1)(att(3, True, 9))
    out = (abi.ReferenceAttempt * 3)()
    n = oracle.library().vt3d_oracle_merge_attempts(tried, 2, fb, 1, out)
    assert n == 2
    merged = [abi._attempt(out[i]) for i in range(n)]
    three = [m for m in merged if m["reference_frame"] == 3][0]
    assert three["accepted"] and three["accepted_points"] == 9 and three["skip_reason"] is None
    assert any(m["reference_frame"] == 5 and not m["accepted"] for m in merged)
    # a rejected fallback never replaces, a new frame is appended
    fb2 = (abi.ReferenceAttempt * 2)(att(5, False, 4, b"x"), att(8, False, 0, b"y"))
    n = oracle.library().vt3d_oracle_merge_attempts(tried, 2, fb2, 2, out)
    assert n == 3 and out[1].accepted_points == 0 and out[2].reference_frame == 8


# ---- estimate_depth_points_limited end to end ----------------------------------------------

def _surface_invariants(m):
    s = m.stats
    pts = m.combined
    assert s["attempted"]
    assert len(s["reference_frames"]) == len(s["reference_patches"])
    assert len(s["reference_patches"]) == sum(a["accepted"] for a in s["reference_attempts"])
    assert all(a["accepted"] or a["skip_reason"] is not None for a in s["reference_attempts"])
    assert len(pts.points) == len(pts.grid_sites) == len(pts.point_attributes)
    attrs = pts.point_attributes
    assert np.all((attrs["reciprocal_consistency"] >= 0) & (attrs["reciprocal_consistency"] <= 1))
    assert np.all((attrs["depth_margin"] >= 0) & (attrs["depth_margin"] <= 1))
    assert len(pts.points) == s["accepted_points"]
    assert s["surface_completion_proposals"] == (
        s["surface_completed_points"] + s["surface_completion_rejected_texture"]
        + s["surface_completion_rejected_cross_view"] + s["surface_completion_rejected_reciprocal"]
        + s["surface_completion_rejected_fusion"] + s["surface_completion_rejected_footprint"])
    assert len(s["source_frames"]) == s["source_views"]
    assert s["reference_frame"] not in s["source_frames"]


def test_reconstructs_multiple_reference_patches_without_relaxing_dense_gates(oracle):
    """dense.rs:1214-1268."""
    frames, cams, sparse, f = _plane3()
    m = oracle.estimate(frames, cams, sparse, f)
    _surface_invariants(m)
    s = m.stats
    assert len(s["reference_frames"]) >= 2
    # the layout of combine_patches: every patch's primary points, then every completion
    patches = s["reference_patches"]
    at = 0
    for p in patches:
        assert p["primary_start"] == at
        at += p["primary_points"]
    assert m.combined.primary_points == at
    for p in patches:
        assert p["completion_start"] == at
        at += p["completed_points"]
    assert at == len(m.combined.points)
    assert s["surface_completed_points"] == sum(p["completed_points"] for p in patches)
    # the participating frames other than the primary, ascending
    part = set()
    for p in patches:
        part.add(p["reference_frame"])
        part.update(p["source_frames"])
    part.discard(s["reference_frame"])
    assert s["source_frames"] == sorted(part)
    # medians and search depths only for a single patch
    if len(patches) > 1:
        assert s["median_photometric_error"] is None and s["search_min_depth"] is None


def test_a_limit_keeps_the_leading_reference_patches_of_the_full_run(oracle):
    """dense.rs:1293-1360 without the budget: reference_views 1 and 2 keep the leading
    patches of the full run, point for point."""
    frames, cams, sparse, f = _plane3()
    full = oracle.estimate(frames, cams, sparse, f)
    fp = full.stats["reference_patches"]
    assert len(fp) >= 2
    one = oracle.estimate(frames, cams, sparse, f, reference_views=1)
    assert one.omitted_references
    lead = fp[0]
    kept = one.stats["reference_patches"]
    assert len(kept) == 1 and kept[0]["reference_frame"] == lead["reference_frame"]
    pts = full.combined.points
    expected = np.concatenate([
        pts[lead["primary_start"]:lead["primary_start"] + lead["primary_points"]],
        pts[lead["completion_start"]:lead["completion_start"] + lead["completed_points"]]])
    assert one.combined.points.tobytes() == expected.tobytes()
    two = oracle.estimate(frames, cams, sparse, f, reference_views=2)
    assert len(two.stats["reference_patches"]) == min(2, len(fp))
    assert [p["reference_frame"] for p in two.stats["reference_patches"]] == \
        [p["reference_frame"] for p in fp[:2]]


def test_a_limit_that_omits_no_reference_is_not_reported(oracle):
    """dense.rs:1445-1470: two cameras, two views allowed, nothing left out."""
    frames, cams, sparse, f = _plane3(centers=(0.0, 0.24))
    m = oracle.estimate(frames, cams, sparse, f, reference_views=2)
    assert not m.omitted_references


def test_fallback_runs_the_legacy_sweep_when_no_patch_is_accepted(oracle):
    """Flat frames: every candidate is attempted and rejected (no points), the legacy sweep
    runs, and its attempt (same reference frame, also rejected) merges into the first."""
    frames, cams, distributed, _, f = _flat240()
    m = oracle.estimate(frames, cams, distributed, f, fine_grid=False)
    s = m.stats
    assert s["reference_frames"] == [0] and len(s["reference_patches"]) == 1
    assert [a["reference_frame"] for a in s["reference_attempts"]] == [0, 1]
    assert all(not a["accepted"] and "at least 3" in a["skip_reason"]
               for a in s["reference_attempts"])


def test_invalid_reference_views_are_rejected(oracle):
    frames, cams, sparse, f = _plane3()
    for views in (0, 4):
        with pytest.raises(ValueError):
            oracle.estimate(frames, cams, sparse, f, reference_views=views)
