"""Multi-reference dense estimate on the device (dense.rs:161-330, 620-1082) against the FP64
oracle: the reference candidates, their attempts and accepted patches, the combined points,
the fine-grid decision, view limits, the legacy fallback, and cancellation at every kind of
poll. The batched path (no poll function: candidates solved together) and the sequential
path (with one) must give the same analysis."""
import numpy as np
import pytest

from vt3d_dense import scenes

from parity import compare_completed
from test_oracle_multi_reference import _flat240, _plane3

pytestmark = pytest.mark.gpu

# search depths are FP64 quantiles of the landmark depths cast to FP32 on both sides
DEPTH_TOL = 1e-6


def _compare_multi(g, o, what):
    gs, os_ = g.stats, o.stats
    assert g.fine_grid == o.fine_grid, what
    assert g.omitted_references == o.omitted_references, what
    for k in ("attempted", "reference_frame", "source_views", "source_frames", "sampled_pixels",
              "depth_hypotheses", "grid_stride", "grid_border", "skip_reason", "reference_frames",
              "reciprocal_checked_points", "reciprocal_rejected_points",
              "reciprocal_consistent_points", "fusion_input_observations",
              "fusion_rejected_observations", "fusion_rejected_points"):
        assert gs[k] == os_[k], (what, k, gs[k], os_[k])
    assert gs["reference_attempts"] == os_["reference_attempts"], what
    assert len(gs["reference_patches"]) == len(os_["reference_patches"]), what
    for gp, op in zip(gs["reference_patches"], os_["reference_patches"]):
        for k in gp:
            if k.startswith("search_"):
                assert (gp[k] is None) == (op[k] is None), (what, k)
                if gp[k] is not None:
                    assert abs(gp[k] - op[k]) <= DEPTH_TOL * abs(op[k]), (what, k)
            else:
                assert gp[k] == op[k], (what, k, gp[k], op[k])
    compare_completed(g.combined, o.combined, what)


def _same(a, b, what):
    """Bit-identical analyses (two GPU paths)."""
    assert a.stats == b.stats, what
    assert a.fine_grid == b.fine_grid and a.omitted_references == b.omitted_references, what
    for f in ("points", "grid_sites", "point_attributes"):
        assert getattr(a.combined, f).tobytes() == getattr(b.combined, f).tobytes(), (what, f)


def test_plane_three_references_match_oracle(product, oracle):
    frames, cams, sparse, f = _plane3()
    o = oracle.estimate(frames, cams, sparse, f)
    batched, omitted = product.estimate_depth_points_limited(frames, cams, sparse, f)
    assert omitted == o.omitted_references
    _compare_multi(batched, o, "plane, batched")
    assert len(o.stats["reference_patches"]) == 3
    polls = []
    seq = product.estimate_depth_points(frames, cams, sparse, f,
                                        control=lambda: polls.append(1) and False)
    _same(seq, batched, "plane, sequential vs batched")
    assert len(polls) > 0


@pytest.mark.parametrize("views", [1, 2])
def test_view_limit_keeps_the_leading_patches(product, oracle, views):
    frames, cams, sparse, f = _plane3()
    o = oracle.estimate(frames, cams, sparse, f, reference_views=views)
    g, omitted = product.estimate_depth_points_limited(frames, cams, sparse, f,
                                                       reference_views=views)
    assert omitted
    _compare_multi(g, o, f"plane, {views} views")


def test_two_cameras_omit_nothing(product, oracle):
    frames, cams, sparse, f = _plane3(centers=(0.0, 0.24))
    o = oracle.estimate(frames, cams, sparse, f, reference_views=2)
    g, omitted = product.estimate_depth_points_limited(frames, cams, sparse, f,
                                                       reference_views=2)
    assert not omitted
    _compare_multi(g, o, "plane, two cameras")


def test_fine_grid_decision_and_legacy_fallback(product, oracle):
    """dense.rs:1113-1157 (unlimited budget): distributed landmarks give stride 4, concentrated
    ones stride 5; the flat frames reject every candidate, so the legacy sweep runs."""
    frames, cams, distributed, concentrated, f = _flat240()
    for sparse, stride in ((distributed, 4), (concentrated, 5)):
        o = oracle.estimate(frames, cams, sparse, f)
        g = product.estimate_depth_points(frames, cams, sparse, f)
        _compare_multi(g, o, f"flat 240, stride {stride}")
        assert g.stats["grid_stride"] == stride and len(g.combined.points) == 0
        assert len(g.stats["reference_attempts"]) == 2


def test_c1_zncc_matches_oracle(product, oracle):
    """Five views on an arc, ZNCC 7x7 over 96 hypotheses, with surface completion."""
    c1 = scenes.config_c1()
    st = dict(c1.settings, surface_completion_passes=2)
    o = oracle.estimate(c1.frames, c1.cameras, c1.sparse, c1.focal, sweep=oracle.params(st),
                        threads=8)
    g, _ = product.estimate_depth_points_limited(c1.frames, c1.cameras, c1.sparse, c1.focal,
                                                 sweep=product.params(st))
    _compare_multi(g, o, "c1 zncc")
    assert len(g.stats["reference_patches"]) == 3


def test_cancellation_at_each_poll(product, oracle):
    """A cancel at poll n: in the candidate loop, in a reference remap, in a sweep row, in a
    completion pass, and in the legacy fallback's sweep. Same poll count and same analysis."""
    frames, cams, sparse, f = _plane3()
    total = []
    oracle.estimate(frames, cams, sparse, f, control=lambda: total.append(1) and False)
    n = len(total)
    assert n > 30
    for cancel_at in sorted({1, 2, 3, 4, 5, 8, 16, n // 3, n // 2, n - 2, n - 1, n}):
        got = []
        for side in (product, oracle):
            polls = []

            def poll(polls=polls):
                polls.append(1)
                return len(polls) >= cancel_at

            if side is product:
                r = product.estimate_depth_points(frames, cams, sparse, f, control=poll)
            else:
                r = oracle.estimate(frames, cams, sparse, f, control=poll)
            got.append((r, len(polls)))
        (g, gn), (o, on) = got
        assert gn == on, (cancel_at, gn, on)
        _compare_multi(g, o, f"cancel at poll {cancel_at}")


@pytest.mark.parametrize("case", ["no-frames", "one-camera", "few-landmarks", "small-frames",
                                  "zero-baseline", "textureless"])
def test_edge_cases_match_oracle(product, oracle, case):
    """Empty candidate orders and rejected candidates end in the legacy sweep's skips."""
    frames, cams, sparse, f = _plane3()
    if case == "no-frames":
        frames = []
    elif case == "one-camera":
        cams = cams[:1]
    elif case == "few-landmarks":
        sparse = sparse[:7]
    elif case == "small-frames":
        frames = [fr[:15, :] for fr in frames]
    elif case == "zero-baseline":
        cams = [scenes.plane_camera(0.0), scenes.plane_camera(0.0)]
    elif case == "textureless":
        frames = [np.full_like(fr, 120) for fr in frames]
    o = oracle.estimate(frames, cams, sparse, f)
    g = product.estimate_depth_points(frames, cams, sparse, f)
    _compare_multi(g, o, case)
    assert len(g.combined.points) == 0
