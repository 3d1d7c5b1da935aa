"""GPU path (through the C ABI) against the FP64 oracle on identical seeded inputs."""
import copy
import os

import numpy as np
import pytest

import parity
from vt3d_dense import abi, scenes

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _gpu(product, problem, settings, **over):
    sw = product.params(settings, **over)
    return product.reconstruct_reference_patch(problem.frames, problem.cameras, problem.sparse,
                                               problem.focal, problem.reference, sweep=sw)


def _oracle(oracle, problem, settings, **over):
    sw = oracle.params(settings, **over)
    return oracle.reference_patch(problem.frames, problem.cameras, problem.sparse, problem.focal,
                                  problem.reference, sweep=sw, threads=THREADS)


def _check_counters(s):
    assert s["reciprocal_checked_points"] == (s["reciprocal_consistent_points"]
                                              + s["reciprocal_rejected_points"])
    assert s["reciprocal_consistent_points"] == s["accepted_points"] + s["fusion_rejected_points"]
    assert s["fusion_input_observations"] >= 2 * s["reciprocal_consistent_points"]


def test_plane_fixture_matches_oracle(product, oracle):
    pf = scenes.plane_fixture()
    st = {"preset": "reference", "surface_completion_passes": 0, "emit_samples": 1}
    g = product.estimate_depth_points_with_control(pf.frames, pf.cameras, pf.sparse, pf.focal,
                                                   sweep=product.params(st))
    o, sl = oracle.plane_sweep(pf.frames, pf.cameras, pf.sparse, pf.focal,
                               sweep=oracle.params(st))
    parity.stats_equal(g.stats, o.stats)
    rep = parity.compare(g, o, sl, abi.COST_ZSAD)
    print("plane fixture", rep)
    assert g.stats["source_frames"] == [1, 2] and len(g.points) >= 12
    z = np.sort(g.points["z"])
    assert abs(float(z[len(z) // 2]) - 4.0) < 0.55
    _check_counters(g.stats)
    for k in ("accepted_points", "reciprocal_checked_points", "reciprocal_consistent_points",
              "fusion_input_observations", "fusion_rejected_observations"):
        assert g.stats[k] == o.stats[k], k


@pytest.mark.parametrize("kernel", ["direct", "tiled"])
@pytest.mark.parametrize("preset", ["reference", "zncc"])
def test_c1_matches_oracle(product, oracle, preset, kernel):
    c1 = scenes.config_c1()
    st = dict(c1.settings) if preset == "zncc" else {"preset": "reference"}
    st.update(surface_completion_passes=0, emit_samples=1)
    g = _gpu(product, c1, st, kernel=kernel)
    o, sl = _oracle(oracle, c1, st)
    parity.stats_equal(g.stats, o.stats)
    rep = parity.compare(g, o, sl, st.get("preset") == "zncc" and 1 or 0)
    print("c1", preset, kernel, rep)
    _check_counters(g.stats)


@pytest.mark.parametrize("kernel", ["direct", "tiled"])
@pytest.mark.parametrize("kind,radius,hyps,sources", [
    ("zncc", 2, 64, 4), ("zncc", 4, 48, 4), ("zncc", 5, 32, 4), ("zncc", 1, 96, 4),
    ("zsad", 2, 40, 4), ("zsad", 3, 32, 4)])
def test_patch_sizes_match_oracle(product, oracle, kind, radius, hyps, sources, kernel):
    c1 = scenes.config_c1()
    st = dict(c1.settings, patch_radius=radius, depth_hypotheses=hyps, max_source_views=sources,
              surface_completion_passes=0, emit_samples=1)
    if kind == "zsad":
        st = {"preset": "reference", "patch_radius": radius, "depth_hypotheses": hyps,
              "max_source_views": sources, "grid_stride": 4, "surface_completion_passes": 0,
              "emit_samples": 1}
    g = _gpu(product, c1, st, kernel=kernel)
    o, sl = _oracle(oracle, c1, st)
    parity.stats_equal(g.stats, o.stats)
    rep = parity.compare(g, o, sl, 1 if kind == "zncc" else 0)
    print(kind, radius, kernel, rep)


@pytest.mark.parametrize("kernel", ["direct", "tiled"])
def test_c2_full_size_band_matches_oracle(product, oracle, kernel):
    """640x360, 8 sources, 128 hypotheses: the full GPU solve against the oracle on a band of
    grid rows (the oracle would take minutes for every row)."""
    c2 = scenes.config_c2()
    st = dict(c2.settings, surface_completion_passes=0, emit_samples=1)
    g = _gpu(product, c2, st, kernel=kernel)
    oracle.set_row_band(80, 3)
    try:
        o, sl = _oracle(oracle, c2, st)
    finally:
        oracle.set_row_band(0, 0)
    ys = np.unique(o.samples["y"])
    assert len(ys) == 3
    gb = copy.copy(g)
    keep = np.isin(g.samples["y"], ys)
    gb.samples = g.samples[keep]
    pk = np.isin(g.grid_sites["y"], ys)
    gb.points, gb.grid_sites = g.points[pk], g.grid_sites[pk]
    gb.stats = dict(g.stats, sampled_pixels=int(keep.sum()))
    rep = parity.compare(gb, o, sl, abi.COST_ZNCC)
    print("c2 band", kernel, rep)
    assert rep["samples"] > 500
    _check_counters(g.stats)


def _band(g, o):
    """The GPU result restricted to the grid rows the oracle evaluated."""
    ys = np.unique(o.samples["y"])
    gb = copy.copy(g)
    keep = np.isin(g.samples["y"], ys)
    gb.samples = g.samples[keep]
    pk = np.isin(g.grid_sites["y"], ys)
    gb.points, gb.grid_sites = g.points[pk], g.grid_sites[pk]
    gb.stats = dict(g.stats, sampled_pixels=int(keep.sum()))
    return gb, ys


@pytest.mark.parametrize("radius", [2, 3, 5])
def test_stride1_box_sums_match_oracle(product, oracle, radius):
    """At stride 1 with ZNCC patches of 5 x 5 and up, the tiled sweep takes each patch's source
    sums from FP64 box sums shared by the tile's overlapping patches. Against the oracle on a
    band of grid rows, and against the direct kernel (per-patch FP32 sums) on every sample."""
    c1 = scenes.config_c1()
    st = dict(c1.settings, patch_radius=radius, depth_hypotheses=48, grid_stride=1,
              surface_completion_passes=0, emit_samples=1)
    g = _gpu(product, c1, st, kernel="tiled")
    oracle.set_row_band(90, 3)
    try:
        o, sl = _oracle(oracle, c1, st)
    finally:
        oracle.set_row_band(0, 0)
    gb, ys = _band(g, o)
    assert len(ys) == 3
    rep = parity.compare(gb, o, sl, abi.COST_ZNCC)
    print("stride 1 box sums", radius, rep)
    assert rep["samples"] > 500
    _check_counters(g.stats)
    d = _gpu(product, c1, st, kernel="direct")
    same = ((d.samples["status"] == g.samples["status"])
            & (d.samples["best_hypothesis"] == g.samples["best_hypothesis"]))
    assert len(g.samples) > 20000 and same.mean() >= 0.999
    # integer luma is exact in FP16 and BF16 texels: the same bits as FP32 storage
    for storage in ("f16", "bf16"):
        r = _gpu(product, c1, st, kernel="tiled", storage=storage)
        assert r.samples.tobytes() == g.samples.tobytes(), storage
        assert r.points.tobytes() == g.points.tobytes(), storage


def test_c2_is_deterministic_and_storage_independent(product):
    """Luma is an integer 0..255, exact in FP16 and BF16: every storage gives the same bits,
    and repeated solves give the same bits."""
    c2 = scenes.config_c2()
    st = dict(c2.settings, surface_completion_passes=0, emit_samples=1)
    base = _gpu(product, c2, st)
    again = _gpu(product, c2, st)
    assert base.points.tobytes() == again.points.tobytes()
    assert base.samples.tobytes() == again.samples.tobytes()
    for storage in ("f16", "bf16"):
        r = _gpu(product, c2, st, storage=storage)
        assert r.points.tobytes() == base.points.tobytes(), storage
        assert r.samples.tobytes() == base.samples.tobytes(), storage
    # the two primary-sweep kernels reach the same decisions (up to FP32 rounding)
    direct = _gpu(product, c2, st, kernel="direct")
    tiled = _gpu(product, c2, st, kernel="tiled")
    same = ((direct.samples["status"] == tiled.samples["status"])
            & (direct.samples["best_hypothesis"] == tiled.samples["best_hypothesis"]))
    assert same.mean() >= 0.999
    pts = base.points
    assert np.all((pts["confidence"] >= 0.05) & (pts["confidence"] <= 1.0))
    at = base.point_attributes
    assert np.all((at["reciprocal_consistency"] >= 0) & (at["reciprocal_consistency"] <= 1))
    assert np.all((at["depth_margin"] >= 0) & (at["depth_margin"] <= 1))
    sites = base.grid_sites
    order = sites["y"].astype(np.int64) * 100000 + sites["x"]
    assert np.all(np.diff(order) > 0), "points are in grid row-major order"


def test_session_matches_one_shot(product):
    c1 = scenes.config_c1()
    st = dict(c1.settings, surface_completion_passes=0, emit_samples=1)
    one = _gpu(product, c1, st)
    sess = product.Session(0)
    try:
        sess.upload(c1.frames, "f32")
        for _ in range(2):
            r = sess.solve(c1.cameras, c1.sparse, c1.focal, c1.reference,
                           sweep=product.params(st))
            assert r.points.tobytes() == one.points.tobytes()
            assert r.samples.tobytes() == one.samples.tobytes()
        t = sess.timing()
        assert t.units == one.stats["sampled_pixels"] * one.stats["depth_hypotheses"]
        assert t.total_ms > 0 and t.sweep_ms > 0
        # plane_sweep entry (reference chosen by visibility) on the session frames
        r2 = sess.solve(c1.cameras, c1.sparse, c1.focal, -1, sweep=product.params(st))
        assert r2.stats["attempted"]
    finally:
        sess.close()


@pytest.mark.parametrize("case,code", [
    ("no-frames", 1), ("one-camera", 2), ("few-landmarks", 3), ("small-frames", 4),
    ("no-reference", 5), ("no-sources", 8)])
def test_skip_reasons_match_oracle(product, oracle, case, code):
    pf = scenes.plane_fixture()
    frames, cams, sparse, f = pf.frames, pf.cameras, pf.sparse, pf.focal
    if case == "no-frames":
        frames = []
    elif case == "one-camera":
        cams = cams[:1]
    elif case == "few-landmarks":
        sparse = sparse[:7]
    elif case == "small-frames":
        frames = [fr[:15, :] for fr in frames]
    elif case == "no-reference":
        sparse = sparse + np.array([100.0, 0.0, 0.0])
    elif case == "no-sources":
        cams = [scenes.plane_camera(0.0), scenes.plane_camera(0.0)]
    st = {"preset": "reference", "surface_completion_passes": 0}
    g = product.estimate_depth_points_with_control(frames, cams, sparse, f,
                                                   sweep=product.params(st))
    o, _ = oracle.plane_sweep(frames, cams, sparse, f, sweep=oracle.params(st))
    assert g.skip_code == o.skip_code == code
    assert g.stats["skip_reason"] == o.stats["skip_reason"]
    assert len(g.points) == 0


def test_cancellation_matches_oracle(product, oracle):
    pf = scenes.plane_fixture()
    st = {"preset": "reference", "surface_completion_passes": 0}
    for cancel_at in (1, 5, 11):
        counts = []
        for run in (product.estimate_depth_points_with_control, None):
            polls = []

            def poll():
                polls.append(1)
                return len(polls) == cancel_at

            if run is None:
                r, _ = oracle.plane_sweep(pf.frames, pf.cameras, pf.sparse, pf.focal,
                                          control=poll, sweep=oracle.params(st))
            else:
                r = run(pf.frames, pf.cameras, pf.sparse, pf.focal, control=poll,
                        sweep=product.params(st))
            assert r.skip_code == 9
            counts.append(len(polls))
        assert counts[0] == counts[1] == cancel_at
    # never canceled: one poll per grid row (11 rows), same on both sides
    polls = []
    g = product.estimate_depth_points_with_control(pf.frames, pf.cameras, pf.sparse, pf.focal,
                                                   control=lambda: polls.append(1) and False,
                                                   sweep=product.params(st))
    assert g.stats["attempted"] and len(polls) == 11


def test_textureless_frames_give_no_points(product):
    W, H, f = 68, 48, 60.0
    flat = np.full((H, W), 120, np.uint8)
    sparse = scenes.plane_sparse_points(W, H, f, 4.0)
    g = product.estimate_depth_points_with_control(
        [flat, flat], [scenes.plane_camera(0.0), scenes.plane_camera(0.2)], sparse, f,
        sweep=product.params({"preset": "reference", "surface_completion_passes": 0}))
    assert g.stats["attempted"] and g.stats["sampled_pixels"] == 0 and len(g.points) == 0


def test_reference_patch_none_cases(product):
    c1 = scenes.config_c1()
    st = product.params(dict(c1.settings, surface_completion_passes=0))
    assert product.reconstruct_reference_patch(c1.frames, c1.cameras, c1.sparse[:5], c1.focal,
                                               2, sweep=st) is None
    assert product.reconstruct_reference_patch(c1.frames, c1.cameras, c1.sparse, c1.focal,
                                               99, sweep=st) is None
