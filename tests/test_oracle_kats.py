"""Pins the FP64 oracle: the reference's own unit-test values (dense_legacy.rs:1184-1515,
dense.rs:1114-1157), closed forms, invariants and brute force on tiny inputs."""
import ctypes as C
import math

import numpy as np
import pytest

from vt3d_dense import abi, scenes


def cam(R, t, idx=0):
    import vt3d_oracle as O
    return O.camera_struct(R, t, idx)


def plane_cam(cx, idx=0):
    return cam(np.eye(3), [-cx, 0.0, 0.0], idx)


# ---- to_luma (dense_legacy.rs:1147-1157) ------------------------------------------------------

def test_to_luma_exhaustive_rgb(oracle):
    lib = oracle.library()
    v = np.arange(1 << 24, dtype=np.uint32)
    rgba = np.zeros((v.size, 4), np.uint8)
    rgba[:, 0] = v & 0xFF
    rgba[:, 1] = (v >> 8) & 0xFF
    rgba[:, 2] = (v >> 16) & 0xFF
    out = np.empty(v.size, np.uint8)
    lib.vt3d_oracle_to_luma(oracle.u8ptr(rgba), v.size, oracle.u8ptr(out))
    expect = ((77 * rgba[:, 0].astype(np.uint32) + 150 * rgba[:, 1].astype(np.uint32)
               + 29 * rgba[:, 2].astype(np.uint32)) >> 8).astype(np.uint8)
    assert np.array_equal(out, expect)
    gray = np.arange(256, dtype=np.uint8)
    assert np.array_equal(expect[gray.astype(np.uint32) * 0x010101], gray)  # gray round trip


# ---- quantile (dense_legacy.rs:1168-1174) -----------------------------------------------------

@pytest.mark.parametrize("n,q,index", [(1, 0.5, 0), (2, 0.5, 1), (3, 0.5, 1), (4, 0.5, 2),
                                       (5, 0.5, 2), (10, 0.10, 1), (10, 0.90, 8), (11, 0.1, 1),
                                       (21, 0.9, 18), (4, 2.0, 3), (4, -1.0, 0)])
def test_quantile_index_rounds_half_away_from_zero(oracle, n, q, index):
    v = np.arange(n, dtype=np.float64) * 10.0
    assert oracle.library().vt3d_oracle_quantile(oracle.f64ptr(v), n, q) == v[index]


def test_quantile_empty_is_nan(oracle):
    v = np.zeros(1)
    assert math.isnan(oracle.library().vt3d_oracle_quantile(oracle.f64ptr(v), 0, 0.5))


# ---- project / unproject (dense_legacy.rs:996-1029) -------------------------------------------

def _random_camera(rng):
    R = scenes.axis_angle(rng.normal(0, 0.3, 3))
    return R, rng.normal(0, 0.5, 3)


def test_project_unproject_round_trip(oracle):
    lib = oracle.library()
    rng = np.random.default_rng(1)
    for _ in range(200):
        R, t = _random_camera(rng)
        c = cam(R, t)
        x, y, d = rng.uniform(0, 640), rng.uniform(0, 360), rng.uniform(0.5, 8.0)
        w = (C.c_double * 3)()
        lib.vt3d_oracle_unproject(C.byref(c), x, y, d, 640, 360, 550.0, w)
        p = (C.c_double * 3)()
        assert lib.vt3d_oracle_project(C.byref(c), w, 640, 360, 550.0, p) == 1
        assert abs(p[0] - x) < 1e-9 and abs(p[1] - y) < 1e-9 and abs(p[2] - d) < 1e-12 * 10


def test_project_rejects_behind_and_near(oracle):
    lib = oracle.library()
    c = cam(np.eye(3), [0, 0, 0])
    out = (C.c_double * 3)()
    for z in (-1.0, 0.0, 1e-4):
        pt = (C.c_double * 3)(0.1, 0.1, z)
        assert lib.vt3d_oracle_project(C.byref(c), pt, 64, 48, 60.0, out) == 0
    pt = (C.c_double * 3)(0.1, 0.1, 1.0001e-4)
    assert lib.vt3d_oracle_project(C.byref(c), pt, 64, 48, 60.0, out) == 1
    pt = (C.c_double * 3)(float("nan"), 0.1, 1.0)
    assert lib.vt3d_oracle_project(C.byref(c), pt, 64, 48, 60.0, out) == 0


def test_product_homography_equals_per_tap_warp(oracle, product):
    """The kernels' FP64 plane homography reproduces unproject + project per tap."""
    lib, plib = oracle.library(), product.library()
    rng = np.random.default_rng(2)
    for _ in range(100):
        Ra, ta = _random_camera(rng)
        Rb = Ra @ scenes.axis_angle(rng.normal(0, 0.1, 3))
        tb = ta + rng.normal(0, 0.2, 3)
        a, b = cam(Ra, ta), cam(Rb, tb, 1)
        wa, ha, wb, hb = 640, 360, 480, 270
        f = 500.0
        d = rng.uniform(0.5, 8.0)
        H = (C.c_double * 9)()
        plib.vt3d_plane_homography(C.byref(a), wa, ha, C.byref(b), wb, hb, f, d, H)
        Hm = np.array(H[:]).reshape(3, 3)
        for _ in range(5):
            u, v = rng.uniform(0, wa), rng.uniform(0, ha)
            w = (C.c_double * 3)()
            lib.vt3d_oracle_unproject(C.byref(a), u, v, d, wa, ha, f, w)
            p = (C.c_double * 3)()
            if not lib.vt3d_oracle_project(C.byref(b), w, wb, hb, f, p):
                continue
            q = Hm @ np.array([u, v, 1.0])
            assert abs(q[2] - p[2]) < 1e-9 * max(1.0, abs(p[2]))
            assert abs(q[0] / q[2] - p[0]) < 1e-7 and abs(q[1] / q[2] - p[1]) < 1e-7


# ---- sample_bilinear / patch_contrast (dense_legacy.rs:1097-1140) -----------------------------

def test_sample_bilinear_bounds_and_values(oracle):
    lib = oracle.library()
    img = np.array([[0, 10, 20], [30, 40, 50], [60, 70, 80]], np.uint8)
    v = C.c_double()
    assert lib.vt3d_oracle_sample_bilinear(oracle.u8ptr(img), 3, 3, 0.5, 0.5, C.byref(v)) == 1
    assert v.value == pytest.approx(20.0)
    assert lib.vt3d_oracle_sample_bilinear(oracle.u8ptr(img), 3, 3, 1.0, 1.0, C.byref(v)) == 1
    assert v.value == 40.0
    # x < width - 1 is exclusive: the last column/row cannot be a bilinear base
    assert lib.vt3d_oracle_sample_bilinear(oracle.u8ptr(img), 3, 3, 2.0, 0.0, C.byref(v)) == 0
    assert lib.vt3d_oracle_sample_bilinear(oracle.u8ptr(img), 3, 3, 1.999, 1.999, C.byref(v)) == 1
    assert lib.vt3d_oracle_sample_bilinear(oracle.u8ptr(img), 3, 3, -1e-9, 0.0, C.byref(v)) == 0


def test_patch_contrast_integer_and_out_of_bounds(oracle):
    lib = oracle.library()
    img = np.arange(64, dtype=np.uint8).reshape(8, 8) * 3
    # 3x3 around (3, 3): max - min = (4*8+4 - (2*8+2)) * 3 = 54
    assert lib.vt3d_oracle_patch_contrast(oracle.u8ptr(img), 8, 8, 3.0, 3.0, 1) == 54.0
    assert lib.vt3d_oracle_patch_contrast(oracle.u8ptr(img), 8, 8, 3.0, 3.0, 2) == 108.0
    assert lib.vt3d_oracle_patch_contrast(oracle.u8ptr(img), 8, 8, 0.0, 3.0, 1) == 0.0


# ---- patch_error (dense_legacy.rs:1031-1095) --------------------------------------------------

def _shift_pair(shift_px, width=40, height=30, focal=60.0, depth=4.0, seed=3):
    """Reference and a source translated along x so a plane at `depth` moves by shift_px."""
    rng = np.random.default_rng(seed)
    ref = rng.integers(0, 256, (height, width)).astype(np.uint8)
    src = np.zeros_like(ref)
    # source camera at +b: a world point seen at x in the reference appears at x - f b / d
    src[:, : width - shift_px] = ref[:, shift_px:]
    baseline = shift_px * depth / focal
    return ref, src, plane_cam(0.0), plane_cam(baseline, 1), focal, depth


@pytest.mark.parametrize("kind,radius", [(0, 1), (0, 3), (1, 1), (1, 3), (1, 5)])
def test_patch_error_zero_at_true_integer_shift(oracle, kind, radius):
    lib = oracle.library()
    ref, src, a, b, f, d = _shift_pair(3)
    h, w = ref.shape
    cost = C.c_double()
    assert lib.vt3d_oracle_patch_error(C.byref(a), C.byref(b), oracle.u8ptr(ref),
                                       oracle.u8ptr(src), w, h, w, h, 15.0, 12.0, d, f, kind,
                                       radius, C.byref(cost)) == 1
    assert abs(cost.value) < 1e-12
    # a wrong depth costs more
    assert lib.vt3d_oracle_patch_error(C.byref(a), C.byref(b), oracle.u8ptr(ref),
                                       oracle.u8ptr(src), w, h, w, h, 15.0, 12.0, d * 1.5, f,
                                       kind, radius, C.byref(cost)) == 1
    assert cost.value > 1e-3


def test_patch_error_offset_and_gain_invariance(oracle):
    """ZSAD ignores an intensity offset; ZNCC ignores offset and positive gain."""
    lib = oracle.library()
    rng = np.random.default_rng(4)
    ref = rng.integers(40, 120, (20, 20)).astype(np.uint8)
    a = plane_cam(0.0)
    cost = C.c_double()
    off = (ref + 30).astype(np.uint8)
    assert lib.vt3d_oracle_patch_error(C.byref(a), C.byref(a), oracle.u8ptr(ref),
                                       oracle.u8ptr(off), 20, 20, 20, 20, 9.0, 9.0, 2.0, 30.0,
                                       0, 1, C.byref(cost)) == 1
    assert cost.value == pytest.approx(0.0, abs=1e-12)
    gain = (ref.astype(np.int32) * 2 + 5).astype(np.uint8)
    assert lib.vt3d_oracle_patch_error(C.byref(a), C.byref(a), oracle.u8ptr(ref),
                                       oracle.u8ptr(gain), 20, 20, 20, 20, 9.0, 9.0, 2.0, 30.0,
                                       1, 2, C.byref(cost)) == 1
    assert cost.value == pytest.approx(0.0, abs=1e-12)


def test_patch_error_hand_computed_zsad(oracle):
    """3x3 identity warp: mean |(r - mean r) - (s - mean s)| by hand."""
    lib = oracle.library()
    ref = np.zeros((5, 5), np.uint8)
    src = np.zeros((5, 5), np.uint8)
    ref[1:4, 1:4] = [[10, 20, 30], [40, 50, 60], [70, 80, 90]]
    src[1:4, 1:4] = [[10, 20, 30], [40, 59, 60], [70, 80, 90]]
    a = plane_cam(0.0)
    cost = C.c_double()
    assert lib.vt3d_oracle_patch_error(C.byref(a), C.byref(a), oracle.u8ptr(ref),
                                       oracle.u8ptr(src), 5, 5, 5, 5, 2.0, 2.0, 1.0, 10.0, 0, 1,
                                       C.byref(cost)) == 1
    # source mean is 51, reference mean 50: |d| = 1 for 8 taps, |9 - 1| = 8 for the centre
    assert cost.value == pytest.approx((8 * 1.0 + 8.0) / 9.0, abs=1e-12)


def test_patch_error_out_of_bounds_is_none(oracle):
    lib = oracle.library()
    ref, src, a, b, f, d = _shift_pair(3)
    h, w = ref.shape
    cost = C.c_double()
    # a tap warps past the source's left edge
    assert lib.vt3d_oracle_patch_error(C.byref(a), C.byref(b), oracle.u8ptr(ref),
                                       oracle.u8ptr(src), w, h, w, h, 3.0, 12.0, d, f, 0, 1,
                                       C.byref(cost)) == 0


# ---- visible depths, envelope (dense_legacy.rs:755-775, 959-973, 1499-1514) -------------------

def test_visible_depths_ignore_offscreen_landmarks(oracle):
    lib = oracle.library()
    c = plane_cam(0.0)
    pts = np.array([[0.0, 0.0, 4.0], [100.0, 0.0, 4.0], [0.0, -100.0, 4.0]])
    out = np.zeros(3)
    n = lib.vt3d_oracle_visible_depths(C.byref(c), oracle.f64ptr(pts), 3, 68, 48, 60.0,
                                       oracle.f64ptr(out))
    assert n == 1 and out[0] == 4.0


def test_depth_search_bounds_closed_form(oracle):
    lib = oracle.library()
    c = plane_cam(0.0)
    depths = np.linspace(2.0, 6.0, 21)
    pts = np.array([[0.0, 0.0, d] for d in depths])
    mn, mx = C.c_double(), C.c_double()
    assert lib.vt3d_oracle_depth_search_bounds(C.byref(c), oracle.f64ptr(pts), 21, 68, 48, 60.0,
                                               C.byref(mn), C.byref(mx)) == 1
    assert mn.value == pytest.approx(depths[2] * 0.8)   # q10 index round(2.0) = 2
    assert mx.value == pytest.approx(depths[18] * 1.25)  # q90 index round(18.0) = 18
    assert lib.vt3d_oracle_depth_search_bounds(C.byref(c), oracle.f64ptr(pts), 7, 68, 48, 60.0,
                                               C.byref(mn), C.byref(mx)) == 0  # < 8 landmarks


# ---- reciprocal check and fusion (dense_legacy.rs:1392-1455) ----------------------------------

def test_reciprocal_consistency_rejects_wrong_depth(oracle):
    lib = oracle.library()
    assert lib.vt3d_oracle_reciprocal_agrees(4.0, 4.2, 0.08) == 1
    assert lib.vt3d_oracle_reciprocal_agrees(4.0, 5.0, 0.08) == 0
    assert lib.vt3d_oracle_reciprocal_agrees(-1.0, 1.0, 0.08) == 0
    assert lib.vt3d_oracle_reciprocal_agrees(float("nan"), 1.0, 0.08) == 0


def test_fusion_rejects_spatially_inconsistent_reciprocal_observations(oracle):
    lib = oracle.library()
    primary = np.array([0.0, 0.0, 4.0])
    recip = np.array([[0.02, -0.01, 4.04], [0.9, 0.0, 4.0]])
    pos = np.zeros(3)
    obs, rej = C.c_uint64(), C.c_uint64()
    lib.vt3d_oracle_fuse(oracle.f64ptr(primary), oracle.f64ptr(recip), 2, 4.0, 0.08,
                         oracle.f64ptr(pos), C.byref(obs), C.byref(rej))
    assert obs.value == 2 and rej.value == 1
    assert abs(pos[2] - 4.02) < 1e-9
    assert abs(pos[0]) < 0.02


def test_reciprocal_depth_search_recovers_source_view_depth(oracle):
    lib = oracle.library()
    W, H, f, depth = 68, 48, 60.0, 4.0
    frames = [scenes.plane_frame(W, H, f, cx, depth) for cx in (0.0, 0.18)]
    cams = [plane_cam(0.0, 0), plane_cam(0.18, 1)]
    sparse = scenes.plane_sparse_points(W, H, f, depth)
    mn, mx = C.c_double(), C.c_double()
    assert lib.vt3d_oracle_depth_search_bounds(C.byref(cams[1]), oracle.f64ptr(sparse),
                                               len(sparse), W, H, f, C.byref(mn), C.byref(mx))
    p = (C.c_double * 3)()
    world = (C.c_double * 3)(0.0, 0.0, depth)
    assert lib.vt3d_oracle_project(C.byref(cams[1]), world, W, H, f, p)
    est = C.c_double()
    assert lib.vt3d_oracle_single_view_depth(C.byref(cams[1]), C.byref(cams[0]),
                                             oracle.u8ptr(frames[1]), oracle.u8ptr(frames[0]),
                                             W, H, W, H, p[0], p[1], mn.value, mx.value, f, None,
                                             C.byref(est)) == 1
    assert lib.vt3d_oracle_reciprocal_agrees(p[2], est.value, 0.08) == 1


def test_surface_depth_prediction_requires_consistent_neighbors(oracle):
    lib = oracle.library()
    out = C.c_double()
    smooth = np.array([4.0, 4.03, 3.98, 4.01])
    assert lib.vt3d_oracle_surface_depth_prediction(oracle.f64ptr(smooth), 4, C.byref(out)) == 1
    assert abs(out.value - 4.01) < 0.05
    disc = np.array([4.0, 4.02, 5.2])
    assert lib.vt3d_oracle_surface_depth_prediction(oracle.f64ptr(disc), 3, C.byref(out)) == 0
    under = np.array([4.0, 4.01])
    assert lib.vt3d_oracle_surface_depth_prediction(oracle.f64ptr(under), 2, C.byref(out)) == 0


# ---- grid stride (dense.rs:215-218, dense.rs:1114-1157) ---------------------------------------

def test_grid_stride_reference_values(oracle, product):
    for lib, fn in ((oracle.library(), "vt3d_oracle_grid_stride"),
                    (product.library(), "vt3d_grid_stride")):
        g = getattr(lib, fn)
        assert g(240, 240, 1) == 4 and g(240, 240, 0) == 5
        assert g(68, 48, 0) == 4 and g(1920, 1080, 0) == 12 and g(640, 360, 1) == 5


# ---- end to end on the reference's analytic plane (dense_legacy.rs:1261-1321) -----------------

def test_estimates_coarse_depth_on_textured_plane(oracle):
    pf = scenes.plane_fixture()
    res, _ = oracle.plane_sweep(pf.frames, pf.cameras, pf.sparse, pf.focal)
    s = res.stats
    assert s["attempted"] and s["skip_reason"] is None
    assert s["reference_frame"] == 0 and s["source_views"] == 2 and s["source_frames"] == [1, 2]
    assert s["reciprocal_checked_points"] == (s["reciprocal_consistent_points"]
                                              + s["reciprocal_rejected_points"])
    assert (s["reciprocal_consistent_points"] + s["surface_completed_points"]
            == s["accepted_points"] + s["fusion_rejected_points"])
    assert s["surface_completion_proposals"] == (
        s["surface_completed_points"] + s["surface_completion_rejected_texture"]
        + s["surface_completion_rejected_cross_view"] + s["surface_completion_rejected_reciprocal"]
        + s["surface_completion_rejected_fusion"] + s["surface_completion_rejected_footprint"])
    assert s["fusion_input_observations"] >= s["reciprocal_consistent_points"] * 2
    assert s["median_fusion_observations"] is not None and s["median_fusion_observations"] >= 2
    assert len(res.points) >= 12
    z = np.sort(res.points["z"].astype(np.float64))
    median = z[int(round((len(z) - 1) * 0.5))]
    assert abs(median - 4.0) < 0.55


def test_rejects_textureless_depth_hypotheses(oracle):
    W, H, f = 68, 48, 60.0
    flat = np.full((H, W), 120, np.uint8)
    sparse = scenes.plane_sparse_points(W, H, f, 4.0)
    res, _ = oracle.plane_sweep([flat, flat], [scenes.plane_camera(0.0), scenes.plane_camera(0.2)],
                                sparse, f)
    s = res.stats
    assert s["attempted"] and s["skip_reason"] is None
    assert len(res.points) == 0 and s["accepted_points"] == 0
    assert s["surface_completed_points"] == 0 and s["surface_completion_proposals"] == 0


def test_does_not_attempt_dense_depth_without_registered_baseline(oracle):
    pf = scenes.plane_fixture(centers=(0.0,))
    res, _ = oracle.plane_sweep(pf.frames, pf.cameras, pf.sparse, pf.focal)
    assert not res.stats["attempted"]
    assert "two accepted registered cameras" in res.stats["skip_reason"]
    assert len(res.points) == 0


def test_flat_frames_have_no_depth_evidence_at_either_spacing(oracle):
    """dense.rs:1114-1157: 240x240 flat frames, coarse stride 5 / fine stride 4, no points."""
    W = H = 240
    f = 200.0
    flat = np.full((H, W), 128, np.uint8)
    pts = []
    for y in (30.0, 90.0, 150.0, 210.0):
        for x in (30.0, 90.0, 150.0, 210.0):
            for dx in (0.0, 2.0):
                pts.append(((x + dx - 120) / f * 4.0, (y - 120) / f * 4.0, 4.0))
    cams = [scenes.plane_camera(0.0), scenes.plane_camera(0.2)]
    for fine, stride in ((True, 4), (False, 5)):
        res, _ = oracle.plane_sweep([flat, flat], cams, np.array(pts), f, fine_grid=fine)
        assert res.stats["grid_stride"] == stride
        assert len(res.points) == 0


@pytest.mark.parametrize("case,code", [
    ("no-frames", 1), ("one-camera", 2), ("few-landmarks", 3), ("small-frames", 4),
    ("no-reference", 5), ("no-sources", 8)])
def test_skip_reasons(oracle, case, code):
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
        cams = [scenes.plane_camera(0.0), scenes.plane_camera(0.0)]  # zero baseline
    res, _ = oracle.plane_sweep(frames, cams, sparse, f)
    assert res.skip_code == code
    assert not res.stats["attempted"]
    assert res.stats["skip_reason"] == oracle.library().vt3d_oracle_skip_reason(code).decode()


def test_cancel_polls_once_per_grid_row_and_is_sticky(oracle):
    pf = scenes.plane_fixture()
    polls = []

    def poll():
        polls.append(1)
        return len(polls) == 3

    res, _ = oracle.plane_sweep(pf.frames, pf.cameras, pf.sparse, pf.focal, control=poll)
    assert res.skip_code == 9 and res.stats["skip_reason"] == "dense reconstruction canceled"
    assert len(polls) == 3
    polls.clear()
    res, _ = oracle.plane_sweep(pf.frames, pf.cameras, pf.sparse, pf.focal,
                                control=lambda: polls.append(1) and False)
    # 11 grid rows, then each completion pass polls once per row
    assert res.stats["attempted"] and len(polls) >= 11


def test_threads_do_not_change_the_output(oracle):
    c1 = scenes.config_c1()
    st = dict(c1.settings, emit_samples=1, surface_completion_passes=0)
    a, sa = oracle.reference_patch(c1.frames, c1.cameras, c1.sparse, c1.focal, c1.reference,
                                   sweep=oracle.params(st), threads=1)
    b, sb = oracle.reference_patch(c1.frames, c1.cameras, c1.sparse, c1.focal, c1.reference,
                                   sweep=oracle.params(st), threads=4)
    assert a.points.tobytes() == b.points.tobytes()
    assert a.samples.tobytes() == b.samples.tobytes()
    assert np.array_equal(sa, sb)


def test_reference_patch_remaps_frames(oracle):
    """dense.rs:674-747: the reference is index 0, the rest nearest-first; indices map back."""
    c1 = scenes.config_c1()
    st = dict(c1.settings, surface_completion_passes=0)
    res, _ = oracle.reference_patch(c1.frames, c1.cameras, c1.sparse, c1.focal, 2,
                                    sweep=oracle.params(st))
    assert res.stats["reference_frame"] == 2
    assert res.stats["source_frames"] == [1, 3, 0, 4]
    none, _ = oracle.reference_patch(c1.frames, c1.cameras, c1.sparse[:5], c1.focal, 2,
                                     sweep=oracle.params(st))
    assert none is None


# ---- hypothesis schedule (dense_legacy.rs:285-293) ---------------------------------------------

@pytest.mark.parametrize("D", [1, 2, 24, 128])
def test_hypothesis_schedule_is_linear_in_inverse_depth(oracle, D):
    lib = oracle.library()
    mn, mx = 0.5, 8.0
    d = np.array([lib.vt3d_oracle_hypothesis_depth(h, D, mn, mx) for h in range(D)])
    assert d[0] == pytest.approx(mn, rel=1e-15)
    if D == 1:
        return
    assert d[-1] == pytest.approx(mx, rel=1e-14)
    assert np.all(np.diff(d) > 0), "depth increases with the hypothesis index"
    inv = 1.0 / d
    assert np.allclose(np.diff(inv, 2), 0.0, atol=1e-12), "inverse depth is affine in h"


# ---- winner-take-all and ambiguity, brute force (dense_legacy.rs:297-362) -----------------------

def _round_half_away(v):
    return int(math.floor(v + 0.5))


def test_wta_and_ambiguity_match_a_brute_force_sort(oracle):
    """Candidates rebuilt from per-patch costs and sorted with Python's stable sort by the
    reference's key (support desc, error asc, depth asc) give the oracle's choices."""
    lib = oracle.library()
    pf = scenes.plane_fixture()
    W, H, f = 68, 48, pf.focal
    st = {"preset": "reference", "surface_completion_passes": 0, "emit_samples": 1}
    res, _ = oracle.plane_sweep(pf.frames, pf.cameras, pf.sparse, f, sweep=oracle.params(st))
    cams = [cam(c.R, c.t, i) for i, c in enumerate(pf.cameras)]
    mn, mx = C.c_double(), C.c_double()
    assert lib.vt3d_oracle_depth_search_bounds(C.byref(cams[0]), oracle.f64ptr(pf.sparse),
                                               len(pf.sparse), W, H, f, C.byref(mn), C.byref(mx))
    D = 24
    depths = [lib.vt3d_oracle_hypothesis_depth(h, D, mn.value, mx.value) for h in range(D)]
    lumas = [np.ascontiguousarray(fr) for fr in pf.frames]
    checked = 0
    for smp in res.samples[::3]:
        x, y = float(smp["x"]), float(smp["y"])
        cands = []
        for h, d in enumerate(depths):
            errs = []
            for si in (1, 2):
                e = C.c_double()
                if lib.vt3d_oracle_patch_error(C.byref(cams[0]), C.byref(cams[si]),
                                               oracle.u8ptr(lumas[0]), oracle.u8ptr(lumas[si]),
                                               W, H, W, H, x, y, d, f, 0, 1, C.byref(e)):
                    if e.value <= 18.0:
                        errs.append(e.value)
            if errs:
                errs.sort()
                cands.append((len(errs), errs[_round_half_away((len(errs) - 1) * 0.5)], d, h))
        cands.sort(key=lambda c: (-c[0], c[1], c[2]))
        if not cands:
            assert smp["status"] == 1
            continue
        best = cands[0]
        assert smp["best_hypothesis"] == best[3]
        assert smp["best_support"] == best[0]
        assert smp["best_cost"] == pytest.approx(best[1], abs=1e-6)
        second = next((c for c in cands[1:] if c[0] == best[0]), None)
        assert smp["second_hypothesis"] == (second[3] if second else -1)
        margin = second[1] - best[1] if second else math.inf
        if margin < 1.0:
            assert smp["status"] == 3
        else:
            assert smp["status"] >= 4
        checked += 1
    assert checked > 20
