"""Wire formats and the dense warning (host functions of libvt3d_dense.so; no GPU).

The packed buffers follow crates/video-to-3d-wasm/src/lib.rs:227-268 and :344-366 and are
checked against Python's own little-endian packing. The warnings are the text of
crates/video-to-3d-core/src/reconstruction/warnings.rs:27-68, checked with the reference's
tests (reconstruction/tests.rs:436-478) and the full text of every branch.
"""
import ctypes as C
import struct

import numpy as np
import pytest

from vt3d_dense import abi


def _points(n, seed=3):
    rng = np.random.default_rng(seed)
    p = np.zeros(n, abi.POINT_DTYPE)
    for f in ("x", "y", "z", "confidence"):
        p[f] = rng.standard_normal(n).astype(np.float32) * 10
    if n >= 3:
        p["x"][:3] = [0.0, -0.0, np.float32(1e-40)]  # signed zero and a subnormal keep bits
    for f in ("r", "g", "b"):
        p[f] = rng.integers(0, 256, n)
    return p


def test_pack_dense_points_is_le_xyzw_then_rgb(product):
    lib = product.library()
    p = _points(37)
    xyzw = (C.c_uint8 * (16 * len(p)))()
    rgb = (C.c_uint8 * (3 * len(p)))()
    assert lib.vt3d_pack_dense_points(p.ctypes.data_as(C.POINTER(abi.Point3)), len(p), xyzw,
                                      rgb) == len(p)
    want = b"".join(struct.pack("<ffff", q["x"], q["y"], q["z"], q["confidence"]) for q in p)
    assert bytes(xyzw) == want
    assert bytes(rgb) == b"".join(bytes([q["r"], q["g"], q["b"]]) for q in p)
    # the round trip of lib.rs:344-366 keeps every float bit and drops the colours
    back = np.zeros(len(p), abi.POINT_DTYPE)
    assert lib.vt3d_unpack_dense_points(xyzw, len(xyzw), len(p),
                                        back.ctypes.data_as(C.POINTER(abi.Point3))) == 0
    for f in ("x", "y", "z", "confidence"):
        assert back[f].tobytes() == p[f].tobytes()
    assert not back["r"].any() and not back["g"].any() and not back["b"].any()


def test_unpack_rejects_inconsistent_lengths(product):
    lib = product.library()
    buf = (C.c_uint8 * 64)()
    out = np.zeros(4, abi.POINT_DTYPE)
    ptr = out.ctypes.data_as(C.POINTER(abi.Point3))
    assert lib.vt3d_unpack_dense_points(buf, 64, 4, ptr) == 0
    for n_bytes, count in ((64, 3), (60, 4), (63, 4), (48, 4)):
        assert lib.vt3d_unpack_dense_points(buf, n_bytes, count, ptr) == abi.VT3D_E_INVALID
    assert lib.vt3d_unpack_dense_points(buf, 0, 0, ptr) == 0


def test_pack_sites_and_attributes(product):
    lib = product.library()
    sites = np.array([(0, 0), (3, 7), (2**32 - 1, 12345)], abi.SITE_DTYPE)
    out = (C.c_uint8 * (8 * len(sites)))()
    lib.vt3d_pack_grid_sites(sites.ctypes.data_as(C.POINTER(abi.GridSite)), len(sites), out)
    assert bytes(out) == b"".join(struct.pack("<II", int(s["x"]), int(s["y"])) for s in sites)
    attrs = np.array([(0.25, 1.0), (1.0, 0.0), (0.5, 0.123456789)], abi.ATTR_DTYPE)
    out = (C.c_uint8 * (8 * len(attrs)))()
    lib.vt3d_pack_point_attributes(attrs.ctypes.data_as(C.POINTER(abi.PointAttributes)),
                                   len(attrs), out)
    assert bytes(out) == b"".join(struct.pack("<ff", a["reciprocal_consistency"],
                                              a["depth_margin"]) for a in attrs)


def _warning(product, **fields):
    st = abi.DenseStats()
    for k, v in fields.items():
        setattr(st, k, v)
    lib = product.library()
    n = lib.vt3d_dense_warning(C.byref(st), None, 0)
    if n == 0:
        return None
    buf = C.create_string_buffer(n + 1)
    assert lib.vt3d_dense_warning(C.byref(st), buf, n + 1) == n
    # a short buffer is truncated and terminated, and the full length is still returned
    small = C.create_string_buffer(8)
    assert lib.vt3d_dense_warning(C.byref(st), small, 8) == n
    assert small.value == buf.value[:7]
    return buf.value.decode()


def test_dense_warning_reports_successful_fusion(product):
    """reconstruction/tests.rs:436-456."""
    w = _warning(product, attempted=1, has_reference_frame=1, reference_frame=1, source_views=3,
                 depth_hypotheses=24, accepted_points=17, reciprocal_consistent_points=19,
                 fusion_rejected_observations=4)
    assert "17 fused scene points" in w
    assert "19 primary candidates passed reciprocal depth consistency" in w
    assert "spatial fusion rejected 4 reverse observations" in w
    for absent in ("fusion remain future", "multi-view depth consistency", "meshing"):
        assert absent not in w
    assert w == (
        "Slice 4 dense point fusion accepted 17 fused scene points from reference frame 2 "
        "using 3 registered source views and 24 inverse-depth hypotheses. 19 primary "
        "candidates passed reciprocal depth consistency; spatial fusion rejected 4 reverse "
        "observations. Fused points remain separate from the sparse map; dense coverage is "
        "bounded to accepted reference patches and does not claim unbounded aggregation or "
        "metric scale.")


def test_dense_warning_distinguishes_fusion_rejection(product):
    """reconstruction/tests.rs:458-478."""
    w = _warning(product, attempted=1, has_reference_frame=1, reference_frame=0, source_views=2,
                 reciprocal_checked_points=9, reciprocal_rejected_points=3,
                 reciprocal_consistent_points=6, fusion_input_observations=14,
                 fusion_rejected_observations=8, fusion_rejected_points=6)
    assert "6 primary candidates passed reciprocal depth consistency" in w
    assert "no fused point survived" in w
    assert "8 reverse observations were rejected" in w
    assert "no sampled pixel passed" not in w
    assert w == (
        "Slice 4 dense point fusion ran from reference frame 1 with 2 registered source views. "
        "6 primary candidates passed reciprocal depth consistency, but no fused point survived "
        "the spatial-consistency and minimum-observation gates; 8 reverse observations were "
        "rejected as spatially inconsistent. No dense geometry was invented.")


def test_dense_warning_earlier_gates_and_none(product):
    w = _warning(product, attempted=1, has_reference_frame=1, reference_frame=4, source_views=4,
                 reciprocal_checked_points=11)
    assert w == (
        "Slice 4 coarse depth estimation ran from reference frame 5 with 4 registered source "
        "views. 11 primary candidates passed the texture, photometric-error, and ambiguity "
        "gates, but reciprocal depth consistency rejected all of them. No dense geometry was "
        "invented.")
    # no reference frame reads as frame 0
    w = _warning(product, attempted=1, source_views=0)
    assert w == (
        "Slice 4 coarse depth estimation ran from reference frame 0 with 0 registered source "
        "views, but no sampled pixel passed the texture, photometric-error, and ambiguity "
        "gates. No dense geometry was invented.")
    assert _warning(product, attempted=0, accepted_points=5) is None


@pytest.mark.parametrize("n", [0, 1, 1000])
def test_pack_sizes(product, n):
    lib = product.library()
    p = _points(max(n, 1))[:n]
    xyzw = (C.c_uint8 * max(1, 16 * n))()
    rgb = (C.c_uint8 * max(1, 3 * n))()
    assert lib.vt3d_pack_dense_points(p.ctypes.data_as(C.POINTER(abi.Point3)), n, xyzw,
                                      rgb) == n


def test_python_mirrors(product, oracle):
    """pack_dense and dense_warning over an analysis (the oracle's, so no GPU is needed)."""
    from vt3d_dense import scenes

    pf = scenes.plane_fixture()
    res, _ = oracle.plane_sweep(pf.frames, pf.cameras, pf.sparse, pf.focal)
    packed = product.pack_dense(res)
    n = len(res.points)
    assert packed["dense_points"]["length"] == n > 0
    vals = np.frombuffer(packed["dense_points"]["values_f32_le"], "<f4").reshape(n, 4)
    assert np.array_equal(vals[:, 2], res.points["z"])
    assert np.frombuffer(packed["dense_grid_sites"]["xy_u32_le"], "<u4").reshape(n, 2)[:, 0] \
        .tolist() == res.grid_sites["x"].tolist()
    w = product.dense_warning(res.stats)
    assert w.startswith(f"Slice 4 dense point fusion accepted {n} fused scene points from "
                        "reference frame 1 using 2 registered source views")


def test_mesh_warning_text(product):
    """reconstruction/warnings.rs:5-25, both branches and None."""
    base = {"attempted": True, "accepted_triangles": 12, "candidate_triangles": 15,
            "candidate_cells": 9, "rejected_discontinuities": 2, "rejected_degenerate": 1}
    assert product.mesh_warning(base) == (
        "Slice 4 bounded mesh reconstruction accepted 12 triangles from 15 candidate triangles "
        "across 9 reference-grid cells. It rejected 2 triangles at depth/spatial "
        "discontinuities and 1 degenerate or orientation-flipped triangles. This is a bounded "
        "multi-reference, non-watertight surface preview. Browser texture projection is "
        "appearance-only over accepted per-reference topology; unsupported holes and metric "
        "scale are not claimed.")
    assert product.mesh_warning(dict(base, accepted_triangles=0)) == (
        "Slice 4 bounded mesh reconstruction evaluated 15 candidate triangles across 9 "
        "reference-grid cells, but no triangle survived the continuity and non-degeneracy "
        "gates. No surface geometry was invented.")
    assert product.mesh_warning(dict(base, attempted=False)) is None
