"""Mesh consumer on the device (mesh_legacy.rs:150-556, mesh.rs:77-304) against the FP64
oracle on the same points: every counter, every triangle (indices and confidence bits) in
the reference's order, the skip reasons, the per-patch wrapper and cancellation. The mesh
gates are FP64 with no contraction on both sides, so the comparison is exact."""
import numpy as np
import pytest

from vt3d_dense import abi, scenes

from test_oracle_mesh import DENSE, IDENTITY, QUAD, point_at_pixel, sites
from test_oracle_multi_reference import _plane3

pytestmark = pytest.mark.gpu


def _same(g, o, what):
    assert g.stats == o.stats, (what, g.stats, o.stats)
    assert g.triangles.tobytes() == o.triangles.tobytes(), what


def _both(product, oracle, pts, st, dense, cams, w, h, f, control=None):
    g = product.reconstruct_dense_mesh(pts, st, dense, cams, w, h, f, control=control)
    o = oracle.mesh(pts, st, dense, cams, w, h, f, patches=dense.get("reference_patches"),
                    control=control)
    return g, o


KATS = {
    "quad": ([(3, 3, 4.0, .8), (7, 3, 4.0, .7), (3, 7, 4.0, .9), (7, 7, 4.0, .6)], QUAD),
    "bridge": ([(x, y, 4.0, .8) for x, y in ((3, 3), (11, 3), (3, 7), (11, 7), (7, 7))],
               ((3, 3), (11, 3), (3, 7), (11, 7), (7, 7))),
    "no-bridge": ([(3, 3, 4.0, .8), (11, 3, 5.0, .8), (3, 7, 4.0, .8), (11, 7, 5.0, .8),
                   (7, 7, 4.0, .8)], ((3, 3), (11, 3), (3, 7), (11, 7), (7, 7))),
    "discontinuity": ([(3, 3, 4.0, .8), (7, 3, 4.0, .8), (3, 7, 4.0, .8), (7, 7, 7.0, .8)], QUAD),
    "shifted": ([(4.8, 3, 4.0, .8), (7, 3, 4.0, .8), (3, 7, 4.0, .8)], QUAD[:3]),
    "left-footprint": ([(5.4, 3, 4.0, .8), (7, 3, 4.0, .8), (3, 7, 4.0, .8), (7, 7, 4.0, .8)],
                       QUAD),
    "flipped": ([(4.9, 1.1, 4.0, .8), (5.1, 4.9, 4.0, .8), (8.9, 5.1, 4.0, .8)],
                ((3, 3), (7, 3), (7, 7))),
    "too-few": ([(3, 3, 4.0, .8), (7, 3, 4.0, .8)], QUAD),
    "admission": ([(3, 3, 4.0, .8), (7, 3, 4.0, .7), (6, 7, 4.0, .9)], QUAD[:3]),
}


@pytest.mark.parametrize("case", sorted(KATS))
def test_reference_cases_match_oracle(product, oracle, case):
    pts, st = KATS[case]
    arr = np.array([point_at_pixel(x, y, d, c) for x, y, d, c in pts], abi.POINT_DTYPE)
    g, o = _both(product, oracle, arr, sites(*st), DENSE, [IDENTITY], 68, 48, 60.0)
    _same(g, o, case)


def test_plane_and_c1_meshes_match_oracle(product, oracle):
    pf = scenes.plane_fixture()
    a = product.estimate_depth_points_with_control(pf.frames, pf.cameras, pf.sparse, pf.focal)
    cams = [(i, c.R, c.t) for i, c in enumerate(pf.cameras)]
    g, o = _both(product, oracle, a.points, a.grid_sites, a.stats, cams, 68, 48, pf.focal)
    _same(g, o, "plane")
    c1 = scenes.config_c1()
    st = dict(c1.settings, surface_completion_passes=2)
    a = product.reconstruct_reference_patch(c1.frames, c1.cameras, c1.sparse, c1.focal,
                                            c1.reference, sweep=product.params(st))
    cams = [(i, c.R, c.t) for i, c in enumerate(c1.cameras)]
    W, H = c1.frames[0].shape[1], c1.frames[0].shape[0]
    g, o = _both(product, oracle, a.points, a.grid_sites, a.stats, cams, W, H, c1.focal)
    _same(g, o, "c1")
    assert g.stats["accepted_triangles"] > 1000


def test_multi_reference_mesh_matches_oracle(product, oracle):
    frames, cams, sparse, f = _plane3()
    m = product.estimate_depth_points(frames, cams, sparse, f)
    ccams = [(i, c.R, c.t) for i, c in enumerate(cams)]
    g, o = _both(product, oracle, m.combined.points, m.combined.grid_sites, m.stats, ccams,
                 68, 48, f)
    _same(g, o, "plane, three patches")
    assert len(g.stats["reference_patches"]) == 3


def test_large_synthetic_grid_matches_oracle(product, oracle):
    """A 500 x 300-site grid (stride 2): many blocks in every launch and in the scan, holes,
    depth steps, duplicate sites with tied and untied confidences."""
    rng = np.random.default_rng(7)
    W, H, f, stride, border = 1006, 606, 800.0, 2, 3
    gx, gy = np.meshgrid(np.arange(500), np.arange(300))
    keep = rng.random(gx.shape) > 0.15
    sx = border + stride * gx[keep]
    sy = border + stride * gy[keep]
    depth = 3.0 + 0.002 * gx[keep] + np.where(gy[keep] > 150, 0.8, 0.0)
    jx = rng.uniform(-0.4, 0.4, sx.shape) * stride
    jy = rng.uniform(-0.4, 0.4, sy.shape) * stride
    n = sx.size
    pts = np.zeros(n, abi.POINT_DTYPE)
    pts["x"] = ((sx + jx - W * 0.5) / f * depth).astype(np.float32)
    pts["y"] = ((sy + jy - H * 0.5) / f * depth).astype(np.float32)
    pts["z"] = depth.astype(np.float32)
    pts["confidence"] = rng.choice([0.5, 0.75, 0.9], n).astype(np.float32)
    st = np.zeros(n, abi.SITE_DTYPE)
    st["x"], st["y"] = sx, sy
    dup = rng.choice(n, n // 20, replace=False)  # second samples at some sites
    pts = np.concatenate([pts, pts[dup]])
    st = np.concatenate([st, st[dup]])
    pts["confidence"][n:] = np.float32(0.75)
    dense = {"attempted": True, "reference_frame": 0, "grid_stride": stride,
             "grid_border": border}
    g, o = _both(product, oracle, pts, st, dense, [IDENTITY], W, H, f)
    _same(g, o, "large grid")
    assert g.stats["accepted_triangles"] > 100000 and g.stats["rejected_discontinuities"] > 0


def test_cancellation_matches_oracle(product, oracle):
    pts, st = KATS["bridge"]
    arr = np.array([point_at_pixel(x, y, d, c) for x, y, d, c in pts], abi.POINT_DTYPE)
    total = []
    oracle.mesh(arr, sites(*st), DENSE, [IDENTITY], 68, 48, 60.0,
                control=lambda: total.append(1) and False)
    for at in sorted({1, 4, 5, 6, 9, len(total) - 1, len(total)}):
        got = []
        for side in ("g", "o"):
            n = []

            def poll(n=n):
                n.append(1)
                return len(n) >= at

            if side == "g":
                r = product.reconstruct_dense_mesh(arr, sites(*st), DENSE, [IDENTITY], 68, 48,
                                                   60.0, control=poll)
            else:
                r = oracle.mesh(arr, sites(*st), DENSE, [IDENTITY], 68, 48, 60.0, control=poll)
            got.append((r, len(n)))
        (g, gn), (o, on) = got
        assert gn == on, (at, gn, on)
        _same(g, o, f"cancel at {at}")
    # the per-patch wrapper polls once before each patch
    frames, cams, sparse, f = _plane3()
    m = product.estimate_depth_points(frames, cams, sparse, f)
    ccams = [(i, c.R, c.t) for i, c in enumerate(cams)]
    base = []
    oracle.mesh(m.combined.points, m.combined.grid_sites, m.stats, ccams, 68, 48, f,
                patches=m.stats["reference_patches"], control=lambda: base.append(1) and False)
    for at in (1, len(base) // 2, len(base)):
        got = []
        for side in ("g", "o"):
            n = []

            def poll(n=n):
                n.append(1)
                return len(n) >= at

            if side == "g":
                r = product.reconstruct_dense_mesh(m.combined.points, m.combined.grid_sites,
                                                   m.stats, ccams, 68, 48, f, control=poll)
            else:
                r = oracle.mesh(m.combined.points, m.combined.grid_sites, m.stats, ccams, 68, 48,
                                f, patches=m.stats["reference_patches"], control=poll)
            got.append((r, len(n)))
        (g, gn), (o, on) = got
        assert gn == on, (at, gn, on)
        _same(g, o, f"multi cancel at {at}")
