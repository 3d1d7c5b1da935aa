"""vt3d_dense_plane_sweep_noalloc (through the C ABI) against vt3d_dense_plane_sweep: the same
bits on two calls in a row, the second smaller; a solve larger than the scratch and arrays
smaller than the result are refused with VT3D_E_BUDGET."""
import numpy as np
import pytest

from vt3d_dense import abi, scenes

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    assert a.skip_code == b.skip_code, what
    assert a.stats == b.stats, what
    assert a.primary_points == b.primary_points, what
    assert a.points.tobytes() == b.points.tobytes(), what
    assert a.grid_sites.tobytes() == b.grid_sites.tobytes(), what
    assert a.point_attributes.tobytes() == b.point_attributes.tobytes(), what
    assert (a.samples is None) == (b.samples is None), what
    if a.samples is not None:
        assert a.samples.tobytes() == b.samples.tobytes(), what


def _cases():
    c1 = scenes.config_c1()
    pf = scenes.plane_fixture()
    # 360x202, 5 views, 7x7 ZNCC, 64 hypotheses, two completion passes; then the 68x48
    # reference-preset plane fixture (smaller in every buffer)
    return [(c1, dict(c1.settings, surface_completion_passes=2, emit_samples=1)),
            (pf, {"preset": "reference", "emit_samples": 1})]


def test_noalloc_matches_plane_sweep(product):
    for rnd in range(2):
        for k, (p, st) in enumerate(_cases()):
            sw = product.params(st)
            want = product.estimate_depth_points_with_control(p.frames, p.cameras, p.sparse,
                                                              p.focal, sweep=sw)
            rc, got = product.estimate_depth_points_noalloc(p.frames, p.cameras, p.sparse,
                                                            p.focal, sweep=sw)
            assert rc == abi.VT3D_OK, (rnd, k, rc)
            assert len(want.points) > 10
            _same(got, want, f"round {rnd} case {k}")


def test_noalloc_limits(product):
    p, st = _cases()[1]
    sw = product.params(st)
    want = product.estimate_depth_points_with_control(p.frames, p.cameras, p.sparse, p.focal,
                                                      sweep=sw)
    n = len(want.points)
    rc, got = product.estimate_depth_points_noalloc(p.frames, p.cameras, p.sparse, p.focal,
                                                    sweep=sw, point_capacity=n - 1)
    assert rc == abi.VT3D_E_BUDGET
    assert got.stats == want.stats and len(got.points) == 0
    # frames beyond the scratch: 5 views of 1920x1080 hold 216 MB as FP32 texels
    c5 = scenes.config_c5(sources=4)
    rc, got = product.estimate_depth_points_noalloc(
        c5.frames, c5.cameras, c5.sparse, c5.focal,
        sweep=product.params(dict(c5.settings, surface_completion_passes=0)))
    assert rc == abi.VT3D_E_BUDGET
    assert "bytes of scratch" in product.library().vt3d_last_error().decode()
    # and the scratch serves a later call as before
    rc, got = product.estimate_depth_points_noalloc(p.frames, p.cameras, p.sparse, p.focal,
                                                    sweep=sw)
    assert rc == abi.VT3D_OK
    _same(got, want, "after a refused call")
