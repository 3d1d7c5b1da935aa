"""The reference's rendered lateral pan (tests/golden/help_pan.npz, true poses) on the device:
sample-level parity with the oracle (the criterion of tests/parity.py), the reference's own
dense accuracy gate on the device's multi-reference estimate, and the device mesh of those
points against the oracle's."""
import pytest

import parity
from golden_data import (MAX_MEDIAN_RELATIVE_ERROR, MIN_ACCURATE_COVERAGE, accuracy,
                         help_pan, patch_points)

pytestmark = pytest.mark.gpu

PRESETS = {
    "reference": {"preset": "reference"},
    "zncc": {"preset": "zncc", "patch_radius": 3, "depth_hypotheses": 64, "grid_stride": 4},
}


@pytest.mark.parametrize("preset", sorted(PRESETS))
def test_lateral_pan_parity(product, oracle, preset):
    frames, cams, landmarks, focal, _ = help_pan()
    st = dict(PRESETS[preset], emit_samples=1, surface_completion_passes=0)
    g = product.reconstruct_reference_patch(frames, cams, landmarks, focal, 2,
                                            sweep=product.params(st))
    o, slack = oracle.reference_patch(frames, cams, landmarks, focal, 2,
                                      sweep=oracle.params(st), threads=8)
    rep = parity.compare(g, o, slack, product.params(st).cost_kind)
    assert rep["samples"] > 3000 and rep["points_compared"] > 1000, rep


@pytest.mark.parametrize("preset", sorted(PRESETS))
def test_lateral_pan_meets_the_reference_dense_gate(product, oracle, preset):
    frames, cams, landmarks, focal, z = help_pan()
    m = product.estimate_depth_points(frames, cams, landmarks, focal,
                                      sweep=product.params(PRESETS[preset]))
    assert m.stats["reference_frames"] == [2, 4, 0]
    for p in m.stats["reference_patches"]:
        pts, sites = patch_points(m, p)
        med, cov = accuracy(pts, sites, p["sampled_pixels"], p["reference_frame"], z)
        assert med <= MAX_MEDIAN_RELATIVE_ERROR and cov >= MIN_ACCURATE_COVERAGE, (p, med, cov)
    # the mesh of the device's points, on the device and on the oracle
    W, H = frames[0].shape[1], frames[0].shape[0]
    gm = product.reconstruct_dense_mesh(m.combined.points, m.combined.grid_sites, m.stats, cams,
                                        W, H, focal)
    om = oracle.mesh(m.combined.points, m.combined.grid_sites, m.stats, cams, W, H, focal,
                     patches=m.stats["reference_patches"])
    assert gm.stats == om.stats and gm.triangles.tobytes() == om.triangles.tobytes()
    assert gm.stats["accepted_triangles"] > 5000
