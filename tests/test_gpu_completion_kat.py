"""verified_surface_completion_adds_expected_plane_geometry
(crates/video-to-3d-core/src/surface_completion_tests.rs:54-110) on the device, and the
device's completion equal to the oracle's on that fixture."""
import pytest

from parity import compare_completed
from test_oracle_completion_kat import check_completion_kat, completion_fixture

pytestmark = pytest.mark.gpu


def test_surface_completion_kat_on_the_device(product, oracle):
    frames, cams, sparse, f, depth = completion_fixture()
    g = product.estimate_depth_points(frames, cams, sparse, f)
    n = check_completion_kat(g.combined, depth)
    o = oracle.estimate(frames, cams, sparse, f)
    assert n == o.combined.stats["surface_completed_points"]
    compare_completed(g.combined, o.combined, "completion KAT fixture")
