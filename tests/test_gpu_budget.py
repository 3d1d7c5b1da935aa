"""The budgeted dense entry on the device (vt3d_dense_estimate_with_budget,
estimate_depth_points_with_budget_and_control, dense.rs:354-420) against the oracle's, at the
budgets of the reference's tests: below the one-view estimate, exactly one, two and three
views, and unlimited."""
import pytest

from test_gpu_multi_reference import _compare_multi
from test_oracle_budget import _plane

pytestmark = pytest.mark.gpu

BUDGET_KEYS = ("working_set_estimate", "working_set_budget_bytes", "full_working_set_bytes",
               "reference_view_limit")


@pytest.mark.parametrize("centers", [(0.0, 0.24, -0.22), (0.0, 0.24)])
def test_budgeted_estimate_matches_oracle(product, oracle, centers):
    frames, cams, sparse, f = _plane(centers)
    views = [oracle.estimate_working_set(frames, 320, 3, v, False)["total_bytes"]
             for v in (1, 2, 3)]
    unbudgeted = product.estimate_depth_points(frames, cams, sparse, f)
    for budget in [1, views[0] - 1] + views + [2 ** 64 - 1]:
        g = product.estimate_depth_points_with_budget(frames, cams, sparse, f, budget, 320, 3)
        o = oracle.estimate_with_budget(frames, cams, sparse, f, budget, 320, 3)
        what = f"centres {centers}, budget {budget}"
        for k in BUDGET_KEYS:
            assert g.stats[k] == o.stats[k], (what, k, g.stats[k], o.stats[k])
        assert g.combined.skip_code == o.combined.skip_code, what
        if not o.stats["attempted"]:
            assert not g.stats["attempted"] and len(g.combined.points) == 0, what
            assert g.stats["skip_reason"] == o.stats["skip_reason"], what
            continue
        _compare_multi(g, o, what)
        if budget >= views[2]:
            # every view allowed: the unbudgeted run, bit for bit
            for k in ("points", "grid_sites", "point_attributes"):
                assert (getattr(g.combined, k).tobytes()
                        == getattr(unbudgeted.combined, k).tobytes()), (what, k)
