"""Comparison of a GPU solve with the FP64 oracle on the same inputs.

A texture-gated sample must reach the same outcome (status) and choose the same hypothesis
index on both sides, unless the oracle's decision slack shows that the sample sat within
FP32 rounding of a decision: a photometric gate, a best/second tie, the ambiguity gate, the
reciprocal texture gate or reverse-sweep tie (cost units, `tau_cost`), a patch tap on the
bilinear-bounds edge (`tau_px` source pixels), or the reciprocal-agreement / fusion-radius
tests (`tau_rel`). Fused points of samples that agree on hypothesis and reciprocal set must
match within 1e-4 relative error.
"""
from __future__ import annotations

import numpy as np

# Decision margin below which a disagreement is a near tie. ZNCC costs (1 - ZNCC) lie in
# [0, 2]: the margin is the north star's 1e-5. ZSAD costs are mean intensity differences on
# the 0..255 scale, 127.5 times the ZNCC range, where the same relative margin would be
# 1.275e-3; 1e-3 is used.
TAU_COST = {0: 1.0e-3, 1: 1.0e-5}
# Reverse-sweep ties that only move a reciprocal observation: the FP32 cost error must
# exceed the gap to flip them, so a tighter bound (a few times the cost error) applies.
TAU_POS = {0: 2.0e-4, 1: 2.0e-5}
TAU_PX = 1.0e-3
TAU_REL = 1.0e-5
POINT_REL = 1.0e-4


def compare(gpu, oracle, slack, cost_kind: int, max_unexplained: int = 0):
    """Returns a report dict; raises AssertionError on a parity violation."""
    gs, os_ = gpu.samples, oracle.samples
    assert gs is not None and os_ is not None, "both sides must emit samples"
    assert gpu.stats["sampled_pixels"] == oracle.stats["sampled_pixels"]
    assert len(gs) == len(os_)
    assert np.array_equal(gs["x"], os_["x"]) and np.array_equal(gs["y"], os_["y"])
    agree = (gs["status"] == os_["status"]) & (gs["best_hypothesis"] == os_["best_hypothesis"])
    near = ((slack["cost"] < TAU_COST[cost_kind]) | (slack["px"] < TAU_PX)
            | (slack["rel"] < TAU_REL))
    unexplained = ~agree & ~near
    report = {"samples": int(len(gs)), "agree": float(agree.mean()) if len(gs) else 1.0,
              "near_ties": int(near.sum()), "disagree": int((~agree).sum()),
              "unexplained": int(unexplained.sum())}
    # Disagreements whose oracle decision margin is at least the north star's 1e-5 and that
    # no edge or relative-test slack explains, and the margins of every disagreement.
    edge = (slack["px"] < TAU_PX) | (slack["rel"] < TAU_REL)
    report["disagree_margin_ge_1e-5"] = int((~agree & ~edge & (slack["cost"] >= 1e-5)).sum())
    report["disagree_cost_slack"] = sorted(float(v) for v in slack["cost"][~agree])[:12]
    report["disagree_edge"] = int((~agree & edge).sum())
    both = agree & (gs["best_hypothesis"] >= 0) & np.isfinite(os_["best_cost"])
    ce = np.abs(gs["best_cost"][both].astype(np.float64) - os_["best_cost"][both].astype(np.float64))
    report["best_cost_abs_err_max"] = float(ce.max()) if ce.size else 0.0
    assert report["unexplained"] <= max_unexplained, (report, _describe(gs, os_, slack,
                                                                        np.nonzero(unexplained)[0]))
    # Points: accepted on both sides with the same hypothesis and reciprocal set, and no near
    # tie (a reverse-sweep tie moves a reciprocal observation without changing any index
    # reported here).
    same = (agree & ~near & (gs["status"] == 6)
            & (gs["reciprocal_mask"] == os_["reciprocal_mask"])
            & (gs["reciprocal_signature"] == os_["reciprocal_signature"]))
    # Reverse-sweep ties below the cost error may pick a neighbouring hypothesis; away from
    # such ties the reverse choices must agree.
    sig_same = gs["reciprocal_signature"] == os_["reciprocal_signature"]
    clear = agree & ~near & (slack["pos"] >= TAU_POS[cost_kind]) & (gs["status"] >= 4)
    report["reverse_agree_clear"] = float(sig_same[clear].mean()) if clear.any() else 1.0
    report["reverse_agree_all"] = float(sig_same[agree & (gs["status"] >= 4)].mean()) \
        if (agree & (gs["status"] >= 4)).any() else 1.0
    assert report["reverse_agree_clear"] == 1.0, report
    report["points_compared"] = int(same.sum())
    gp = _points_by_site(gpu)
    op = _points_by_site(oracle)
    worst = 0.0
    rgb_diff = 0
    for i in np.nonzero(same)[0]:
        key = (int(gs["x"][i]), int(gs["y"][i]))
        a, b = gp[key], op[key]
        va = np.array([a["x"], a["y"], a["z"]], dtype=np.float64)
        vb = np.array([b["x"], b["y"], b["z"]], dtype=np.float64)
        rel = np.linalg.norm(va - vb) / max(np.linalg.norm(vb), 1e-12)
        worst = max(worst, rel)
        # sample_rgb (dense_legacy.rs:1159-1166): the reference frame's bytes at the site
        rgb_diff += int((a["r"], a["g"], a["b"]) != (b["r"], b["g"], b["b"]))
    report["point_rel_max"] = worst
    report["rgb_mismatch"] = rgb_diff
    assert worst <= POINT_REL, report
    assert rgb_diff == 0, report
    return report


def _points_by_site(res):
    return {(int(s["x"]), int(s["y"])): p for s, p in zip(res.grid_sites, res.points)}


def _describe(gs, os_, slack, idx):
    rows = []
    for i in idx[:6]:
        rows.append(f"i={i} ({gs['x'][i]},{gs['y'][i]}) gpu st={gs['status'][i]} "
                    f"h={gs['best_hypothesis'][i]} c={gs['best_cost'][i]:.6g} | oracle "
                    f"st={os_['status'][i]} h={os_['best_hypothesis'][i]} "
                    f"c={os_['best_cost'][i]:.6g} slack={tuple(slack[i])}")
    return "\n".join(rows)


def stats_equal(a: dict, b: dict, keys=None):
    keys = keys or ["attempted", "reference_frame", "source_views", "source_frames",
                    "sampled_pixels", "depth_hypotheses", "grid_stride", "grid_border",
                    "skip_reason"]
    for k in keys:
        assert a[k] == b[k], (k, a[k], b[k])


COMPLETION_KEYS = ("accepted_points", "surface_completion_proposals", "surface_completed_points",
                   "surface_completion_rejected_texture",
                   "surface_completion_rejected_cross_view",
                   "surface_completion_rejected_reciprocal",
                   "surface_completion_rejected_fusion",
                   "surface_completion_rejected_footprint")


def compare_completed(g, o, what):
    """Counters, sites and order exactly; points and attributes within FP32 error of the
    oracle. For fixtures without near ties (surface completion appends points in a fixed
    order, so the arrays are compared element by element)."""
    for k in COMPLETION_KEYS:
        assert g.stats[k] == o.stats[k], (what, k, g.stats[k], o.stats[k])
    assert g.primary_points == o.primary_points, what
    assert np.array_equal(g.grid_sites, o.grid_sites), what
    gp = np.stack([g.points["x"], g.points["y"], g.points["z"]], 1).astype(np.float64)
    op = np.stack([o.points["x"], o.points["y"], o.points["z"]], 1).astype(np.float64)
    if len(gp):
        rel = np.linalg.norm(gp - op, axis=1) / np.maximum(np.linalg.norm(op, axis=1), 1e-12)
        assert rel.max() <= 1e-4, (what, rel.max())
        assert np.abs(g.points["confidence"] - o.points["confidence"]).max() <= 1e-4, what
        assert np.array_equal(g.points[["r", "g", "b"]], o.points[["r", "g", "b"]]), what
        # depth_margin is (second - best cost) / saturation: FP32 costs (|error| ~1e-5)
        # over the ZNCC saturation 0.15
        for f in ("reciprocal_consistency", "depth_margin"):
            assert np.abs(g.point_attributes[f] - o.point_attributes[f]).max() <= 1e-4, (what, f)
    for k in ("median_supporting_views", "median_fusion_observations"):
        assert g.stats[k] == o.stats[k], (what, k)
    ge, oe = g.stats["median_photometric_error"], o.stats["median_photometric_error"]
    assert (ge is None) == (oe is None) and (ge is None or abs(ge - oe) <= 1e-3), (what, ge, oe)


def band(g, o):
    """The GPU result restricted to the grid rows the oracle evaluated (a row band)."""
    import copy

    ys = np.unique(o.samples["y"])
    gb = copy.copy(g)
    keep = np.isin(g.samples["y"], ys)
    gb.samples = g.samples[keep]
    pk = np.isin(g.grid_sites["y"], ys)
    gb.points, gb.grid_sites = g.points[pk], g.grid_sites[pk]
    gb.stats = dict(g.stats, sampled_pixels=int(keep.sum()))
    return gb, ys


def check_counters(s):
    """Counter identities of dense_legacy.rs:1282-1308."""
    assert s["reciprocal_checked_points"] == (s["reciprocal_consistent_points"]
                                              + s["reciprocal_rejected_points"])
    assert s["reciprocal_consistent_points"] == s["accepted_points"] + s["fusion_rejected_points"]
    assert s["fusion_input_observations"] >= 2 * s["reciprocal_consistent_points"]
