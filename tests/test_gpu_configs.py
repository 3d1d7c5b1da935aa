"""The batched, served and large-solve workloads against the FP64 oracle.

* 120 keyframes of one room clip (640x360 / 360x640 / 480x270, 8 nearest sources, 96
  hypotheses, 7x7 ZNCC, stride 2, FP16 texels) as one graph-replayed batch: a band of grid
  rows of one keyframe per frame size against the oracle, and those keyframes equal to the
  FP32 single solve bit for bit.
* Served requests at the serving stream's extremes (9x9 patches, 12 sources, 192
  hypotheses at stride 2 and 1; 5x5, 4 sources, 64 hypotheses; 7x7 with surface completion)
  admitted by the continuous-batching server: each equal to the request alone, and bands of
  grid rows against the oracle.
* The 1920x1080 reference + 16 sources, 256 hypotheses, 11x11 ZNCC, stride 1 solve (box
  sums) with FP16 and FP32 texels: equal bit for bit, and a band of grid rows that crosses
  the flat +-2 patches against the oracle.
"""
import os

import numpy as np
import pytest

import parity
from vt3d_dense import abi, scenes, serving

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _oracle_band(oracle, frames, cameras, sparse, focal, reference, settings, first, rows):
    oracle.set_row_band(first, rows)
    try:
        return oracle.reference_patch(frames, cameras, sparse, focal, reference,
                                      sweep=oracle.params(settings), threads=THREADS)
    finally:
        oracle.set_row_band(0, 0)


def _grid_rows(height, settings):
    r = settings["patch_radius"]
    return len(range(r + 2, height - (r + 2), settings["grid_stride"]))


def _same(a, b, what):
    assert (a is None) == (b is None), what
    if a is None:
        return
    assert a.stats == b.stats, what
    assert a.points.tobytes() == b.points.tobytes(), what
    assert a.grid_sites.tobytes() == b.grid_sites.tobytes(), what
    assert a.point_attributes.tobytes() == b.point_attributes.tobytes(), what
    if a.samples is not None or b.samples is not None:
        assert a.samples.tobytes() == b.samples.tobytes(), what


def test_batched_keyframes_match_oracle(product, oracle):
    problems = scenes.config_c3(count=120)
    frames, maps = product.pack_frames([p.frames for p in problems])
    sess = product.Session(0)
    try:
        sess.upload(frames, "f16")
        reqs = []
        for p, m in zip(problems, maps):
            st = dict(p.settings, surface_completion_passes=0, emit_samples=1)
            reqs.append({"cameras": p.cameras, "sparse_points": p.sparse, "focal": p.focal,
                         "reference_camera_index": p.reference, "frame_map": m,
                         "sweep": product.params(st)})
        batch = sess.batch(reqs, use_graph=True)
        try:
            captured = batch.run()
            res = batch.run()  # a replay of the captured graph
        finally:
            batch.close()
    finally:
        sess.close()
    assert len(res) == 120 and all(r is not None for r in res)
    for k, (a, b) in enumerate(zip(captured, res)):
        _same(a, b, f"keyframe {k}: replay")
        parity.check_counters(b.stats)
    total = sum(r.stats["sampled_pixels"] for r in res)
    assert total > 4_000_000, total
    print("batched keyframes: sampled", total, "accepted",
          sum(r.stats["accepted_points"] for r in res), flush=True)
    for W, H in scenes.C3_SIZES:
        k = next(i for i, p in enumerate(problems) if p.frames[0].shape == (H, W))
        p = problems[k]
        st = dict(p.settings, surface_completion_passes=0, emit_samples=1)
        first = _grid_rows(H, st) // 2 - 2
        o, sl = _oracle_band(oracle, p.frames, p.cameras, p.sparse, p.focal, p.reference, st,
                             first, 4)
        gb, ys = parity.band(res[k], o)
        assert len(ys) == 4
        rep = parity.compare(gb, o, sl, abi.COST_ZNCC)
        print(f"batched keyframe {k} ({W}x{H})", rep, flush=True)
        assert rep["samples"] > 300 and rep["points_compared"] > 200
        # the FP16-texel batch equals the FP32 single solve under the same launch shape
        alone = product.reconstruct_reference_patch(
            p.frames, p.cameras, p.sparse, p.focal, p.reference,
            sweep=product.params(st, storage="f32"))
        _same(res[k], alone, f"keyframe {k}: batched FP16 vs single FP32")


def _serving_requests(product):
    spec, arrivals = scenes.c4_stream(12, clips=4)
    # the stream's extremes first: 9x9 / 12 sources / 192 hypotheses (stride 2, then 1),
    # 5x5 / 4 sources / 64 hypotheses, 7x7 / 8 / 128 with two completion passes
    extremes = [(4, 12, 192, 2, 0), (4, 12, 192, 1, 0), (2, 4, 64, 2, 0), (3, 8, 128, 2, 2)]
    for a, (r, s, d, _, _) in zip(arrivals, extremes):
        a.update(patch_radius=r, sources=s, hypotheses=d)
    clips = scenes.c4_clips(spec, n_frames=16)
    frames, base = [], []
    for c in clips:
        base.append(len(frames))
        frames += c["frames"]
    requests = []
    for i, a in enumerate(arrivals):
        rq = scenes.c4_request(a, clips, base)
        st = dict(rq["settings"], emit_samples=1)
        if i < len(extremes):
            st.update(grid_stride=extremes[i][3], surface_completion_passes=extremes[i][4])
        rq["settings"] = st
        rq["sweep"] = product.params(st)
        requests.append(rq)
    return frames, requests, len(extremes)


def test_served_requests_match_oracle(product, oracle):
    frames, requests, n_checked = _serving_requests(product)
    sess = product.Session(0)
    try:
        sess.upload(frames, "f16")
        alone = []
        for rq in requests:
            b = sess.batch([rq])
            alone.append(b.run()[0])
            b.close()
        server = product.Server(sess, max_batch=8, max_inflight=2, hbm_budget_bytes=64 << 30)
        try:
            out = serving.saturate(server, requests, keep_results=True)
            st = server.stats()
        finally:
            server.close()
    finally:
        sess.close()
    done = out["completions"]
    assert sorted(done) == list(range(len(requests)))
    for i, a in enumerate(alone):
        assert done[i]["status"] == abi.VT3D_OK, (i, done[i]["status"])
        _same(done[i]["result"], a, f"request {i}: served vs alone")
    assert st["max_batch_requests_seen"] > 1
    assert requests[3]["settings"]["surface_completion_passes"] == 2
    assert alone[3].stats["surface_completed_points"] > 0
    for i in range(n_checked):
        rq, g = requests[i], done[i]["result"]
        fr = [frames[int(m)] for m in rq["frame_map"]]
        s = rq["settings"]
        H = fr[rq["reference_camera_index"]].shape[0]
        rows = 2 if s["grid_stride"] == 1 else 3
        st = dict(s, surface_completion_passes=0)
        o, sl = _oracle_band(oracle, fr, rq["cameras"], rq["sparse_points"], rq["focal"],
                             rq["reference_camera_index"], st, _grid_rows(H, st) // 2, rows)
        gp = g
        if s["surface_completion_passes"]:
            # primary samples and points only: the band oracle has no completion pass
            import copy

            gp = copy.copy(g)
            n = g.primary_points
            gp.points, gp.grid_sites = g.points[:n], g.grid_sites[:n]
        gb, ys = parity.band(gp, o)
        assert len(ys) == rows
        rep = parity.compare(gb, o, sl, abi.COST_ZNCC)
        print(f"served request {i} (P={2 * s['patch_radius'] + 1}, S={s['max_source_views']}, "
              f"D={s['depth_hypotheses']}, stride {s['grid_stride']})", rep, flush=True)
        assert rep["samples"] > 200 and rep["points_compared"] > 100


def _flat_band(gray, settings, rows):
    """First grid row of a band of `rows` grid rows in which 10-60% of the grid sites fail
    the texture gate (the flat +-2 patches), nearest the image centre."""
    from scipy.ndimage import maximum_filter, minimum_filter

    r, s = settings["patch_radius"], settings["grid_stride"]
    P = 2 * r + 1
    g = gray.astype(np.int32)
    contrast = maximum_filter(g, size=P) - minimum_filter(g, size=P)
    H, W = gray.shape
    ys = list(range(r + 2, H - (r + 2), s))
    xs = np.arange(r + 2, W - (r + 2), s)
    frac = np.array([np.mean(contrast[y, xs] < 14) for y in ys])
    ok = [i for i in range(len(ys) - rows)
          if np.all((frac[i:i + rows] >= 0.10) & (frac[i:i + rows] <= 0.60))]
    assert ok, "no band crosses the flat patches"
    mid = len(ys) // 2
    return min(ok, key=lambda i: abs(i - mid))


def test_large_solve_matches_oracle(product, oracle):
    p = scenes.config_c5()
    st = dict(p.settings, surface_completion_passes=0, emit_samples=1)
    res = {}
    sess = product.Session(0)
    try:
        for storage in ("f16", "f32"):
            sess.upload(p.frames, storage)
            res[storage] = sess.solve(p.cameras, p.sparse, p.focal, p.reference,
                                      sweep=product.params(st, storage=storage))
    finally:
        sess.close()
    _same(res["f16"], res["f32"], "FP16 vs FP32 texels")
    g = res["f16"]
    parity.check_counters(g.stats)
    first = _flat_band(p.frames[p.reference], st, 4)
    o, sl = _oracle_band(oracle, p.frames, p.cameras, p.sparse, p.focal, p.reference, st,
                         first, 4)
    gb, ys = parity.band(g, o)
    assert len(ys) == 4
    nx = len(range(st["patch_radius"] + 2, p.frames[0].shape[1] - (st["patch_radius"] + 2)))
    gated = rep_samples = len(o.samples)
    assert gated < 0.9 * 4 * nx, "the band holds sites that fail the texture gate"
    rep = parity.compare(gb, o, sl, abi.COST_ZNCC)
    print(f"large solve band rows {first}..{first + 3}", rep, flush=True)
    assert rep_samples > 3000 and rep["points_compared"] > 2000
