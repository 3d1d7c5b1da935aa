"""Batches: uneven solves in one kernel pipeline (and its captured HIP graph) give the same
bits as each solve run alone under the same launch shape, and agree within the parity
tolerance across launch shapes (the tiled sweep warps each tile's region in FP32 about the
region's centre, so the tile width moves the last bits of a cost)."""
import numpy as np
import pytest

from vt3d_dense import scenes

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    assert (a is None) == (b is None), what
    if a is None:
        return
    assert a.skip_code == b.skip_code, what
    assert a.stats == b.stats, what
    assert a.points.tobytes() == b.points.tobytes(), what
    assert a.grid_sites.tobytes() == b.grid_sites.tobytes(), what
    assert a.point_attributes.tobytes() == b.point_attributes.tobytes(), what
    if a.samples is not None or b.samples is not None:
        assert a.samples.tobytes() == b.samples.tobytes(), what


def _mix(product):
    """Six batched-config keyframes (three frame sizes, 8 sources, 96 hypotheses, 7x7 ZNCC),
    the plain CPU-size config (4 sources, 64 hypotheses), the same at stride 1 with 5x5
    patches (the box-sum sweep, a launch group of its own), the 2-source ZSAD plane fixture
    on the plain-sweep entry, a remap that returns None and a skipped solve."""
    c3 = scenes.config_c3(count=6)
    c1 = scenes.config_c1()
    pf = scenes.plane_fixture()
    jobs = []
    for p in c3 + [c1]:
        jobs.append(dict(frames=p.frames, cameras=p.cameras, sparse_points=p.sparse,
                         focal=p.focal, reference_camera_index=p.reference,
                         settings=dict(p.settings)))
    jobs.append(dict(frames=c1.frames, cameras=c1.cameras, sparse_points=c1.sparse,
                     focal=c1.focal, reference_camera_index=c1.reference,
                     settings=dict(c1.settings, grid_stride=1, patch_radius=2,
                                   depth_hypotheses=32)))
    ref = {"preset": "reference"}
    jobs.append(dict(frames=pf.frames, cameras=pf.cameras, sparse_points=pf.sparse,
                     focal=pf.focal, reference_camera_index=-1, settings=ref))
    jobs.append(dict(frames=c1.frames, cameras=c1.cameras, sparse_points=c1.sparse[:5],
                     focal=c1.focal, reference_camera_index=c1.reference,
                     settings=dict(c1.settings)))
    jobs.append(dict(frames=pf.frames, cameras=pf.cameras[:1], sparse_points=pf.sparse,
                     focal=pf.focal, reference_camera_index=-1, settings=ref))
    for j in jobs:
        j["settings"].update(surface_completion_passes=0, emit_samples=1, storage="f16")
        j["sweep"] = product.params(j["settings"])
    return jobs


def _alone(product, j):
    if j["reference_camera_index"] >= 0:
        return product.reconstruct_reference_patch(j["frames"], j["cameras"], j["sparse_points"],
                                                   j["focal"], j["reference_camera_index"],
                                                   sweep=j["sweep"])
    return product.estimate_depth_points_with_control(j["frames"], j["cameras"],
                                                      j["sparse_points"], j["focal"],
                                                      sweep=j["sweep"])


@pytest.fixture(scope="module")
def mix(product):
    jobs = _mix(product)
    alone = [_alone(product, j) for j in jobs]
    return jobs, alone


def _batch(product, jobs, use_graph, tuning=None):
    frames, maps = product.pack_frames([j["frames"] for j in jobs])
    sess = product.Session(0)
    sess.upload(frames, "f16")
    if tuning:
        sess.set_tuning(**tuning)
    reqs = [dict(j, frame_map=m) for j, m in zip(jobs, maps)]
    return sess, sess.batch(reqs, use_graph=use_graph)


def _alone_tuned(product, jobs, tuning):
    """Each job as a batch of one on a session with `tuning`."""
    out = []
    for j in jobs:
        frames, maps = product.pack_frames([j["frames"]])
        sess = product.Session(0)
        try:
            sess.upload(frames, "f16")
            sess.set_tuning(**tuning)
            b = sess.batch([dict(j, frame_map=maps[0])])
            out.append(b.run()[0])
            b.close()
        finally:
            sess.close()
    return out


@pytest.mark.parametrize("tuning", [None, {"tile_width": 16, "target_waves": 16384},
                                    {"tile_width": 64, "target_waves": 1}])
def test_batch_matches_each_solve_alone(product, mix, tuning):
    jobs, alone = mix
    if tuning is not None:
        alone = _alone_tuned(product, jobs, tuning)
    sess, b = _batch(product, jobs, False, tuning)
    try:
        got = b.run()
        assert len(got) == len(jobs)
        for i, (g, a) in enumerate(zip(got, alone)):
            _same(g, a, f"request {i}")
        assert got[-2] is None and got[-1].skip_code == 2
        assert sum(len(g.points) for g in got[:7]) > 1000
        t = b.timing()
        assert t.units == sum(g.stats["sampled_pixels"] * g.stats["depth_hypotheses"]
                              for g in got if g is not None)
        assert t.sweep_ms > 0 and t.total_ms > 0
    finally:
        b.close()
        sess.close()


def test_graph_replays_match_eager(product, mix):
    jobs, alone = mix
    sess, b = _batch(product, jobs, True)
    try:
        for replay in range(3):
            got = b.run()
            for i, (g, a) in enumerate(zip(got, alone)):
                _same(g, a, f"replay {replay} request {i}")
            # eager work of the caller between replays
            one = _alone(product, jobs[0])
            _same(one, alone[0], "eager between replays")
    finally:
        b.close()
        sess.close()


def test_launch_shapes_agree_within_tolerance(product, mix):
    """Tile widths 16 and 64 against the default 32: the same samples, the same outcome and
    hypothesis wherever the best/second cost gap exceeds the FP32 cost error, costs within
    it, and the same hypothesis blocking never changes a bit (target waves only split the
    hypotheses between work items)."""
    jobs, alone = mix
    for tuning in ({"tile_width": 16, "target_waves": 4096},
                   {"tile_width": 64, "target_waves": 4096}):
        sess, b = _batch(product, jobs, False, tuning)
        try:
            got = b.run()
        finally:
            b.close()
            sess.close()
        for i, (g, a) in enumerate(zip(got, alone)):
            if a is None or a.samples is None:
                assert (g is None) == (a is None)
                continue
            tau = 1.0e-3 if jobs[i]["sweep"].cost_kind == 0 else 1.0e-4
            gs, as_ = g.samples, a.samples
            assert np.array_equal(gs[["x", "y"]], as_[["x", "y"]]), i
            tie = np.abs(as_["second_cost"] - as_["best_cost"]) < tau
            tie |= np.abs(gs["second_cost"] - gs["best_cost"]) < tau
            same = (gs["status"] == as_["status"]) & (gs["best_hypothesis"] ==
                                                       as_["best_hypothesis"])
            # Other thresholds (the ambiguity ratio, the photometric gate, a reverse-sweep or
            # fusion test) can sit within the cost error too; at 1e-7 relative they are rare.
            odd = ~same & ~tie
            assert odd.sum() <= max(1, len(gs) // 500), (i, tuning, int(odd.sum()))
            both = same & np.isfinite(as_["best_cost"]) & np.isfinite(gs["best_cost"])
            assert np.abs(gs["best_cost"][both] - as_["best_cost"][both]).max(initial=0.0) \
                < tau, (i, tuning)
            assert abs(len(g.points) - len(a.points)) <= max(2, len(a.points) // 200), i
        blocked = _alone_tuned(product, jobs[:1], {"tile_width": 32, "target_waves": 1})
        _same(blocked[0], alone[0], "hypothesis blocking")


def test_batches_in_flight_together(product, mix):
    """Two batches of one session submitted together (each on its own stream) give the same
    bits as each solve alone, whichever is collected first; misuse is refused."""
    jobs, alone = mix
    frames, maps = product.pack_frames([j["frames"] for j in jobs])
    sess = product.Session(0)
    try:
        sess.upload(frames, "f16")
        reqs = [dict(j, frame_map=m) for j, m in zip(jobs, maps)]
        first, second = list(range(0, 6)), list(range(6, len(jobs)))
        a = sess.batch([reqs[i] for i in first], use_graph=True)
        b = sess.batch([reqs[i] for i in second])
        with pytest.raises(product.Vt3dError):
            a.collect()  # nothing submitted
        for order in ("ab", "ba"):
            a.submit()
            b.submit()
            with pytest.raises(product.Vt3dError):
                a.submit()  # already in flight
            with pytest.raises(product.Vt3dError):
                a.run()
            with pytest.raises(product.Vt3dError):
                sess.upload(frames, "f16")  # frames in use by batches in flight
            got = {}
            for k in order:
                got[k] = (a if k == "a" else b).collect()
            for i, g in zip(first, got["a"]):
                _same(g, alone[i], f"{order}: request {i}")
            for i, g in zip(second, got["b"]):
                _same(g, alone[i], f"{order}: request {i}")
        sess.upload(frames, "f16")  # nothing in flight
        _same(a.run()[0], alone[0], "run after collect")
        a.close()
        b.close()
    finally:
        sess.close()


def test_pooled_buffers_serve_later_batches(product, mix):
    """A destroyed batch's buffers serve the session's next batches: a small batch after a
    large one, a large one after a small one, and a third batch while two are alive."""
    jobs, alone = mix
    frames, maps = product.pack_frames([j["frames"] for j in jobs])
    sess = product.Session(0)
    try:
        sess.upload(frames, "f16")
        reqs = [dict(j, frame_map=m) for j, m in zip(jobs, maps)]
        for sel in (list(range(len(jobs))), [6, 7], [0, 1, 2, 3, 4, 5, 7], [8]):
            b = sess.batch([reqs[i] for i in sel])
            for i, g in zip(sel, b.run()):
                _same(g, alone[i], f"{sel}: request {i}")
            b.close()
        x, y = sess.batch([reqs[0]]), sess.batch(reqs[1:3])
        z = sess.batch(reqs[3:7])
        for bt, sel in ((x, [0]), (y, [1, 2]), (z, [3, 4, 5, 6])):
            bt.submit()
        for bt, sel in ((z, [3, 4, 5, 6]), (x, [0]), (y, [1, 2])):
            for i, g in zip(sel, bt.collect()):
                _same(g, alone[i], f"three alive: request {i}")
        for bt in (x, y, z):
            bt.close()
    finally:
        sess.close()
