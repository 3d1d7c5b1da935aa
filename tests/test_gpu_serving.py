"""Serving loop: requests of uneven size, sources, hypotheses and patch size admitted by
continuous batching give the same bits as each request solved alone; HBM-budget admission;
open-loop latency accounting."""
import numpy as np
import pytest

from vt3d_dense import abi, scenes, serving

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c4(product):
    spec, arrivals = scenes.c4_stream(40, clips=5)
    clips = scenes.c4_clips(spec, n_frames=14)
    frames, base = [], []
    for c in clips:
        base.append(len(frames))
        frames += c["frames"]
    sess = product.Session(0)
    sess.upload(frames, "f16")
    requests = []
    for i, a in enumerate(arrivals):
        rq = scenes.c4_request(a, clips, base)
        if i % 8 == 3:  # some requests at stride 1: the box-sum sweep in the same batches
            rq["settings"] = dict(rq["settings"], grid_stride=1)
        if i % 8 == 5:  # and some with the reference's two surface-completion passes
            rq["settings"] = dict(rq["settings"], surface_completion_passes=2)
        rq["sweep"] = product.params(rq["settings"])
        requests.append(rq)
    # each request alone: a batch of one on the session (before any server runs on it)
    alone = []
    for rq in requests:
        b = sess.batch([rq])
        alone.append(b.run()[0])
        b.close()
    yield sess, requests, arrivals, alone
    sess.close()


def _same(a, b, what):
    assert (a is None) == (b is None), what
    if a is None:
        return
    assert a.stats == b.stats, what
    assert a.points.tobytes() == b.points.tobytes(), what
    assert a.grid_sites.tobytes() == b.grid_sites.tobytes(), what
    assert a.point_attributes.tobytes() == b.point_attributes.tobytes(), what


def test_served_results_match_alone(product, c4):
    sess, requests, arrivals, alone = c4
    sizes = {(r["settings"]["patch_radius"], r["settings"]["max_source_views"]) for r in requests}
    assert len(sizes) > 3, "the mix spans patch sizes and source counts"
    assert {r["settings"]["grid_stride"] for r in requests} == {1, 2}
    server = product.Server(sess, max_batch=8, max_inflight=2)
    try:
        out = serving.saturate(server, requests, keep_results=True)
        st = server.stats()
    finally:
        server.close()
    done = out["completions"]
    assert sorted(done) == list(range(len(requests)))
    for i, a in enumerate(alone):
        assert done[i]["status"] == abi.VT3D_OK, done[i]["status"]
        _same(done[i]["result"], a, f"request {i}")
        assert done[i]["units"] == a.stats["sampled_pixels"] * a.stats["depth_hypotheses"]
    assert sum(len(a.points) for a in alone) > 1000
    assert st["batches"] >= len(requests) // 8 and st["max_batch_requests_seen"] > 1
    assert st["completed"] == st["submitted"] == len(requests)
    assert out["latency_ms"]["p99"] >= out["latency_ms"]["p50"] > 0


def test_launches_share_the_backlog(product, c4):
    # A launch takes its share of the queue over two launches per in-flight slot (at most
    # max_batch): 40 requests at once on 2 slots run as launches of at most 10, not one of 40.
    sess, requests, arrivals, alone = c4
    server = product.Server(sess, max_batch=64, max_inflight=2)
    try:
        out = serving.saturate(server, requests, keep_results=True)
        st = server.stats()
    finally:
        server.close()
    n = len(requests)
    assert st["completed"] == n
    assert st["max_batch_requests_seen"] <= (n + 3) // 4
    assert st["batches"] >= 4
    done = out["completions"]
    for i in range(0, n, 7):
        assert done[i]["status"] == abi.VT3D_OK
        _same(done[i]["result"], alone[i], f"request {i}")


def test_budget_admission(product, c4):
    """A slot budget that holds only the smaller requests: those are served (and match),
    the rest complete with VT3D_E_BUDGET."""
    sess, requests, arrivals, alone = c4
    probe = product.Server(sess)
    frame_bytes = probe.stats()["frame_bytes"]
    probe.close()
    server = product.Server(sess, max_batch=8, max_inflight=1,
                            hbm_budget_bytes=frame_bytes + (24 << 20))
    try:
        out = serving.saturate(server, requests, keep_results=True)
        st = server.stats()
    finally:
        server.close()
    done = out["completions"]
    ok = [i for i in done if done[i]["status"] == abi.VT3D_OK]
    over = [i for i in done if done[i]["status"] == abi.VT3D_E_BUDGET]
    assert ok and over and len(ok) + len(over) == len(requests)
    assert st["rejected_budget"] == len(over)
    # frames + the one slot's buffers never exceed the budget
    assert st["workspace_bytes"] <= 24 << 20, st
    for i in ok:
        _same(done[i]["result"], alone[i], f"request {i}")
    px = lambda i: alone[i].stats["sampled_pixels"] * len(requests[i]["cameras"])  # noqa: E731
    assert np.mean([px(i) for i in ok]) < np.mean([px(i) for i in over])


def test_open_loop_latency(product, c4):
    sess, requests, arrivals, alone = c4
    server = product.Server(sess)
    try:
        t = [a["arrival"] for a in arrivals]
        out = serving.open_loop(server, requests, t, warmup=5)
    finally:
        server.close()
    assert out["requests"] == len(requests) - 5
    assert out["statuses"] == {str(abi.VT3D_OK): len(requests) - 5}
    lat = out["latency_ms"]
    assert 0 < lat["p50"] <= lat["p99"] <= lat["max"]
    assert out["units_per_s"] > 0
