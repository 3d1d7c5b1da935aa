"""The ZNCC reverse sweep's partial-patch bound (vt3d_tuning.reverse_bound) drops only
hypotheses that cannot be the sweep's first minimum, so every result is bit-identical with
the bound on (the default) and off: chosen reverse hypotheses (the samples' reciprocal
signatures), reciprocal counters, points and attributes. Covered: 3x3 (ZSAD, where the bound
does not apply), 3x3 to 11x11 ZNCC, 1 to 130 hypotheses, short pair lists, the 120-keyframe
batch and the 1920x1080 solve."""
import pytest

from vt3d_dense import scenes

from test_gpu_batch import _mix, _same

pytestmark = pytest.mark.gpu


def _run(product, frames, reqs, storage, bound):
    sess = product.Session(0)
    try:
        sess.upload(frames, storage)
        sess.set_tuning(reverse_bound=bound)
        b = sess.batch(reqs)
        try:
            out = b.run()
            t = b.timing()
        finally:
            b.close()
    finally:
        sess.close()
    return out, t


def test_mix_same_with_and_without_bound(product):
    jobs = _mix(product)
    frames, maps = product.pack_frames([j["frames"] for j in jobs])
    reqs = [dict(j, frame_map=m) for j, m in zip(jobs, maps)]
    on, _ = _run(product, frames, reqs, "f16", 0)
    off, _ = _run(product, frames, reqs, "f16", -1)
    for i, (a, b) in enumerate(zip(on, off)):
        _same(a, b, f"request {i}")
    assert sum(a.stats["reciprocal_consistent_points"] for a in on if a is not None) > 1000


def test_batched_keyframes_same_with_and_without_bound(product):
    problems = scenes.config_c3(count=120)
    frames, maps = product.pack_frames([p.frames for p in problems])
    reqs = []
    for p, m in zip(problems, maps):
        st = dict(p.settings, surface_completion_passes=0, emit_samples=1)
        reqs.append({"cameras": p.cameras, "sparse_points": p.sparse, "focal": p.focal,
                     "reference_camera_index": p.reference, "frame_map": m,
                     "sweep": product.params(st)})
    on, t_on = _run(product, frames, reqs, "f16", 0)
    off, t_off = _run(product, frames, reqs, "f16", -1)
    for i, (a, b) in enumerate(zip(on, off)):
        _same(a, b, f"keyframe {i}")
    print(f"reverse sweep: bound {t_on.reverse_sweep_ms:.1f} ms, none {t_off.reverse_sweep_ms:.1f} ms")


def test_large_solve_same_with_and_without_bound(product):
    p = scenes.config_c5()
    st = dict(p.settings, surface_completion_passes=0, emit_samples=1, storage="f16")
    req = {"cameras": p.cameras, "sparse_points": p.sparse, "focal": p.focal,
           "reference_camera_index": p.reference, "frame_map": list(range(len(p.frames))),
           "sweep": product.params(st)}
    on, t_on = _run(product, p.frames, [req], "f16", 0)
    off, t_off = _run(product, p.frames, [req], "f16", -1)
    _same(on[0], off[0], "1920x1080")
    assert on[0].stats["reciprocal_checked_points"] > 100000
    print(f"reverse sweep: bound {t_on.reverse_sweep_ms:.1f} ms, none {t_off.reverse_sweep_ms:.1f} ms")


def test_queue_edges_same_with_and_without_bound(product):
    # The bounded sweep's continuation queue at its edges: 1 to 130 hypotheses (a queue
    # drained only after the last step, partly filled launches of 64 entries, hypothesis
    # counts on either side of 64), 3x3 to 11x11 patches (bound after 1 to 3 rows), and
    # lists shorter than a wave (stride 8 leaves a few hundred samples per source).
    c1 = scenes.config_c1()
    frames, maps = product.pack_frames([c1.frames])
    reqs = []
    for r in (1, 2, 3, 4, 5):
        for d in (1, 2, 3, 63, 65, 130):
            for stride in ((4, 8) if d in (3, 65) else (4,)):
                st = dict(c1.settings, patch_radius=r, depth_hypotheses=d, grid_stride=stride,
                          surface_completion_passes=0, emit_samples=1, storage="f16")
                reqs.append({"cameras": c1.cameras, "sparse_points": c1.sparse,
                             "focal": c1.focal, "reference_camera_index": c1.reference,
                             "frame_map": maps[0], "sweep": product.params(st)})
    on, _ = _run(product, frames, reqs, "f16", 0)
    off, _ = _run(product, frames, reqs, "f16", -1)
    checked = 0
    for i, (a, b) in enumerate(zip(on, off)):
        _same(a, b, f"request {i}")
        checked += a.stats["reciprocal_checked_points"] if a is not None else 0
    assert checked > 10000
