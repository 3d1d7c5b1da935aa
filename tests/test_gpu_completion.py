"""Surface completion on the device (dense_legacy.rs:416-432, 519-753) against the FP64
oracle: counters, the appended points and their order, the reference's default parameters
(two passes) through the drop-in entry points, cancellation inside a completion pass, and
completion inside batches."""
import numpy as np
import pytest

from vt3d_dense import scenes

from parity import compare_completed as _compare

pytestmark = pytest.mark.gpu


def test_plane_fixture_default_parameters(product, oracle):
    """params = NULL: the reference constants, two completion passes."""
    pf = scenes.plane_fixture()
    g = product.estimate_depth_points_with_control(pf.frames, pf.cameras, pf.sparse, pf.focal)
    o, _ = oracle.plane_sweep(pf.frames, pf.cameras, pf.sparse, pf.focal)
    _compare(g, o, "plane fixture")
    assert g.stats["attempted"] and len(g.points) >= 12


@pytest.mark.parametrize("preset", ["reference", "zncc"])
def test_c1_completion_matches_oracle(product, oracle, preset):
    c1 = scenes.config_c1()
    st = dict(c1.settings) if preset == "zncc" else {"preset": "reference"}
    st.update(surface_completion_passes=2)
    g = product.reconstruct_reference_patch(c1.frames, c1.cameras, c1.sparse, c1.focal,
                                            c1.reference, sweep=product.params(st))
    o, _ = oracle.reference_patch(c1.frames, c1.cameras, c1.sparse, c1.focal, c1.reference,
                                  sweep=oracle.params(st), threads=8)
    _compare(g, o, f"c1 {preset}")
    s = g.stats
    assert s["surface_completion_proposals"] == (
        s["surface_completed_points"] + s["surface_completion_rejected_texture"]
        + s["surface_completion_rejected_cross_view"] + s["surface_completion_rejected_reciprocal"]
        + s["surface_completion_rejected_fusion"] + s["surface_completion_rejected_footprint"])
    assert s["accepted_points"] == g.primary_points + s["surface_completed_points"]


def test_completion_cancellation_matches_oracle(product, oracle):
    """Cancels inside the first and (when it runs) the second completion pass; the polls and
    the partial counters match the oracle's."""
    c1 = scenes.config_c1()
    st = dict(c1.settings, surface_completion_passes=2)
    ny = len(range(st["patch_radius"] + 2, c1.frames[0].shape[0] - (st["patch_radius"] + 2),
                   st["grid_stride"]))
    for cancel_at in (ny + 1, ny + 7, 2 * ny + 5, 3 * ny + 3):
        got = []
        for side in ("gpu", "oracle"):
            polls = []

            def poll():
                polls.append(1)
                return len(polls) >= cancel_at

            if side == "gpu":
                r = product.reconstruct_reference_patch(c1.frames, c1.cameras, c1.sparse,
                                                        c1.focal, c1.reference, control=poll,
                                                        sweep=product.params(st))
            else:
                r, _ = oracle.reference_patch(c1.frames, c1.cameras, c1.sparse, c1.focal,
                                              c1.reference, control=poll,
                                              sweep=oracle.params(st))
            got.append((r, len(polls)))
        (g, gn), (o, on) = got
        assert gn == on, (cancel_at, gn, on)
        _compare(g, o, f"cancel at poll {cancel_at}")


def test_batch_with_completion_matches_alone(product):
    c1 = scenes.config_c1()
    pf = scenes.plane_fixture()
    jobs = [(c1, dict(c1.settings, surface_completion_passes=2)),
            (c1, dict(c1.settings, surface_completion_passes=0)),
            (c1, dict(c1.settings, surface_completion_passes=1)),
            (c1, {"preset": "reference"})]
    frames, maps = product.pack_frames([c1.frames, pf.frames])
    sess = product.Session(0)
    try:
        sess.upload(frames, "f32")
        reqs = [{"cameras": p.cameras, "sparse_points": p.sparse, "focal": p.focal,
                 "reference_camera_index": p.reference, "frame_map": maps[0],
                 "sweep": product.params(st)} for p, st in jobs]
        reqs.append({"cameras": pf.cameras, "sparse_points": pf.sparse, "focal": pf.focal,
                     "reference_camera_index": -1, "frame_map": maps[1], "sweep": None})
        alone = []
        for rq in reqs:
            b = sess.batch([rq])
            alone.append(b.run()[0])
            b.close()
        b = sess.batch(reqs, use_graph=True)
        for _ in range(2):
            out = b.run()
            for i, (x, y) in enumerate(zip(out, alone)):
                assert x.stats == y.stats, i
                assert x.points.tobytes() == y.points.tobytes(), i
                assert x.grid_sites.tobytes() == y.grid_sites.tobytes(), i
                assert x.point_attributes.tobytes() == y.point_attributes.tobytes(), i
        b.close()
        assert alone[1].stats["surface_completion_proposals"] == 0
        assert alone[0].stats["surface_completion_proposals"] >= \
            alone[2].stats["surface_completion_proposals"]
    finally:
        sess.close()
