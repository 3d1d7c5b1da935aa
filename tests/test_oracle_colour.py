"""The colour scene pins the luma weights and the colour read (dense_legacy.rs:1147-1166) on
the oracle: its luma, not the channels' average or a swapped order, decides the samples,
and the points' colours are the reference frame's bytes. The device is held to the oracle
on the same frames in test_gpu_colour.py, so a channel swap there fails too."""
import numpy as np

from vt3d_dense import scenes


def _solve(oracle, p, frames):
    st = {"preset": "reference", "surface_completion_passes": 0, "emit_samples": 1}
    r, _ = oracle.reference_patch(frames, p.cameras, p.sparse, p.focal, p.reference,
                                  sweep=oracle.params(st), threads=8)
    return r


def test_colour_scene_pins_luma_and_colour(oracle):
    p = scenes.config_c1_colour()
    base = _solve(oracle, p, p.frames)
    assert base.stats["accepted_points"] > 500
    ref = p.frames[p.reference]
    s = base.grid_sites
    for c, name in enumerate("rgb"):
        assert np.array_equal(base.points[name], ref[s["y"], s["x"], c])
    # the luma as gray frames: the same decisions and geometry, gray colours
    luma = [((77 * f[..., 0].astype(np.uint32) + 150 * f[..., 1].astype(np.uint32)
              + 29 * f[..., 2].astype(np.uint32)) >> 8).astype(np.uint8) for f in p.frames]
    gray = _solve(oracle, p, luma)
    assert gray.samples.tobytes() == base.samples.tobytes()
    assert np.array_equal(gray.points[["x", "y", "z", "confidence"]],
                          base.points[["x", "y", "z", "confidence"]])
    assert not np.array_equal(gray.points["r"], base.points["r"])
    # planted mistakes: a red/blue swap, or alpha read as a channel, move the samples
    for order in ([2, 1, 0, 3], [0, 1, 3, 2]):
        wrong = _solve(oracle, p, [np.ascontiguousarray(f[..., order]) for f in p.frames])
        assert (wrong.stats["sampled_pixels"] != base.stats["sampled_pixels"]
                or wrong.samples.tobytes() != base.samples.tobytes()), order
