"""Colour input on the device (to_luma, dense_legacy.rs:1147-1157; sample_rgb,
dense_legacy.rs:1159-1166): RGBA frames whose three channels carry independent textures and
whose alpha is noise, solved on the GPU and by the FP64 oracle. A swapped channel or a wrong
luma weight changes the luma, and with it the gated samples and every decision; a wrong
colour read changes the points' r, g, b."""
import os

import numpy as np
import pytest

import parity
from vt3d_dense import abi, scenes

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def colour():
    return scenes.config_c1_colour()


@pytest.mark.parametrize("kernel", ["direct", "tiled"])
@pytest.mark.parametrize("preset", ["reference", "zncc"])
def test_colour_frames_match_oracle(product, oracle, colour, preset, kernel):
    p = colour
    st = dict(p.settings) if preset == "zncc" else {"preset": "reference"}
    st.update(surface_completion_passes=2, emit_samples=1)
    g = product.reconstruct_reference_patch(p.frames, p.cameras, p.sparse, p.focal,
                                            p.reference, sweep=product.params(st, kernel=kernel))
    o, sl = oracle.reference_patch(p.frames, p.cameras, p.sparse, p.focal, p.reference,
                                   sweep=oracle.params(st), threads=THREADS)
    parity.stats_equal(g.stats, o.stats)
    kind = abi.COST_ZNCC if preset == "zncc" else abi.COST_ZSAD
    gp, op = g, o
    # primary points only for the sample comparison; completion is compared below
    rep = parity.compare(_primary(gp), _primary(op), sl, kind)
    print("colour", preset, kernel, rep)
    assert rep["points_compared"] > 200
    # every point's colour is the reference frame's RGB at its grid site
    ref = p.frames[p.reference]
    s = g.grid_sites
    for c, name in enumerate("rgb"):
        assert np.array_equal(g.points[name], ref[s["y"], s["x"], c]), name
    if rep["disagree"] == 0:
        parity.compare_completed(g, o, f"colour {preset} {kernel}")


def _primary(r):
    import copy

    q = copy.copy(r)
    n = r.primary_points
    q.points, q.grid_sites = r.points[:n], r.grid_sites[:n]
    return q
