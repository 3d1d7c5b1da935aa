"""The dense stage's artifact over a device estimate (reconstruct_dense_with_cancel,
reconstruction_stages.rs:549-589): the budgeted estimate runs on the GPU, and its stored
document must hold exactly that estimate, the oracle's counters and skip reasons, the dense
stage's warnings, and read back to the same bytes (tests/reconstruction_stages.rs:425-501)."""
import hashlib
import json

import numpy as np
import pytest

from test_oracle_budget import _plane
from vt3d_dense import stages

pytestmark = pytest.mark.gpu

H_CAMERAS = "sha256:" + "5a" * 32


def _same(a, b):
    """Equal values, floats compared as f32 (the document's text reads back to the f32)."""
    if isinstance(b, dict):
        return isinstance(a, dict) and a.keys() == b.keys() and all(
            _same(a[k], b[k]) for k in b)
    if isinstance(b, list):
        return isinstance(a, list) and len(a) == len(b) and all(map(_same, a, b))
    if isinstance(b, float):
        return isinstance(a, (int, float)) and np.float32(a) == np.float32(b)
    return a == b and type(a) is type(b)


@pytest.mark.parametrize("views_allowed", [0, 2, 3])
def test_artifact_of_the_device_estimate(product, oracle, views_allowed):
    frames, cams, sparse, f = _plane((0.0, 0.24, -0.22))
    f = float(np.float32(f))  # the camera stage stores the focal as f32
    sizes = [oracle.estimate_working_set(frames, 320, 3, v, False)["total_bytes"]
             for v in (1, 2, 3)]
    budget = 1 if views_allowed == 0 else sizes[views_allowed - 1]
    opts = product.reconstruction_options(max_features=320, descriptor_radius=3,
                                          max_dense_working_set_bytes=budget)
    art = stages.reconstruct_dense(frames, cams, sparse, f, H_CAMERAS, options=opts)
    print(f"artifact: views {views_allowed}, budget {budget}, points "
          f"{len(art.dense.combined.points)}", flush=True)
    g = product.estimate_depth_points_with_budget(frames, cams, sparse, f, budget, 320, 3)
    o = oracle.estimate_with_budget(frames, cams, sparse, f, budget, 320, 3)

    doc = art.to_json()
    d = json.loads(doc)
    assert d["schema_version"] == 1 and d["cameras_hash"] == H_CAMERAS
    assert d["request_hash"] == product.request_content_hash(frames, opts)
    s = d["dense"]["stats"]
    # the document holds the device estimate exactly (floats as the same f32)
    for k, v in g.stats.items():
        assert _same(s[k], v), (k, s[k], v)
    pts = d["dense"]["points"]
    assert len(pts) == len(g.combined.points)
    for fld in ("x", "y", "z", "confidence"):
        got = np.array([q[fld] for q in pts], np.float32)
        assert got.tobytes() == np.ascontiguousarray(g.combined.points[fld]).tobytes(), fld
    # and the oracle's counters, envelope flags and skip reason
    for k in ("attempted", "skip_reason", "reference_frame", "source_frames", "sampled_pixels",
              "accepted_points", "reference_frames", "working_set_budget_bytes",
              "full_working_set_bytes", "reference_view_limit"):
        assert s[k] == o.stats[k], (k, s[k], o.stats[k])
    assert s["working_set_estimate"] == o.stats["working_set_estimate"]
    # the dense stage's warnings (reconstruction_stages.rs:772-782)
    w = d["dense"]["warnings"]
    if o.stats["reference_view_limit"] is not None:
        assert w[0] == (
            f"The dense working set for every reference view "
            f"({o.stats['full_working_set_bytes'] // 1048576} MiB) exceeds the "
            f"{o.stats['working_set_budget_bytes'] // 1048576} MiB budget, so dense "
            f"reconstruction kept the first {o.stats['reference_view_limit']} reference view(s) "
            f"in its usual order ({o.stats['working_set_estimate']['total_bytes'] // 1048576} "
            f"MiB). Each kept view passed the same gates as a full run; the remaining views "
            f"were not reconstructed.")
        w = w[1:]
    want = product.dense_warning(o.stats)
    assert w == ([want] if want else [])
    # identity and round trip
    assert art.content_hash() == "sha256:" + hashlib.sha256(doc.encode()).hexdigest()
    back = stages.DenseStageArtifact.from_json(doc)
    assert back.to_json() == doc and back.content_hash() == art.content_hash()


def test_dense_stage_checks_the_request_identity(product):
    # reconstruction_stages.rs:564-569: a camera artifact of another request is refused
    frames, cams, sparse, f = _plane((0.0, 0.24))
    other = "sha256:" + "0" * 64
    with pytest.raises(stages.InputMismatch) as e:
        stages.reconstruct_dense(frames, cams, sparse, f, H_CAMERAS, camera_request_hash=other)
    assert (e.value.stage, e.value.input, e.value.recorded) == ("Cameras", "Request", other)
    assert str(e.value).startswith(f"Cameras stage artifact was computed from Request {other}, "
                                   "not sha256:")
