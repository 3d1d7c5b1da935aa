"""The dense stage's stored document (host functions of libvt3d_dense.so; no GPU).

DenseStageArtifact (crates/video-to-3d-core/src/reconstruction_stages.rs:150-160) is written
by serde_json's compact encoder, f32 values by the ryu crate's pretty printer. What pins it:
- serde_json's own f32 known answers (test_write_f32 of serde_json 1.x: 3.0, 3.1, -1.5, 0.5,
  f32::MIN, f32::MAX, f32::EPSILON) and the ryu crate's f32 printer tests (f2s_test.rs, whose
  expected text is the literal itself). These fix the decimal form for values of 1e-4 up to
  1e11 and the exponent form at 1e14 and above and below 1e-7; that the switch is at 1e13 and
  at 1e-6 is restated from the ryu crate (format32) and otherwise unpinned;
- mathematics: every float reads back as the same f32 bits, with the fewest digits that do
  (numpy's shortest unique digits), every string reads back unchanged with the escapes
  Python's encoder also chooses, and the content hash is SHA-256 of the document (hashlib);
- the reference's field order (dense.rs:14-139, types.rs:68-77, reconstruction_evidence.rs:
  107-115, reconstruction_stages.rs:150-160, 244-252), and its tests of the stage artifacts
  (tests/reconstruction_stages.rs:425-456, 492-501, 570-640; reconstruction_stages.rs:
  1300-1320): the round trip, determinism, schema rejection and the non-finite field path.
"""
import ctypes as C
import hashlib
import json
import struct

import numpy as np
import pytest

from vt3d_dense import abi
from vt3d_dense import stages

H1 = "sha256:" + "0123456789abcdef" * 4
H2 = "sha256:" + "fedcba9876543210" * 4

STATS_FIELDS = [  # dense.rs:100-139
    "attempted", "skip_reason", "reference_frame", "source_views", "source_frames",
    "sampled_pixels", "depth_hypotheses", "accepted_points", "surface_completion_proposals",
    "surface_completed_points", "surface_completion_rejected_texture",
    "surface_completion_rejected_cross_view", "surface_completion_rejected_reciprocal",
    "surface_completion_rejected_fusion", "surface_completion_rejected_footprint",
    "grid_stride", "grid_border", "reciprocal_checked_points", "reciprocal_rejected_points",
    "reciprocal_consistent_points", "fusion_input_observations",
    "fusion_rejected_observations", "fusion_rejected_points", "median_fusion_observations",
    "median_supporting_views", "median_photometric_error", "search_min_depth",
    "search_max_depth", "reference_frames", "reference_attempts", "reference_patches",
    "working_set_estimate", "working_set_budget_bytes", "full_working_set_bytes",
    "reference_view_limit"]
ATTEMPT_FIELDS = [  # dense.rs:62-81
    "reference_frame", "attempted", "accepted", "skip_reason", "sampled_pixels",
    "accepted_points", "reciprocal_checked_points", "reciprocal_rejected_points",
    "reciprocal_consistent_points", "surface_completion_proposals", "surface_completed_points",
    "surface_completion_rejected_texture", "surface_completion_rejected_cross_view",
    "surface_completion_rejected_reciprocal", "surface_completion_rejected_fusion",
    "surface_completion_rejected_footprint"]
PATCH_FIELDS = [  # dense.rs:83-98
    "reference_frame", "source_frames", "primary_start", "primary_points", "completion_start",
    "completed_points", "sampled_pixels", "accepted_points", "reciprocal_consistent_points",
    "grid_stride", "grid_border", "search_min_depth", "search_max_depth"]


class _Handle:
    def __init__(self, m):
        self.m = m
        self.ptr = C.pointer(m)


def _multi(n=40, seed=5, limit=False, reason=None, floats=None):
    """A hand-made multi-reference result (as vt3d_dense_estimate_with_budget returns) with
    `n` points of seeded values; `floats` overrides the first points' x values."""
    rng = np.random.default_rng(seed)
    m = abi.MultiResult()
    st = m.combined.stats
    st.attempted = 1
    st.has_reference_frame = 1
    st.reference_frame = 5
    st.source_views = 3
    for i, f in enumerate((4, 6, 17)):
        st.source_frames[i] = f
    for k, name in enumerate(("sampled_pixels", "depth_hypotheses", "accepted_points",
                              "surface_completion_proposals", "surface_completed_points",
                              "reciprocal_checked_points", "reciprocal_rejected_points",
                              "reciprocal_consistent_points", "fusion_input_observations",
                              "fusion_rejected_observations", "fusion_rejected_points")):
        setattr(st, name, 1000 + 7 * k)
    st.accepted_points = n
    st.surface_completed_points = n // 4
    st.grid_stride, st.grid_border = 4, 3
    st.has_median_fusion_observations, st.median_fusion_observations = 1, 3.0
    st.has_median_supporting_views, st.median_supporting_views = 1, 2.5
    st.has_median_photometric_error, st.median_photometric_error = 0, 0.0
    st.has_search_depth, st.search_min_depth, st.search_max_depth = 1, 5.086011, 8.2099085
    if reason is not None:
        m.combined.skip_code = 11  # VT3D_SKIP_OTHER
        m.combined.skip_reason = reason.encode()
    m.n_reference_frames = 2
    m.reference_frames[0], m.reference_frames[1] = 5, 17
    m.n_attempts = 3
    for i in range(3):
        a = m.attempts[i]
        a.reference_frame, a.attempted, a.accepted = (5, 9, 17)[i], 1, int(i != 1)
        if i == 1:
            a.skip_reason = b"remapped reference rejected"
        a.sampled_pixels, a.accepted_points = 3904 + i, 1174 + i
    m.n_patches = 2
    for i in range(2):
        p = m.patches[i]
        p.reference_frame = (5, 17)[i]
        p.n_source_frames = 2
        p.source_frames[0], p.source_frames[1] = 4 + i, 6 + i
        p.primary_start, p.primary_points = 10 * i, 10
        p.grid_stride, p.grid_border = 4, 3
        p.has_search_depth, p.search_min_depth, p.search_max_depth = int(i == 0), 1.25, 7.5
    for k, name in enumerate(abi.WORKING_SET_FIELDS):
        setattr(m.working_set_estimate, name, (k + 1) * 1048576 + k)
    m.working_set_budget_bytes = 64 << 20
    m.full_working_set_bytes = 200 << 20
    if limit:
        m.has_reference_view_limit, m.reference_view_limit = 1, 2
    pts = np.zeros(n, abi.POINT_DTYPE)
    for f in ("x", "y", "z", "confidence"):
        pts[f] = (rng.standard_normal(n) * 3).astype(np.float32)
    if floats is not None:
        pts["x"][:len(floats)] = np.asarray(floats, np.float32)
    for f in ("r", "g", "b"):
        pts[f] = rng.integers(0, 256, n)
    sites = np.zeros(n, abi.SITE_DTYPE)
    sites["x"] = rng.integers(0, 2000, n)
    sites["y"] = rng.integers(0, 2000, n)
    attrs = np.zeros(n, abi.ATTR_DTYPE)
    attrs["reciprocal_consistency"] = rng.random(n).astype(np.float32)
    attrs["depth_margin"] = rng.random(n).astype(np.float32) * np.float32(1e-3)
    m.combined.n_points = n
    m.combined.points = pts.ctypes.data_as(C.POINTER(abi.Point3))
    m.combined.grid_sites = sites.ctypes.data_as(C.POINTER(abi.GridSite))
    m.combined.attributes = attrs.ctypes.data_as(C.POINTER(abi.PointAttributes))
    h = _Handle(m)
    h.keep = (pts, sites, attrs)
    return abi.unpack_multi(m, h), (pts, sites, attrs)


def _f32(v):
    return struct.unpack("<f", struct.pack("<f", v))[0]


def _shortest(v):
    """numpy's shortest unique digits and decimal exponent of an f32."""
    mant, exp = np.format_float_scientific(np.float32(v), unique=True, trim="-").split("e")
    return mant.lstrip("-").replace(".", ""), int(exp)


def _x_text(doc):
    """The text of every point's x value, in order."""
    doc = doc[doc.index('"points":['):doc.index('"grid_sites":[')]
    out, i = [], 0
    while True:
        i = doc.find('{"x":', i)
        if i < 0:
            return out
        j = doc.index(",", i)
        out.append(doc[i + 5:j])
        i = j


def _doc(product, floats):
    m, _ = _multi(n=len(floats), floats=floats)
    return stages.dense_stage_artifact(m, H1, H2).to_json()


def test_serde_json_f32_known_answers(product):
    # serde_json's test_write_f32 (tests/test.rs): the text serde_json writes for these f32s
    kat = [(3.0, "3.0"), (3.1, "3.1"), (-1.5, "-1.5"), (0.5, "0.5"),
           (-3.4028235e38, "-3.4028235e38"), (3.4028235e38, "3.4028235e38"),
           (1.1920929e-7, "1.1920929e-7"), (0.0, "0.0"), (-0.0, "-0.0")]
    doc = _doc(product, [v for v, _ in kat])
    assert _x_text(doc)[:len(kat)] == [t for _, t in kat]


def test_ryu_f32_printer_known_answers(product):
    # the ryu crate's f2s_test.rs: the printed text of each f32 is the literal itself
    lits = ["1.1754944e-38", "3.4028235e38", "1e-45", "33554450.0", "9000000000.0",
            "34366720000.0", "8099.0312", "0.00024414062", "0.0024414062", "0.0043945312",
            "0.0063476562", "4.7223665e21", "8388608.0", "16777216.0", "33554436.0",
            "67131496.0", "1.9310392e-38", "-2.47e-43", "1.993244e-38", "4103.9004",
            "5339999700.0", "6.0898e-39", "0.0010310042", "2.882326e17", "9.223404e17",
            "67108870.0", "1e-44", "2.816025e14", "9.223372e18", "1.5846086e29",
            "1.1811161e19", "5.368709e18", "4.6143166e18", "0.007812537", "1.18697725e20",
            "1.00014165e-36", "200.0", "33554432.0"]
    doc = _doc(product, [float(s) for s in lits])
    assert _x_text(doc) == lits


def test_floats_read_back_with_the_fewest_digits(product):
    rng = np.random.default_rng(11)
    bits = rng.integers(0, 2 ** 32, 4000, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32)
    vals = vals[np.isfinite(vals)]
    vals = np.concatenate([vals, (rng.random(2000) * 10.0 ** rng.integers(-8, 15, 2000))
                           .astype(np.float32)])
    doc = _doc(product, vals)
    texts = _x_text(doc)
    assert len(texts) == len(vals)
    for v, t in zip(vals, texts):
        assert np.float32(float(t)).view(np.uint32) == np.float32(v).view(np.uint32), t
        digits, exp = _shortest(v)
        if "e" in t:
            m, e = t.lstrip("-").split("e")
            got = (m.replace(".", ""), int(e))
        else:
            # decimal form: the digits without leading zeros or the ".0" suffix
            s = t.lstrip("-")
            ip, fp = s.split(".")
            ds = (ip + fp).lstrip("0")
            got = ((ds.rstrip("0") or "0"), len(ip.lstrip("0")) - 1 if ip.strip("0") else
                   -(len(fp) - len(fp.lstrip("0"))) - 1)
        assert got[0].rstrip("0") == digits.rstrip("0") and got[1] == exp, (t, digits, exp)
        # ryu's layout: the decimal form for 1e-6 <= |v| < 1e13, the exponent form outside
        if float(v) != 0.0:
            assert ("e" in t) == (not -6 <= exp <= 12), t


def test_document_fields_order_and_values(product):
    m, (pts, sites, attrs) = _multi(limit=True)
    art = stages.dense_stage_artifact(m, H1, H2)
    doc = art.to_json()
    d = json.loads(doc, object_pairs_hook=lambda kv: kv)
    top = [k for k, _ in d]
    assert top == ["schema_version", "request_hash", "cameras_hash", "dense"]
    d = json.loads(doc)
    assert d["schema_version"] == 1 and d["request_hash"] == H1 and d["cameras_hash"] == H2
    pairs = json.loads(doc, object_pairs_hook=lambda kv: kv)
    dense = dict(pairs)["dense"]
    assert [k for k, _ in dense] == ["stats", "points", "grid_sites", "point_attributes",
                                     "warnings"]
    stats_pairs = dict(dense)["stats"]
    assert [k for k, _ in stats_pairs] == STATS_FIELDS
    sd = dict(stats_pairs)
    assert [k for k, _ in sd["reference_attempts"][0]] == ATTEMPT_FIELDS
    assert [k for k, _ in sd["reference_patches"][0]] == PATCH_FIELDS
    assert [k for k, _ in sd["working_set_estimate"]] == list(abi.WORKING_SET_FIELDS)
    assert [k for k, _ in dict(dense)["points"][0]] == ["x", "y", "z", "confidence", "r", "g",
                                                        "b"]
    s = d["dense"]["stats"]
    st = m.combined.stats
    assert s["attempted"] is True and s["skip_reason"] is None and s["reference_frame"] == 5
    assert s["source_views"] == 3 and s["source_frames"] == [4, 6, 17]
    assert s["median_photometric_error"] is None
    assert np.float32(s["search_min_depth"]) == np.float32(5.086011)
    assert s["reference_frames"] == [5, 17]
    assert [a["skip_reason"] for a in s["reference_attempts"]] == [
        None, "remapped reference rejected", None]
    assert s["reference_patches"][1]["search_min_depth"] is None
    assert s["reference_patches"][0]["source_frames"] == [4, 6]
    assert s["reference_view_limit"] == 2
    assert s["working_set_estimate"]["total_bytes"] == 12 * 1048576 + 11
    for k in ("sampled_pixels", "fusion_rejected_points", "grid_stride", "grid_border"):
        assert s[k] == st[k]
    p = d["dense"]["points"]
    assert len(p) == len(pts)
    for f in ("x", "y", "z", "confidence"):
        assert np.array_equal(np.array([q[f] for q in p], np.float32).view(np.uint32),
                              pts[f].view(np.uint32))
    for f in ("r", "g", "b"):
        assert [q[f] for q in p] == pts[f].tolist()
    assert [(q["x"], q["y"]) for q in d["dense"]["grid_sites"]] == list(
        zip(sites["x"].tolist(), sites["y"].tolist()))
    a = d["dense"]["point_attributes"]
    assert np.array_equal(np.array([q["depth_margin"] for q in a], np.float32),
                          attrs["depth_margin"])
    # the dense stage's warnings (reconstruction_stages.rs:772-782): the budget notice, then
    # dense_warning of the stats
    w = d["dense"]["warnings"]
    assert w[0] == ("The dense working set for every reference view (200 MiB) exceeds the 64 "
                    "MiB budget, so dense reconstruction kept the first 2 reference view(s) in "
                    "its usual order (12 MiB). Each kept view passed the same gates as a full "
                    "run; the remaining views were not reconstructed.")
    assert w[1:] == [product.dense_warning(m.stats)] and w[1].startswith("Slice 4 dense")
    assert art.warnings == w


def test_content_hash_is_sha256_of_the_document_and_deterministic(product):
    m, _ = _multi()
    a1 = stages.dense_stage_artifact(m, H1, H2)
    a2 = stages.dense_stage_artifact(m, H1, H2)
    doc = a1.to_json()
    assert a1.content_hash() == "sha256:" + hashlib.sha256(doc.encode()).hexdigest()
    assert a1.to_json() == a2.to_json() and a1.content_hash() == a2.content_hash()
    assert product.content_hash(doc.encode()) == a1.content_hash()
    # another cameras identity is another artifact
    assert stages.dense_stage_artifact(m, H1, H1).content_hash() != a1.content_hash()


def test_round_trip_through_the_document(product):
    # tests/reconstruction_stages.rs:492-501: from_json(to_json) is the same artifact and
    # writes the same bytes
    m, (pts, _, _) = _multi(n=300, limit=True)
    a = stages.dense_stage_artifact(m, H1, H2)
    doc = a.to_json()
    b = stages.DenseStageArtifact.from_json(doc)
    assert b.to_json() == doc and b == a
    assert b.request_hash == H1 and b.cameras_hash == H2 and b.warnings == a.warnings
    assert np.array_equal(b.dense.combined.points.view(np.uint8), pts.view(np.uint8))
    assert b.dense.stats["reference_attempts"] == m.stats["reference_attempts"]
    assert b.dense.stats["reference_patches"] == m.stats["reference_patches"]
    assert b.dense.combined.primary_points == 300 - 300 // 4
    # whitespace, key order and escapes in a stored document read the same
    pretty = json.dumps(json.loads(doc), indent=2)
    assert stages.DenseStageArtifact.from_json(pretty).to_json() == doc
    d = json.loads(doc)
    d["dense"] = dict(reversed(list(d["dense"].items())))
    assert stages.DenseStageArtifact.from_json(json.dumps(d)).to_json() == doc


def test_strings_are_escaped_as_json_encoders_do(product):
    reason = 'quote " slash \\ nl \n tab \t bell \x07 del \x7f unicode é中 /'
    m, _ = _multi(n=3, reason=reason)
    m.combined.stats["skip_reason"] = reason
    doc = stages.dense_stage_artifact(m, H1, H2).to_json()
    want = json.dumps(reason, ensure_ascii=False)
    assert '"skip_reason":' + want in doc
    assert json.loads(doc)["dense"]["stats"]["skip_reason"] == reason
    back = stages.DenseStageArtifact.from_json(doc)
    assert back.to_json() == doc and back.dense.stats["skip_reason"] == reason


def test_non_finite_values_name_their_field(product):
    # reconstruction_stages.rs:1300-1320: an infinite median error is reported at
    # "dense.stats.median_photometric_error" and no artifact is made
    m, _ = _multi(n=6)
    st = m.handle.m.combined.stats
    st.has_median_photometric_error, st.median_photometric_error = 1, float("inf")
    with pytest.raises(stages.NonFinite) as e:
        stages.dense_stage_artifact(m, H1, H2)
    assert e.value.field == "dense.stats.median_photometric_error"
    assert str(e.value) == ("Dense stage produced a non-finite value at "
                            "`dense.stats.median_photometric_error`, which its stored document "
                            "cannot represent")
    m, (pts, _, attrs) = _multi(n=6)
    pts["z"][3] = np.nan
    with pytest.raises(stages.NonFinite) as e:
        stages.dense_stage_artifact(m, H1, H2)
    assert e.value.field == "dense.points.3.z"
    m, (pts, _, attrs) = _multi(n=6)
    attrs["depth_margin"][5] = -np.inf
    with pytest.raises(stages.NonFinite) as e:
        stages.dense_stage_artifact(m, H1, H2)
    assert e.value.field == "dense.point_attributes.5.depth_margin"
    # a document whose number overflows f32 reads as infinity, which is refused
    doc = stages.dense_stage_artifact(_multi(n=2)[0], H1, H2).to_json()
    d = json.loads(doc)
    d["dense"]["points"][1]["y"] = 1e39
    with pytest.raises(stages.NonFinite) as e:
        stages.DenseStageArtifact.from_json(json.dumps(d))
    assert e.value.field == "dense.points.1.y"


def test_readers_reject_unknown_schema_versions_and_malformed_documents(product):
    # tests/reconstruction_stages.rs:570-640
    doc = stages.dense_stage_artifact(_multi(n=4)[0], H1, H2).to_json()
    d = json.loads(doc)
    d["schema_version"] = 999
    with pytest.raises(stages.UnsupportedSchemaVersion) as e:
        stages.DenseStageArtifact.from_json(json.dumps(d))
    assert (e.value.found, e.value.supported) == (999, 1)
    assert str(e.value) == "Dense stage artifact has schema version 999; this build reads version 1"

    def malformed(text):
        with pytest.raises(stages.Malformed) as e:
            stages.DenseStageArtifact.from_json(text)
        assert str(e.value).startswith("malformed Dense stage artifact: ")
        return e.value.message

    del d["schema_version"]
    assert "schema_version" in malformed(json.dumps(d))
    malformed("not json")
    malformed(doc + "x")                 # trailing characters
    malformed(doc[:-1])                  # truncated
    d = json.loads(doc)
    # a camera document is not a dense document (deny_unknown_fields on the artifact)
    camera = {"schema_version": 1, "request_hash": H1, "sparse": {}}
    assert "sparse" in malformed(json.dumps(camera))
    d2 = dict(d, extra=1)
    assert "extra" in malformed(json.dumps(d2))
    d2 = json.loads(doc)
    d2["dense"]["extra"] = []
    assert "extra" in malformed(json.dumps(d2))
    d2 = json.loads(doc)
    del d2["dense"]["warnings"]
    assert "warnings" in malformed(json.dumps(d2))
    d2 = json.loads(doc)
    d2["dense"]["points"][0]["r"] = 256                  # u8 out of range
    malformed(json.dumps(d2))
    d2 = json.loads(doc)
    d2["dense"]["stats"]["sampled_pixels"] = 1.5         # a float for an integer
    malformed(json.dumps(d2))
    d2 = json.loads(doc)
    d2["dense"]["stats"]["sampled_pixels"] = -1          # negative for usize
    malformed(json.dumps(d2))
    d2 = json.loads(doc)
    d2["request_hash"] = "sha256:" + "A" * 64            # not lowercase hex
    assert "content hash" in malformed(json.dumps(d2))
    d2 = json.loads(doc)
    d2["dense"]["grid_sites"].pop()                      # not one site per point
    malformed(json.dumps(d2))
    dup = doc.replace('"attempted":true,', '"attempted":true,"attempted":true,', 1)
    assert "duplicate" in malformed(dup)
    # structs below DenseEvidence do not deny unknown fields: serde skips them
    d2 = json.loads(doc)
    d2["dense"]["stats"]["future_field"] = {"a": [1, 2.5, None]}
    d2["dense"]["points"][0]["w"] = 1
    assert stages.DenseStageArtifact.from_json(json.dumps(d2)).to_json() == doc
    # a missing Option field reads as None
    d2 = json.loads(doc)
    del d2["dense"]["stats"]["median_photometric_error"]
    assert stages.DenseStageArtifact.from_json(json.dumps(d2)).to_json() == doc


def test_invalid_identities_are_refused(product):
    m, _ = _multi(n=2)
    for bad in ("sha256:" + "0" * 63, "sha1:" + "0" * 64, "sha256:" + "G" * 64):
        with pytest.raises(stages.ReconstructionStageError):
            stages.dense_stage_artifact(m, bad, H2)
        with pytest.raises(stages.ReconstructionStageError):
            stages.dense_stage_artifact(m, H1, bad)


def test_library_exports_the_artifact_entry_points(product):
    lib = product.library()
    for name in ("vt3d_dense_artifact_create", "vt3d_dense_artifact_to_json",
                 "vt3d_dense_artifact_content_hash", "vt3d_dense_artifact_from_json",
                 "vt3d_dense_artifact_free"):
        assert hasattr(lib, name)
