"""ctypes mirror of include/vt3d_dense.h (the C ABI of the dense plane sweep).

The structs here are byte-for-byte the header's; `bind` declares the argument and return
types of every exported function on a loaded library. The oracle library shares these
structs (it implements the same data contract), so the helpers that marshal inputs and
unpack results work for both.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Dict, List, Optional, Sequence

import numpy as np

VT3D_OK = 0
VT3D_NO_PATCH = 1
VT3D_E_INVALID = -1
VT3D_E_DEVICE = -2
VT3D_E_NO_DEVICE = -3
VT3D_E_BUDGET = -4
VT3D_E_BUSY = -5

COST_ZSAD, COST_ZNCC = 0, 1
AMBIGUITY_MARGIN, AMBIGUITY_RATIO = 0, 1
STORAGE = {"f32": 0, "f16": 1, "bf16": 2}
KERNEL = {"auto": 0, "direct": 1, "tiled": 2}

SAMPLE_NO_CANDIDATE = 1
SAMPLE_REJECT_COST = 2
SAMPLE_REJECT_AMBIGUITY = 3
SAMPLE_REJECT_RECIPROCAL = 4
SAMPLE_REJECT_FUSION = 5
SAMPLE_ACCEPTED = 6


class Frame(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgba", C.POINTER(C.c_uint8))]


class Camera(C.Structure):
    _fields_ = [("frame_index", C.c_uint64), ("rotation_colmajor", C.c_double * 9),
                ("translation", C.c_double * 3)]


class SweepParams(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("cost_kind", C.c_int32),
                ("patch_radius", C.c_int32), ("depth_hypotheses", C.c_int32),
                ("max_source_views", C.c_int32), ("grid_stride", C.c_int32),
                ("ambiguity_kind", C.c_int32), ("min_fusion_observations", C.c_int32),
                ("surface_completion_passes", C.c_int32), ("storage", C.c_int32),
                ("emit_samples", C.c_int32), ("sweep_kernel", C.c_int32),
                ("min_reference_contrast", C.c_double), ("max_photometric_error", C.c_double),
                ("ambiguity_threshold", C.c_double),
                ("max_reciprocal_relative_depth_error", C.c_double),
                ("max_fusion_relative_position_error", C.c_double),
                ("error_confidence_scale", C.c_double), ("depth_margin_saturation", C.c_double),
                ("depth_min_override", C.c_double), ("depth_max_override", C.c_double)]


class Point3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("confidence", C.c_float),
                ("r", C.c_uint8), ("g", C.c_uint8), ("b", C.c_uint8), ("pad", C.c_uint8)]


class GridSite(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32)]


class PointAttributes(C.Structure):
    _fields_ = [("reciprocal_consistency", C.c_float), ("depth_margin", C.c_float)]


class DenseStats(C.Structure):
    _fields_ = [("attempted", C.c_int32), ("has_reference_frame", C.c_int32),
                ("reference_frame", C.c_uint64), ("source_views", C.c_uint64),
                ("source_frames", C.c_uint64 * 32), ("sampled_pixels", C.c_uint64),
                ("depth_hypotheses", C.c_uint64), ("accepted_points", C.c_uint64),
                ("surface_completion_proposals", C.c_uint64),
                ("surface_completed_points", C.c_uint64),
                ("surface_completion_rejected_texture", C.c_uint64),
                ("surface_completion_rejected_cross_view", C.c_uint64),
                ("surface_completion_rejected_reciprocal", C.c_uint64),
                ("surface_completion_rejected_fusion", C.c_uint64),
                ("surface_completion_rejected_footprint", C.c_uint64),
                ("grid_stride", C.c_uint64), ("grid_border", C.c_uint32), ("pad0", C.c_uint32),
                ("reciprocal_checked_points", C.c_uint64),
                ("reciprocal_rejected_points", C.c_uint64),
                ("reciprocal_consistent_points", C.c_uint64),
                ("fusion_input_observations", C.c_uint64),
                ("fusion_rejected_observations", C.c_uint64),
                ("fusion_rejected_points", C.c_uint64),
                ("has_median_fusion_observations", C.c_int32),
                ("median_fusion_observations", C.c_float),
                ("has_median_supporting_views", C.c_int32),
                ("median_supporting_views", C.c_float),
                ("has_median_photometric_error", C.c_int32),
                ("median_photometric_error", C.c_float),
                ("has_search_depth", C.c_int32), ("search_min_depth", C.c_float),
                ("search_max_depth", C.c_float), ("pad1", C.c_int32)]


class SampleDiag(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("status", C.c_int32),
                ("best_hypothesis", C.c_int32), ("best_support", C.c_int32),
                ("second_hypothesis", C.c_int32), ("best_cost", C.c_float),
                ("second_cost", C.c_float), ("reciprocal_mask", C.c_uint32),
                ("fused_observations", C.c_uint32), ("reciprocal_signature", C.c_uint32),
                ("pad", C.c_uint32)]


class DenseResult(C.Structure):
    _fields_ = [("stats", DenseStats), ("skip_code", C.c_int32), ("skip_reason", C.c_char * 256),
                ("n_points", C.c_size_t), ("primary_points", C.c_size_t),
                ("points", C.POINTER(Point3)), ("grid_sites", C.POINTER(GridSite)),
                ("attributes", C.POINTER(PointAttributes)), ("n_samples", C.c_size_t),
                ("samples", C.POINTER(SampleDiag))]


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("upload_ms", C.c_double), ("gate_ms", C.c_double),
                ("sweep_ms", C.c_double), ("reciprocal_ms", C.c_double), ("fuse_ms", C.c_double),
                ("completion_ms", C.c_double), ("download_ms", C.c_double),
                ("units", C.c_uint64), ("reverse_sweep_ms", C.c_double),
                ("reciprocal_pairs", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("tile_width", C.c_int32), ("target_waves", C.c_int32),
                ("reverse_bound", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("cameras", C.POINTER(Camera)), ("n_cameras", C.c_size_t),
                ("sparse_xyz", C.POINTER(C.c_double)), ("n_sparse", C.c_size_t),
                ("frame_map", C.POINTER(C.c_uint64)), ("n_frames", C.c_size_t),
                ("focal", C.c_double), ("reference_camera_index", C.c_int64),
                ("fine_grid", C.c_int32), ("pad", C.c_int32),
                ("params", C.POINTER(SweepParams))]


class ServerConfig(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("max_batch_requests", C.c_int32),
                ("max_inflight_batches", C.c_int32), ("pad", C.c_int32),
                ("hbm_budget_bytes", C.c_uint64)]


class Completion(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("status", C.c_int32), ("batch_requests", C.c_int32),
                ("result", C.POINTER(DenseResult)), ("queue_ms", C.c_double),
                ("latency_ms", C.c_double), ("batch_device_ms", C.c_double),
                ("units", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("batches", C.c_uint64),
                ("rejected_budget", C.c_uint64), ("max_batch_requests_seen", C.c_uint64),
                ("frame_bytes", C.c_uint64), ("workspace_bytes", C.c_uint64),
                ("budget_bytes", C.c_uint64)]


MAX_REFERENCE_VIEWS = 3      # dense.rs:8
MAX_REFERENCE_ATTEMPTS = 6   # dense.rs:9
MIN_REFERENCE_PATCH_POINTS = 3  # dense.rs:10


class ReferenceAttempt(C.Structure):
    _fields_ = [("reference_frame", C.c_uint64), ("attempted", C.c_int32),
                ("accepted", C.c_int32), ("skip_reason", C.c_char * 256)] + [
        (n, C.c_uint64) for n in (
            "sampled_pixels", "accepted_points", "reciprocal_checked_points",
            "reciprocal_rejected_points", "reciprocal_consistent_points",
            "surface_completion_proposals", "surface_completed_points",
            "surface_completion_rejected_texture", "surface_completion_rejected_cross_view",
            "surface_completion_rejected_reciprocal", "surface_completion_rejected_fusion",
            "surface_completion_rejected_footprint")]


class ReferencePatchStats(C.Structure):
    _fields_ = [("reference_frame", C.c_uint64), ("n_source_frames", C.c_uint64),
                ("source_frames", C.c_uint64 * 32)] + [
        (n, C.c_uint64) for n in (
            "primary_start", "primary_points", "completion_start", "completed_points",
            "sampled_pixels", "accepted_points", "reciprocal_consistent_points",
            "grid_stride")] + [
        ("grid_border", C.c_uint32), ("has_search_depth", C.c_int32),
        ("search_min_depth", C.c_float), ("search_max_depth", C.c_float)]


class ReconstructionOptions(C.Structure):
    """ReconstructionOptions (crates/video-to-3d-core/src/reconstruction/types.rs:25-35)."""
    _fields_ = [("max_features", C.c_uint64), ("min_feature_distance", C.c_uint32),
                ("descriptor_radius", C.c_uint32), ("match_radius", C.c_uint32),
                ("max_descriptor_distance", C.c_float), ("ratio_threshold", C.c_float),
                ("has_focal_length_pixels", C.c_int32), ("focal_length_pixels", C.c_float),
                ("pad", C.c_uint32), ("max_dense_working_set_bytes", C.c_uint64)]


WORKING_SET_FIELDS = ("frame_bytes", "retry_frame_bytes", "remapped_frame_bytes",
                      "luminance_bytes", "sparse_pipeline_bytes", "dense_sample_bytes",
                      "topology_bytes", "mesh_builder_bytes", "surface_fusion_bytes",
                      "retained_candidate_bytes", "packed_output_bytes", "total_bytes")


class WorkingSet(C.Structure):
    """DenseWorkingSetEstimate (crates/video-to-3d-core/src/dense.rs:14-29)."""
    _fields_ = [(n, C.c_uint64) for n in WORKING_SET_FIELDS]

    def as_dict(self) -> Dict[str, int]:
        return {n: int(getattr(self, n)) for n in WORKING_SET_FIELDS}


class MultiResult(C.Structure):
    _fields_ = [("combined", DenseResult), ("n_reference_frames", C.c_size_t),
                ("reference_frames", C.c_uint64 * MAX_REFERENCE_VIEWS),
                ("n_attempts", C.c_size_t),
                ("attempts", ReferenceAttempt * (MAX_REFERENCE_ATTEMPTS + 1)),
                ("n_patches", C.c_size_t),
                ("patches", ReferencePatchStats * MAX_REFERENCE_VIEWS),
                ("omitted_references", C.c_int32), ("fine_grid", C.c_int32),
                ("working_set_estimate", WorkingSet), ("working_set_budget_bytes", C.c_uint64),
                ("full_working_set_bytes", C.c_uint64),
                ("has_reference_view_limit", C.c_int32), ("reference_view_limit", C.c_int32)]


CONTENT_HASH_LEN = 72
DENSE_STAGE_ARTIFACT_SCHEMA_VERSION = 1  # reconstruction_stages.rs:37
VT3D_E_NONFINITE = -7
VT3D_E_SCHEMA = -8
VT3D_E_MALFORMED = -9


class DenseArtifact(C.Structure):
    """vt3d_dense_artifact: DenseStageArtifact (reconstruction_stages.rs:150-160)."""
    _fields_ = [("schema_version", C.c_uint32), ("pad", C.c_uint32),
                ("request_hash", C.c_char * CONTENT_HASH_LEN),
                ("cameras_hash", C.c_char * CONTENT_HASH_LEN),
                ("dense", C.POINTER(MultiResult)), ("n_warnings", C.c_size_t),
                ("warnings", C.POINTER(C.c_char_p))]


class MeshTriangle(C.Structure):
    _fields_ = [("a", C.c_uint64), ("b", C.c_uint64), ("c", C.c_uint64),
                ("confidence", C.c_float), ("pad", C.c_uint32)]


_MESH_COUNTS = ("grid_vertices", "rejected_grid_vertices", "candidate_cells",
                "candidate_triangles", "accepted_triangles", "rejected_discontinuities",
                "rejected_degenerate")


class MeshPatchStats(C.Structure):
    _fields_ = [("reference_frame", C.c_uint64), ("point_count", C.c_uint64),
                ("attempted", C.c_int32), ("has_skip_reason", C.c_int32),
                ("skip_reason", C.c_char * 128)] + [(n, C.c_uint64) for n in _MESH_COUNTS]


class MeshStats(C.Structure):
    _fields_ = [("attempted", C.c_int32), ("has_skip_reason", C.c_int32),
                ("skip_reason", C.c_char * 128), ("has_reference_frame", C.c_int32),
                ("pad", C.c_int32), ("reference_frame", C.c_uint64)] + [
        (n, C.c_uint64) for n in _MESH_COUNTS] + [
        ("n_reference_frames", C.c_size_t), ("reference_frames", C.c_uint64 * 3),
        ("n_patches", C.c_size_t), ("patches", MeshPatchStats * 3)]


class MeshResult(C.Structure):
    _fields_ = [("stats", MeshStats), ("n_triangles", C.c_size_t),
                ("triangles", C.POINTER(MeshTriangle))]


TRIANGLE_DTYPE = np.dtype([("a", "<u8"), ("b", "<u8"), ("c", "<u8"), ("confidence", "<f4"),
                           ("pad", "<u4")])


POLL_FN = C.CFUNCTYPE(C.c_int32, C.c_void_p)

POINT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("confidence", "<f4"),
                        ("r", "u1"), ("g", "u1"), ("b", "u1"), ("pad", "u1")])
SITE_DTYPE = np.dtype([("x", "<u4"), ("y", "<u4")])
ATTR_DTYPE = np.dtype([("reciprocal_consistency", "<f4"), ("depth_margin", "<f4")])
SAMPLE_DTYPE = np.dtype([("x", "<u4"), ("y", "<u4"), ("status", "<i4"), ("best_hypothesis", "<i4"),
                         ("best_support", "<i4"), ("second_hypothesis", "<i4"),
                         ("best_cost", "<f4"), ("second_cost", "<f4"),
                         ("reciprocal_mask", "<u4"), ("fused_observations", "<u4"),
                         ("reciprocal_signature", "<u4"), ("pad", "<u4")])

assert C.sizeof(Point3) == POINT_DTYPE.itemsize
assert C.sizeof(SampleDiag) == SAMPLE_DTYPE.itemsize


@dataclasses.dataclass
class DenseAnalysis:
    """DenseAnalysis (crates/video-to-3d-core/src/dense_legacy.rs:61-68) as numpy arrays."""
    stats: Dict[str, object]
    points: np.ndarray        # POINT_DTYPE
    grid_sites: np.ndarray    # SITE_DTYPE
    point_attributes: np.ndarray  # ATTR_DTYPE
    primary_points: int
    samples: Optional[np.ndarray]  # SAMPLE_DTYPE, texture-gated samples in row-major order
    skip_code: int


def _array(ptr, n: int, dtype: np.dtype, owner=None) -> np.ndarray:
    """n records at ptr. With an owner, a view of the C buffer that keeps the owner (which
    frees the buffer) alive; otherwise a copy."""
    if n == 0 or not ptr:
        return np.zeros(0, dtype)
    raw = (C.c_uint8 * (n * dtype.itemsize)).from_address(C.cast(ptr, C.c_void_p).value)
    if owner is None:
        return np.frombuffer(raw, dtype=dtype).copy()
    raw._owner = owner
    return np.frombuffer(raw, dtype=dtype)


def unpack_result(res: DenseResult, owner=None) -> DenseAnalysis:
    """DenseAnalysis of a C result. With `owner` (an object that frees the result when it is
    collected) the point arrays are views of the C buffers; without, copies."""
    st = res.stats
    stats: Dict[str, object] = {}
    for name, _ in DenseStats._fields_:
        if name.startswith("pad") or name.startswith("has_"):
            continue
        stats[name] = getattr(st, name)
    stats["attempted"] = bool(st.attempted)
    stats["source_frames"] = [int(st.source_frames[i]) for i in range(min(int(st.source_views), 32))]
    stats["reference_frame"] = int(st.reference_frame) if st.has_reference_frame else None
    for opt in ("median_fusion_observations", "median_supporting_views", "median_photometric_error"):
        stats[opt] = float(getattr(st, opt)) if getattr(st, "has_" + opt) else None
    if st.has_search_depth:
        stats["search_min_depth"] = float(st.search_min_depth)
        stats["search_max_depth"] = float(st.search_max_depth)
    else:
        stats["search_min_depth"] = stats["search_max_depth"] = None
    reason = res.skip_reason.decode()
    stats["skip_reason"] = reason if res.skip_code else None
    n = int(res.n_points)
    samples = (_array(res.samples, int(res.n_samples), SAMPLE_DTYPE, owner) if res.samples
               else None)
    return DenseAnalysis(stats, _array(res.points, n, POINT_DTYPE, owner),
                         _array(res.grid_sites, n, SITE_DTYPE, owner),
                         _array(res.attributes, n, ATTR_DTYPE, owner), int(res.primary_points),
                         samples, int(res.skip_code))


@dataclasses.dataclass
class MultiAnalysis:
    """The multi-reference DenseAnalysis of crates/video-to-3d-core/src/dense.rs:100-146
    (without the working-set fields): `combined` with stats["reference_frames"],
    stats["reference_attempts"] and stats["reference_patches"] filled in."""
    combined: DenseAnalysis
    omitted_references: bool
    fine_grid: bool
    # an object whose `ptr` is the C result (POINTER(MultiResult)) these arrays view, kept
    # alive with them; None for copies
    handle: object = dataclasses.field(default=None, repr=False, compare=False)

    @property
    def stats(self) -> Dict[str, object]:
        return self.combined.stats


def _attempt(a: ReferenceAttempt) -> Dict[str, object]:
    d = {n: int(getattr(a, n)) for n, _ in ReferenceAttempt._fields_
         if n not in ("attempted", "accepted", "skip_reason")}
    d["attempted"] = bool(a.attempted)
    d["accepted"] = bool(a.accepted)
    d["skip_reason"] = None if a.accepted else a.skip_reason.decode()
    return d


def _patch(p: ReferencePatchStats) -> Dict[str, object]:
    d = {n: int(getattr(p, n)) for n, _ in ReferencePatchStats._fields_
         if n not in ("n_source_frames", "source_frames", "has_search_depth",
                      "search_min_depth", "search_max_depth")}
    d["source_frames"] = [int(p.source_frames[i]) for i in range(min(int(p.n_source_frames), 32))]
    d["search_min_depth"] = float(p.search_min_depth) if p.has_search_depth else None
    d["search_max_depth"] = float(p.search_max_depth) if p.has_search_depth else None
    return d


def unpack_multi(m: MultiResult, owner=None) -> MultiAnalysis:
    """MultiAnalysis of a C multi result (views of its buffers with `owner`, else copies)."""
    combined = unpack_result(m.combined, owner)
    st = combined.stats
    st["reference_frames"] = [int(m.reference_frames[i]) for i in range(int(m.n_reference_frames))]
    st["reference_attempts"] = [_attempt(m.attempts[i]) for i in range(int(m.n_attempts))]
    st["reference_patches"] = [_patch(m.patches[i]) for i in range(int(m.n_patches))]
    # dense.rs:132-139 (set by the budgeted entry only)
    st["working_set_estimate"] = m.working_set_estimate.as_dict()
    st["working_set_budget_bytes"] = int(m.working_set_budget_bytes)
    st["full_working_set_bytes"] = int(m.full_working_set_bytes)
    st["reference_view_limit"] = (int(m.reference_view_limit) if m.has_reference_view_limit
                                  else None)
    return MultiAnalysis(combined, bool(m.omitted_references), bool(m.fine_grid), owner)


@dataclasses.dataclass
class MeshAnalysis:
    """MeshAnalysis (crates/video-to-3d-core/src/mesh.rs:45-49): stats dict and triangles."""
    stats: Dict[str, object]
    triangles: np.ndarray  # TRIANGLE_DTYPE


def _mesh_patch(p: "MeshPatchStats") -> Dict[str, object]:
    d = {n: int(getattr(p, n)) for n in ("reference_frame", "point_count") + _MESH_COUNTS}
    d["attempted"] = bool(p.attempted)
    d["skip_reason"] = p.skip_reason.decode() if p.has_skip_reason else None
    return d


def unpack_mesh(m: "MeshResult") -> MeshAnalysis:
    st = m.stats
    d = {n: int(getattr(st, n)) for n in _MESH_COUNTS}
    d["attempted"] = bool(st.attempted)
    d["skip_reason"] = st.skip_reason.decode() if st.has_skip_reason else None
    d["reference_frame"] = int(st.reference_frame) if st.has_reference_frame else None
    d["reference_frames"] = [int(st.reference_frames[i]) for i in range(int(st.n_reference_frames))]
    d["reference_patches"] = [_mesh_patch(st.patches[i]) for i in range(int(st.n_patches))]
    return MeshAnalysis(d, _array(m.triangles, int(m.n_triangles), TRIANGLE_DTYPE))


def dense_stats_struct(stats: Dict[str, object]) -> "DenseStats":
    """The DenseStats fields a mesh call reads, from an analysis' stats dict."""
    st = DenseStats()
    st.attempted = int(bool(stats["attempted"]))
    if stats.get("reference_frame") is not None:
        st.has_reference_frame = 1
        st.reference_frame = int(stats["reference_frame"])
    st.grid_stride = int(stats.get("grid_stride") or 0)
    st.grid_border = int(stats.get("grid_border") or 0)
    return st


def patch_structs(patches) -> "C.Array":
    """vt3d_reference_patch_stats[] from stats["reference_patches"] dicts."""
    arr = (ReferencePatchStats * max(1, len(patches)))()
    for i, p in enumerate(patches):
        q = arr[i]
        for n in ("reference_frame", "primary_start", "primary_points", "completion_start",
                  "completed_points", "sampled_pixels", "accepted_points",
                  "reciprocal_consistent_points", "grid_stride", "grid_border"):
            setattr(q, n, int(p[n]))
        q.n_source_frames = len(p["source_frames"])
        for j, f in enumerate(p["source_frames"][:32]):
            q.source_frames[j] = int(f)
        if p.get("search_min_depth") is not None:
            q.has_search_depth = 1
            q.search_min_depth = float(p["search_min_depth"])
            q.search_max_depth = float(p["search_max_depth"])
    return arr


class Inputs:
    """Frames, cameras and landmarks marshalled for one call (keeps the buffers alive)."""

    def __init__(self, frames_rgba: Sequence[np.ndarray], cameras: Sequence, sparse: np.ndarray):
        self._rgba = [np.ascontiguousarray(f, dtype=np.uint8) for f in frames_rgba]
        self.frames = (Frame * max(1, len(self._rgba)))()
        for i, f in enumerate(self._rgba):
            h, w = f.shape[:2]
            self.frames[i] = Frame(w, h, f.ctypes.data_as(C.POINTER(C.c_uint8)))
        self.n_frames = len(self._rgba)
        self.cameras = (Camera * max(1, len(cameras)))()
        for i, cam in enumerate(cameras):
            if isinstance(cam, tuple):
                idx, R, t = cam
            else:
                idx, R, t = i, cam.R, cam.t
            self.cameras[i].frame_index = int(idx)
            Rc = np.asarray(R, dtype=np.float64).reshape(3, 3).T.reshape(-1)  # column-major
            for k in range(9):
                self.cameras[i].rotation_colmajor[k] = float(Rc[k])
            for k in range(3):
                self.cameras[i].translation[k] = float(np.asarray(t, dtype=np.float64)[k])
        self.n_cameras = len(cameras)
        self.sparse = np.ascontiguousarray(np.asarray(sparse, dtype=np.float64).reshape(-1, 3))
        self.sparse_ptr = self.sparse.ctypes.data_as(C.POINTER(C.c_double))
        self.n_sparse = self.sparse.shape[0]


def make_params(lib_reference, lib_zncc, settings: Optional[dict] = None, **overrides) -> SweepParams:
    """vt3d_sweep_params from a settings dict ({"preset": "reference"|"zncc", ...})."""
    s = dict(settings or {})
    s.update(overrides)
    p = SweepParams()
    if s.get("preset", "reference") == "zncc":
        lib_zncc(C.byref(p), int(s.get("patch_radius", 3)), int(s.get("depth_hypotheses", 64)),
                 int(s.get("max_source_views", 4)))
    else:
        lib_reference(C.byref(p))
    mapping = {"patch_radius": "patch_radius", "depth_hypotheses": "depth_hypotheses",
               "max_source_views": "max_source_views", "grid_stride": "grid_stride",
               "surface_completion_passes": "surface_completion_passes",
               "emit_samples": "emit_samples", "min_fusion_observations": "min_fusion_observations",
               "depth_min": "depth_min_override", "depth_max": "depth_max_override",
               "max_photometric_error": "max_photometric_error",
               "ambiguity_threshold": "ambiguity_threshold",
               "min_reference_contrast": "min_reference_contrast"}
    for key, field in mapping.items():
        if key in s:
            setattr(p, field, type(getattr(p, field))(s[key]))
    if "storage" in s:
        p.storage = STORAGE[s["storage"]] if isinstance(s["storage"], str) else int(s["storage"])
    if "cost_kind" in s:
        p.cost_kind = int(s["cost_kind"])
    if "kernel" in s:
        p.sweep_kernel = KERNEL[s["kernel"]] if isinstance(s["kernel"], str) else int(s["kernel"])
    return p


def bind_common(lib, prefix: str) -> None:
    """Declares the entry points shared by the product and the oracle libraries."""
    pfx = prefix
    getattr(lib, f"{pfx}_result_free").argtypes = [C.POINTER(DenseResult)]
    getattr(lib, f"{pfx}_result_free").restype = None
