"""vt3d_dense -- MI355X dense plane-sweep depth pass (Python mirror of the C ABI).

The functions below mirror the reference's Rust interface for this path:

* `estimate_depth_points_with_control(frames, cameras, sparse_points, focal, fine_grid,
  control)` -- crates/video-to-3d-core/src/dense_legacy.rs:149-156
* `reconstruct_reference_patch(frames, cameras, sparse_points, focal, reference_camera_index,
  fine_grid, control)` -- crates/video-to-3d-core/src/dense.rs:674-747
* `estimate_depth_points_limited(frames, cameras, sparse_points, focal, reference_views,
  fine_grid, control)` -- crates/video-to-3d-core/src/dense.rs:244-330

They run the HIP kernels of libvt3d_dense.so on the current GPU. There is no CPU path: if the
library is missing or no GPU is present, they raise.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Callable, Optional, Sequence

import numpy as np

from . import abi
from .abi import DenseAnalysis

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libvt3d_dense.so")

_lib = None


class Vt3dError(RuntimeError):
    pass


def library() -> C.CDLL:
    """Loads libvt3d_dense.so from this package directory (built by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY_PATH):
        raise Vt3dError(f"{LIBRARY_PATH} is missing: build it with __graft_entry__.build()")
    lib = C.CDLL(LIBRARY_PATH)
    _bind(lib)
    _lib = lib
    return lib


def _bind(lib: C.CDLL) -> None:
    P = C.POINTER
    lib.vt3d_sweep_params_reference.argtypes = [P(abi.SweepParams)]
    lib.vt3d_sweep_params_reference.restype = None
    lib.vt3d_sweep_params_zncc.argtypes = [P(abi.SweepParams), C.c_int32, C.c_int32, C.c_int32]
    lib.vt3d_sweep_params_zncc.restype = None
    lib.vt3d_skip_reason.argtypes = [C.c_int32]
    lib.vt3d_skip_reason.restype = C.c_char_p
    lib.vt3d_grid_stride.argtypes = [C.c_uint32, C.c_uint32, C.c_int32]
    lib.vt3d_grid_stride.restype = C.c_uint32
    lib.vt3d_abi_version.argtypes = []
    lib.vt3d_abi_version.restype = C.c_int32
    lib.vt3d_plane_homography.argtypes = [P(abi.Camera), C.c_uint32, C.c_uint32, P(abi.Camera),
                                          C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                          P(C.c_double)]
    lib.vt3d_plane_homography.restype = None
    lib.vt3d_last_error.argtypes = []
    lib.vt3d_last_error.restype = C.c_char_p
    common = [P(abi.Frame), C.c_size_t, P(abi.Camera), C.c_size_t, P(C.c_double), C.c_size_t,
              C.c_double]
    lib.vt3d_dense_plane_sweep.argtypes = common + [C.c_int32, P(abi.SweepParams), abi.POLL_FN,
                                                    C.c_void_p, P(P(abi.DenseResult))]
    lib.vt3d_dense_plane_sweep.restype = C.c_int32
    lib.vt3d_dense_reference_patch.argtypes = common + [C.c_size_t, C.c_int32, P(abi.SweepParams),
                                                        abi.POLL_FN, C.c_void_p,
                                                        P(P(abi.DenseResult))]
    lib.vt3d_dense_reference_patch.restype = C.c_int32
    lib.vt3d_dense_plane_sweep_noalloc.argtypes = common + [
        C.c_int32, P(abi.SweepParams), abi.POLL_FN, C.c_void_p, P(abi.DenseResult), C.c_size_t,
        C.c_size_t]
    lib.vt3d_dense_plane_sweep_noalloc.restype = C.c_int32
    lib.vt3d_dense_result_free.argtypes = [P(abi.DenseResult)]
    lib.vt3d_dense_result_free.restype = None
    lib.vt3d_dense_estimate.argtypes = common + [C.c_int32, C.c_int32, P(abi.SweepParams),
                                                 abi.POLL_FN, C.c_void_p, P(P(abi.MultiResult))]
    lib.vt3d_dense_estimate.restype = C.c_int32
    lib.vt3d_estimate_working_set.argtypes = [P(abi.Frame), C.c_size_t, C.c_uint64, C.c_uint32,
                                              C.c_size_t, C.c_int32, P(abi.WorkingSet)]
    lib.vt3d_estimate_working_set.restype = None
    lib.vt3d_dense_estimate_with_budget.argtypes = common + [
        C.c_uint64, C.c_uint64, C.c_uint32, P(abi.SweepParams), abi.POLL_FN, C.c_void_p,
        P(P(abi.MultiResult))]
    lib.vt3d_dense_estimate_with_budget.restype = C.c_int32
    lib.vt3d_reconstruction_options_default.argtypes = [P(abi.ReconstructionOptions)]
    lib.vt3d_reconstruction_options_default.restype = None
    lib.vt3d_request_content_hash.argtypes = [P(abi.Frame), C.c_size_t,
                                              P(abi.ReconstructionOptions), abi.POLL_FN,
                                              C.c_void_p, C.c_char_p]
    lib.vt3d_request_content_hash.restype = C.c_int32
    lib.vt3d_content_hash.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p]
    lib.vt3d_content_hash.restype = None
    lib.vt3d_multi_result_free.argtypes = [P(abi.MultiResult)]
    lib.vt3d_multi_result_free.restype = None
    lib.vt3d_dense_artifact_create.argtypes = [P(abi.MultiResult), C.c_char_p, C.c_char_p,
                                               P(P(abi.DenseArtifact)), C.c_char_p, C.c_size_t]
    lib.vt3d_dense_artifact_create.restype = C.c_int32
    lib.vt3d_dense_artifact_to_json.argtypes = [P(abi.DenseArtifact), C.c_void_p, C.c_size_t,
                                                P(C.c_size_t), C.c_char_p, C.c_size_t]
    lib.vt3d_dense_artifact_to_json.restype = C.c_int32
    lib.vt3d_dense_artifact_content_hash.argtypes = [P(abi.DenseArtifact), C.c_char_p]
    lib.vt3d_dense_artifact_content_hash.restype = C.c_int32
    lib.vt3d_dense_artifact_from_json.argtypes = [C.c_char_p, C.c_size_t,
                                                  P(P(abi.DenseArtifact)), P(C.c_uint32),
                                                  C.c_char_p, C.c_size_t]
    lib.vt3d_dense_artifact_from_json.restype = C.c_int32
    lib.vt3d_dense_artifact_free.argtypes = [P(abi.DenseArtifact)]
    lib.vt3d_dense_artifact_free.restype = None
    u8p = P(C.c_uint8)
    lib.vt3d_pack_dense_points.argtypes = [P(abi.Point3), C.c_size_t, u8p, u8p]
    lib.vt3d_pack_dense_points.restype = C.c_size_t
    lib.vt3d_pack_grid_sites.argtypes = [P(abi.GridSite), C.c_size_t, u8p]
    lib.vt3d_pack_grid_sites.restype = C.c_size_t
    lib.vt3d_pack_point_attributes.argtypes = [P(abi.PointAttributes), C.c_size_t, u8p]
    lib.vt3d_pack_point_attributes.restype = C.c_size_t
    lib.vt3d_unpack_dense_points.argtypes = [u8p, C.c_size_t, C.c_size_t, P(abi.Point3)]
    lib.vt3d_unpack_dense_points.restype = C.c_int32
    lib.vt3d_dense_warning.argtypes = [P(abi.DenseStats), C.c_char_p, C.c_size_t]
    lib.vt3d_dense_warning.restype = C.c_size_t
    lib.vt3d_dense_mesh.argtypes = [P(abi.Point3), C.c_size_t, P(abi.GridSite), C.c_size_t,
                                    P(abi.DenseStats), P(abi.ReferencePatchStats), C.c_size_t,
                                    P(abi.Camera), C.c_size_t, C.c_uint32, C.c_uint32,
                                    C.c_double, abi.POLL_FN, C.c_void_p, P(P(abi.MeshResult))]
    lib.vt3d_dense_mesh.restype = C.c_int32
    lib.vt3d_mesh_result_free.argtypes = [P(abi.MeshResult)]
    lib.vt3d_mesh_result_free.restype = None
    lib.vt3d_mesh_warning.argtypes = [P(abi.MeshStats), C.c_char_p, C.c_size_t]
    lib.vt3d_mesh_warning.restype = C.c_size_t
    lib.vt3d_session_create.argtypes = [C.c_int32, P(C.c_void_p)]
    lib.vt3d_session_create.restype = C.c_int32
    lib.vt3d_session_destroy.argtypes = [C.c_void_p]
    lib.vt3d_session_destroy.restype = None
    lib.vt3d_session_upload_frames.argtypes = [C.c_void_p, P(abi.Frame), C.c_size_t, C.c_int32]
    lib.vt3d_session_upload_frames.restype = C.c_int32
    lib.vt3d_session_solve.argtypes = [C.c_void_p, P(abi.Camera), C.c_size_t, P(C.c_double),
                                       C.c_size_t, C.c_double, C.c_int64, C.c_int32,
                                       P(abi.SweepParams), abi.POLL_FN, C.c_void_p,
                                       P(P(abi.DenseResult))]
    lib.vt3d_session_solve.restype = C.c_int32
    lib.vt3d_session_last_timing.argtypes = [C.c_void_p, P(abi.Timing)]
    lib.vt3d_session_last_timing.restype = C.c_int32
    lib.vt3d_session_set_tuning.argtypes = [C.c_void_p, P(abi.Tuning)]
    lib.vt3d_session_set_tuning.restype = C.c_int32
    lib.vt3d_batch_create.argtypes = [C.c_void_p, P(abi.Request), C.c_size_t, C.c_int32,
                                      P(C.c_void_p)]
    lib.vt3d_batch_create.restype = C.c_int32
    lib.vt3d_batch_run.argtypes = [C.c_void_p, P(P(abi.DenseResult))]
    lib.vt3d_batch_run.restype = C.c_int32
    lib.vt3d_batch_submit.argtypes = [C.c_void_p]
    lib.vt3d_batch_submit.restype = C.c_int32
    lib.vt3d_batch_collect.argtypes = [C.c_void_p, P(P(abi.DenseResult))]
    lib.vt3d_batch_collect.restype = C.c_int32
    lib.vt3d_batch_last_timing.argtypes = [C.c_void_p, P(abi.Timing)]
    lib.vt3d_batch_last_timing.restype = C.c_int32
    lib.vt3d_batch_destroy.argtypes = [C.c_void_p]
    lib.vt3d_batch_destroy.restype = None
    lib.vt3d_server_create.argtypes = [C.c_void_p, P(abi.ServerConfig), P(C.c_void_p)]
    lib.vt3d_server_create.restype = C.c_int32
    lib.vt3d_server_submit.argtypes = [C.c_void_p, P(abi.Request), C.c_uint64]
    lib.vt3d_server_submit.restype = C.c_int32
    lib.vt3d_server_poll.argtypes = [C.c_void_p, P(abi.Completion), C.c_size_t, C.c_int32,
                                     P(C.c_size_t)]
    lib.vt3d_server_poll.restype = C.c_int32
    lib.vt3d_server_get_stats.argtypes = [C.c_void_p, P(abi.ServerStats)]
    lib.vt3d_server_get_stats.restype = C.c_int32
    lib.vt3d_server_destroy.argtypes = [C.c_void_p]
    lib.vt3d_server_destroy.restype = None


EXPORTED_SYMBOLS = (
    "vt3d_sweep_params_reference", "vt3d_sweep_params_zncc", "vt3d_skip_reason",
    "vt3d_grid_stride", "vt3d_abi_version", "vt3d_dense_plane_sweep",
    "vt3d_dense_reference_patch", "vt3d_dense_result_free", "vt3d_session_create",
    "vt3d_session_destroy", "vt3d_session_upload_frames", "vt3d_session_solve",
    "vt3d_session_last_timing", "vt3d_last_error", "vt3d_plane_homography",
    "vt3d_session_set_tuning", "vt3d_batch_create", "vt3d_batch_run",
    "vt3d_batch_submit", "vt3d_batch_collect",
    "vt3d_batch_last_timing", "vt3d_batch_destroy", "vt3d_server_create",
    "vt3d_server_submit", "vt3d_server_poll", "vt3d_server_get_stats", "vt3d_server_destroy",
    "vt3d_dense_estimate", "vt3d_multi_result_free", "vt3d_pack_dense_points",
    "vt3d_pack_grid_sites", "vt3d_pack_point_attributes", "vt3d_unpack_dense_points",
    "vt3d_dense_warning", "vt3d_dense_mesh", "vt3d_mesh_result_free", "vt3d_mesh_warning",
    "vt3d_dense_plane_sweep_noalloc", "vt3d_estimate_working_set",
    "vt3d_dense_estimate_with_budget", "vt3d_reconstruction_options_default",
    "vt3d_request_content_hash", "vt3d_content_hash", "vt3d_dense_artifact_create",
    "vt3d_dense_artifact_to_json", "vt3d_dense_artifact_content_hash",
    "vt3d_dense_artifact_from_json", "vt3d_dense_artifact_free",
)


def params(settings: Optional[dict] = None, **overrides) -> abi.SweepParams:
    lib = library()
    return abi.make_params(lib.vt3d_sweep_params_reference, lib.vt3d_sweep_params_zncc,
                           settings, **overrides)


def _poll(control: Optional[Callable[[], bool]]):
    if control is None:
        return abi.POLL_FN(0)
    return abi.POLL_FN(lambda _user: 1 if control() else 0)


def _check(rc: int) -> None:
    if rc < 0:
        msg = library().vt3d_last_error().decode()
        raise Vt3dError(f"vt3d call failed ({rc}): {msg}")


class _ResultOwner:
    """Frees a C result once no array viewing its buffers is left."""

    __slots__ = ("ptr", "free")

    def __init__(self, ptr, free):
        self.ptr = ptr
        self.free = free

    def __del__(self):
        if self.ptr:
            self.free(self.ptr)
            self.ptr = None


def _take(ptr) -> DenseAnalysis:
    """Unpacks a C result; its arrays view the C buffers, which are freed with the last
    array (or at once if the result has none)."""
    owner = _ResultOwner(ptr, library().vt3d_dense_result_free)
    return abi.unpack_result(ptr.contents, owner)


def _rgba(frames: Sequence[np.ndarray]):
    from .scenes import to_rgba
    return [f if f.ndim == 3 else to_rgba(f) for f in frames]


def estimate_depth_points_with_control(frames, cameras, sparse_points, focal: float,
                                       fine_grid: bool = False,
                                       control: Optional[Callable[[], bool]] = None,
                                       sweep: Optional[abi.SweepParams] = None) -> DenseAnalysis:
    """dense_legacy.rs:149-156 on the GPU. `frames`: gray (H, W) or RGBA (H, W, 4) uint8;
    `cameras`: objects with .R/.t (frame_index = position) or (frame_index, R, t) tuples."""
    lib = library()
    inp = abi.Inputs(_rgba(frames), cameras, sparse_points)
    out = C.POINTER(abi.DenseResult)()
    poll = _poll(control)
    rc = lib.vt3d_dense_plane_sweep(inp.frames, inp.n_frames, inp.cameras, inp.n_cameras,
                                    inp.sparse_ptr, inp.n_sparse, float(focal), int(fine_grid),
                                    C.byref(sweep) if sweep is not None else None, poll, None,
                                    C.byref(out))
    _check(rc)
    return _take(out)


def estimate_depth_points_noalloc(frames, cameras, sparse_points, focal: float,
                                  fine_grid: bool = False,
                                  control: Optional[Callable[[], bool]] = None,
                                  sweep: Optional[abi.SweepParams] = None,
                                  point_capacity: int = 0, sample_capacity: int = 0):
    """estimate_depth_points_with_control through vt3d_dense_plane_sweep_noalloc: the library
    allocates no device memory; the result arrays are numpy arrays of the given capacities
    (default: one entry per grid site of the first frame). Returns (return code,
    DenseAnalysis); a code below zero leaves the stats and the counts needed."""
    lib = library()
    inp = abi.Inputs(_rgba(frames), cameras, sparse_points)
    if not point_capacity:
        h, w = (frames[0].shape[:2] if len(frames) else (0, 0))
        point_capacity = max(1, h * w)
    if not sample_capacity and sweep is not None and sweep.emit_samples:
        sample_capacity = point_capacity
    pts = np.zeros(point_capacity, abi.POINT_DTYPE)
    sites = np.zeros(point_capacity, abi.SITE_DTYPE)
    attrs = np.zeros(point_capacity, abi.ATTR_DTYPE)
    samples = np.zeros(max(1, sample_capacity), abi.SAMPLE_DTYPE)
    res = abi.DenseResult()
    res.points = pts.ctypes.data_as(C.POINTER(abi.Point3))
    res.grid_sites = sites.ctypes.data_as(C.POINTER(abi.GridSite))
    res.attributes = attrs.ctypes.data_as(C.POINTER(abi.PointAttributes))
    res.samples = samples.ctypes.data_as(C.POINTER(abi.SampleDiag)) if sample_capacity \
        else C.POINTER(abi.SampleDiag)()
    poll = _poll(control)
    rc = lib.vt3d_dense_plane_sweep_noalloc(
        inp.frames, inp.n_frames, inp.cameras, inp.n_cameras, inp.sparse_ptr, inp.n_sparse,
        float(focal), int(fine_grid), C.byref(sweep) if sweep is not None else None, poll, None,
        C.byref(res), point_capacity, sample_capacity)
    if rc < 0:
        return rc, abi.unpack_result(_counts_only(res))
    return rc, abi.unpack_result(res)  # copies of the caller-side arrays


def _counts_only(res: abi.DenseResult) -> abi.DenseResult:
    """The stats and counts of a result whose arrays were not written."""
    q = abi.DenseResult()
    C.pointer(q)[0] = res
    q.points = C.POINTER(abi.Point3)()
    q.grid_sites = C.POINTER(abi.GridSite)()
    q.attributes = C.POINTER(abi.PointAttributes)()
    q.samples = C.POINTER(abi.SampleDiag)()
    return q


def reconstruct_reference_patch(frames, cameras, sparse_points, focal: float,
                                reference_camera_index: int, fine_grid: bool = False,
                                control: Optional[Callable[[], bool]] = None,
                                sweep: Optional[abi.SweepParams] = None) -> Optional[DenseAnalysis]:
    """dense.rs:674-747 on the GPU; None where the reference returns None."""
    lib = library()
    inp = abi.Inputs(_rgba(frames), cameras, sparse_points)
    out = C.POINTER(abi.DenseResult)()
    poll = _poll(control)
    rc = lib.vt3d_dense_reference_patch(inp.frames, inp.n_frames, inp.cameras, inp.n_cameras,
                                        inp.sparse_ptr, inp.n_sparse, float(focal),
                                        int(reference_camera_index), int(fine_grid),
                                        C.byref(sweep) if sweep is not None else None, poll, None,
                                        C.byref(out))
    if rc == abi.VT3D_NO_PATCH:
        return None
    _check(rc)
    return _take(out)


def estimate_depth_points_limited(frames, cameras, sparse_points, focal: float,
                                  reference_views: int = abi.MAX_REFERENCE_VIEWS,
                                  fine_grid: Optional[bool] = None,
                                  control: Optional[Callable[[], bool]] = None,
                                  sweep: Optional[abi.SweepParams] = None):
    """dense.rs:244-330 on the GPU: returns (MultiAnalysis, omitted_references). fine_grid None
    decides the grid as the budgeted entry does (dense.rs:366-372, no working-set budget)."""
    lib = library()
    inp = abi.Inputs(_rgba(frames), cameras, sparse_points)
    out = C.POINTER(abi.MultiResult)()
    poll = _poll(control)
    fine = -1 if fine_grid is None else int(bool(fine_grid))
    rc = lib.vt3d_dense_estimate(inp.frames, inp.n_frames, inp.cameras, inp.n_cameras,
                                 inp.sparse_ptr, inp.n_sparse, float(focal), int(reference_views),
                                 fine, C.byref(sweep) if sweep is not None else None, poll, None,
                                 C.byref(out))
    _check(rc)
    owner = _ResultOwner(out, lib.vt3d_multi_result_free)
    m = abi.unpack_multi(out.contents, owner)
    return m, m.omitted_references


def reconstruction_options(**overrides) -> abi.ReconstructionOptions:
    """ReconstructionOptions::default (reconstruction/types.rs:37-50) with overrides;
    focal_length_pixels=None clears the focal."""
    o = abi.ReconstructionOptions()
    library().vt3d_reconstruction_options_default(C.byref(o))
    for k, v in overrides.items():
        if k == "focal_length_pixels":
            o.has_focal_length_pixels = 0 if v is None else 1
            o.focal_length_pixels = 0.0 if v is None else float(v)
        else:
            setattr(o, k, v)
    return o


def request_content_hash(frames, options: Optional[abi.ReconstructionOptions] = None,
                         control: Optional[Callable[[], bool]] = None) -> str:
    """request_content_hash (reconstruction_stages.rs:458-510): "sha256:..." of the frames
    ((H, W) gray or (H, W, 4) RGBA uint8) and options. Raises Vt3dError on a cancel."""
    lib = library()
    inp = abi.Inputs(_rgba(frames), [], np.zeros((0, 3)))
    opts = options if options is not None else reconstruction_options()
    buf = C.create_string_buffer(72)
    _check(lib.vt3d_request_content_hash(inp.frames, inp.n_frames, C.byref(opts),
                                         _poll(control), None, buf))
    return buf.value.decode()


def content_hash(data: bytes) -> str:
    """ContentHash::of_bytes (scene_project.rs:463-465)."""
    buf = C.create_string_buffer(72)
    library().vt3d_content_hash(bytes(data), len(data), buf)
    return buf.value.decode()


def estimate_working_set(frames, max_features: int = 320, descriptor_radius: int = 3,
                         reference_views: int = abi.MAX_REFERENCE_VIEWS,
                         fine_grid: bool = False) -> dict:
    """estimate_working_set (dense.rs:423-605) of frames ((H, W) or (H, W, 4) arrays, or
    (width, height) pairs): the DenseWorkingSetEstimate as a dict. Host arithmetic."""
    dims = [f if isinstance(f, tuple) else (f.shape[1], f.shape[0]) for f in frames]
    fr = (abi.Frame * max(1, len(dims)))()
    for i, (w_, h_) in enumerate(dims):
        fr[i].width, fr[i].height = int(w_), int(h_)
    w = abi.WorkingSet()
    library().vt3d_estimate_working_set(fr, len(dims), int(max_features),
                                        int(descriptor_radius), int(reference_views),
                                        int(bool(fine_grid)), C.byref(w))
    return w.as_dict()


def estimate_depth_points_with_budget(frames, cameras, sparse_points, focal: float,
                                      working_set_budget_bytes: int, max_features: int = 320,
                                      descriptor_radius: int = 3,
                                      control: Optional[Callable[[], bool]] = None,
                                      sweep: Optional[abi.SweepParams] = None):
    """estimate_depth_points_with_budget_and_control (dense.rs:354-420) on the GPU."""
    lib = library()
    inp = abi.Inputs(_rgba(frames), cameras, sparse_points)
    out = C.POINTER(abi.MultiResult)()
    poll = _poll(control)
    rc = lib.vt3d_dense_estimate_with_budget(
        inp.frames, inp.n_frames, inp.cameras, inp.n_cameras, inp.sparse_ptr, inp.n_sparse,
        float(focal), int(working_set_budget_bytes), int(max_features), int(descriptor_radius),
        C.byref(sweep) if sweep is not None else None, poll, None, C.byref(out))
    _check(rc)
    owner = _ResultOwner(out, lib.vt3d_multi_result_free)
    return abi.unpack_multi(out.contents, owner)


def estimate_depth_points(frames, cameras, sparse_points, focal: float,
                          control: Optional[Callable[[], bool]] = None,
                          sweep: Optional[abi.SweepParams] = None) -> "abi.MultiAnalysis":
    """dense.rs:222-244: all three reference views, the grid decided from the landmarks."""
    return estimate_depth_points_limited(frames, cameras, sparse_points, focal,
                                         abi.MAX_REFERENCE_VIEWS, None, control, sweep)[0]


def pack_dense(analysis: DenseAnalysis) -> dict:
    """The WASM adapter's packed buffers (crates/video-to-3d-wasm/src/lib.rs:227-268) of an
    analysis: {"dense_points": {"values_f32_le", "rgb", "length"}, "dense_grid_sites":
    {"xy_u32_le", "length"}, "dense_point_attributes": {"values_f32_le", "length"}}."""
    lib = library()
    pts = np.ascontiguousarray(analysis.points)
    sites = np.ascontiguousarray(analysis.grid_sites)
    attrs = np.ascontiguousarray(analysis.point_attributes)
    n, ns, na = len(pts), len(sites), len(attrs)
    xyzw = (C.c_uint8 * max(1, 16 * n))()
    rgb = (C.c_uint8 * max(1, 3 * n))()
    xy = (C.c_uint8 * max(1, 8 * ns))()
    av = (C.c_uint8 * max(1, 8 * na))()
    lib.vt3d_pack_dense_points(pts.ctypes.data_as(C.POINTER(abi.Point3)), n, xyzw, rgb)
    lib.vt3d_pack_grid_sites(sites.ctypes.data_as(C.POINTER(abi.GridSite)), ns, xy)
    lib.vt3d_pack_point_attributes(attrs.ctypes.data_as(C.POINTER(abi.PointAttributes)), na, av)
    return {"dense_points": {"values_f32_le": bytes(xyzw)[:16 * n], "rgb": bytes(rgb)[:3 * n],
                             "length": n},
            "dense_grid_sites": {"xy_u32_le": bytes(xy)[:8 * ns], "length": ns},
            "dense_point_attributes": {"values_f32_le": bytes(av)[:8 * na], "length": na}}


def dense_warning(stats: dict) -> Optional[str]:
    """dense_warning (reconstruction/warnings.rs:27-68) of an analysis' stats dict."""
    st = abi.DenseStats()
    for k in ("accepted_points", "source_views", "depth_hypotheses",
              "reciprocal_consistent_points", "reciprocal_checked_points",
              "fusion_rejected_observations"):
        setattr(st, k, int(stats[k]))
    st.attempted = int(bool(stats["attempted"]))
    if stats.get("reference_frame") is not None:
        st.has_reference_frame = 1
        st.reference_frame = int(stats["reference_frame"])
    lib = library()
    n = lib.vt3d_dense_warning(C.byref(st), None, 0)
    if n == 0:
        return None
    buf = C.create_string_buffer(n + 1)
    lib.vt3d_dense_warning(C.byref(st), buf, n + 1)
    return buf.value.decode()


def reconstruct_dense_mesh(points, grid_sites, dense_stats: dict, cameras, width: int,
                           height: int, focal: float,
                           control: Optional[Callable[[], bool]] = None) -> "abi.MeshAnalysis":
    """mesh.rs:77-215 on the GPU: reference-grid triangulation of an analysis' points.
    `dense_stats` is the analysis' stats dict; its "reference_patches" (a multi-reference
    estimate's) select the per-patch triangulation. `cameras` as for the sweeps."""
    lib = library()
    pts = np.ascontiguousarray(points, dtype=abi.POINT_DTYPE)
    sites = np.ascontiguousarray(grid_sites, dtype=abi.SITE_DTYPE)
    inp = abi.Inputs([], cameras, np.zeros((0, 3)))
    st = abi.dense_stats_struct(dense_stats)
    patches = dense_stats.get("reference_patches") or []
    pa = abi.patch_structs(patches)
    out = C.POINTER(abi.MeshResult)()
    poll = _poll(control)
    rc = lib.vt3d_dense_mesh(pts.ctypes.data_as(C.POINTER(abi.Point3)), len(pts),
                             sites.ctypes.data_as(C.POINTER(abi.GridSite)), len(sites),
                             C.byref(st), pa, len(patches), inp.cameras, inp.n_cameras,
                             int(width), int(height), float(focal), poll, None, C.byref(out))
    _check(rc)
    try:
        return abi.unpack_mesh(out.contents)
    finally:
        lib.vt3d_mesh_result_free(out)


def mesh_warning(mesh_stats: dict) -> Optional[str]:
    """mesh_warning (reconstruction/warnings.rs:5-25) of a mesh analysis' stats dict."""
    st = abi.MeshStats()
    st.attempted = int(bool(mesh_stats["attempted"]))
    for k in ("accepted_triangles", "candidate_triangles", "candidate_cells",
              "rejected_discontinuities", "rejected_degenerate"):
        setattr(st, k, int(mesh_stats[k]))
    lib = library()
    n = lib.vt3d_mesh_warning(C.byref(st), None, 0)
    if n == 0:
        return None
    buf = C.create_string_buffer(n + 1)
    lib.vt3d_mesh_warning(C.byref(st), buf, n + 1)
    return buf.value.decode()


class Session:
    """Frames resident in HBM, solved many times (benchmarks, batching, serving)."""

    def __init__(self, device: int = 0):
        lib = library()
        self._h = C.c_void_p()
        _check(lib.vt3d_session_create(int(device), C.byref(self._h)))
        self._inputs = None

    def upload(self, frames, storage: str = "f32") -> None:
        lib = library()
        self._frames = abi.Inputs(_rgba(frames), [], np.zeros((0, 3)))
        _check(lib.vt3d_session_upload_frames(self._h, self._frames.frames, self._frames.n_frames,
                                              abi.STORAGE[storage]))

    def solve(self, cameras, sparse_points, focal: float, reference_camera_index: int = -1,
              fine_grid: bool = False, sweep: Optional[abi.SweepParams] = None,
              control: Optional[Callable[[], bool]] = None) -> Optional[DenseAnalysis]:
        lib = library()
        inp = abi.Inputs([], cameras, sparse_points)
        out = C.POINTER(abi.DenseResult)()
        poll = _poll(control)
        rc = lib.vt3d_session_solve(self._h, inp.cameras, inp.n_cameras, inp.sparse_ptr,
                                    inp.n_sparse, float(focal), int(reference_camera_index),
                                    int(fine_grid), C.byref(sweep) if sweep is not None else None,
                                    poll, None, C.byref(out))
        if rc == abi.VT3D_NO_PATCH:
            return None
        _check(rc)
        return _take(out)

    def timing(self) -> abi.Timing:
        t = abi.Timing()
        _check(library().vt3d_session_last_timing(self._h, C.byref(t)))
        return t

    def set_tuning(self, tile_width: int = 0, target_waves: int = 0,
                   reverse_bound: int = 0) -> None:
        """Launch shape of later solves and batches (0 keeps a default); reverse_bound -1
        turns off the ZNCC reverse sweep's partial-patch bound (same results, more work)."""
        t = abi.Tuning(int(tile_width), int(target_waves), int(reverse_bound))
        _check(library().vt3d_session_set_tuning(self._h, C.byref(t)))

    def batch(self, requests: Sequence[dict], use_graph: bool = False) -> "Batch":
        """Prepares many solves on the uploaded frames as one pipeline. Each request is a
        dict with cameras, sparse_points, focal and optionally frame_map (request frame ->
        session frame, default identity), reference_camera_index (-1: plain sweep),
        fine_grid and sweep."""
        return Batch(self, requests, use_graph)

    def close(self) -> None:
        if self._h:
            library().vt3d_session_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fill_request(r: abi.Request, rq: dict, n_frames: int):
    """Fills a vt3d_request from a request dict; returns the objects that must outlive it."""
    inp = abi.Inputs([], rq["cameras"], rq["sparse_points"])
    fmap = rq.get("frame_map")
    fmap = np.arange(n_frames, dtype=np.uint64) if fmap is None else \
        np.ascontiguousarray(np.asarray(fmap, dtype=np.uint64))
    sweep = rq.get("sweep")
    r.cameras = C.cast(inp.cameras, C.POINTER(abi.Camera))
    r.n_cameras = inp.n_cameras
    r.sparse_xyz = inp.sparse_ptr
    r.n_sparse = inp.n_sparse
    r.frame_map = fmap.ctypes.data_as(C.POINTER(C.c_uint64))
    r.n_frames = fmap.shape[0]
    r.focal = float(rq["focal"])
    r.reference_camera_index = int(rq.get("reference_camera_index", -1))
    r.fine_grid = int(bool(rq.get("fine_grid", False)))
    r.params = C.pointer(sweep) if sweep is not None else C.POINTER(abi.SweepParams)()
    return (inp, fmap, sweep)


class Batch:
    """Many solves on one session's frames, run as one kernel pipeline (optionally replayed
    from a captured HIP graph) on a stream of the batch's own. Build once, run many times;
    submit() and collect() split a run so that several batches can be in flight."""

    def __init__(self, session: Session, requests: Sequence[dict], use_graph: bool = False):
        lib = library()
        self._session = session  # the session must outlive the batch
        self._keep = []
        n = len(requests)
        self._reqs = (abi.Request * max(1, n))()
        n_frames = len(session._frames._rgba) if getattr(session, "_frames", None) else 0
        for i, rq in enumerate(requests):
            self._keep.append(_fill_request(self._reqs[i], rq, n_frames))
        self.n = n
        self._h = C.c_void_p()
        _check(lib.vt3d_batch_create(session._h, self._reqs, n, int(bool(use_graph)),
                                     C.byref(self._h)))

    def run(self) -> list:
        """One DenseAnalysis per request (None where the reference remap returns None)."""
        lib = library()
        outs = (C.POINTER(abi.DenseResult) * max(1, self.n))()
        _check(lib.vt3d_batch_run(self._h, outs))
        return [_take(outs[i]) if outs[i] else None for i in range(self.n)]

    def submit(self) -> None:
        """Enqueues the batch's launches on its stream and returns (see vt3d_batch_submit)."""
        _check(library().vt3d_batch_submit(self._h))

    def collect(self) -> list:
        """Waits for the submitted launch and returns what run() returns."""
        lib = library()
        outs = (C.POINTER(abi.DenseResult) * max(1, self.n))()
        _check(lib.vt3d_batch_collect(self._h, outs))
        return [_take(outs[i]) if outs[i] else None for i in range(self.n)]

    def timing(self) -> abi.Timing:
        t = abi.Timing()
        _check(library().vt3d_batch_last_timing(self._h, C.byref(t)))
        return t

    def close(self) -> None:
        if self._h:
            library().vt3d_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pack_frames(frame_lists: Sequence[Sequence[np.ndarray]]):
    """Deduplicates the frames of many solves (by array identity) into one list for a session
    upload; returns (frames, maps) with maps[i][k] the session index of solve i's frame k."""
    frames, index, maps = [], {}, []
    for fl in frame_lists:
        m = []
        for f in fl:
            key = id(f)
            if key not in index:
                index[key] = len(frames)
                frames.append(f)
            m.append(index[key])
        maps.append(np.asarray(m, dtype=np.uint64))
    return frames, maps


class Server:
    """Continuous-batching request loop on a session's resident frames (see
    vt3d_server_create). submit() queues a request dict (as for Session.batch) under a tag;
    poll() returns finished requests as dicts with tag, status, result (DenseAnalysis or None)
    and latency / queue / batch device times in ms."""

    def __init__(self, session: Session, max_batch: int = 0, max_inflight: int = 0,
                 hbm_budget_bytes: int = 0):
        lib = library()
        self._session = session  # the session must outlive the server
        cfg = abi.ServerConfig(C.sizeof(abi.ServerConfig), int(max_batch), int(max_inflight), 0,
                               int(hbm_budget_bytes))
        self._h = C.c_void_p()
        self._n_frames = len(session._frames._rgba) if getattr(session, "_frames", None) else 0
        _check(lib.vt3d_server_create(session._h, C.byref(cfg), C.byref(self._h)))

    def submit(self, request: dict, tag: int) -> None:
        r = abi.Request()
        keep = _fill_request(r, request, self._n_frames)  # noqa: F841 (copied by the server)
        _check(library().vt3d_server_submit(self._h, C.byref(r), int(tag)))

    def poll(self, timeout_ms: int = 0, capacity: int = 256) -> list:
        buf = (abi.Completion * capacity)()
        n = C.c_size_t()
        _check(library().vt3d_server_poll(self._h, buf, capacity, int(timeout_ms), C.byref(n)))
        out = []
        for i in range(n.value):
            c = buf[i]
            res = _take(c.result) if c.result else None
            out.append({"tag": int(c.tag), "status": int(c.status), "result": res,
                        "batch_requests": int(c.batch_requests), "queue_ms": c.queue_ms,
                        "latency_ms": c.latency_ms, "batch_device_ms": c.batch_device_ms,
                        "units": int(c.units)})
        return out

    def stats(self) -> dict:
        st = abi.ServerStats()
        _check(library().vt3d_server_get_stats(self._h, C.byref(st)))
        return {name: int(getattr(st, name)) for name, _ in abi.ServerStats._fields_}

    def close(self) -> None:
        if self._h:
            library().vt3d_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
