"""ctypes binding of the FP64 CPU oracle (libvt3d_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the CPU-baseline
leg of bench.py, as the reference the HIP path is tested against. The product never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys
from typing import Callable, Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_REPO = os.path.dirname(_HERE)
sys.path.insert(0, os.path.join(_REPO, "video-to-3d_amd"))

from vt3d_dense import abi  # noqa: E402  (shared data contract)

LIBRARY_PATH = os.path.join(_HERE, "libvt3d_oracle.so")
SLACK_DTYPE = np.dtype([("cost", "<f8"), ("px", "<f8"), ("rel", "<f8"), ("pos", "<f8")])

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def library() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY_PATH):
        build()
    lib = C.CDLL(LIBRARY_PATH)
    P = C.POINTER
    lib.vt3d_oracle_params_reference.argtypes = [P(abi.SweepParams)]
    lib.vt3d_oracle_params_zncc.argtypes = [P(abi.SweepParams), C.c_int32, C.c_int32, C.c_int32]
    lib.vt3d_oracle_skip_reason.restype = C.c_char_p
    lib.vt3d_oracle_skip_reason.argtypes = [C.c_int32]
    lib.vt3d_oracle_grid_stride.restype = C.c_uint32
    lib.vt3d_oracle_grid_stride.argtypes = [C.c_uint32, C.c_uint32, C.c_int32]
    common = [P(abi.Frame), C.c_size_t, P(abi.Camera), C.c_size_t, P(C.c_double), C.c_size_t,
              C.c_double]
    lib.vt3d_oracle_plane_sweep.argtypes = common + [C.c_int32, P(abi.SweepParams), abi.POLL_FN,
                                                     C.c_void_p, C.c_int32,
                                                     P(P(abi.DenseResult)), P(C.c_void_p)]
    lib.vt3d_oracle_plane_sweep.restype = C.c_int32
    lib.vt3d_oracle_reference_patch.argtypes = common + [C.c_size_t, C.c_int32,
                                                         P(abi.SweepParams), abi.POLL_FN,
                                                         C.c_void_p, C.c_int32,
                                                         P(P(abi.DenseResult)), P(C.c_void_p)]
    lib.vt3d_oracle_reference_patch.restype = C.c_int32
    lib.vt3d_oracle_estimate.argtypes = common + [C.c_int32, C.c_int32, P(abi.SweepParams),
                                                  abi.POLL_FN, C.c_void_p, C.c_int32,
                                                  P(P(abi.MultiResult))]
    lib.vt3d_oracle_estimate.restype = C.c_int32
    lib.vt3d_oracle_multi_result_free.argtypes = [P(abi.MultiResult)]
    lib.vt3d_oracle_multi_result_free.restype = None
    lib.vt3d_oracle_estimate_working_set.argtypes = [P(abi.Frame), C.c_size_t, C.c_uint64,
                                                     C.c_uint32, C.c_size_t, C.c_int32,
                                                     P(abi.WorkingSet)]
    lib.vt3d_oracle_estimate_working_set.restype = None
    lib.vt3d_oracle_estimate_with_budget.argtypes = common + [
        C.c_uint64, C.c_uint64, C.c_uint32, P(abi.SweepParams), abi.POLL_FN, C.c_void_p,
        C.c_int32, P(P(abi.MultiResult))]
    lib.vt3d_oracle_estimate_with_budget.restype = C.c_int32
    lib.vt3d_oracle_distributed_landmarks.argtypes = common + [abi.POLL_FN, C.c_void_p]
    lib.vt3d_oracle_distributed_landmarks.restype = C.c_int32
    lib.vt3d_oracle_reference_order.argtypes = common + [P(C.c_size_t)]
    lib.vt3d_oracle_reference_order.restype = C.c_size_t
    lib.vt3d_oracle_reference_attempt.argtypes = [C.c_uint64, P(abi.DenseResult), C.c_int32,
                                                   P(abi.ReferenceAttempt)]
    lib.vt3d_oracle_reference_attempt.restype = None
    lib.vt3d_oracle_merge_attempts.argtypes = [P(abi.ReferenceAttempt), C.c_size_t,
                                               P(abi.ReferenceAttempt), C.c_size_t,
                                               P(abi.ReferenceAttempt)]
    lib.vt3d_oracle_merge_attempts.restype = C.c_size_t
    lib.vt3d_oracle_dense_mesh.argtypes = [P(abi.Point3), C.c_size_t, P(abi.GridSite),
                                           C.c_size_t, P(abi.DenseStats),
                                           P(abi.ReferencePatchStats), C.c_size_t,
                                           P(abi.Camera), C.c_size_t, C.c_uint32, C.c_uint32,
                                           C.c_double, abi.POLL_FN, C.c_void_p,
                                           P(P(abi.MeshResult))]
    lib.vt3d_oracle_dense_mesh.restype = C.c_int32
    lib.vt3d_oracle_mesh_result_free.argtypes = [P(abi.MeshResult)]
    lib.vt3d_oracle_mesh_result_free.restype = None
    lib.vt3d_oracle_set_row_band.argtypes = [C.c_int32, C.c_int32]
    lib.vt3d_oracle_set_row_band.restype = None
    lib.vt3d_oracle_result_free.argtypes = [P(abi.DenseResult)]
    lib.vt3d_oracle_free.argtypes = [C.c_void_p]
    lib.vt3d_oracle_to_luma.argtypes = [P(C.c_uint8), C.c_size_t, P(C.c_uint8)]
    lib.vt3d_oracle_quantile.argtypes = [P(C.c_double), C.c_size_t, C.c_double]
    lib.vt3d_oracle_quantile.restype = C.c_double
    lib.vt3d_oracle_project.argtypes = [P(abi.Camera), P(C.c_double), C.c_uint32, C.c_uint32,
                                        C.c_double, P(C.c_double)]
    lib.vt3d_oracle_project.restype = C.c_int32
    lib.vt3d_oracle_unproject.argtypes = [P(abi.Camera), C.c_double, C.c_double, C.c_double,
                                          C.c_uint32, C.c_uint32, C.c_double, P(C.c_double)]
    lib.vt3d_oracle_sample_bilinear.argtypes = [P(C.c_uint8), C.c_uint32, C.c_uint32, C.c_double,
                                                C.c_double, P(C.c_double)]
    lib.vt3d_oracle_sample_bilinear.restype = C.c_int32
    lib.vt3d_oracle_patch_contrast.argtypes = [P(C.c_uint8), C.c_uint32, C.c_uint32, C.c_double,
                                               C.c_double, C.c_int32]
    lib.vt3d_oracle_patch_contrast.restype = C.c_double
    lib.vt3d_oracle_patch_error.argtypes = [P(abi.Camera), P(abi.Camera), P(C.c_uint8),
                                            P(C.c_uint8), C.c_uint32, C.c_uint32, C.c_uint32,
                                            C.c_uint32, C.c_double, C.c_double, C.c_double,
                                            C.c_double, C.c_int32, C.c_int32, P(C.c_double)]
    lib.vt3d_oracle_patch_error.restype = C.c_int32
    lib.vt3d_oracle_visible_depths.argtypes = [P(abi.Camera), P(C.c_double), C.c_size_t,
                                               C.c_uint32, C.c_uint32, C.c_double, P(C.c_double)]
    lib.vt3d_oracle_visible_depths.restype = C.c_size_t
    lib.vt3d_oracle_depth_search_bounds.argtypes = [P(abi.Camera), P(C.c_double), C.c_size_t,
                                                    C.c_uint32, C.c_uint32, C.c_double,
                                                    P(C.c_double), P(C.c_double)]
    lib.vt3d_oracle_depth_search_bounds.restype = C.c_int32
    lib.vt3d_oracle_single_view_depth.argtypes = [P(abi.Camera), P(abi.Camera), P(C.c_uint8),
                                                  P(C.c_uint8), C.c_uint32, C.c_uint32,
                                                  C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                                  C.c_double, C.c_double, C.c_double,
                                                  P(abi.SweepParams), P(C.c_double)]
    lib.vt3d_oracle_single_view_depth.restype = C.c_int32
    lib.vt3d_oracle_hypothesis_depth.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double]
    lib.vt3d_oracle_hypothesis_depth.restype = C.c_double
    lib.vt3d_oracle_reciprocal_agrees.argtypes = [C.c_double, C.c_double, C.c_double]
    lib.vt3d_oracle_reciprocal_agrees.restype = C.c_int32
    lib.vt3d_oracle_fuse.argtypes = [P(C.c_double), P(C.c_double), C.c_size_t, C.c_double,
                                     C.c_double, P(C.c_double), P(C.c_uint64), P(C.c_uint64)]
    lib.vt3d_oracle_surface_depth_prediction.argtypes = [P(C.c_double), C.c_size_t,
                                                         P(C.c_double)]
    lib.vt3d_oracle_surface_depth_prediction.restype = C.c_int32
    _lib = lib
    return lib


def params(settings: Optional[dict] = None, **overrides) -> abi.SweepParams:
    lib = library()
    return abi.make_params(lib.vt3d_oracle_params_reference, lib.vt3d_oracle_params_zncc,
                           settings, **overrides)


def _rgba(frames):
    from vt3d_dense.scenes import to_rgba
    return [f if f.ndim == 3 else to_rgba(f) for f in frames]


def _poll(control):
    if control is None:
        return abi.POLL_FN(0)
    return abi.POLL_FN(lambda _user: 1 if control() else 0)


def _take(out, slack) -> Tuple[abi.DenseAnalysis, Optional[np.ndarray]]:
    lib = library()
    res = abi.unpack_result(out.contents)
    sl = None
    if slack:
        n = int(out.contents.n_samples)
        sl = np.frombuffer(C.string_at(slack, n * SLACK_DTYPE.itemsize), SLACK_DTYPE).copy()
        lib.vt3d_oracle_free(slack)
    lib.vt3d_oracle_result_free(out)
    return res, sl


def plane_sweep(frames, cameras, sparse_points, focal: float, fine_grid: bool = False,
                control: Optional[Callable[[], bool]] = None,
                sweep: Optional[abi.SweepParams] = None, threads: int = 1):
    """Oracle of estimate_depth_points_with_control; returns (DenseAnalysis, slack)."""
    lib = library()
    inp = abi.Inputs(_rgba(frames), cameras, sparse_points)
    out = C.POINTER(abi.DenseResult)()
    slack = C.c_void_p()
    rc = lib.vt3d_oracle_plane_sweep(inp.frames, inp.n_frames, inp.cameras, inp.n_cameras,
                                     inp.sparse_ptr, inp.n_sparse, float(focal), int(fine_grid),
                                     C.byref(sweep) if sweep is not None else None,
                                     _poll(control), None, int(threads), C.byref(out),
                                     C.byref(slack))
    if rc < 0:
        raise ValueError(f"oracle rejected the input ({rc})")
    return _take(out, slack.value)


def reference_patch(frames, cameras, sparse_points, focal: float, reference_camera_index: int,
                    fine_grid: bool = False, control: Optional[Callable[[], bool]] = None,
                    sweep: Optional[abi.SweepParams] = None, threads: int = 1):
    """Oracle of reconstruct_reference_patch; (None, None) where the reference returns None."""
    lib = library()
    inp = abi.Inputs(_rgba(frames), cameras, sparse_points)
    out = C.POINTER(abi.DenseResult)()
    slack = C.c_void_p()
    rc = lib.vt3d_oracle_reference_patch(inp.frames, inp.n_frames, inp.cameras, inp.n_cameras,
                                         inp.sparse_ptr, inp.n_sparse, float(focal),
                                         int(reference_camera_index), int(fine_grid),
                                         C.byref(sweep) if sweep is not None else None,
                                         _poll(control), None, int(threads), C.byref(out),
                                         C.byref(slack))
    if rc == abi.VT3D_NO_PATCH:
        return None, None
    if rc < 0:
        raise ValueError(f"oracle rejected the input ({rc})")
    return _take(out, slack.value)


def estimate(frames, cameras, sparse_points, focal: float, reference_views: int = 3,
             fine_grid: Optional[bool] = None, control: Optional[Callable[[], bool]] = None,
             sweep: Optional[abi.SweepParams] = None, threads: int = 1) -> abi.MultiAnalysis:
    """Oracle of estimate_depth_points_limited (fine_grid None: decided as the budgeted entry
    decides it, without a working-set budget)."""
    lib = library()
    inp = abi.Inputs(_rgba(frames), cameras, sparse_points)
    out = C.POINTER(abi.MultiResult)()
    fine = -1 if fine_grid is None else int(bool(fine_grid))
    rc = lib.vt3d_oracle_estimate(inp.frames, inp.n_frames, inp.cameras, inp.n_cameras,
                                  inp.sparse_ptr, inp.n_sparse, float(focal),
                                  int(reference_views), fine,
                                  C.byref(sweep) if sweep is not None else None, _poll(control),
                                  None, int(threads), C.byref(out))
    if rc < 0:
        raise ValueError(f"oracle rejected the input ({rc})")
    m = abi.unpack_multi(out.contents)
    lib.vt3d_oracle_multi_result_free(out)
    return m


def estimate_working_set(frames, max_features: int = 320, descriptor_radius: int = 3,
                         reference_views: int = 3, fine_grid: bool = False) -> dict:
    """Oracle of estimate_working_set (dense.rs:423-605); frames as arrays or (w, h)."""
    dims = [f if isinstance(f, tuple) else (f.shape[1], f.shape[0]) for f in frames]
    fr = (abi.Frame * max(1, len(dims)))()
    for i, (w_, h_) in enumerate(dims):
        fr[i].width, fr[i].height = int(w_), int(h_)
    w = abi.WorkingSet()
    library().vt3d_oracle_estimate_working_set(fr, len(dims), int(max_features),
                                               int(descriptor_radius), int(reference_views),
                                               int(bool(fine_grid)), C.byref(w))
    return w.as_dict()


def estimate_with_budget(frames, cameras, sparse_points, focal: float,
                         working_set_budget_bytes: int, max_features: int = 320,
                         descriptor_radius: int = 3,
                         control: Optional[Callable[[], bool]] = None,
                         sweep: Optional[abi.SweepParams] = None,
                         threads: int = 1) -> abi.MultiAnalysis:
    """Oracle of estimate_depth_points_with_budget_and_control (dense.rs:354-420)."""
    lib = library()
    inp = abi.Inputs(_rgba(frames), cameras, sparse_points)
    out = C.POINTER(abi.MultiResult)()
    rc = lib.vt3d_oracle_estimate_with_budget(
        inp.frames, inp.n_frames, inp.cameras, inp.n_cameras, inp.sparse_ptr, inp.n_sparse,
        float(focal), int(working_set_budget_bytes), int(max_features), int(descriptor_radius),
        C.byref(sweep) if sweep is not None else None, _poll(control), None, int(threads),
        C.byref(out))
    if rc < 0:
        raise ValueError(f"oracle rejected the input ({rc})")
    m = abi.unpack_multi(out.contents)
    lib.vt3d_oracle_multi_result_free(out)
    return m


def distributed_landmarks(frames, cameras, sparse_points, focal: float,
                          control: Optional[Callable[[], bool]] = None) -> bool:
    """Oracle of has_distributed_landmarks (frames may be (H, W) arrays; only sizes count)."""
    inp = abi.Inputs(_rgba(frames), cameras, sparse_points)
    return bool(library().vt3d_oracle_distributed_landmarks(
        inp.frames, inp.n_frames, inp.cameras, inp.n_cameras, inp.sparse_ptr, inp.n_sparse,
        float(focal), _poll(control), None))


def reference_order(frames, cameras, sparse_points, focal: float):
    """Oracle of reference_candidate_order: camera indices."""
    inp = abi.Inputs(_rgba(frames), cameras, sparse_points)
    buf = (C.c_size_t * max(1, inp.n_cameras))()
    n = library().vt3d_oracle_reference_order(inp.frames, inp.n_frames, inp.cameras,
                                              inp.n_cameras, inp.sparse_ptr, inp.n_sparse,
                                              float(focal), buf)
    return [int(buf[i]) for i in range(n)]


def mesh(points, grid_sites, dense_stats, cameras, width: int, height: int, focal: float,
         patches=None, control: Optional[Callable[[], bool]] = None) -> abi.MeshAnalysis:
    """Oracle of reconstruct_dense_mesh_with_control (mesh.rs:77-215). `points`/`grid_sites`:
    POINT_DTYPE / SITE_DTYPE arrays; `dense_stats`: an analysis' stats dict; `patches`: its
    stats["reference_patches"] (None or [] for a single-reference analysis)."""
    lib = library()
    pts = np.ascontiguousarray(points, dtype=abi.POINT_DTYPE)
    sites = np.ascontiguousarray(grid_sites, dtype=abi.SITE_DTYPE)
    inp = abi.Inputs([], cameras, np.zeros((0, 3)))
    st = abi.dense_stats_struct(dense_stats)
    pa = abi.patch_structs(patches or [])
    out = C.POINTER(abi.MeshResult)()
    rc = lib.vt3d_oracle_dense_mesh(pts.ctypes.data_as(C.POINTER(abi.Point3)), len(pts),
                                    sites.ctypes.data_as(C.POINTER(abi.GridSite)), len(sites),
                                    C.byref(st), pa, len(patches or []), inp.cameras,
                                    inp.n_cameras, int(width), int(height), float(focal),
                                    _poll(control), None, C.byref(out))
    if rc < 0:
        raise ValueError(f"oracle rejected the input ({rc})")
    m = abi.unpack_mesh(out.contents)
    lib.vt3d_oracle_mesh_result_free(out)
    return m


def camera_struct(R, t, frame_index: int = 0) -> abi.Camera:
    c = abi.Camera()
    c.frame_index = frame_index
    Rc = np.asarray(R, dtype=np.float64).reshape(3, 3).T.reshape(-1)
    for k in range(9):
        c.rotation_colmajor[k] = float(Rc[k])
    for k in range(3):
        c.translation[k] = float(t[k])
    return c


def u8ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def f64ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def set_row_band(first: int = 0, count: int = 0) -> None:
    """Restricts the oracle sweep to grid rows [first, first + count); count 0 = all rows."""
    library().vt3d_oracle_set_row_band(int(first), int(count))
