"""vt3d_dense -- MI355X dense plane-sweep depth pass (Python mirror of the C ABI).

The functions below mirror the reference's Rust interface for this path:

* `estimate_depth_points_with_control(frames, cameras, sparse_points, focal, fine_grid,
  control)` -- crates/video-to-3d-core/src/dense_legacy.rs:149-156
* `reconstruct_reference_patch(frames, cameras, sparse_points, focal, reference_camera_index,
  fine_grid, control)` -- crates/video-to-3d-core/src/dense.rs:674-747

Both run the HIP kernels of libvt3d_dense.so on the current GPU. There is no CPU path: if the
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
    lib.vt3d_dense_result_free.argtypes = [P(abi.DenseResult)]
    lib.vt3d_dense_result_free.restype = None
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


EXPORTED_SYMBOLS = (
    "vt3d_sweep_params_reference", "vt3d_sweep_params_zncc", "vt3d_skip_reason",
    "vt3d_grid_stride", "vt3d_abi_version", "vt3d_dense_plane_sweep",
    "vt3d_dense_reference_patch", "vt3d_dense_result_free", "vt3d_session_create",
    "vt3d_session_destroy", "vt3d_session_upload_frames", "vt3d_session_solve",
    "vt3d_session_last_timing", "vt3d_last_error", "vt3d_plane_homography",
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


def _take(ptr) -> DenseAnalysis:
    try:
        return abi.unpack_result(ptr.contents)
    finally:
        library().vt3d_dense_result_free(ptr)


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

    def close(self) -> None:
        if self._h:
            library().vt3d_session_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
