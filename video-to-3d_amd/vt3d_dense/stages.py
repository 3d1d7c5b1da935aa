"""The dense stage's stored artifact (crates/video-to-3d-core/src/reconstruction_stages.rs).

* `DenseStageArtifact` (:150-160): `to_json` (:287-292, with the checks of
  to_json_with_cancel, :319-330), `content_hash` (:359-363) and `DenseStageArtifact.from_json`
  (:295-317), over libvt3d_dense.so's host code.
* `dense_stage_artifact(dense, request_hash, cameras_hash)`: the artifact step of
  reconstruct_dense_with_cancel (:570-589) over a dense estimate.
* `reconstruct_dense(...)`: reconstruct_dense_with_cancel (:549-589) from the camera stage's
  outputs (registered cameras, dense landmarks, focal, and the camera artifact's recorded
  request hash and content hash), with the dense estimate on the GPU. The camera stage itself
  is not part of this library.

Errors are `ReconstructionStageError` subclasses named and worded as the reference's enum
variants (:62-125).
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, List, Optional

import numpy as np

from . import abi


class ReconstructionStageError(Exception):
    """ReconstructionStageError (reconstruction_stages.rs:62-97)."""


class UnsupportedSchemaVersion(ReconstructionStageError):
    def __init__(self, artifact: str, found: int, supported: int):
        self.artifact, self.found, self.supported = artifact, found, supported
        super().__init__(f"{artifact} stage artifact has schema version {found}; this build "
                         f"reads version {supported}")


class InputMismatch(ReconstructionStageError):
    def __init__(self, stage: str, input: str, recorded: str, actual: str):
        self.stage, self.input, self.recorded, self.actual = stage, input, recorded, actual
        super().__init__(f"{stage} stage artifact was computed from {input} {recorded}, not "
                         f"{actual}")


class Malformed(ReconstructionStageError):
    def __init__(self, artifact: str, message: str):
        self.artifact, self.message = artifact, message
        super().__init__(f"malformed {artifact} stage artifact: {message}")


class NonFinite(ReconstructionStageError):
    def __init__(self, artifact: str, field: str):
        self.artifact, self.field = artifact, field
        super().__init__(f"{artifact} stage produced a non-finite value at `{field}`, which its "
                         "stored document cannot represent")


def _lib():
    from . import library
    return library()


class _ArtifactOwner:
    """Frees a C artifact once no view of its buffers is left."""

    __slots__ = ("ptr", "free")

    def __init__(self, ptr):
        self.ptr = ptr
        self.free = _lib().vt3d_dense_artifact_free

    def __del__(self):
        if self.ptr:
            self.free(self.ptr)
            self.ptr = None


class _DenseView:
    """The artifact's C multi result (`ptr`), keeping the artifact alive."""

    __slots__ = ("ptr", "owner")

    def __init__(self, ptr, owner):
        self.ptr, self.owner = ptr, owner


class DenseStageArtifact:
    """DenseStageArtifact (reconstruction_stages.rs:150-160): schema_version, request_hash,
    cameras_hash, and the DenseEvidence (:244-252) as `dense` (a MultiAnalysis whose stats hold
    the multi-reference DenseStats) and `warnings`."""

    def __init__(self, ptr):
        self._owner = _ArtifactOwner(ptr)
        a = ptr.contents
        self.schema_version = int(a.schema_version)
        self.request_hash = a.request_hash.decode()
        self.cameras_hash = a.cameras_hash.decode()
        self.warnings: List[str] = [a.warnings[i].decode() for i in range(int(a.n_warnings))]
        self.dense = abi.unpack_multi(a.dense.contents, _DenseView(a.dense, self._owner))

    @property
    def _ptr(self):
        return self._owner.ptr

    def to_json(self) -> str:
        """The stored document (to_json_with_cancel's checks: schema version, finite floats)."""
        lib = _lib()
        n = C.c_size_t(0)
        field = C.create_string_buffer(256)
        rc = lib.vt3d_dense_artifact_to_json(self._ptr, None, 0, C.byref(n), field, 256)
        if rc == abi.VT3D_E_SCHEMA:
            raise UnsupportedSchemaVersion("Dense", self._ptr.contents.schema_version,
                                           abi.DENSE_STAGE_ARTIFACT_SCHEMA_VERSION)
        if rc == abi.VT3D_E_NONFINITE:
            raise NonFinite("Dense", field.value.decode())
        if rc != 0:
            raise ReconstructionStageError(f"vt3d_dense_artifact_to_json failed ({rc})")
        buf = C.create_string_buffer(n.value + 1)
        rc = lib.vt3d_dense_artifact_to_json(self._ptr, buf, n.value + 1, C.byref(n), field,
                                             256)
        if rc != 0:
            raise ReconstructionStageError(f"vt3d_dense_artifact_to_json failed ({rc})")
        return buf.raw[:n.value].decode()

    def content_hash(self) -> str:
        """ContentHash::of_bytes(to_json) -- the identity a mesh artifact records."""
        out = C.create_string_buffer(abi.CONTENT_HASH_LEN)
        rc = _lib().vt3d_dense_artifact_content_hash(self._ptr, out)
        if rc != 0:
            raise ReconstructionStageError(f"vt3d_dense_artifact_content_hash failed ({rc})")
        return out.value.decode()

    @classmethod
    def from_json(cls, document: str) -> "DenseStageArtifact":
        data = document.encode() if isinstance(document, str) else bytes(document)
        out = C.POINTER(abi.DenseArtifact)()
        found = C.c_uint32(0)
        msg = C.create_string_buffer(512)
        rc = _lib().vt3d_dense_artifact_from_json(data, len(data), C.byref(out), C.byref(found),
                                                  msg, 512)
        if rc == abi.VT3D_E_SCHEMA:
            raise UnsupportedSchemaVersion("Dense", int(found.value),
                                           abi.DENSE_STAGE_ARTIFACT_SCHEMA_VERSION)
        if rc == abi.VT3D_E_MALFORMED:
            raise Malformed("Dense", msg.value.decode(errors="replace"))
        if rc == abi.VT3D_E_NONFINITE:
            raise NonFinite("Dense", msg.value.decode())
        if rc != 0:
            raise ReconstructionStageError(f"vt3d_dense_artifact_from_json failed ({rc})")
        return cls(out)

    def __eq__(self, other) -> bool:
        return isinstance(other, DenseStageArtifact) and self.to_json() == other.to_json()


def dense_stage_artifact(dense: abi.MultiAnalysis, request_hash: str,
                         cameras_hash: str) -> DenseStageArtifact:
    """The artifact of a dense estimate (vt3d_dense_estimate_with_budget's MultiAnalysis):
    its stats, points, sites and attributes, the dense stage's warnings and the identities."""
    if dense.handle is None or not getattr(dense.handle, "ptr", None):
        raise ValueError("the dense estimate holds no C result (a copy)")
    out = C.POINTER(abi.DenseArtifact)()
    field = C.create_string_buffer(256)
    rc = _lib().vt3d_dense_artifact_create(dense.handle.ptr, request_hash.encode(),
                                           cameras_hash.encode(), C.byref(out), field, 256)
    if rc == abi.VT3D_E_NONFINITE:
        raise NonFinite("Dense", field.value.decode())
    if rc != 0:
        raise ReconstructionStageError(
            f"vt3d_dense_artifact_create failed ({rc}): a content hash must be `sha256:` "
            "followed by 64 lowercase hex digits")
    return DenseStageArtifact(out)


def reconstruct_dense(frames, cameras, sparse_points, focal: float, cameras_hash: str,
                      camera_request_hash: Optional[str] = None,
                      options: Optional[abi.ReconstructionOptions] = None,
                      control: Optional[Callable[[], bool]] = None,
                      sweep: Optional[abi.SweepParams] = None) -> DenseStageArtifact:
    """reconstruct_dense_with_cancel (reconstruction_stages.rs:549-589) given the camera stage's
    outputs: the registered cameras (dense_cameras), the dense landmarks, the focal (stored as
    f32 by the camera stage, :186), the camera artifact's content hash and its recorded request
    hash. The request identity is checked first (InputMismatch, :564-569), then the budgeted
    dense estimate runs on the GPU under the options' working-set budget (:750-771)."""
    from . import estimate_depth_points_with_budget, reconstruction_options, request_content_hash
    opts = options if options is not None else reconstruction_options()
    request_hash = request_content_hash(frames, opts, control)
    if camera_request_hash is not None and camera_request_hash != request_hash:
        raise InputMismatch("Cameras", "Request", camera_request_hash, request_hash)
    dense = estimate_depth_points_with_budget(
        frames, cameras, sparse_points, float(np.float32(focal)),
        int(opts.max_dense_working_set_bytes), int(opts.max_features),
        int(opts.descriptor_radius), control=control, sweep=sweep)
    return dense_stage_artifact(dense, request_hash, cameras_hash)
