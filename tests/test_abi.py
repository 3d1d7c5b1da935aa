"""The C ABI: every function declared in include/vt3d_dense.h is exported by the built
library, and the ctypes mirror has the C compiler's struct layout. No GPU needed."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from vt3d_dense import abi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "vt3d_dense.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(vt3d_\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if not n.endswith("_fn")))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("vt3d_dense_plane_sweep", "vt3d_dense_reference_patch", "vt3d_dense_result_free",
              "vt3d_session_solve"):
        assert n in names


def test_library_exports_every_declared_symbol(product):
    lib = product.library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_nm_lists_the_symbols(product):
    out = subprocess.run(["nm", "-D", "--defined-only", product.LIBRARY_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    assert set(declared_functions()) <= exported


STRUCTS = {
    "vt3d_frame": abi.Frame, "vt3d_camera": abi.Camera, "vt3d_sweep_params": abi.SweepParams,
    "vt3d_point3": abi.Point3, "vt3d_grid_site": abi.GridSite,
    "vt3d_point_attributes": abi.PointAttributes, "vt3d_dense_stats": abi.DenseStats,
    "vt3d_sample_diag": abi.SampleDiag, "vt3d_dense_result": abi.DenseResult,
    "vt3d_timing": abi.Timing, "vt3d_tuning": abi.Tuning, "vt3d_request": abi.Request,
    "vt3d_server_config": abi.ServerConfig, "vt3d_completion": abi.Completion,
    "vt3d_server_stats": abi.ServerStats, "vt3d_reference_attempt": abi.ReferenceAttempt,
    "vt3d_reference_patch_stats": abi.ReferencePatchStats, "vt3d_multi_result": abi.MultiResult,
    "vt3d_mesh_triangle": abi.MeshTriangle, "vt3d_mesh_patch_stats": abi.MeshPatchStats,
    "vt3d_mesh_stats": abi.MeshStats, "vt3d_mesh_result": abi.MeshResult,
    "vt3d_working_set": abi.WorkingSet,
    "vt3d_reconstruction_options": abi.ReconstructionOptions,
}


def test_ctypes_layout_matches_the_c_compiler():
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-std=c99", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.splitlines())
    for cname, cls in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, (cname, fname)


def test_presets_and_strings(product, oracle):
    lib, olib = product.library(), oracle.library()
    for fn, ofn in (("vt3d_sweep_params_reference", "vt3d_oracle_params_reference"),):
        a, b = abi.SweepParams(), abi.SweepParams()
        getattr(lib, fn)(C.byref(a))
        getattr(olib, ofn)(C.byref(b))
        assert bytes(a) == bytes(b)
    a, b = abi.SweepParams(), abi.SweepParams()
    lib.vt3d_sweep_params_zncc(C.byref(a), 3, 128, 8)
    olib.vt3d_oracle_params_zncc(C.byref(b), 3, 128, 8)
    assert bytes(a) == bytes(b)
    assert a.struct_size == C.sizeof(abi.SweepParams)
    for code in range(0, 10):
        assert lib.vt3d_skip_reason(code) == olib.vt3d_oracle_skip_reason(code)
    assert lib.vt3d_skip_reason(9) == b"dense reconstruction canceled"
    assert lib.vt3d_abi_version() == 1


def test_invalid_params_are_rejected_without_a_device(product):
    """Argument validation happens before any device call."""
    from vt3d_dense import scenes
    pf = scenes.plane_fixture()
    bad = product.params({"preset": "reference"}, patch_radius=9)
    with pytest.raises(product.Vt3dError):
        product.estimate_depth_points_with_control(pf.frames, pf.cameras, pf.sparse, pf.focal,
                                                   sweep=bad)
