"""Writes the committed fixtures of tests/golden/ (run once, in a checkout that has the
reference tree at /root/reference; nothing at test time reads the reference).

help_pan.npz -- the reference's rendered "slow-lateral-pan" case (--help/, truth.json
    schema video-to-3d/reference-fixture-truth/v1): frames 0-4 at 360x203 as the luma the
    dense path computes from them (to_luma, dense_legacy.rs:1147-1157; a gray frame
    (L, L, L) gives back L exactly), their true world-to-camera poses, landmarks unprojected
    from the true depth maps of frames 0, 2 and 4, and the true depth of every frame at
    every odd coordinate (which holds the stride-4 grid sites of 3x3 and 7x7 patches;
    grid_stride(360, 203) is 4 either way), bilinear in the depth map's half-integer pixel
    centres at the dense path's integer coordinates.
plane_frames.npz -- the analytic plane fixture's frames (dense_legacy.rs:1205-1224), as
    scenes.plane_frame renders them.
oracle_vectors.npz -- per-function input/output vectors of the FP64 oracle on seeded inputs
    (regression vectors: each function's parity pin is its known-answer test).
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "video-to-3d_amd"), os.path.join(REPO, "oracle")]

HELP = "/root/reference/--help"
FRAMES = [0, 1, 2, 3, 4]


def read_ppm(path):
    raw = open(path, "rb").read()
    magic, size, maxval, data = raw.split(b"\n", 3)
    assert magic == b"P6" and int(maxval) == 255
    w, h = map(int, size.split())
    return np.frombuffer(data[:w * h * 3], np.uint8).reshape(h, w, 3)


def read_pfm(path):
    raw = open(path, "rb").read()
    magic, size, scale, data = raw.split(b"\n", 3)
    assert magic == b"Pf"
    w, h = map(int, size.split())
    a = np.frombuffer(data[:w * h * 4], "<f4" if float(scale) < 0 else ">f4").reshape(h, w)
    return a[::-1].astype(np.float64)  # rows bottom to top


def to_luma(rgb):
    r, g, b = (rgb[..., k].astype(np.uint32) for k in range(3))
    return ((77 * r + 150 * g + 29 * b) >> 8).astype(np.uint8)


def bilinear(img, x, y):
    """img sampled at continuous pixel-centre coordinates (x, y) (centre of pixel i at i)."""
    x0 = np.clip(np.floor(x).astype(int), 0, img.shape[1] - 2)
    y0 = np.clip(np.floor(y).astype(int), 0, img.shape[0] - 2)
    fx, fy = x - x0, y - y0
    return ((img[y0, x0] * (1 - fx) + img[y0, x0 + 1] * fx) * (1 - fy)
            + (img[y0 + 1, x0] * (1 - fx) + img[y0 + 1, x0 + 1] * fx) * fy)


def help_pan():
    truth = json.load(open(os.path.join(HELP, "truth", "truth.json")))
    intr = truth["intrinsics"]
    W, H, f = intr["width"], intr["height"], intr["fx"]
    assert intr["cx"] == W * 0.5 and intr["cy"] == H * 0.5 and intr["fy"] == f
    luma, R, t, depth_grid = [], [], [], []
    # every odd coordinate: the sites of stride 4 with border 3 (3x3 patches) or 5 (7x7)
    gx = np.arange(1, W, 2)
    gy = np.arange(1, H, 2)
    GX, GY = np.meshgrid(gx.astype(float), gy.astype(float))
    rng = np.random.default_rng(20260)
    landmarks = []
    for i in FRAMES:
        fr = truth["frames"][i]
        assert fr["index"] == i
        luma.append(to_luma(read_ppm(os.path.join(HELP, fr["image"]))))
        Ri = np.array(fr["rotation_world_to_camera"], float).reshape(3, 3)
        ti = np.array(fr["translation"], float)
        R.append(Ri)
        t.append(ti)
        d = read_pfm(os.path.join(HELP, fr["depth"]))
        # the dense path's coordinate x is the ray through COLMAP coordinate x, which lies
        # between the pixel centres x - 0.5 and x + 0.5
        depth_grid.append(bilinear(d, GX - 0.5, GY - 0.5))
        if i in (0, 2, 4):
            # 150 landmarks on a jittered grid, unprojected with the true depth and pose
            u = rng.uniform(8, W - 8, 150)
            v = rng.uniform(8, H - 8, 150)
            z = bilinear(d, u - 0.5, v - 0.5)
            cam = np.stack([(u - W * 0.5) / f * z, (v - H * 0.5) / f * z, z], 1)
            landmarks.append((cam - ti) @ Ri)  # R^T (x_cam - t), row by row
    np.savez_compressed(os.path.join(HERE, "help_pan.npz"), luma=np.stack(luma),
                        rotation=np.stack(R), translation=np.stack(t),
                        landmarks=np.concatenate(landmarks), depth_grid=np.stack(depth_grid),
                        grid_x=gx.astype(np.uint32), grid_y=gy.astype(np.uint32),
                        focal=np.float64(f), frames=np.array(FRAMES))


def plane_frames():
    from vt3d_dense import scenes

    pf = scenes.plane_fixture()
    np.savez_compressed(os.path.join(HERE, "plane_frames.npz"), frames=np.stack(pf.frames))


def oracle_vectors():
    import ctypes as C

    import vt3d_oracle as O

    lib = O.library()
    rng = np.random.default_rng(7)
    out = {}
    # quantile (dense_legacy.rs:1168-1174)
    qs = []
    for n in (1, 2, 3, 7, 20):
        v = np.sort(rng.uniform(0.5, 8.0, n))
        for q in (0.1, 0.5, 0.9):
            qs.append(lib.vt3d_oracle_quantile(O.f64ptr(v), n, q))
    out["quantile"] = np.array(qs)
    # hypothesis schedule (dense_legacy.rs:285-293)
    out["hypothesis_depth"] = np.array([lib.vt3d_oracle_hypothesis_depth(h, D, 0.5, 8.0)
                                        for D in (1, 24, 96) for h in range(D)])
    # patch_error on the plane fixture frames (dense_legacy.rs:1031-1095)
    from vt3d_dense import scenes

    pf = scenes.plane_fixture()
    r = [O.camera_struct(c.R, c.t, i) for i, c in enumerate(pf.cameras)]
    f0, f1 = pf.frames[0], pf.frames[1]
    pts = rng.uniform([8, 8, 3.0], [60, 40, 5.0], (40, 3))
    errs = []
    for kind in (0, 1):
        for rad in (1, 2, 3):
            for x, y, d in pts:
                c = C.c_double()
                ok = lib.vt3d_oracle_patch_error(C.byref(r[0]), C.byref(r[1]), O.u8ptr(f0),
                                                 O.u8ptr(f1), 68, 48, 68, 48, float(x), float(y),
                                                 float(d), 60.0, kind, rad, C.byref(c))
                errs.append(c.value if ok else np.nan)
    out["patch_error_inputs"] = pts
    out["patch_error"] = np.array(errs)
    # reciprocal agreement (dense_legacy.rs:951-957)
    ex = rng.uniform(1, 8, 64)
    es = ex * rng.uniform(0.85, 1.15, 64)
    out["reciprocal_agrees"] = np.array([lib.vt3d_oracle_reciprocal_agrees(a, b, 0.08)
                                         for a, b in zip(ex, es)], np.int32)
    np.savez_compressed(os.path.join(HERE, "oracle_vectors.npz"), **out)


if __name__ == "__main__":
    help_pan()
    plane_frames()
    oracle_vectors()
    for n in ("help_pan.npz", "plane_frames.npz", "oracle_vectors.npz"):
        print(n, os.path.getsize(os.path.join(HERE, n)))
