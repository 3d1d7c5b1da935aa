"""Seeded synthetic inputs for the dense plane sweep.

Scenes are textured planes ray-cast at integer pixel coordinates in the reference's
camera convention (pixel i <-> coordinate i, principal point (W/2, H/2), one focal length,
world->camera x_c = R x_w + t; crates/video-to-3d-core/src/dense_legacy.rs:996-1029).
Texture is value noise with a smoothstep lattice and a SplitMix64 lattice hash, the recipe
of crates/video-to-3d-core/examples/synthetic_room_fixture.rs:573-601, summed over three
octaves (cells c, c/2, c/4; weights 1/2, 1/4, 1/8) and mapped to round(255 v).

`plane_fixture` regenerates the reference's own analytic unit-test scene
(dense_legacy.rs:1189-1238) exactly.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _u64(v: int) -> np.uint64:
    return np.uint64(v & 0xFFFFFFFFFFFFFFFF)


def lattice_hash(a: np.ndarray, b: np.ndarray, salt: int, seed: int) -> np.ndarray:
    """SplitMix64 finaliser of a lattice cell to [0, 1) (synthetic_room_fixture.rs:579-588)."""
    with np.errstate(over="ignore"):
        z = (_u64(seed)
             ^ (a.astype(np.int64).view(np.uint64) * _u64(0x9E3779B97F4A7C15))
             ^ (b.astype(np.int64).view(np.uint64) * _u64(0xC2B2AE3D27D4EB4F))
             ^ (_u64(salt) * _u64(0x165667B19E3779F9)))
        z = (z ^ (z >> np.uint64(30))) * _u64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * _u64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def _smoothstep(t: np.ndarray) -> np.ndarray:
    t = np.clip(t, 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def value_noise(x: np.ndarray, y: np.ndarray, cell: float, salt: int, seed: int) -> np.ndarray:
    """Smooth value noise on a lattice of `cell` metres (synthetic_room_fixture.rs:590-601)."""
    gx = x / cell
    gy = y / cell
    ix = np.floor(gx)
    iy = np.floor(gy)
    fx = _smoothstep(gx - ix)
    fy = _smoothstep(gy - iy)
    ix = ix.astype(np.int64)
    iy = iy.astype(np.int64)
    top = lattice_hash(ix, iy, salt, seed) * (1.0 - fx) + lattice_hash(ix + 1, iy, salt, seed) * fx
    bottom = (lattice_hash(ix, iy + 1, salt, seed) * (1.0 - fx)
              + lattice_hash(ix + 1, iy + 1, salt, seed) * fx)
    return top * (1.0 - fy) + bottom * fy


def fractal_noise(x: np.ndarray, y: np.ndarray, cell: float, salt: int, seed: int,
                  octaves: int = 3) -> np.ndarray:
    """Octave sum normalised to [0, 1]: cells c, c/2, c/4 with weights 1/2, 1/4, 1/8."""
    acc = np.zeros_like(x)
    total = 0.0
    for k in range(octaves):
        w = 0.5 ** (k + 1)
        acc += w * value_noise(x, y, cell / (2 ** k), salt * 16 + k, seed)
        total += w
    return acc / total


@dataclasses.dataclass
class Plane:
    """A textured rectangle: origin, orthonormal in-plane axes u, v, extent in metres."""
    origin: np.ndarray
    u: np.ndarray
    v: np.ndarray
    extent: Tuple[float, float, float, float]  # (u_min, u_max, v_min, v_max)
    salt: int
    cell: float = 0.06
    low_texture_fraction: float = 0.0  # share of the plane replaced by flat +-2 patches
    contrast: float = 1.6              # texture contrast stretch about mid-grey

    @property
    def normal(self) -> np.ndarray:
        return np.cross(self.u, self.v)


@dataclasses.dataclass
class Camera:
    R: np.ndarray  # world->camera rotation, row-major 3x3
    t: np.ndarray  # world->camera translation

    @property
    def center(self) -> np.ndarray:
        return -self.R.T @ self.t


@dataclasses.dataclass
class Problem:
    """One dense solve: gray frames, cameras, landmarks, focal and sweep settings."""
    name: str
    frames: List[np.ndarray]           # uint8 (H, W) per frame
    cameras: List[Camera]              # cameras[i] images frames[i] (frame_index i)
    sparse: np.ndarray                 # (L, 3) float64 landmarks
    focal: float
    reference: int                     # camera index used as the reference patch
    settings: dict                     # sweep parameters for vt3d_sweep_params
    depths: Optional[List[np.ndarray]] = None  # true depth per frame (float64, inf = miss)


def look_at(center: Sequence[float], target: Sequence[float],
            down: Sequence[float] = (0.0, 1.0, 0.0)) -> Camera:
    c = np.asarray(center, dtype=np.float64)
    z = np.asarray(target, dtype=np.float64) - c
    z /= np.linalg.norm(z)
    x = np.cross(np.asarray(down, dtype=np.float64), z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    R = np.stack([x, y, z])
    return Camera(R, -R @ c)


def axis_angle(w: np.ndarray) -> np.ndarray:
    th = float(np.linalg.norm(w))
    if th < 1e-15:
        return np.eye(3)
    k = w / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * (K @ K)


def texture_value(plane: Plane, a: np.ndarray, b: np.ndarray, seed: int) -> np.ndarray:
    v = fractal_noise(a, b, plane.cell, plane.salt, seed)
    v = np.clip(0.5 + (v - 0.5) * plane.contrast, 0.0, 1.0)
    val = np.floor(255.0 * v + 0.5)
    if plane.low_texture_fraction > 0.0:
        # Flat patches: coarse noise above its (1 - fraction) level marks a flat region with a
        # per-plane grey and +-2 integer noise on a 1 cm lattice (multi-view consistent).
        mask_noise = value_noise(a, b, 0.35, plane.salt * 16 + 9, seed)
        level = 1.0 - plane.low_texture_fraction
        # value noise is concentrated near 0.5: map the fraction through its rough spread
        thr = 0.5 + (level - 0.5) * 0.55
        flat = mask_noise > thr
        grey = 60.0 + 130.0 * lattice_hash(np.zeros(1, np.int64), np.zeros(1, np.int64),
                                           plane.salt * 16 + 10, seed)[0]
        jitter = np.floor(lattice_hash(np.floor(a / 0.01).astype(np.int64),
                                       np.floor(b / 0.01).astype(np.int64),
                                       plane.salt * 16 + 11, seed) * 5.0) - 2.0
        val = np.where(flat, np.round(grey) + jitter, val)
    return np.clip(val, 0, 255)


def render(planes: Sequence[Plane], cam: Camera, width: int, height: int, focal: float,
           seed: int, background: int = 96) -> Tuple[np.ndarray, np.ndarray]:
    """Ray-cast one gray view; returns (uint8 image (H, W), depth (H, W), inf on a miss)."""
    xs = np.arange(width, dtype=np.float64)
    ys = np.arange(height, dtype=np.float64)
    X, Y = np.meshgrid(xs, ys)
    dc = np.stack([(X - width * 0.5) / focal, (Y - height * 0.5) / focal, np.ones_like(X)], -1)
    dw = dc @ cam.R  # R^T d for each pixel (row vectors)
    C = cam.center
    best = np.full((height, width), np.inf)
    img = np.full((height, width), float(background))
    for pl in planes:
        n = pl.normal
        denom = dw @ n
        with np.errstate(divide="ignore", invalid="ignore"):
            s = ((pl.origin - C) @ n) / denom
            P = C + s[..., None] * dw
            rel = P - pl.origin
            a = rel @ pl.u
            b = rel @ pl.v
        ok = (np.isfinite(s) & (s > 1e-6) & (a >= pl.extent[0]) & (a <= pl.extent[1])
              & (b >= pl.extent[2]) & (b <= pl.extent[3]) & (s < best))
        if not ok.any():
            continue
        best = np.where(ok, s, best)
        img[ok] = texture_value(pl, a[ok], b[ok], seed)
    return img.astype(np.uint8), best  # s along d with d_z = 1 is the camera depth


def cast_points(planes: Sequence[Plane], cam: Camera, px: np.ndarray, py: np.ndarray,
                width: int, height: int, focal: float) -> np.ndarray:
    """World points hit by the rays through pixels (px, py); rows of NaN on a miss."""
    dc = np.stack([(px - width * 0.5) / focal, (py - height * 0.5) / focal, np.ones_like(px)], -1)
    dw = dc @ cam.R
    C = cam.center
    best = np.full(px.shape, np.inf)
    for pl in planes:
        n = pl.normal
        with np.errstate(divide="ignore", invalid="ignore"):
            s = ((pl.origin - C) @ n) / (dw @ n)
            P = C + s[..., None] * dw
            rel = P - pl.origin
            a = rel @ pl.u
            b = rel @ pl.v
        ok = (np.isfinite(s) & (s > 1e-6) & (a >= pl.extent[0]) & (a <= pl.extent[1])
              & (b >= pl.extent[2]) & (b <= pl.extent[3]) & (s < best))
        best = np.where(ok, s, best)
    out = C + best[..., None] * dw
    out[~np.isfinite(best)] = np.nan
    return out


def landmarks(planes: Sequence[Plane], cams: Sequence[Camera], sizes: Sequence[Tuple[int, int]],
              focal: float, count: int, seed: int) -> np.ndarray:
    """`count` seeded surface points, an equal share ray-cast from random pixels of each view."""
    rng = np.random.default_rng(seed ^ 0x5A17)
    per = int(math.ceil(count / len(cams)))
    pts = []
    for cam, (w, h) in zip(cams, sizes):
        px = rng.uniform(4.0, w - 5.0, per)
        py = rng.uniform(4.0, h - 5.0, per)
        p = cast_points(planes, cam, px, py, w, h, focal)
        pts.append(p[np.isfinite(p).all(axis=1)])
    allp = np.concatenate(pts)[:count]
    return np.ascontiguousarray(allp, dtype=np.float64)


# --------------------------------------------------------------------------------------------
# The reference's analytic unit-test scene (dense_legacy.rs:1189-1238).
# --------------------------------------------------------------------------------------------

def _round_half_away(v: np.ndarray) -> np.ndarray:
    return np.where(v >= 0.0, np.floor(v + 0.5), np.ceil(v - 0.5))


def plane_texture(wx: np.ndarray, wy: np.ndarray) -> np.ndarray:
    """texture(), dense_legacy.rs:1197-1203."""
    v = (128.0 + 55.0 * np.sin(wx * 15.0 + wy * 2.7) + 40.0 * np.cos(wy * 17.0 - wx * 3.1)
         + 25.0 * np.sin((wx + wy) * 11.0))
    return np.clip(_round_half_away(v), 0.0, 255.0).astype(np.uint8)


def plane_frame(width: int, height: int, focal: float, center_x: float, depth: float) -> np.ndarray:
    """plane_frame(), dense_legacy.rs:1205-1224 (gray; the reference stores v in R, G and B)."""
    xs = np.arange(width, dtype=np.float64)
    ys = np.arange(height, dtype=np.float64)
    X, Y = np.meshgrid(xs, ys)
    wx = center_x + (X - width * 0.5) / focal * depth
    wy = (Y - height * 0.5) / focal * depth
    return plane_texture(wx, wy)


def plane_camera(center_x: float) -> Camera:
    """camera(), dense_legacy.rs:1189-1195: identity rotation at x = center_x."""
    return Camera(np.eye(3), np.array([-center_x, 0.0, 0.0]))


def plane_sparse_points(width: int, height: int, focal: float, depth: float) -> np.ndarray:
    """plane_sparse_points(), dense_legacy.rs:1226-1238."""
    pts = []
    for y in (10, 18, 26, 34):
        for x in (12, 24, 36, 48, 56):
            pts.append(((x - width * 0.5) / focal * depth, (y - height * 0.5) / focal * depth, depth))
    return np.array(pts, dtype=np.float64)


def plane_fixture(centers: Sequence[float] = (0.0, 0.18, -0.16), width: int = 68, height: int = 48,
                  focal: float = 60.0, depth: float = 4.0) -> Problem:
    """estimates_coarse_depth_on_textured_plane, dense_legacy.rs:1261-1275."""
    frames = [plane_frame(width, height, focal, c, depth) for c in centers]
    cams = [plane_camera(c) for c in centers]
    return Problem("plane-fixture", frames, cams, plane_sparse_points(width, height, focal, depth),
                   focal, 0, {"preset": "reference"})


# --------------------------------------------------------------------------------------------
# Benchmark and parity configurations.
# --------------------------------------------------------------------------------------------

SEED_C1 = 0xC0F10001
SEED_C2 = 0xC0F10002
SEED_C3 = 0xC0F10003
SEED_C4 = 0xC0F10004
SEED_C5 = 0xC0F10005


def arc_planes(seed: int) -> List[Plane]:
    """Three textured planes 1.5-4.5 m in front of the arc centre camera."""
    ux = np.array([1.0, 0.0, 0.0])
    uy = np.array([0.0, 1.0, 0.0])
    th = math.radians(25.0)
    tilted_u = np.array([math.cos(th), 0.0, -math.sin(th)])
    return [
        # near panel on the left, 1.5 m from the centre camera (which sits at z = -2)
        Plane(np.array([-0.55, 0.0, -0.5]), ux, uy, (-0.6, 0.45, -0.5, 0.45), salt=1, cell=0.05),
        # tilted panel in the middle distance
        Plane(np.array([0.45, -0.1, 0.8]), tilted_u, uy, (-0.6, 0.9, -0.8, 0.8), salt=2, cell=0.07),
        # back wall at 4.5 m
        Plane(np.array([0.0, 0.0, 2.5]), ux, uy, (-6.0, 6.0, -4.0, 4.0), salt=3, cell=0.1),
    ]


def config_c1(seed: int = SEED_C1) -> Problem:
    """1 reference + 4 sources at 360x202 on a 2 m arc at 5 degree steps, f = 300."""
    W, H, f = 360, 202, 300.0
    planes = arc_planes(seed)
    cams = []
    for k in range(5):
        th = math.radians(5.0 * (k - 2))
        cams.append(look_at((2.0 * math.sin(th), 0.0, -2.0 * math.cos(th)), (0.0, 0.0, 0.0)))
    frames, depths = zip(*[render(planes, c, W, H, f, seed) for c in cams])
    sparse = landmarks(planes, cams, [(W, H)] * 5, f, 1024, seed)
    return Problem("c1-arc-360x202", list(frames), cams, sparse, f, 2,
                   {"preset": "zncc", "patch_radius": 3, "depth_hypotheses": 64,
                    "max_source_views": 4, "grid_stride": 4, "depth_min": 1.0, "depth_max": 5.0},
                   list(depths))


def colour_view(planes: Sequence[Plane], cam: Camera, width: int, height: int, focal: float,
                seed: int) -> np.ndarray:
    """One RGBA view whose channels carry independent value-noise textures (seeds seed,
    seed + 1, seed + 2, each multi-view consistent) and whose alpha is per-pixel noise that
    the dense pass must ignore: a swapped channel or a wrong luma weight changes the luma."""
    out = np.empty((height, width, 4), np.uint8)
    for c in range(3):
        out[..., c] = render(planes, cam, width, height, focal, seed + c,
                             background=(40, 96, 170)[c])[0]
    rng = np.random.default_rng(seed ^ 0xA1FA)
    out[..., 3] = rng.integers(0, 256, (height, width), dtype=np.uint8)
    return out


def config_c1_colour(seed: int = SEED_C1) -> Problem:
    """The arc scene of `config_c1` rendered in colour (RGBA frames, see `colour_view`)."""
    c1 = config_c1(seed)
    W, H = c1.frames[0].shape[1], c1.frames[0].shape[0]
    planes = arc_planes(seed)
    frames = [colour_view(planes, cam, W, H, c1.focal, seed) for cam in c1.cameras]
    return Problem("c1-arc-360x202-rgba", frames, c1.cameras, c1.sparse, c1.focal,
                   c1.reference, dict(c1.settings), c1.depths)


def room_planes(seed: int, low_texture: float = 0.0) -> List[Plane]:
    """Inside of a 5 x 2.6 x 5 m box room (x right, y down, z forward; floor y = 0)."""
    ex, ey, ez = np.eye(3)
    hw, hh = 2.5, 2.6
    front, back = -1.0, 4.0
    mz = (front + back) * 0.5
    hd = (back - front) * 0.5
    lt = low_texture
    return [
        Plane(np.array([0.0, -hh * 0.5, back]), ex, ey, (-hw, hw, -hh * 0.5, hh * 0.5), 11, 0.08, lt),
        Plane(np.array([0.0, -hh * 0.5, front]), ex, ey, (-hw, hw, -hh * 0.5, hh * 0.5), 12, 0.08, lt),
        Plane(np.array([-hw, -hh * 0.5, mz]), ez, ey, (-hd, hd, -hh * 0.5, hh * 0.5), 13, 0.07, lt),
        Plane(np.array([hw, -hh * 0.5, mz]), ez, ey, (-hd, hd, -hh * 0.5, hh * 0.5), 14, 0.07, lt),
        Plane(np.array([0.0, 0.0, mz]), ex, ez, (-hw, hw, -hd, hd), 15, 0.06, lt),
        Plane(np.array([0.0, -hh, mz]), ex, ez, (-hw, hw, -hd, hd), 16, 0.09, lt),
    ]


def random_walk(n: int, seed: int, start=(0.0, -1.3, 0.2), sigma_t: float = 0.02,
                sigma_r_deg: float = 1.0, yaw0: float = 0.0) -> List[Camera]:
    """Handheld-style path: per frame, N(0, sigma_t) translation per axis and an axis-angle
    rotation step with N(0, sigma_r) per axis, starting from a camera looking down +z."""
    rng = np.random.default_rng(seed)
    base = look_at(start, (start[0] + math.sin(yaw0), start[1], start[2] + math.cos(yaw0)))
    Rcw = base.R.T.copy()  # camera->world
    c = np.array(start, dtype=np.float64)
    cams = []
    for _ in range(n):
        R = Rcw.T
        cams.append(Camera(R.copy(), -R @ c))
        c = c + rng.normal(0.0, sigma_t, 3)
        Rcw = Rcw @ axis_angle(np.radians(rng.normal(0.0, sigma_r_deg, 3)))
    return cams


def config_c2(seed: int = SEED_C2) -> Problem:
    """1 reference + 8 sources at 640x360 in the box room, 128 hypotheses 0.5-8 m, 7x7 ZNCC."""
    W, H = 640, 360
    f = 0.86 * W
    planes = room_planes(seed)
    cams = random_walk(17, seed)
    frames, depths = zip(*[render(planes, c, W, H, f, seed) for c in cams])
    sparse = landmarks(planes, cams, [(W, H)] * len(cams), f, 1024, seed)
    return Problem("c2-room-640x360", list(frames), cams, sparse, f, 8,
                   {"preset": "zncc", "patch_radius": 3, "depth_hypotheses": 128,
                    "max_source_views": 8, "grid_stride": 2, "depth_min": 0.5, "depth_max": 8.0},
                   list(depths))


C3_SIZES = ((640, 360), (360, 640), (480, 270))


def _render_pool(jobs, threads: int = 0):
    """render(*job) for each job on a thread pool; numpy releases the GIL in the large array
    operations of render(), so threads scale on the host cores."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    if threads <= 0:
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda j: render(*j), jobs))


def _render_many(planes, jobs, seed: int, threads: int):
    """Images of (camera, W, H, f) jobs over one scene."""
    return [img for img, _ in _render_pool([(planes, c, W, H, f, seed) for c, W, H, f in jobs],
                                           threads)]


def config_c3(count: int = 120, seed: int = SEED_C3, sources: int = 8,
              threads: int = 0) -> List[Problem]:
    """`count` keyframes of one room clip, each the reference with its `sources` nearest
    frames; each keyframe solve has a seeded size class (landscape, portrait, small). Solves
    of the same size class share frame arrays."""
    rng = np.random.default_rng(seed)
    n_frames = count + sources
    cams = random_walk(n_frames, seed, sigma_t=0.02)
    planes = room_planes(seed)
    size_of = [C3_SIZES[int(k)] for k in rng.integers(0, len(C3_SIZES), count)]
    half = sources // 2
    keys = []
    for k in range(count):
        W, H = size_of[k]
        for i in range(k, k + sources + 1):
            if (i, W, H) not in keys:
                keys.append((i, W, H))
    imgs = _render_many(planes, [(cams[i], W, H, 0.86 * max(W, H)) for i, W, H in keys], seed,
                        threads)
    cache = dict(zip(keys, imgs))
    problems = []
    for k in range(count):
        W, H = size_of[k]
        f = 0.86 * max(W, H)
        ref = k + half
        idx = list(range(ref - half, ref - half + sources + 1))
        frames = [cache[(i, W, H)] for i in idx]
        sub = [cams[i] for i in idx]
        sparse = landmarks(planes, sub, [(W, H)] * len(sub), f, 512, seed + k)
        problems.append(Problem(f"c3-key{k}-{W}x{H}", frames, sub, sparse, f, ref - idx[0],
                                {"preset": "zncc", "patch_radius": 3, "depth_hypotheses": 96,
                                 "max_source_views": sources, "grid_stride": 2,
                                 "depth_min": 0.5, "depth_max": 8.0, "storage": "f16"}))
    return problems


def config_c5(seed: int = SEED_C5, width: int = 1920, height: int = 1080,
              sources: int = 16) -> Problem:
    """1920x1080 reference + 16 sources, 12 planes with 10% flat +-2 patches, 11x11 ZNCC."""
    W, H = width, height
    f = 0.86 * max(W, H)
    planes = room_planes(seed, low_texture=0.10)
    rng = np.random.default_rng(seed)
    ex, ey, ez = np.eye(3)
    for k in range(6):  # six free-standing panels inside the room
        c = np.array([rng.uniform(-1.5, 1.5), rng.uniform(-2.0, -0.6), rng.uniform(1.2, 3.4)])
        th = rng.uniform(-0.6, 0.6)
        u = np.array([math.cos(th), 0.0, -math.sin(th)])
        planes.append(Plane(c, u, ey, (-0.35, 0.35, -0.3, 0.3), 21 + k, 0.05, 0.10))
    cams = random_walk(sources + 1, seed)
    frames, depths = zip(*_render_pool([(planes, c, W, H, f, seed) for c in cams]))
    sparse = landmarks(planes, cams, [(W, H)] * len(cams), f, 1024, seed)
    return Problem(f"c5-{W}x{H}", list(frames), cams, sparse, f, sources // 2,
                   {"preset": "zncc", "patch_radius": 5, "depth_hypotheses": 256,
                    "max_source_views": sources, "grid_stride": 1, "depth_min": 0.5,
                    "depth_max": 8.0}, list(depths))


def c4_stream(n_requests: int, seed: int = SEED_C4, rate: float = 200.0, clips: int = 16):
    """Poisson arrivals (mean `rate`/s) of solves from `clips` seeded clips: long edge
    U[360, 640] (16:9 or 9:16), 4-12 sources, 64-192 hypotheses, 5x5/7x7/9x9 patches."""
    rng = np.random.default_rng(seed)
    clip_spec = []
    for c in range(clips):
        long_edge = int(rng.integers(360, 641)) // 8 * 8
        short = int(round(long_edge * 9 / 16)) // 2 * 2
        portrait = bool(rng.integers(0, 2))
        W, H = (short, long_edge) if portrait else (long_edge, short)
        clip_spec.append({"clip": c, "size": (W, H), "seed": seed * 131 + c})
    t = 0.0
    reqs = []
    for i in range(n_requests):
        t += float(rng.exponential(1.0 / rate))
        c = int(rng.integers(0, clips))
        reqs.append({"arrival": t, "clip": c,
                     "sources": int(rng.integers(4, 13)),
                     "hypotheses": int(rng.integers(64, 193)),
                     "patch_radius": int(rng.choice([2, 3, 4])),
                     "key": int(rng.integers(0, 1 << 30))})
    return clip_spec, reqs


def c4_clips(clip_spec, n_frames: int = 16, threads: int = 0) -> List[dict]:
    """Frames, cameras and 1024 landmarks of each serving clip: a handheld random walk of
    `n_frames` views through a box room whose texture is seeded per clip."""
    clips = []
    jobs = []
    for spec in clip_spec:
        W, H = spec["size"]
        seed = spec["seed"]
        f = 0.86 * max(W, H)
        cams = random_walk(n_frames, seed, sigma_t=0.02)
        planes = room_planes(seed)
        clips.append({"clip": spec["clip"], "size": (W, H), "focal": f, "cameras": cams,
                      "sparse": landmarks(planes, cams, [(W, H)] * n_frames, f, 1024, seed)})
        jobs += [(planes, c, W, H, f, seed) for c in cams]
    imgs = [img for img, _ in _render_pool(jobs, threads)]
    for k, c in enumerate(clips):
        c["frames"] = imgs[k * n_frames:(k + 1) * n_frames]
    return clips


def c4_request(req: dict, clips: List[dict], frame_base: List[int]) -> dict:
    """Request dict (for Session.batch / Server.submit) of one serving arrival: the
    reference is the middle of a window of `sources + 1` consecutive clip frames chosen by
    the request key; frame_map points into a session holding every clip's frames, clip c
    starting at frame_base[c]."""
    c = clips[req["clip"]]
    S = req["sources"]
    n = len(c["cameras"])
    lo = req["key"] % (n - S)
    idx = list(range(lo, lo + S + 1))
    return {"cameras": [c["cameras"][i] for i in idx], "sparse_points": c["sparse"],
            "focal": c["focal"], "reference_camera_index": S // 2,
            "frame_map": np.array([frame_base[req["clip"]] + i for i in idx], np.uint64),
            "settings": {"preset": "zncc", "patch_radius": req["patch_radius"],
                         "depth_hypotheses": req["hypotheses"], "max_source_views": S,
                         "grid_stride": 2, "depth_min": 0.5, "depth_max": 8.0,
                         "storage": "f16", "surface_completion_passes": 0}}


def to_rgba(gray: np.ndarray) -> np.ndarray:
    """Gray v -> RGBA (v, v, v, 255); to_luma returns v exactly since 77 + 150 + 29 = 256."""
    h, w = gray.shape
    out = np.empty((h, w, 4), np.uint8)
    out[..., 0] = gray
    out[..., 1] = gray
    out[..., 2] = gray
    out[..., 3] = 255
    return out
