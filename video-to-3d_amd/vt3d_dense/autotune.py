"""Launch-shape autotuner for batches of dense plane-sweep solves.

The knobs it searches change how a batch maps onto the 256 CUs and their LDS, not what is
computed. Candidates agree within FP32 rounding: the tiled sweep warps each tile's region
about the region's centre, and the direct sweep about each patch's centre, so the kernel
and tile width move the last bits of a cost, while hypothesis blocking changes none:

* the primary-sweep kernel: the LDS-tiled sweep (neighbouring patches share warped source
  pixels) or the direct sweep (one patch per lane, gathers straight from the texture);
* tile width of the tiled sweep (16 x 16, 32 x 8 or 64 x 4 grid sites per workgroup);
* target_waves: hypothesis blocking -- each solve's hypotheses are split into blocks until
  the sweep launch holds at least this many waves.

Each candidate builds the batch, runs it once (graph capture, first-touch) and then times
`reps` runs with the batch's own HIP events; the fastest is applied to the session and
cached in a JSON file under a key made of the device name and a signature of the requests.
"""
from __future__ import annotations

import copy
import hashlib
import json
import os
from typing import Dict, List, Optional, Sequence

from . import abi

TILE_WIDTHS = (16, 32, 64)
TARGET_WAVES = (2048, 4096, 8192)


def candidates() -> List[Dict]:
    """Tiled sweep over tile widths x hypothesis blocking, plus the direct sweep at the
    default blocking."""
    out = [{"kernel": "tiled", "tile_width": tw, "target_waves": tg}
           for tw in TILE_WIDTHS for tg in TARGET_WAVES]
    out.append({"kernel": "direct", "tile_width": 32, "target_waves": 4096})
    return out


def default_cache_path() -> str:
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "vt3d_dense", "autotune.json")


def signature(requests: Sequence[dict]) -> str:
    """Workload signature: the per-request shape (sizes are part of the cameras' frames, so
    the count, sources, hypotheses, radius, stride and storage stand for it)."""
    h = hashlib.sha1()
    for r in requests:
        sw = r.get("sweep")
        fm = r.get("frame_map")
        fields = (len(r["cameras"]), 0 if fm is None else len(fm),
                  int(r.get("reference_camera_index", -1)))
        if sw is not None:
            fields += (sw.cost_kind, sw.patch_radius, sw.depth_hypotheses, sw.max_source_views,
                       sw.grid_stride, sw.storage)
        h.update(repr(fields).encode())
    return h.hexdigest()[:16]


def _device_name() -> str:
    try:
        import torch

        if torch.cuda.is_available():
            return torch.cuda.get_device_name(0)
    except Exception:
        pass
    return "hip-device"


def _with_kernel(requests: Sequence[dict], kernel: str) -> List[dict]:
    code = abi.KERNEL[kernel]
    out = []
    for r in requests:
        q = dict(r)
        sw = copy.copy(r["sweep"]) if r.get("sweep") is not None else None
        if sw is not None:
            sw.sweep_kernel = code
            q["sweep"] = sw
        out.append(q)
    return out


def apply(session, requests: Sequence[dict], choice: Dict) -> List[dict]:
    """Sets the session tuning of `choice`; returns the requests with its kernel selected."""
    session.set_tuning(tile_width=choice["tile_width"], target_waves=choice["target_waves"])
    return _with_kernel(requests, choice["kernel"])


def tune(session, requests: Sequence[dict], use_graph: bool = True, reps: int = 3,
         cands: Optional[List[Dict]] = None, cache_path: Optional[str] = None,
         refresh: bool = False, log=None) -> Dict:
    """Returns {"choice", "ms", "table", "cached", "requests"} and leaves the session tuned
    to the choice; "requests" are the caller's requests with the chosen kernel selected."""
    cache_path = cache_path or default_cache_path()
    key = f"{_device_name()}|{signature(requests)}|graph={int(bool(use_graph))}"
    cache = {}
    if os.path.exists(cache_path):
        try:
            with open(cache_path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if not refresh and key in cache:
        ent = cache[key]
        return dict(ent, cached=True, requests=apply(session, requests, ent["choice"]))
    table = []
    best = None
    for cand in cands or candidates():
        reqs = apply(session, requests, cand)
        b = session.batch(reqs, use_graph=use_graph)
        try:
            b.run()
            ms = []
            for _ in range(reps):
                b.run()
                ms.append(b.timing().total_ms)
        finally:
            b.close()
        m = sorted(ms)[len(ms) // 2]
        table.append(dict(cand, ms=m))
        if log:
            log(f"autotune {cand} -> {m:.3f} ms")
        if best is None or m < best[1]:
            best = (cand, m)
    chosen = apply(session, requests, best[0])
    ent = {"choice": best[0], "ms": best[1], "table": table}
    cache[key] = ent
    try:
        os.makedirs(os.path.dirname(cache_path), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    except OSError:
        pass
    return dict(ent, cached=False, requests=chosen)
