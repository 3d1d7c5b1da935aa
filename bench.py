#!/usr/bin/env python3
"""Depth-hypothesis x sample throughput of the dense plane sweep on one MI355X.

Workload (default): the plain FP32 GPU kernel at native size -- 1 reference + 8 sources at
640x360 rendered from a seeded 6-plane textured box room, handheld random-walk poses,
128 inverse-depth hypotheses over 0.5-8 m, 7x7 ZNCC, stride 2, reciprocal check and fusion.
A step is one solve of that reference (frames resident in HBM): selection, table upload, the
kernel pipeline and the download of the results. Units = texture-gated samples x hypotheses.

Prints one JSON line. `roofline` prices the sweep kernel (k_sweep) against the FP32 vector
peak with the naive 24*S*P^2 FLOPs per unit; `cpu_baseline` times the FP64 oracle on a band
of grid rows of the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "video-to-3d_amd"))

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def _host_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def _traffic(workload: str):
    """HBM bytes per k_sweep launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(REPO, "profiles", f"k_sweep_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def cpu_baseline(problem, settings, rows: int, threads: int, gpu_samples=None):
    """FP64 oracle on grid rows [first, first + rows) of the workload; units per second."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np

    import vt3d_oracle as O

    sw = O.params(settings)
    ny = len(range(sw.patch_radius + 2, problem.frames[problem.reference].shape[0]
                   - (sw.patch_radius + 2), sw.grid_stride))
    first = max(0, ny // 2 - rows // 2)
    O.set_row_band(first, rows)
    try:
        t0 = time.perf_counter()
        res, _ = O.reference_patch(problem.frames, problem.cameras, problem.sparse, problem.focal,
                                   problem.reference, sweep=sw, threads=threads)
        dt = time.perf_counter() - t0
        # Parity of the GPU solve on a few rows of the same band (oracle with diagnostics).
        O.set_row_band(first, min(rows, 4))
        chk, _ = O.reference_patch(problem.frames, problem.cameras, problem.sparse,
                                   problem.focal, problem.reference,
                                   sweep=O.params(settings, emit_samples=1), threads=threads)
    finally:
        O.set_row_band(0, 0)
    units = res.stats["sampled_pixels"] * res.stats["depth_hypotheses"]
    out = {"value": units / dt, "unit": "depth-hyp*samples/s", "cores": threads,
           "kind": "port",
           "sample": f"FP64 oracle, grid rows {first}..{first + rows - 1} of {ny} "
                     f"({res.stats['sampled_pixels']} samples x {res.stats['depth_hypotheses']} "
                     f"hypotheses, reciprocal + fusion) in {dt:.2f} s"}
    agreement = None
    if gpu_samples is not None and chk.samples is not None:
        ys = set(int(y) for y in chk.samples["y"])
        g = gpu_samples[np.isin(gpu_samples["y"], list(ys))]
        o = chk.samples
        if len(g) == len(o):
            agreement = {
                "samples": int(len(o)),
                "status_agree": float(np.mean(g["status"] == o["status"])),
                "best_hypothesis_agree": float(np.mean(g["best_hypothesis"] == o["best_hypothesis"])),
            }
    return out, agreement


def copy_calibration(mib: int) -> int:
    """A plain device-to-device copy of known size: under rocprofv3 --pmc its FETCH_SIZE and
    WRITE_SIZE per dispatch give the counters' bytes-per-unit scale."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.arange(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--cpu-rows", type=int, default=0,
                    help="grid rows in the CPU-baseline sample (0: twice the host threads)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration for rocprof)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the sweep shards by request or by row band, see the design notes).",
              file=sys.stderr)
        return 2

    import numpy as np

    import vt3d_dense as V
    from vt3d_dense import scenes

    problem = scenes.config_c2()
    settings = dict(problem.settings, surface_completion_passes=0)
    sweep = V.params(settings, storage=args.storage)
    sess = V.Session(0)
    sess.upload(problem.frames, args.storage)

    def solve():
        return sess.solve(problem.cameras, problem.sparse, problem.focal, problem.reference,
                          sweep=sweep)

    for _ in range(args.warmup):
        solve()
    walls, units, sweep_ms, recip_ms, dev_ms = [], [], [], [], []
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        res = solve()
        walls.append(time.perf_counter() - t0)
        tm = sess.timing()
        units.append(tm.units)
        sweep_ms.append(tm.sweep_ms)
        recip_ms.append(tm.reciprocal_ms)
        dev_ms.append(tm.total_ms)
    elapsed = time.perf_counter() - t_all
    value = sum(units) / sum(walls)
    S = res.stats["source_views"]
    D = res.stats["depth_hypotheses"]
    P = 2 * sweep.patch_radius + 1
    flops_per_unit = 24 * S * P * P
    sweep_s = statistics.mean(sweep_ms) * 1e-3
    achieved = units[-1] * flops_per_unit / sweep_s / 1e12
    walls_ms = sorted(w * 1e3 for w in walls)

    def pct(q):
        return walls_ms[min(len(walls_ms) - 1, int(round(q * (len(walls_ms) - 1))))]

    line = {
        "metric": "depth-hyp×samples/s batched vs plain GPU + p50/p99 solve latency",
        "value": value,
        "unit": "depth-hyp*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sum(walls) / len(walls),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else f"{args.storage}-storage/f32",
        "data": "synthetic (seeded 6-plane box room, value-noise texture)",
        "config": {"workload": "plain GPU kernel, 640x360 ref + 8 sources, 128 hyps 0.5-8 m, "
                               "7x7 ZNCC, stride 2, reciprocal + fusion",
                   "width": 640, "height": 360, "sources": int(S), "hypotheses": int(D),
                   "patch": P, "stride": int(sweep.grid_stride), "storage": args.storage,
                   "sampled_pixels": int(res.stats["sampled_pixels"]),
                   "accepted_points": int(res.stats["accepted_points"])},
        "latency_ms": {"p50": pct(0.5), "p99": pct(0.99)},
        "device_ms": {"total": statistics.mean(dev_ms), "k_sweep": statistics.mean(sweep_ms),
                      "reciprocal": statistics.mean(recip_ms)},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": _traffic("c2"),
                     "kernel": "k_sweep", "flops_per_unit": flops_per_unit},
        "timed_wall_s": elapsed,
    }
    if not args.no_cpu_baseline:
        diag = sess.solve(problem.cameras, problem.sparse, problem.focal, problem.reference,
                          sweep=V.params(settings, storage=args.storage, emit_samples=1))
        threads = _host_threads()
        rows = args.cpu_rows if args.cpu_rows > 0 else 2 * threads
        cpu, agreement = cpu_baseline(problem, settings, rows, threads, diag.samples)
        line["cpu_baseline"] = cpu
        if agreement:
            line["parity_band"] = agreement
    else:
        line["cpu_baseline"] = None
    sess.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
