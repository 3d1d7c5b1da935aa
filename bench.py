#!/usr/bin/env python3
"""Depth-hypothesis x sample throughput of the dense plane sweep on one MI355X: batched
FP16-storage solves against the plain FP32 kernel.

Four legs, all with frames resident in HBM; units = texture-gated samples x hypotheses.

* batched (the `value`): 120 keyframes of one seeded box-room clip, each the reference of a
  solve with its 8 nearest frames; frame sizes a seeded mix of 640x360, 360x640 and 480x270;
  96 inverse-depth hypotheses over 0.5-8 m, 7x7 ZNCC, stride 2, FP16 texel storage with FP32
  accumulation, reciprocal check and fusion. All 120 solves form one batch; the launch shape
  is autotuned. A step plans a new batch of the 120 (host selection, layout, FP64 tables,
  upload), runs it and downloads and unpacks all 120 results; four batches are in flight,
  each submitted as soon as it is planned. The same batch planned once and replayed from a
  captured HIP graph (one at a time, and four in flight), and planned afresh but run one at a
  time, are reported beside it.
* plain: 1 reference + 8 sources at 640x360 from a seeded 6-plane box room, 128 hypotheses,
  7x7 ZNCC, stride 2, FP32 storage; a step is one solve (selection, table upload, kernel
  pipeline, download).
* serving: a Poisson stream (200/s) of solves from 16 clips of uneven size, sources,
  hypotheses and patch size, admitted by the continuous-batching server under a 64 GB HBM
  budget; `latency_ms` is its per-request latency (submit -> result on the host), then the
  same requests submitted at once give the served throughput.
* large solve: 1920x1080 reference + 16 sources, 256 hypotheses, 11x11 ZNCC, stride 1, in
  FP32 and FP16 storage, with the FP16 result's error against the FP64 oracle on a band of
  grid rows (chosen hypothesis, photometric cost, fused points).

Prints one JSON line. `roofline` prices the batched leg's dominant kernel, the reverse sweep
(k_recip), at 24*P^2 FLOPs per (listed pair, hypothesis) patch cost over its launch time
(HIP events on the batch's stream), against the FP32 vector peak (157.3 TFLOP/s, packed) and
the non-packed issue limit this build runs at (78.6); the primary sweep's figure is labelled
effective (naive-count), and the whole step's is given too. `cpu_baseline` times the FP64
oracle on whole keyframes of the batched workload (about 10 s of work) on the process's
cores, and on one thread.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

# Each batch in flight and each serving slot runs on a stream of its own. The HIP runtime
# maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 unless set), and two
# in-flight batches whose streams share a queue run one after the other: a trace of the
# three-in-flight leg showed two of its streams on one queue and the device idle 20 ms every
# second step. Eight queues cover the session's own stream, four batches and the slots. The
# runtime reads the variable at its first call, which comes after this.
_QUEUES_NEEDED = 8
if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < _QUEUES_NEEDED:
    os.environ["GPU_MAX_HW_QUEUES"] = str(_QUEUES_NEEDED)

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "video-to-3d_amd"))

from vt3d_dense.serving import frozen_heap  # noqa: E402

PEAK_FP32_TFLOPS = 157.3          # MI355X vector FP32, packed (v_pk_fma_f32)
PEAK_FP32_NON_PACKED_TFLOPS = 78.6  # one wave64 FMA per SIMD per 4 cycles: this build's ceiling
PEAK_HBM_GBS = 8000.0


def _host_threads() -> int:
    """The process's cores: its CPU affinity, capped by OMP_NUM_THREADS where set (the GPU
    box gives each GPU's processes a share of 16)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def _profile(name: str):
    """A committed rocprofv3 --pmc summary (profiles/<name>.json), or None."""
    path = os.path.join(REPO, "profiles", f"{name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def _traffic(name: str):
    """HBM bytes per launch of a kernel from the committed rocprofv3 --pmc summary, if any."""
    prof = _profile(f"{name}_traffic")
    return prof.get("bytes_per_launch") if prof else None


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(problem, settings, rows: int, threads: int, gpu_samples=None):
    """FP64 oracle on grid rows [first, first + rows) of the workload; units per second."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np

    import vt3d_oracle as O

    sw = O.params(settings)
    ny = len(range(sw.patch_radius + 2, problem.frames[problem.reference].shape[0]
                   - (sw.patch_radius + 2), sw.grid_stride))
    rows = min(rows, ny) if rows > 0 else ny
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
           "kind": "port", "units": units, "seconds": dt,
           "sample": f"FP64 oracle, grid rows {first}..{first + rows - 1} of {ny}, "
                     f"{res.stats['sampled_pixels']} samples x {res.stats['depth_hypotheses']} "
                     f"hypotheses, reciprocal + fusion, {dt:.2f} s"}
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


def _pct(vals, q):
    v = sorted(vals)
    return v[min(len(v) - 1, int(round(q * (len(v) - 1))))]


def plain_leg(V, scenes, steps: int, warmup: int, storage: str):
    """One C2-size solve per step on a session (the plain FP32 kernel)."""
    problem = scenes.config_c2()
    settings = dict(problem.settings, surface_completion_passes=0)
    sweep = V.params(settings, storage=storage)
    sess = V.Session(0)
    sess.upload(problem.frames, storage)

    def solve():
        return sess.solve(problem.cameras, problem.sparse, problem.focal, problem.reference,
                          sweep=sweep)

    for _ in range(warmup):
        solve()
    walls, units, sweep_ms, recip_ms, rev_ms, dev_ms = [], [], [], [], [], []
    pairs = 0
    with frozen_heap():
        for _ in range(steps):
            t0 = time.perf_counter()
            res = solve()
            walls.append(time.perf_counter() - t0)
            tm = sess.timing()
            units.append(tm.units)
            sweep_ms.append(tm.sweep_ms)
            recip_ms.append(tm.reciprocal_ms)
            rev_ms.append(tm.reverse_sweep_ms)
            dev_ms.append(tm.total_ms)
            pairs = int(tm.reciprocal_pairs)
    sess.close()
    walls_ms = [w * 1e3 for w in walls]
    return {
        "value": sum(units) / sum(walls),
        "device_value": sum(units) / (1e-3 * sum(dev_ms)),
        "ms_per_step": statistics.mean(walls_ms),
        "latency_ms": {"p50": _pct(walls_ms, 0.5), "p99": _pct(walls_ms, 0.99)},
        "device_ms": {"total": statistics.mean(dev_ms), "sweep": statistics.mean(sweep_ms),
                      "reciprocal": statistics.mean(recip_ms),
                      "reverse_sweep": statistics.mean(rev_ms)},
        "reciprocal_pairs": pairs,
        "source_views": int(res.stats["source_views"]),
        "workload": "640x360 ref + 8 sources, 128 hyps 0.5-8 m, 7x7 ZNCC, stride 2, "
                    f"{storage.upper()} storage, reciprocal + fusion",
        "sampled_pixels": int(res.stats["sampled_pixels"]),
        "accepted_points": int(res.stats["accepted_points"]),
    }


# Batches in flight in the batched legs. With two, the pair shared the device, ended together
# and left it idle while the host built the older one's results; three or more batches that
# start together also end together, and the host then replaces them one by one while the
# device runs out of work: 109.8-110.2 ms per step with three, 106.8-108.7 with four, 106.2
# with five.
IN_FLIGHT = 4


def _pipelined(sess, reqs, steps: int, use_graph: bool, depth: int = IN_FLIGHT):
    """`steps` batches of `reqs`, each planned afresh (host selection, layout, FP64 tables,
    upload) and submitted as soon as it is planned, `depth` in flight: batch k + depth - 1 is
    planned and submitted before batch k is collected, so the host's planning and result
    building overlap device work (see IN_FLIGHT). Returns (seconds per step, batch latencies
    in ms from planning to collected results, the last batch's results)."""
    from collections import deque
    with frozen_heap():
        t0 = time.perf_counter()
        live = deque()  # (planning start, batch), oldest first, all submitted
        lat, res = [], None
        planned = 0
        for k in range(steps):
            while planned < steps and len(live) < depth:
                born = time.perf_counter()
                b = sess.batch(reqs, use_graph=use_graph)
                b.submit()
                live.append((born, b))
                planned += 1
            born, cur = live.popleft()
            res = cur.collect()
            lat.append(1e3 * (time.perf_counter() - born))
            cur.close()
        return (time.perf_counter() - t0) / steps, lat, res


def batched_leg(V, scenes, steps: int, warmup: int, count: int, tune: bool, log):
    """All `count` keyframe solves of the batched clip as one batch per step. The headline
    plans every step's batch afresh and keeps four in flight; the same batch replayed from a
    captured graph (planned once), serially and four in flight, and planned afresh but run
    one at a time (as the plain leg's solve is) are reported beside it."""
    from vt3d_dense import autotune

    t0 = time.perf_counter()
    problems = scenes.config_c3(count=count)
    gen_s = time.perf_counter() - t0
    frames, maps = V.pack_frames([p.frames for p in problems])
    sess = V.Session(0)
    sess.upload(frames, "f16")
    reqs = []
    for p, m in zip(problems, maps):
        st = dict(p.settings, surface_completion_passes=0)
        reqs.append({"cameras": p.cameras, "sparse_points": p.sparse, "focal": p.focal,
                     "reference_camera_index": p.reference, "frame_map": m,
                     "sweep": V.params(st)})
    tuned = None
    if tune:
        tuned = autotune.tune(sess, reqs, use_graph=True, log=log)
        reqs = tuned["requests"]
    # The headline: planned afresh every step, IN_FLIGHT batches in flight.
    _pipelined(sess, reqs, max(IN_FLIGHT, warmup + 1), False)
    step_s, lat, res = _pipelined(sess, reqs, steps, False)
    # One batch planned once and replayed from its graph: serially, then IN_FLIGHT in flight.
    batch = sess.batch(reqs, use_graph=True)
    for _ in range(max(1, warmup)):
        batch.run()
    walls, dev_ms = [], []
    units = 0
    for _ in range(steps):
        t1 = time.perf_counter()
        rep = batch.run()
        walls.append(time.perf_counter() - t1)
        tm = batch.timing()
        dev_ms.append(tm.total_ms)
        units += tm.units
    batch.close()
    per_step_units = units // max(1, steps)
    pipe = [sess.batch(reqs, use_graph=True) for _ in range(IN_FLIGHT)]
    for pb in pipe:
        pb.run()
    t1 = time.perf_counter()
    for k in range(min(IN_FLIGHT - 1, steps)):
        pipe[k].submit()
    for k in range(steps):
        if k + IN_FLIGHT - 1 < steps:
            pipe[(k + IN_FLIGHT - 1) % IN_FLIGHT].submit()
        pipe[k % IN_FLIGHT].collect()
    replay_pipelined_s = (time.perf_counter() - t1) / steps
    for pb in pipe:
        pb.close()
    # Stage times of the same batch launched eagerly (HIP events between the stages on the
    # batch's stream; the graph replays the same launches).
    eager = sess.batch(reqs, use_graph=False)
    eager.run()
    sweep_ms, recip_ms, rev_ms, eager_ms, pairs = [], [], [], [], 0
    for _ in range(max(3, steps // 2)):
        eager.run()
        tm = eager.timing()
        sweep_ms.append(tm.sweep_ms)
        recip_ms.append(tm.reciprocal_ms)
        rev_ms.append(tm.reverse_sweep_ms)
        eager_ms.append(tm.total_ms)
        pairs = int(tm.reciprocal_pairs)
    eager.close()
    # Planned afresh and run one at a time: build, run, collect, destroy every step, as the
    # plain leg's solve does for its single request (the session keeps destroyed batches'
    # buffers, so no step allocates device memory once warm).
    planned = []
    for _ in range(max(3, steps // 2)):
        t1 = time.perf_counter()
        b2 = sess.batch(reqs, use_graph=False)
        b2.run()
        b2.close()
        planned.append(time.perf_counter() - t1)
    sess.close()
    same_bits = all((a is None and b is None) or (a is not None and b is not None and
                                                  a.points.tobytes() == b.points.tobytes())
                    for a, b in zip(res, rep))
    walls_ms = [w * 1e3 for w in walls]
    ok = [r for r in res if r is not None]
    # (sample, hypothesis, source) triples of one step: each is one P x P patch cost of the
    # primary sweep, whatever the number of sources each keyframe kept.
    sample_hyp_sources = sum(int(r.stats["sampled_pixels"]) * int(r.stats["depth_hypotheses"])
                             * int(r.stats["source_views"]) for r in ok)
    return {
        "value": per_step_units / step_s,
        "device_value": units / (1e-3 * sum(dev_ms)),
        "ms_per_step": 1e3 * step_s,
        "what": "each step plans a new batch of the keyframes (host selection, layout, FP64 "
                f"tables, upload) and runs it eagerly; {IN_FLIGHT} batches in flight, each "
                "submitted as soon as it is planned",
        "latency_ms": {"p50": _pct(lat, 0.5), "p99": _pct(lat, 0.99),
                       "what": "one batch, planning to collected results"},
        "device_ms": {"graph_total": statistics.mean(dev_ms),
                      "eager_total": statistics.mean(eager_ms),
                      "sweep": statistics.mean(sweep_ms),
                      "reciprocal": statistics.mean(recip_ms),
                      "reverse_sweep": statistics.mean(rev_ms)},
        "points_equal_replay": same_bits,
        "replay": {"value": units / sum(walls),
                   "ms_per_step": statistics.mean(walls_ms),
                   "latency_ms": {"p50": _pct(walls_ms, 0.5), "p99": _pct(walls_ms, 0.99)},
                   "what": "one batch planned once, its graph replayed and collected every "
                           "step, one at a time"},
        "replay_pipelined": {"value": per_step_units / replay_pipelined_s,
                             "ms_per_step": 1e3 * replay_pipelined_s,
                             "what": f"{IN_FLIGHT} replayed batches in flight: steps k + 1 "
                                     f"to k + {IN_FLIGHT - 1} submitted before step k is "
                                     "collected"},
        "planning_included": {"value": per_step_units / statistics.mean(planned),
                              "ms_per_step": 1e3 * statistics.mean(planned),
                              "what": "planned afresh and run one at a time: build, run, "
                                      "collect, destroy every step, as the plain leg's "
                                      "solve is"},
        "reciprocal_pairs": pairs,
        "units_per_step": per_step_units,
        "solves": len(problems),
        "sampled_pixels": int(sum(r.stats["sampled_pixels"] for r in ok)),
        "accepted_points": int(sum(r.stats["accepted_points"] for r in ok)),
        "frames_resident": len(frames),
        "scene_gen_s": gen_s,
        "tuning": None if tuned is None else {"choice": tuned["choice"],
                                             "cached": tuned["cached"],
                                             "table": tuned.get("table")},
        "cpu_problems": problems[:4],
        "sample_hyp_sources": sample_hyp_sources, "P": 7,
    }


def serving_leg(V, scenes, n_requests: int, warmup: int, log, inflight: int = 2):
    """The serving stream: Poisson arrivals at 200/s from 16 clips (long edge 360-640,
    4-12 sources, 64-192 hypotheses, 5x5 / 7x7 / 9x9 ZNCC, stride 2, FP16 storage), admitted by
    continuous batching under a 64 GB HBM budget; then the same requests all at once."""
    from vt3d_dense import serving

    t0 = time.perf_counter()
    spec, arrivals = scenes.c4_stream(n_requests + warmup)
    clips = scenes.c4_clips(spec)
    gen_s = time.perf_counter() - t0
    frames, base = [], []
    for c in clips:
        base.append(len(frames))
        frames += c["frames"]
    sess = V.Session(0)
    sess.upload(frames, "f16")
    requests = []
    for a in arrivals:
        rq = scenes.c4_request(a, clips, base)
        rq["sweep"] = V.params(rq["settings"])
        requests.append(rq)
    server = V.Server(sess, max_batch=64, max_inflight=inflight, hbm_budget_bytes=64 << 30)
    try:
        opened = serving.open_loop(server, requests, [a["arrival"] for a in arrivals],
                                   warmup=warmup)
        log(f"serving open loop: p50 {opened['latency_ms']['p50']:.2f} ms "
            f"p99 {opened['latency_ms']['p99']:.2f} ms, {opened['units_per_s']:.4g} units/s")
        sat = serving.saturate(server, requests)
        log(f"serving saturated: {sat['units_per_s']:.4g} units/s")
        st = server.stats()
    finally:
        server.close()
        sess.close()
    return {"open_loop": opened, "saturated": sat, "server": st, "clips": len(clips),
            "frames_resident": len(frames), "scene_gen_s": gen_s,
            "workload": f"{n_requests} requests after {warmup} warm-up, Poisson 200/s from "
                        "16 clips (long edge 360-640, 4-12 sources, 64-192 hyps, 5x5/7x7/9x9 "
                        "ZNCC, stride 2, FP16), 64 GB HBM budget, up to 64 requests per "
                        f"launch, {inflight} launches in flight"}


def precision_leg(V, scenes, rows: int, threads: int, log):
    """Large single solve (1920x1080 reference + 16 sources, 12 planes with 10% flat +-2
    patches, 256 hypotheses, 11x11 ZNCC, stride 1, reciprocal + fusion) with FP32 and FP16
    texel storage, and the error of the FP16 solve against the FP64 oracle on a band of
    `rows` grid rows: chosen hypothesis (depth), photometric cost and fused points."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np

    import vt3d_oracle as O

    t0 = time.perf_counter()
    p = scenes.config_c5()
    gen_s = time.perf_counter() - t0
    st = dict(p.settings, surface_completion_passes=0, emit_samples=1)
    sess = V.Session(0)
    res, out = {}, {"workload": "1920x1080 ref + 16 sources, 12 planes (10% flat +-2 patches), "
                                "256 hyps 0.5-8 m, 11x11 ZNCC, stride 1, reciprocal + fusion",
                    "scene_gen_s": gen_s}
    try:
        for storage in ("f32", "f16"):
            sess.upload(p.frames, storage)
            sw = V.params(st, storage=storage)
            sess.solve(p.cameras, p.sparse, p.focal, p.reference, sweep=sw)
            t1 = time.perf_counter()
            r = sess.solve(p.cameras, p.sparse, p.focal, p.reference, sweep=sw)
            wall = time.perf_counter() - t1
            tm = sess.timing()
            res[storage] = r
            out[storage] = {"wall_ms": wall * 1e3, "device_ms": tm.total_ms,
                            "sweep_ms": tm.sweep_ms, "reciprocal_ms": tm.reciprocal_ms,
                            "reverse_sweep_ms": tm.reverse_sweep_ms,
                            "reciprocal_pairs": int(tm.reciprocal_pairs),
                            "units": int(tm.units), "units_per_s": tm.units / wall,
                            "sampled_pixels": int(r.stats["sampled_pixels"]),
                            "accepted_points": int(r.stats["accepted_points"])}
            log(f"C5 {storage}: {wall * 1e3:.1f} ms wall, {tm.units / wall:.4g} units/s")
    finally:
        sess.close()
    a, b = res["f32"], res["f16"]
    out["f16_equals_f32_bitwise"] = bool(a.points.tobytes() == b.points.tobytes()
                                         and a.samples.tobytes() == b.samples.tobytes())
    # FP64 oracle on a band of grid rows in the middle of the reference
    ny = len(range(st["patch_radius"] + 2, p.frames[p.reference].shape[0]
                   - (st["patch_radius"] + 2), st["grid_stride"]))
    first = max(0, ny // 2 - rows // 2)
    O.set_row_band(first, rows)
    try:
        t1 = time.perf_counter()
        o, slack = O.reference_patch(p.frames, p.cameras, p.sparse, p.focal, p.reference,
                                     sweep=O.params(st), threads=threads)
        oracle_s = time.perf_counter() - t1
    finally:
        O.set_row_band(0, 0)
    g = b.samples[np.isin(b.samples["y"], np.unique(o.samples["y"]))]
    os_ = o.samples
    n = len(os_)
    err = {"band_rows": [first, first + rows - 1], "samples": int(n),
           "oracle_s": oracle_s, "oracle_threads": threads}
    if len(g) == n and n > 0 and np.array_equal(g["x"], os_["x"]):
        agree = (g["status"] == os_["status"]) & (g["best_hypothesis"] == os_["best_hypothesis"])
        # the parity tests' criterion: a decision may differ only where the oracle's margin
        # to it is below the FP32 error (cost 1e-5 for ZNCC, 1e-3 px at a texture edge, 1e-5
        # relative in the 8% depth test)
        near = (slack["cost"] < 1e-5) | (slack["px"] < 1e-3) | (slack["rel"] < 1e-5)
        err["hypothesis_agree"] = float(agree.mean())
        err["disagree_explained_by_near_tie"] = int((~agree & near).sum())
        err["disagree_unexplained"] = int((~agree & ~near).sum())
        has_cost = agree & np.isfinite(os_["best_cost"]) & (os_["best_hypothesis"] >= 0)
        ce = np.abs(g["best_cost"][has_cost].astype(np.float64)
                    - os_["best_cost"][has_cost].astype(np.float64))
        err["cost_abs_err"] = {"max": float(ce.max()) if ce.size else 0.0,
                               "p99": float(np.percentile(ce, 99)) if ce.size else 0.0}
        D = st["depth_hypotheses"]
        inv = lambda h: (1 / st["depth_min"]) * (1 - h / (D - 1)) + (1 / st["depth_max"]) * (h / (D - 1))  # noqa: E731
        dis = ~agree & (g["best_hypothesis"] >= 0) & (os_["best_hypothesis"] >= 0)
        dr = np.abs(1 / inv(g["best_hypothesis"][dis]) - 1 / inv(os_["best_hypothesis"][dis])) \
            * inv(os_["best_hypothesis"][dis])
        err["depth_rel_err_where_hypothesis_differs"] = {
            "count": int(dis.sum()), "max": float(dr.max()) if dr.size else 0.0}
        gp = {(int(q["x"]), int(q["y"])): v for q, v in zip(b.grid_sites, b.points)}
        op = {(int(q["x"]), int(q["y"])): v for q, v in zip(o.grid_sites, o.points)}
        same = (agree & ~near & (g["status"] == 6) & (g["reciprocal_signature"]
                                                      == os_["reciprocal_signature"]))
        rel = []
        for i in np.nonzero(same)[0]:
            k = (int(g["x"][i]), int(g["y"][i]))
            if k in gp and k in op:
                va = np.array([gp[k]["x"], gp[k]["y"], gp[k]["z"]], np.float64)
                vb = np.array([op[k]["x"], op[k]["y"], op[k]["z"]], np.float64)
                rel.append(np.linalg.norm(va - vb) / max(np.linalg.norm(vb), 1e-12))
        err["points_compared"] = len(rel)
        err["point_rel_err"] = {"max": float(max(rel)) if rel else 0.0,
                                "p99": float(np.percentile(rel, 99)) if rel else 0.0}
    out["fp16_vs_fp64_oracle"] = err
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--keyframes", type=int, default=120,
                    help="solves in the batched leg (the batched workload is 120)")
    ap.add_argument("--plain-storage", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--no-tune", action="store_true", help="default launch shape, no autotuner")
    ap.add_argument("--serving-requests", type=int, default=2000,
                    help="measured requests of the serving stream (0: skip the serving leg)")
    ap.add_argument("--serving-warmup", type=int, default=200)
    ap.add_argument("--serving-inflight", type=int, default=3,
                    help="batches the server keeps in flight (1-8)")
    ap.add_argument("--precision-rows", type=int, default=2,
                    help="grid rows of the large solve checked against the FP64 oracle "
                         "(0: skip the large-solve leg)")
    ap.add_argument("--cpu-rows", type=int, default=0,
                    help="grid rows per keyframe in the CPU-baseline sample (0: all rows)")
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

    import vt3d_dense as V
    from vt3d_dense import scenes

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    t_all = time.perf_counter()
    plain = plain_leg(V, scenes, args.steps, args.warmup, args.plain_storage)
    log(f"plain leg: {plain['value']:.4g} units/s")
    bat = batched_leg(V, scenes, args.steps, args.warmup, args.keyframes, not args.no_tune, log)
    log(f"batched leg: {bat['value']:.4g} units/s")
    serv = None
    if args.serving_requests > 0:
        serv = serving_leg(V, scenes, args.serving_requests, args.serving_warmup, log,
                           args.serving_inflight)
    prec = None
    if args.precision_rows > 0:
        prec = precision_leg(V, scenes, args.precision_rows, _host_threads(), log)
    triples, P = bat.pop("sample_hyp_sources"), bat.pop("P")
    cpu_problems = bat.pop("cpu_problems")
    D = 96
    mean_sources = triples / max(1, bat["units_per_step"])
    kernel = ("k_sweep_tile<3, ZNCC, F16, 8>"
              if not bat["tuning"] or bat["tuning"]["choice"]["kernel"] == "tiled"
              else "k_sweep<3, ZNCC, F16, 8>")
    # 24 P^2 FLOPs per patch cost (9 warp, 10 bilinear, 5 cost per tap). The reverse sweep
    # evaluates one patch cost per (listed pair, hypothesis) and shares nothing; the primary
    # sweep's count is naive: its tiles warp each pixel once for all the patches covering it.
    fpc = 24 * P * P
    rev_s = bat["device_ms"]["reverse_sweep"] * 1e-3
    sweep_s = bat["device_ms"]["sweep"] * 1e-3
    step_s = bat["device_ms"]["graph_total"] * 1e-3
    rev_flops = fpc * D * bat["reciprocal_pairs"]
    sweep_flops = fpc * triples
    rev_tf = rev_flops / rev_s / 1e12
    sweep_tf = sweep_flops / sweep_s / 1e12
    step_tf = (rev_flops + sweep_flops) / step_s / 1e12
    rev_traffic = _traffic("k_recip_c3")
    sweep_traffic = _traffic("k_sweep_tile_c3")
    issue = _profile("k_recip_k_sweep_tile_c3_valu_issue")

    def frac(tf):
        return {"achieved": tf, "frac": tf / PEAK_FP32_TFLOPS,
                "frac_non_packed": tf / PEAK_FP32_NON_PACKED_TFLOPS}

    roofline = {
        "bound": "valu", "kernel": "k_recip<3, ZNCC, F16> (reverse sweep)",
        "achieved": rev_tf, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
        "frac": rev_tf / PEAK_FP32_TFLOPS, "traffic": rev_traffic,
        "peak_non_packed": PEAK_FP32_NON_PACKED_TFLOPS,
        "frac_non_packed": rev_tf / PEAK_FP32_NON_PACKED_TFLOPS,
        "share_of_step": rev_s / step_s,
        "flops_per_pair_hypothesis": fpc, "pairs_per_launch": bat["reciprocal_pairs"],
        "hypotheses": D, "launch_ms": rev_s * 1e3,
        "hbm_gbs": rev_traffic / rev_s / 1e9 if rev_traffic else None,
        "hbm_frac": rev_traffic / rev_s / 1e9 / PEAK_HBM_GBS if rev_traffic else None,
        "valu_issue": issue.get("k_recip") if issue else None,
        "primary_sweep": dict(frac(sweep_tf), kernel=kernel,
                              label="effective (naive-count): tiles share each warp",
                              patch_costs_per_launch=triples, launch_ms=sweep_s * 1e3,
                              traffic=sweep_traffic,
                              hbm_gbs=sweep_traffic / sweep_s / 1e9 if sweep_traffic else None,
                              valu_issue=issue.get("k_sweep_tile") if issue else None),
        "whole_step": dict(frac(step_tf), device_ms=step_s * 1e3,
                           label="both sweeps' naive FLOPs over the graph replay"),
    }
    line = {
        "metric": "depth-hyp×samples/s batched vs plain GPU + p50/p99 solve latency",
        "value": bat["value"],
        "unit": "depth-hyp*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": bat["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16-storage/f32",
        "data": "synthetic (seeded box-room clips, value-noise texture)",
        "config": {"workload": f"batched: {bat['solves']} keyframes (640x360 / 360x640 / "
                               "480x270 mix) each with its 8 nearest frames, 96 hyps 0.5-8 m, "
                               "7x7 ZNCC, stride 2, FP16 storage, reciprocal + fusion; each "
                               f"step plans a new batch and runs it, {IN_FLIGHT} batches in "
                               "flight",
                   "solves": bat["solves"], "sources": 8,
                   "mean_sources_kept": round(mean_sources, 3), "hypotheses": D, "patch": P,
                   "stride": 2, "storage": "f16",
                   "sampled_pixels": bat["sampled_pixels"],
                   "accepted_points": bat["accepted_points"]},
        "batched": bat,
        "plain": plain,
        "batched_vs_plain": bat["value"] / plain["value"],
        "batched_vs_plain_device": bat["device_value"] / plain["device_value"],
        "batched_serial_vs_plain": bat["planning_included"]["value"] / plain["value"],
        "latency_ms": (serv["open_loop"]["latency_ms"] if serv else plain["latency_ms"]),
        "latency_source": ("serving open loop, submit -> result on host" if serv
                           else "plain solve wall time"),
        "serving": serv,
        "large_solve": prec,
        "served_vs_plain": (serv["saturated"]["units_per_s"] / plain["value"] if serv else None),
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        threads = _host_threads()
        # All of this process's cores: whole keyframes of the batched workload, one after
        # another, until about 10 s of CPU work; the first also checks the GPU's decisions on
        # 4 of its rows.
        units, secs, names, agreement = 0.0, 0.0, [], None
        for k, prob in enumerate(cpu_problems):
            settings = dict(prob.settings, surface_completion_passes=0, storage="f32")
            diag = None
            if k == 0:
                sess = V.Session(0)
                sess.upload(prob.frames, "f16")
                diag = sess.solve(prob.cameras, prob.sparse, prob.focal, prob.reference,
                                  sweep=V.params(settings, storage="f16", emit_samples=1))
                sess.close()
            cpu, agr = cpu_baseline(prob, settings, args.cpu_rows, threads,
                                    diag.samples if diag is not None else None)
            agreement = agreement or agr
            units += cpu["units"]
            secs += cpu["seconds"]
            names.append(f"{prob.name} ({cpu['sample']})")
            if secs >= 10.0:
                break
        # One thread (the reference's own structure): bands of 2 grid rows of the first
        # keyframe from its middle outwards, until about 10 s.
        prob = cpu_problems[0]
        settings = dict(prob.settings, surface_completion_passes=0, storage="f32")
        u1, s1, bands = 0.0, 0.0, 0
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import vt3d_oracle as O

        sw = O.params(settings)
        ny = len(range(sw.patch_radius + 2, prob.frames[prob.reference].shape[0]
                       - (sw.patch_radius + 2), sw.grid_stride))
        first = ny // 2
        while s1 < 10.0 and first + 2 <= ny:
            O.set_row_band(first, 2)
            try:
                t0 = time.perf_counter()
                res, _ = O.reference_patch(prob.frames, prob.cameras, prob.sparse, prob.focal,
                                           prob.reference, sweep=sw, threads=1)
                s1 += time.perf_counter() - t0
            finally:
                O.set_row_band(0, 0)
            u1 += res.stats["sampled_pixels"] * res.stats["depth_hypotheses"]
            first += 2
            bands += 1
        try:
            affinity = len(os.sched_getaffinity(0))
        except AttributeError:
            affinity = os.cpu_count() or 1
        line["cpu_baseline"] = {
            "value": units / secs, "unit": "depth-hyp*samples/s", "cores": threads,
            "kind": "port",
            "sample": f"{len(names)} keyframes of the batched workload in {secs:.1f} s: "
                      + "; ".join(names),
            "single_thread": {"value": u1 / s1, "cores": 1,
                              "sample": f"{prob.name}: {bands} bands of 2 grid rows from row "
                                        f"{ny // 2} ({int(u1)} units, {s1:.1f} s)"},
            "nproc": os.cpu_count(), "affinity": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": _cpu_model(),
            "thread_scaling": (units / secs) / (u1 / s1) if s1 > 0 else None}
        if agreement:
            line["parity_band"] = agreement
    else:
        line["cpu_baseline"] = None
    line["timed_wall_s"] = time.perf_counter() - t_all
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
