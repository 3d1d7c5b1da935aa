"""Load generation and latency accounting for the serving loop (vt3d_dense.Server).

`open_loop` replays a request stream at its arrival times (a submitting thread sleeps until
each arrival) while the caller's thread collects completions; `saturate` submits everything
at once. Both return per-request latency percentiles and throughput in depth-hypothesis x
sample units per second.
"""
from __future__ import annotations

import contextlib
import gc
import threading
import time
from typing import Dict, List, Sequence

import numpy as np


@contextlib.contextmanager
def frozen_heap():
    """Moves every object alive at the start out of the collector's reach while a timed run
    goes on: a full collection over a large heap (a caller's earlier results) holds the
    interpreter lock for tens of milliseconds, which would hold back a submitting thread and
    turn an arrival schedule into bursts, or stall the host side of a pipeline. Objects made
    during the run are still collected."""
    gc.collect()
    gc.freeze()
    try:
        yield
    finally:
        gc.unfreeze()


def _percentiles(v: Sequence[float]) -> Dict[str, float]:
    a = np.sort(np.asarray(v, dtype=np.float64))
    if a.size == 0:
        return {"p50": None, "p99": None, "max": None}
    pick = lambda q: float(a[min(a.size - 1, int(round(q * (a.size - 1))))])  # noqa: E731
    return {"p50": pick(0.5), "p99": pick(0.99), "max": float(a[-1])}


def _collect(server, n: int, done: Dict[int, dict], timeout_s: float):
    deadline = time.perf_counter() + timeout_s
    while len(done) < n:
        if time.perf_counter() > deadline:
            raise TimeoutError(f"{n - len(done)} requests unfinished after {timeout_s} s")
        for c in server.poll(timeout_ms=50):
            c["t_done"] = time.perf_counter()
            done[c["tag"]] = c


def open_loop(server, requests: Sequence[dict], arrivals: Sequence[float], warmup: int = 0,
              timeout_s: float = 300.0, keep_results: bool = False) -> dict:
    """Submits requests[i] at arrivals[i] seconds after the start; statistics cover the
    requests after the first `warmup`."""
    n = len(requests)
    done: Dict[int, dict] = {}
    t0 = time.perf_counter()
    submit_t = [0.0] * n
    errors: List[BaseException] = []

    def feed():
        try:
            for i, (rq, at) in enumerate(zip(requests, arrivals)):
                delay = t0 + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                submit_t[i] = time.perf_counter()
                server.submit(rq, i)
        except BaseException as e:  # surfaced after the join
            errors.append(e)

    with frozen_heap():
        th = threading.Thread(target=feed, daemon=True)
        th.start()
        try:
            _collect(server, n, done, timeout_s)
        finally:
            th.join()
    if errors:
        raise errors[0]
    return _summary(done, submit_t, warmup, keep_results)


def saturate(server, requests: Sequence[dict], timeout_s: float = 300.0,
             keep_results: bool = False) -> dict:
    """Submits every request at once and waits for all of them."""
    n = len(requests)
    done: Dict[int, dict] = {}
    submit_t = [0.0] * n
    with frozen_heap():
        for i, rq in enumerate(requests):
            submit_t[i] = time.perf_counter()
            server.submit(rq, i)
        _collect(server, n, done, timeout_s)
    return _summary(done, submit_t, 0, keep_results)


def _summary(done: Dict[int, dict], submit_t: List[float], warmup: int, keep: bool) -> dict:
    tags = sorted(done)
    meas = [t for t in tags if t >= warmup]
    lat = [done[t]["latency_ms"] for t in meas]
    queue = [done[t]["queue_ms"] for t in meas]
    units = sum(done[t]["units"] for t in meas)
    t_first = min(submit_t[t] for t in meas) if meas else 0.0
    t_last = max(done[t]["t_done"] for t in meas) if meas else 0.0
    wall = max(t_last - t_first, 1e-9)
    statuses: Dict[int, int] = {}
    for t in meas:
        statuses[done[t]["status"]] = statuses.get(done[t]["status"], 0) + 1
    out = {
        "requests": len(meas),
        "units": int(units),
        "wall_s": wall,
        "units_per_s": units / wall,
        "requests_per_s": len(meas) / wall,
        "latency_ms": _percentiles(lat),
        "queue_ms": _percentiles(queue),
        "mean_batch_requests": float(np.mean([done[t]["batch_requests"] for t in meas]))
        if meas else 0.0,
        "statuses": {str(k): v for k, v in statuses.items()},
    }
    if keep:
        out["completions"] = done
    return out
