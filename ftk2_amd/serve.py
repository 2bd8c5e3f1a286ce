"""Request loop of the serving mode: Poisson arrivals of streaming-tracking requests,
admission against the server's device-memory budget (first come, first served),
continuous batching (every server step advances each admitted stream by one slab,
all slabs of one predicate template in one batched pass), per-request latency from
arrival to result.

The loop talks to any object with the ftk2_amd.engine.Server methods (open,
append, close_stream, step, take), so its scheduling is testable without a GPU.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    arrival: float               # seconds after the loop starts
    spatial: tuple               # spatial extents (x fastest)
    kind: int
    m: int
    dtype: int
    timesteps: list              # per timestep: list of component arrays (device or host)
    thresholds: tuple = (0.0, 0.0)
    simplices: int = 0           # m-simplices over all slabs (the work unit)
    meta: dict = field(default_factory=dict)


@dataclass
class Outcome:
    rid: int
    arrival: float
    admitted: float
    done: float
    result: object = None

    @property
    def latency(self):
        return self.done - self.arrival


def poisson_arrivals(n: int, rate: float, rng: np.random.Generator) -> np.ndarray:
    """n arrival times of a Poisson process of the given rate (per second), from 0."""
    if n == 0:
        return np.zeros(0)
    gaps = rng.exponential(1.0 / rate, size=n)
    gaps[0] = 0.0
    return np.cumsum(gaps)


def run(server, requests, keep_results=False, clock=time.perf_counter, sleep=time.sleep, on_result=None):
    """Serve every request; returns (outcomes sorted by rid, wall seconds, rejected rids).

    A request is admitted when server.open() accepts its reservation; all its
    timesteps are then queued and the stream closed, and each server.step()
    extracts its next slab until it drains.  Requests that can never fit the
    budget are rejected (server.open raises)."""
    pending = sorted(requests, key=lambda r: r.arrival)
    waiting = deque()
    active = {}
    outcomes = {}
    rejected = []
    i = 0
    t0 = clock()
    while i < len(pending) or waiting or active:
        now = clock() - t0
        while i < len(pending) and pending[i].arrival <= now:
            waiting.append(pending[i])
            i += 1
        while waiting:
            rq = waiting[0]
            ncomp = len(rq.timesteps[0]) if rq.timesteps else 1
            try:
                sid = server.open(rq.spatial, rq.kind, rq.m, rq.dtype, ncomp, max(1, len(rq.timesteps)),
                                  rq.thresholds)
            except Exception:
                waiting.popleft()
                rejected.append(rq.rid)
                continue
            if sid is None:
                break  # budget in use: first come, first served
            waiting.popleft()
            for comps in rq.timesteps:
                server.append(sid, comps)
            server.close_stream(sid)
            active[sid] = (rq, clock() - t0)
        if not active:
            if i < len(pending):
                wait = pending[i].arrival - (clock() - t0)
                if wait > 0:
                    sleep(min(wait, 0.01))
            continue
        for sid in server.step():
            res = server.take(sid)
            rq, adm = active.pop(sid)
            oc = Outcome(rq.rid, rq.arrival, adm, clock() - t0, res if keep_results else None)
            if on_result is not None:
                on_result(rq, res)
            outcomes[rq.rid] = oc
    wall = clock() - t0
    return [outcomes[k] for k in sorted(outcomes)], wall, rejected


def summarize(outcomes, requests, wall):
    """p50/p99 latency (ms), throughput (simplices/s, requests/s)."""
    by = {r.rid: r for r in requests}
    lat = np.array([o.latency for o in outcomes]) * 1e3
    work = sum(by[o.rid].simplices for o in outcomes)
    return {
        "requests": len(outcomes),
        "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
        "mean_ms": float(lat.mean()) if len(lat) else None,
        "throughput_simplices_per_s": work / wall if wall > 0 else None,
        "requests_per_s": len(outcomes) / wall if wall > 0 else None,
        "wall_s": wall,
    }
