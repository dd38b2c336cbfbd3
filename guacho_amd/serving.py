"""Continuous-batching server for MHD solves over the batch ABI (gm_batch_*).

Requests arrive over time (Poisson, seeded).  A request is admitted as soon
as it has arrived and its box fits the device-memory budget, also while a
launch runs; admitted boxes stay resident on the GPU and advance together,
`steps_per_launch` steps per graph launch, each with its own CFL dt.
Completed boxes are downloaded and released while the next launch runs,
which frees memory for waiting requests.
Reported: time-to-solution per request (completion minus arrival),
p50/p99, and aggregate zone-updates/s.  The baseline serves the same
requests one at a time with the plain FP64 single-box path.

Per-box arithmetic does not depend on the other boxes, so each result is
bitwise the same alone and in the mix.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import _capi, problems
from .host import timestep_clamp


@dataclass
class Request:
    rid: int
    kind: str
    params: _capi.GmParams
    u0: np.ndarray
    t_end: float = math.inf
    max_steps: int = 0
    arrival: float = 0.0          # seconds after the stream starts
    cfl: float = 0.4
    # filled by the server
    admitted: float = -1.0
    finished: float = -1.0
    rejected: bool = False        # its box exceeds the whole memory budget
    steps: int = 0
    result: Optional[np.ndarray] = field(default=None, repr=False)

    @property
    def cells(self) -> int:
        return self.params.nx * self.params.ny * self.params.nz


def make_request(rid: int, kind: str, rng, precision: int) -> Request:
    """The serving mix: Brio-Wu 1024x4x4 to t = 0.1, Orszag-Tang 256^2 x 4 to
    t = 0.05, seeded blasts n^3 (n ~ U{64..192}) for 40 steps; all HLLD with
    minmod and flux-CD."""
    if kind == "briowu":
        u, dx, dy, dz = problems.brio_wu(1024, 4, 4)
        p = _capi.make_params(1024, 4, 4, dx, dy, dz, cv=1.0, bc=(1, 1, 3, 3, 3, 3), precision=precision)
        return Request(rid, kind, p, u, t_end=0.1)
    if kind == "ot":
        u, dx, dy, dz = problems.orszag_tang(256, 256, 4)
        p = _capi.make_params(256, 256, 4, dx, dy, dz, precision=precision)
        return Request(rid, kind, p, u, t_end=0.05)
    n = int(rng.integers(64, 193))
    u, dx, dy, dz = problems.blast(n, seed=int(rng.integers(1 << 30)))
    p = _capi.make_params(n, n, n, dx, dy, dz, bc=(1,) * 6, precision=precision)
    return Request(rid, kind, p, u, max_steps=40)


def make_stream(n: int, rate: float, seed: int, precision: int = _capi.GM_PRECISION_MIXED) -> List[Request]:
    """n requests, kinds drawn uniformly, Poisson arrivals at `rate` per second."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    kinds = ("briowu", "ot", "blast")
    for i in range(n):
        t += float(rng.exponential(1.0 / rate))
        r = make_request(i, kinds[int(rng.integers(0, 3))], rng, precision)
        r.arrival = t
        reqs.append(r)
    return reqs


def serve(requests: List[Request], steps_per_launch: int = 4, memory_budget: int = 0, keep_results: bool = False,
          max_boxes: int = 256, clock: Callable[[], float] = time.perf_counter,
          batch_factory: Optional[Callable[..., object]] = None, poll_s: float = 2e-4) -> Dict[str, float]:
    """Serve the requests with continuous batching (mixed precision).

    The host never waits for the GPU while there is work to queue: while a
    launch runs it admits arrived requests (their uploads run on the batch's
    transfer stream beside the launch) and polls the launch without
    blocking; once the launch has finished it queues the next one right
    away, then downloads and releases the boxes that completed (their
    buffers are out of the new launch, so the download overlaps it).

    Admission: every arrived request whose box fits the free memory and a
    free slot is admitted, in arrival order, and a request that does not fit
    does not hold back smaller ones behind it.  A request whose box exceeds
    the whole budget can never run: it is rejected on arrival (r.rejected).
    """
    pending = sorted(requests, key=lambda r: r.arrival)
    precision = pending[0].params.precision
    active: Dict[int, Request] = {}
    make = batch_factory or _capi.Batch
    with make(precision=precision, max_boxes=max_boxes, steps_per_launch=steps_per_launch,
              memory_budget=memory_budget) as B:
        budget = B.free_bytes()
        waiting: List[Request] = []
        in_flight = False
        t0 = clock()
        qi = 0

        def admit():
            nonlocal qi, waiting
            now = clock() - t0
            while qi < len(pending) and pending[qi].arrival <= now:
                r = pending[qi]
                qi += 1
                if B.box_bytes(r.params) > budget:
                    r.rejected = True
                    r.finished = now
                    continue
                waiting.append(r)
            still: List[Request] = []
            for r in waiting:
                if len(active) >= max_boxes or B.box_bytes(r.params) > B.free_bytes():
                    still.append(r)
                    continue
                try:
                    box = B.add(r.params, r.u0, cfl=r.cfl, t_end=r.t_end, max_steps=r.max_steps)
                except _capi.GmError as e:
                    if e.code != _capi.GM_ERR_BUSY:
                        raise
                    still.append(r)
                    continue
                r.admitted = clock() - t0
                active[box] = r
            waiting = still

        while qi < len(pending) or waiting or active:
            admit()
            if in_flight:
                if not B.query():
                    time.sleep(poll_s)
                    continue
                in_flight = False
            done = [box for box in active if B.clock(box)[2]]
            if len(done) < len(active):
                B.advance(1)              # the boxes still running, without the completed ones
                in_flight = True
            for box in done:
                r = active.pop(box)
                r.steps = B.clock(box)[1]
                u = B.download_u(box)
                if keep_results:
                    r.result = u
                B.remove(box)
                r.finished = clock() - t0
            if not active and not in_flight:
                if waiting and not done:
                    raise RuntimeError("waiting requests fit the budget but none can be admitted with no box active")
                if not waiting and qi < len(pending):
                    time.sleep(max(0.0, pending[qi].arrival - (clock() - t0)))
        makespan = clock() - t0
    return summarize(requests, makespan)


def serve_sequential_fp64(requests: List[Request], keep_results: bool = False,
                          clock: Callable[[], float] = time.perf_counter) -> Dict[str, float]:
    """The baseline: one request at a time, in arrival order, plain FP64
    single-box path with the reference's host loop."""
    t0 = clock()
    for r in sorted(requests, key=lambda q: q.arrival):
        now = clock() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        p = _capi.make_params(r.params.nx, r.params.ny, r.params.nz, r.params.dx, r.params.dy, r.params.dz,
                              cv=r.params.cv, gamma=r.params.gamma, bc=list(r.params.bc),
                              precision=_capi.GM_PRECISION_FP64)
        r.admitted = clock() - t0
        with _capi.Subdomain(p) as dev:
            dev.upload_u(r.u0)
            t, it = 0.0, 1
            while True:
                dt, dump = timestep_clamp(it, 10, r.cfl, dev.cfl_dtp(), t, r.t_end)
                dev.tstep(dt)
                t = t + dt
                if dump or (r.max_steps and it >= r.max_steps):
                    break
                it += 1
            u = dev.download_u()
        r.steps = it
        if keep_results:
            r.result = u
        r.finished = clock() - t0
    return summarize(requests, clock() - t0)


def summarize(requests: List[Request], makespan: float) -> Dict[str, float]:
    served = [r for r in requests if not r.rejected]
    tts = np.array([r.finished - r.arrival for r in served]) if served else np.zeros(1)
    work = float(sum(r.cells * r.steps for r in served))
    return {"requests": len(served), "rejected": len(requests) - len(served), "makespan_s": makespan,
            "zone_updates": work,
            "zone_updates_per_s": work / makespan if makespan > 0 else 0.0,
            "p50_s": float(np.percentile(tts, 50)), "p99_s": float(np.percentile(tts, 99)),
            "mean_s": float(tts.mean())}
