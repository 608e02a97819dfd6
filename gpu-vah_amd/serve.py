"""Serving form: a continuous-batching request loop over the batch engine.

Requests (events of different lattice sizes) arrive over time.  The loop
admits waiting requests first-come first-served whenever the batch has a free
slot and room in its HBM budget, advances every resident event together
(``Batch.step``: one persistent launch per Heun stage, graph-captured), and
retires each event at its freeze-out (the engine stops an event after the
first step whose hottest cell is below T_f), recording its latency from
arrival to retirement.  Requests that arrive while others are on the GPU
join them at the next poll; nothing waits for the batch to drain.

The batch object only needs admit / step / poll / state / retire, so the
scheduling logic is exercised without a GPU by a stand-in with the same
methods.
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import Callable, Dict, List, Optional

import numpy as np

EVENT_LIVE = 0
REJECTED = -2  # larger than the whole budget (or the batch refuses it when empty)


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float            # seconds after the start of the run
    nx: int
    ny: int
    nz: int
    dx: float
    dz: float
    dt: float
    t0: float
    etabar: float
    e: np.ndarray             # initial energy density, interior cells (x fastest)
    max_steps: int = 100000
    # filled by the loop
    admitted: Optional[float] = None
    finished: Optional[float] = None
    nstep: int = 0
    tau: float = 0.0
    state: int = -1
    fields: Optional[dict] = None

    @property
    def cells(self):
        return self.nx * self.ny * self.nz

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


def poisson_arrivals(n, rate, rng):
    """n arrival times of a Poisson stream of the given rate [1/s]."""
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def serve(batch, requests: List[Request], steps_per_poll: int = 16, keep_fields: bool = False,
          clock: Callable[[], float] = time.perf_counter, sleep: Callable[[float], None] = time.sleep,
          on_poll: Optional[Callable[[float, int, int], None]] = None) -> Dict[str, float]:
    """Run the request loop until every request has been retired.

    Fills admitted / finished / nstep / tau / state (and fields when
    keep_fields) of each request and returns summary figures: events/s over
    the run, p50 / p99 latency, and the peak number of resident events."""
    pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
    waiting: collections.deque = collections.deque()
    resident: Dict[int, Request] = {}
    peak_resident = 0
    polls = 0
    start = clock()
    while pending or waiting or resident:
        now = clock() - start
        while pending and pending[0].arrival <= now:
            waiting.append(pending.popleft())
        # first come, first served: the head of the queue waits for room
        while waiting:
            r = waiting[0]
            slot = batch.admit(r.nx, r.ny, r.nz, r.dx, r.dx, r.dz, r.dt, r.t0, r.etabar, r.e, pl=r.e / 3,
                               max_steps=r.max_steps, tag=r.rid, raise_on_full=False)
            if slot is None:
                break
            waiting.popleft()
            r.admitted = clock() - start
            resident[slot] = r
        peak_resident = max(peak_resident, len(resident))
        if not resident:
            if waiting:
                # nothing resident and the head still does not fit: it never will
                r = waiting.popleft()
                r.finished = clock() - start
                r.state = REJECTED
                continue
            if pending:
                sleep(max(0.0, pending[0].arrival - (clock() - start)))
            continue
        batch.step(steps_per_poll)
        statuses = batch.poll()
        polls += 1
        t = clock() - start
        for s in statuses:
            if s["state"] == EVENT_LIVE or s["slot"] not in resident:
                continue
            r = resident.pop(s["slot"])
            r.finished = t
            r.nstep = s["nstep"]
            r.tau = s["tau"]
            r.state = s["state"]
            if keep_fields:
                r.fields = batch.state(s["slot"])
            batch.retire(s["slot"])
        if on_poll is not None:
            on_poll(t, len(resident), len(waiting))
    total = clock() - start
    served = [r for r in requests if r.state != REJECTED]
    lat = np.array([r.latency for r in served]) if served else np.zeros(1)
    return {
        "events": len(served),
        "rejected": len(requests) - len(served),
        "seconds": total,
        "events_per_s": len(served) / total if total > 0 else float("inf"),
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "peak_resident": peak_resident,
        "polls": polls,
    }


def fifo_latencies(arrivals, service_times):
    """Latencies of a first-come first-served single server (one event at a
    time, the plain per-event path) for the given arrival and service times."""
    free = 0.0
    out = []
    for a, s in sorted(zip(arrivals, service_times)):
        begin = max(a, free)
        free = begin + s
        out.append(free - a)
    return np.array(out)
