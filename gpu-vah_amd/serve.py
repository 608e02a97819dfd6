"""Serving form: the continuous-batching request loop, from Python.

The loop itself is C++ behind the C ABI (include/vah_serve.h,
gpu-vah_amd/host/serve.cpp): requests (events of different lattice sizes)
arrive over time; waiting requests are admitted first come, first served
whenever the batch has a free slot and room in its HBM budget; every
resident event advances together (one persistent launch per Heun stage,
graph-captured); each is retired when the engine stops it (freeze-out, step
limit or NaN), its latency counted from arrival to retirement.  Requests
that arrive while others are on the GPU join them at the next poll.  A
request either carries its initial energy density or a Monte-Carlo Glauber
geometry and seed, from which the engine generates it on the device at
admission.

``serve`` drives the batch engine through ``vah_serve``, or any object with
admit / step / poll / retire (and a clock) through ``vah_serve_run``, which
is how the scheduling is tested on the CPU against a stand-in.
"""
from __future__ import annotations

import ctypes
import dataclasses
import importlib
from typing import Callable, Dict, List, Optional

import numpy as np

EVENT_LIVE = 0
REJECTED = -2  # VAH_REQUEST_REJECTED: larger than the whole budget


def _vah():
    return importlib.import_module(__name__.rsplit(".", 1)[0])


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
    e: Optional[np.ndarray]   # initial energy density, interior cells (x fastest); None: from `glauber`
    max_steps: int = 100000
    glauber: Optional[object] = None   # a vah_glauber (gpu-vah_amd.Glauber), generated on the device
    pl_ratio: float = 1.0 / 3.0        # P_L = pl_ratio e
    # filled by the loop
    admitted: Optional[float] = None
    finished: Optional[float] = None
    nstep: int = 0
    tau: float = 0.0
    state: int = -1
    slot: int = -1
    fields: Optional[dict] = None

    @property
    def cells(self):
        return self.nx * self.ny * self.nz

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival

    @property
    def pl(self):
        """The initial P_L the engine builds from e (pl_ratio * e)."""
        return None if self.e is None else self.pl_ratio * self.e


def poisson_arrivals(n, rate, rng):
    """n arrival times of a Poisson stream of the given rate [1/s]."""
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


class _Errors:
    """Python exceptions raised inside callbacks, re-raised after the C call."""

    def __init__(self):
        self.exc = None

    def guard(self, fn, fail):
        def wrapped(*a):
            try:
                return fn(*a)
            except BaseException as exc:  # noqa: BLE001 -- carried across the C boundary
                if self.exc is None:
                    self.exc = exc
                return fail
        return wrapped


def serve(batch, requests: List[Request], steps_per_poll: int = 16, keep_fields: bool = False,
          clock: Optional[Callable[[], float]] = None, sleep: Optional[Callable[[float], None]] = None,
          on_poll: Optional[Callable[[float, int, int], None]] = None) -> Dict[str, float]:
    """Run the request loop until every request has been retired or rejected.

    Fills admitted / finished / nstep / tau / state / slot (and fields when
    keep_fields) of each request and returns summary figures: events/s over
    the run, p50 / p99 latency, and the peak number of resident events.
    With a gpu-vah_amd Batch and no clock, sleep or on_poll the loop runs
    entirely in C++ (vah_serve); otherwise its C++ scheduling drives `batch`
    through Python callbacks."""
    V = _vah()
    L = V.lib()
    n = len(requests)
    reqs = (V.ServeRequest * max(n, 1))()
    keep = []
    by_tag = {}
    for i, r in enumerate(requests):
        q = reqs[i]
        q.ev = V.BatchEvent(r.nx, r.ny, r.nz, r.dx, r.dx, r.dz, r.dt, r.t0, r.etabar, r.max_steps, r.rid)
        q.arrival_s = float(r.arrival)
        q.pl_ratio = r.pl_ratio
        if r.e is not None:
            e = np.ascontiguousarray(r.e, dtype=np.float64).ravel()
            keep.append(e)
            q.e = e.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        else:
            if r.glauber is None:
                raise ValueError(f"request {r.rid}: neither e nor glauber given")
            q.e = None
            q.glauber = r.glauber
        by_tag[r.rid] = r
    errs = _Errors()

    def retired(_user, res):
        if keep_fields:
            res = res[0]
            r = by_tag[res.tag]
            batch.dims[res.slot] = (r.nx, r.ny, r.nz)
            r.fields = batch.state(res.slot)

    retire_cb = V.RETIRE_FN(errs.guard(retired, None))
    results = (V.ServeResult * max(n, 1))()
    stats = V.ServeStats()
    native = isinstance(batch, V.Batch) and clock is None and sleep is None and on_poll is None
    if native:
        rc = L.vah_serve(batch.h, reqs, n, steps_per_poll, retire_cb, None, results, ctypes.byref(stats))
    else:
        def admit(_user, rq, slot):
            r = by_tag[rq[0].ev.tag]
            if isinstance(batch, V.Batch) and r.e is None:
                s = batch.admit_glauber(r.nx, r.ny, r.nz, r.dx, r.dx, r.dz, r.dt, r.t0, r.etabar, r.glauber,
                                        pl_ratio=r.pl_ratio, max_steps=r.max_steps, tag=r.rid, raise_on_full=False)
            else:
                s = batch.admit(r.nx, r.ny, r.nz, r.dx, r.dx, r.dz, r.dt, r.t0, r.etabar, r.e, pl=r.pl,
                                max_steps=r.max_steps, tag=r.rid, raise_on_full=False)
            if s is None:
                return V.VAH_ERR_BUDGET
            slot[0] = s
            return V.VAH_OK

        def step(_user, k):
            batch.step(k)
            return V.VAH_OK

        def poll(_user, out, cap, got):
            st = batch.poll()
            for i, d in enumerate(st[:cap]):
                for k, _ in V.BatchStatus._fields_:
                    setattr(out[i], k, d.get(k, 0))
            got[0] = min(len(st), cap)
            return V.VAH_OK

        def retire(_user, slot):
            batch.retire(slot)
            return V.VAH_OK

        be = V.ServeBackend(None, V.ADMIT_FN(errs.guard(admit, V.VAH_ERR_ARG)),
                            V.STEP_FN(errs.guard(step, V.VAH_ERR_ARG)), V.POLL_FN(errs.guard(poll, V.VAH_ERR_ARG)),
                            V.RETIRE_SLOT_FN(errs.guard(retire, V.VAH_ERR_ARG)),
                            V.NOW_FN(errs.guard(lambda _u: clock(), 0.0)) if clock else V.NOW_FN(),
                            V.SLEEP_FN(errs.guard(lambda _u, s: sleep(s), None)) if sleep else V.SLEEP_FN(),
                            V.ON_POLL_FN(errs.guard(lambda _u, t, a, b: on_poll(t, a, b), None)) if on_poll
                            else V.ON_POLL_FN())
        rc = L.vah_serve_run(ctypes.byref(be), reqs, n, steps_per_poll, retire_cb, None, results, ctypes.byref(stats))
    if errs.exc is not None:
        raise errs.exc
    if rc != V.VAH_OK:
        raise V.VahError(rc, L.vah_last_error().decode())
    for i, r in enumerate(requests):
        res = results[i]
        r.slot = res.slot
        r.state = res.state
        r.nstep = res.nstep
        r.tau = res.tau
        r.admitted = res.admitted_s if res.admitted_s >= 0 else None
        r.finished = res.finished_s if res.finished_s >= 0 else None
    return {
        "events": stats.served,
        "rejected": stats.rejected,
        "seconds": stats.seconds,
        "events_per_s": stats.events_per_s,
        "p50_latency_s": stats.p50_latency_s,
        "p99_latency_s": stats.p99_latency_s,
        "peak_resident": stats.peak_resident,
        "polls": stats.polls,
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
