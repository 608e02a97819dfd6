"""Scheduling logic of the serving loop on the CPU: the C++ loop
(vah_serve_run, include/vah_serve.h, called through gpu-vah_amd/serve.py)
drives a stand-in batch with the engine's admission rules -- a byte budget,
a slot limit, per-event step counts to freeze-out -- and a simulated clock,
both supplied as callbacks."""
import importlib

import numpy as np
import pytest

serve_mod = importlib.import_module("gpu-vah_amd.serve")


class FakeBatch:
    """admit / step / poll / state / retire with the engine's semantics;
    the clock advances by `cost` seconds per resident cell-step."""

    def __init__(self, budget, max_events, need, clock, cost=1e-9):
        self.budget, self.max_events, self.need = budget, max_events, need
        self.clock, self.cost = clock, cost
        self.ev = {}
        self.used = 0
        self.admit_order = []
        self.max_used = 0

    def admit(self, nx, ny, nz, dx, dy, dz, dt, t0, etabar, e, pl=None, max_steps=0, tag=0, raise_on_full=True):
        b = nx * ny * nz * 8 * 61
        if len(self.ev) >= self.max_events or self.used + b > self.budget:
            assert not raise_on_full
            return None
        slot = min(set(range(self.max_events)) - set(self.ev))
        self.ev[slot] = dict(tag=tag, n=0, live=True, bytes=b, cells=nx * ny * nz, t0=t0, dt=dt,
                             stop=min(self.need[tag], max_steps))
        self.used += b
        self.max_used = max(self.max_used, self.used)
        self.admit_order.append(tag)
        return slot

    def step(self, n):
        cells = 0
        for v in self.ev.values():
            if v["live"]:
                k = min(n, v["stop"] - v["n"])
                v["n"] += k
                cells += v["cells"] * k
                if v["n"] >= v["stop"]:
                    v["live"] = False
        self.clock.t += cells * self.cost + 1e-4

    def poll(self):
        return [dict(slot=s, state=0 if v["live"] else 1, tag=v["tag"], nstep=v["n"], tau=v["t0"] + v["n"] * v["dt"])
                for s, v in sorted(self.ev.items())]

    def state(self, slot):
        return {"nstep": self.ev[slot]["n"]}

    def retire(self, slot):
        self.used -= self.ev.pop(slot)["bytes"]


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


def make_requests(n, rate, seed):
    rng = np.random.default_rng(seed)
    arr = serve_mod.poisson_arrivals(n, rate, rng)
    reqs = []
    for i, a in enumerate(arr):
        L = int(rng.integers(20, 60))
        nz = 1 if rng.random() < 0.8 else int(rng.integers(5, 12))
        reqs.append(serve_mod.Request(i, float(a), L, L, nz, 0.1, 0.1, 0.02, 0.5, 0.16, np.ones(L * L * nz)))
    need = {r.rid: int(rng.integers(30, 400)) for r in reqs}
    return reqs, need


@pytest.mark.parametrize("rate", [5.0, 50.0, 5000.0])
def test_every_request_is_served_in_arrival_order_within_budget(rate):
    reqs, need = make_requests(60, rate, seed=int(rate))
    clk = Clock()
    budget = 3 * 59 * 59 * 11 * 8 * 61  # a few events at a time; the largest fits
    B = FakeBatch(budget, max_events=6, need=need, clock=clk)
    out = serve_mod.serve(B, reqs, steps_per_poll=16, clock=clk, sleep=clk.sleep)
    assert out["events"] == 60 and out["rejected"] == 0
    assert all(r.finished is not None and r.nstep == need[r.rid] for r in reqs)
    assert B.admit_order == sorted(B.admit_order, key=lambda t: reqs[t].arrival)  # FIFO
    assert B.max_used <= budget and not B.ev and B.used == 0
    assert all(r.admitted >= r.arrival - 1e-12 and r.finished > r.admitted for r in reqs)
    lat = np.array([r.latency for r in reqs])
    assert out["p50_latency_s"] == pytest.approx(np.percentile(lat, 50))
    assert out["p99_latency_s"] == pytest.approx(np.percentile(lat, 99))
    assert 1 <= out["peak_resident"] <= 6


def test_requests_arriving_during_a_run_join_it():
    """Continuous batching: a request that arrives while another is resident
    is admitted at the next poll, not after the first one finishes."""
    clk = Clock()
    reqs = [serve_mod.Request(0, 0.0, 30, 30, 1, 0.1, 0.1, 0.02, 0.5, 0.2, np.ones(900)),
            serve_mod.Request(1, 0.002, 30, 30, 1, 0.1, 0.1, 0.02, 0.5, 0.2, np.ones(900))]
    B = FakeBatch(1 << 30, 4, need={0: 1000, 1: 50}, clock=clk)
    seen = []
    serve_mod.serve(B, reqs, steps_per_poll=10, clock=clk, sleep=clk.sleep,
                    on_poll=lambda t, nres, nwait: seen.append(nres))
    assert max(seen) == 2
    assert reqs[1].finished < reqs[0].finished


def test_fifo_latencies_of_the_sequential_path():
    lat = serve_mod.fifo_latencies([0.0, 1.0, 1.5, 10.0], [2.0, 1.0, 3.0, 0.5])
    # finishes: 2, 3, 6, 10.5
    assert np.allclose(lat, [2.0, 2.0, 4.5, 0.5])


def test_a_request_larger_than_the_budget_is_rejected_not_waited_on():
    clk = Clock()
    reqs = [serve_mod.Request(0, 0.0, 100, 100, 1, 0.1, 0.1, 0.02, 0.5, 0.2, np.ones(10000)),
            serve_mod.Request(1, 0.001, 10, 10, 1, 0.1, 0.1, 0.02, 0.5, 0.2, np.ones(100))]
    B = FakeBatch(50 * 50 * 8 * 61, 4, need={0: 10, 1: 20}, clock=clk)
    out = serve_mod.serve(B, reqs, clock=clk, sleep=clk.sleep)
    assert reqs[0].state == serve_mod.REJECTED and reqs[1].nstep == 20
    assert out["events"] == 1 and out["rejected"] == 1
