#!/usr/bin/env python3
"""Throughput of the FP64 VAH Kurganov-Tadmor Heun-RK2 step on one MI355X,
and of the batched and serving forms.

Workload (default): one 3+1D event at real shape -- Monte-Carlo Glauber Pb+Pb
at b = 6 fm (Woods-Saxon R = 6.62 fm, a = 0.546 fm, sigma_NN = 70 mb, wounded
nucleons as Gaussians of width 0.5 fm, seed 42, peak 60 GeV/fm^3), eta plateau
half-width 1.5 with Gaussian tails 1.4, lattice 201 x 201 x 101,
dx = deta = 0.1, tau0 = 0.5 fm/c, dtau = 0.02 fm/c, lattice-QCD-like EOS
tabulated at 10^4 nodes, eta/s = 0.16.  A step is one Heun-RK2 step of the
whole lattice (two fused stage launches); the state is resident in HBM.  The
value is the median over --reps repetitions of --steps timed steps.  The
event's 0.5 fm hot spots break the scheme down at the dilute edge near
step 80 (tau ~ 2.1 fm/c) at this dtau (the CPU oracle does the same), so a
repetition restarts the event from its initial state every 80 steps, outside
the timed region; the cost of each 10-step block of one evolution from
tau = 0.5 to 2.1 fm/c is reported beside it.

The same line carries a sample of the batched form (boost-invariant
MC-Glauber events of uneven sizes in one batch, each run to freeze-out) and
of the serving form (a Poisson request stream with 3+1D requests, continuous
batching), both against the plain per-event FP64 path.

Prints one JSON line.  --gpus above 1 is refused: the benchmark runs on one GPU.
"""
import argparse
import importlib
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
BYTES_PER_CELL_UPDATE = 856  # FP64: 107 doubles per cell and step (stage 1: 46, stage 2: 61)
BYTES_PER_CELL_STAGE = (368, 488)  # algorithmic bytes per cell of the stage-1 / stage-2 launch
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "stage_kernel_hbm_traffic_201x201x101_fp64.md")

WORKLOADS = {
    "mc3d": dict(nx=201, ny=201, nz=101, dx=0.1, dz=0.1, dt=0.02, t0=0.5, etabar=0.16, eos="table", ic="mc", restart=80,
               name="3+1D MC-Glauber Pb+Pb b=6 fm, 201x201x101, dx=deta=0.1, tau0=0.5, dtau=0.02, "
                    "LQCD-like EOS table (1e4 nodes), eta/s=0.16"),
    "optical2d": dict(nx=81, ny=81, nz=1, dx=0.1, dz=0.1, dt=0.01, t0=0.5, etabar=0.2, eos="conformal", ic="optical",
                      restart=None,
               name="2+1D optical Glauber Au+Au b=7 fm, 81x81x1, dx=0.1, tau0=0.5, dtau=0.01, conformal EOS, "
                    "eta/s=0.2"),
}


def initial_state(V, w):
    if w["ic"] == "mc":
        g = V.glauber(A=208, R=6.62, a=0.546, sigma_nn_mb=70.0, b=6.0, seed=42, e_peak=60.0 / V.HBARC,
                      nucleon_width=0.5, eta_flat=1.5, eta_sigma=1.4, e_floor=1e-3)
        e, _ = V.ic_mc_glauber(w["nx"], w["ny"], w["nz"], w["dx"], w["dz"], g)
    else:
        g = V.glauber()  # Au+Au: R 6.38, a 0.535, 42 mb, b 7, alpha 0.5, peak 30 GeV/fm^3
        e = V.ic_optical_glauber(w["nx"], w["ny"], w["nz"], w["dx"], w["dz"], g)
    return e


def progress(msg):
    """One line on stderr per phase (the JSON line stays alone on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """CPU model and the number of cores this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    return {"model": model, "nproc": nproc}


def cpu_baseline(w, e, eos_table, seconds=15.0):
    """The CPU oracle (this project's C restatement of the reference CPU step,
    OpenMP over the host cores) on the same lattice for a few steps."""
    from oracle.oracle import Oracle
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    n = e.size
    o = Oracle(w["nx"], w["ny"], w["nz"], w["dx"], w["dx"], w["dz"], w["dt"], w["etabar"], w["t0"],
               threads=cores, eos_table=eos_table)
    o.set_primitives(e, np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n), pl=e / 3)
    steps, t_used = 0, 0.0
    while t_used < seconds and steps < 50:
        a = time.perf_counter()
        o.step(1)
        t_used += time.perf_counter() - a
        steps += 1
    # the reference's step ends with the checkValidity sweep
    # (FullyDiscreteKurganovTadmorScheme.cpp:791): one sweep timed alone
    a = time.perf_counter()
    o.validity()
    t_val = time.perf_counter() - a
    o.close()
    out = {"value": n * steps / t_used, "unit": "cell-updates/s", "cores": cores, "kind": "port",
           "sample": f"{steps} full-lattice steps of the same workload on the CPU oracle ({t_used:.1f} s), "
                     "without the validity sweep",
           "with_validity_sweep": n / (t_used / steps + t_val), "host": host_cpu()}
    # one core, as the reference's own single-threaded loops run: the middle
    # 12 eta planes of the same initial state (same per-cell work), 2 steps
    if w["nz"] > 12:
        k0 = (w["nz"] - 12) // 2
        sub = e.reshape(w["nz"], -1)[k0:k0 + 12].ravel()
        m = sub.size
        o1 = Oracle(w["nx"], w["ny"], 12, w["dx"], w["dx"], w["dz"], w["dt"], w["etabar"], w["t0"], threads=1,
                    eos_table=eos_table)
        o1.set_primitives(sub, np.ones(m), np.zeros(m), np.zeros(m), np.zeros(m), pl=sub / 3)
        a = time.perf_counter()
        o1.step(2)
        t1 = time.perf_counter() - a
        o1.close()
        out["serial"] = {"value": m * 2 / t1, "cores": 1,
                         "sample": f"2 steps of the middle 12 eta planes ({m} cells) on one core ({t1:.1f} s)"}
    return out


# ---------------------------------------------------------------------------
# Batched and serving forms
# ---------------------------------------------------------------------------
HBARC = 0.197326938
T_FREEZE_GEV = 0.150
# A request retires at freeze-out or after this many steps, the cap of the
# batched workload: an event whose dilute edge turns unstable under the
# reference scheme (the oracle shows the same breakdown) stays hot and would
# otherwise run until it produces NaN.
SERVE_MAX_STEPS = 1000


def mc_event(V, L, nz, dx, dz, b, seed, width=1.0):
    """MC-Glauber Pb+Pb energy density (peak 60 GeV/fm^3) on an L x L x nz lattice;
    an empty draw (no wounded nucleon at large b) is redrawn with the next seed.
    Wounded nucleons deposit Gaussians of width 1 fm: with 0.5 fm hot spots
    the reference's scheme loses the Newton recovery in the dilute, fast edge
    of many events after tau ~ 3 fm/c (the CPU oracle shows the same), and
    such events never cool below the freeze-out temperature."""
    while True:
        g = V.glauber(A=208, R=6.62, a=0.546, sigma_nn_mb=70.0, b=b, seed=seed, e_peak=60.0 / HBARC,
                      nucleon_width=width, eta_flat=1.5, eta_sigma=1.4, e_floor=1e-3)
        e, nw = V.ic_mc_glauber(L, L, nz, dx, dz, g)
        if nw > 0:
            return e
        seed += 7919


def parallel_map(fn, items, workers=16):
    from concurrent.futures import ThreadPoolExecutor  # ctypes calls release the GIL
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(fn, items))


def batch_events(V, n, seed):
    """n boost-invariant events: side ~ U{101..261}, dx ~ U[0.08, 0.12],
    dtau = 0.2 dx, b ~ U[0, 12] fm, tau0 = 0.5, eta/s = 0.16."""
    rng = np.random.default_rng(seed)
    evs = []
    for i in range(n):
        L = int(rng.integers(101, 262))
        dx = float(rng.uniform(0.08, 0.12))
        evs.append(dict(nx=L, ny=L, nz=1, dx=dx, dz=0.1, dt=0.2 * dx, t0=0.5, etabar=0.16,
                        b=float(rng.uniform(0.0, 12.0)), seed=seed * 100003 + i))
    es = parallel_map(lambda ev: mc_event(V, ev["nx"], 1, ev["dx"], ev["dz"], ev["b"], ev["seed"]), evs)
    for ev, e in zip(evs, es):
        ev["e"] = e
    return evs


def admit_event(B, ev, max_steps=0, tag=0):
    return B.admit(ev["nx"], ev["ny"], ev["nz"], ev["dx"], ev["dx"], ev["dz"], ev["dt"], ev["t0"], ev["etabar"],
                   ev["e"], pl=ev["e"] / 3, max_steps=max_steps, tag=tag)


def run_batched(V, evs, tab, precision, freeze, max_steps=None, autotune_steps=0, poll_steps=32):
    """All events resident in one batch until each stops; returns per-event
    stop step, stop time, and the batch (events still resident)."""
    budget = sum(V.batch_event_bytes(ev["nx"], ev["ny"], ev["nz"]) for ev in evs) + (64 << 20)
    B = V.Batch(budget_bytes=budget, max_events=len(evs), eos_table=tab, precision=precision,
                freezeout_T_GeV=freeze, steps_per_graph=16)
    t0 = time.perf_counter()
    slots = [admit_event(B, ev, 0 if max_steps is None else max_steps[i], i) for i, ev in enumerate(evs)]
    B.synchronize()
    t_admit = time.perf_counter() - t0
    tune = None
    if autotune_steps:
        best, ms = B.autotune(autotune_steps)
        tune = {"best": V.batch_variants()[best], "ms_per_step": [round(float(x), 4) for x in ms],
                "seconds": time.perf_counter() - t0 - t_admit}
    stop_step, stop_time, stop_state, left = {}, {}, {}, set(slots)
    while left:
        B.step(poll_steps)
        now = time.perf_counter() - t0
        for st in B.poll():
            if st["slot"] in left and st["state"] != V.EVENT_LIVE:
                left.discard(st["slot"])
                stop_step[st["tag"]] = st["nstep"]
                stop_time[st["tag"]] = now
                stop_state[st["tag"]] = st["state"]
    wall = time.perf_counter() - t0
    n = len(evs)
    return dict(B=B, slots=slots, wall=wall, admit_s=t_admit, stop_step=[stop_step[i] for i in range(n)],
                stop_time=[stop_time[i] for i in range(n)], stop_state=[stop_state[i] for i in range(n)],
                autotune=tune)


def run_plain(V, ev, steps, tab):
    """The plain per-event FP64 path: one lattice context, set-up to teardown.
    Returns the wall time and whether the run stayed free of NaN."""
    t0 = time.perf_counter()
    g = V.Vah(ev["nx"], ev["ny"], ev["nz"], ev["dx"], ev["dx"], ev["dz"], ev["dt"], ev["etabar"], ev["t0"],
              eos_table=tab)
    n = ev["nx"] * ev["ny"] * ev["nz"]
    g.set_primitives(ev["e"], np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n), pl=ev["e"] / 3)
    g.step(steps)
    clean = True
    try:
        g.synchronize()
    except V.VahError as err:
        if err.code != V.VAH_ERR_NAN:
            raise
        clean = False
    g.close()
    return time.perf_counter() - t0, clean


def per_cell_errors(a, b):
    """Per-cell relative errors of state a against FP64 state b: e, and u^mu
    relative to u^tau (the bar of the batched and serving forms)."""
    e = np.abs(a["e"] - b["e"]) / b["e"]
    u = np.max(np.abs(a["u"] - b["u"]), axis=0) / b["u"][0]
    return float(e.max()), float(u.max())


def batch_sample(V, n_events, tab, autotune_steps=0, max_steps=1000):
    """The batched form on n_events events drawn like the batched workload
    (boost-invariant MC-Glauber Pb+Pb, side U{101..261}, dx U[0.08,0.12],
    dtau = 0.2 dx, b U[0,12] fm), each run to freeze-out (hottest cell below
    150 MeV) in one FP64 batch; the same events through the plain per-event
    FP64 path one after the other; and the split-precision batch, with its
    per-cell error against the FP64 batch at the FP64 freeze-out steps."""
    evs = batch_events(V, n_events, seed=2024)
    cells = np.array([ev["nx"] * ev["ny"] for ev in evs])
    progress("batched FP64 to freeze-out")
    r64 = run_batched(V, evs, tab, V.FP64, T_FREEZE_GEV, max_steps=[max_steps] * n_events,
                      autotune_steps=autotune_steps)
    progress("batched split precision")
    nsteps = r64["stop_step"]
    frozen = int(np.sum(np.array(r64["stop_state"]) == V.EVENT_FROZEN))
    r32t = run_batched(V, evs, tab, V.FP32_SPLIT, T_FREEZE_GEV, max_steps=[max_steps] * n_events,
                       autotune_steps=autotune_steps)
    r32t["B"].close()
    r32 = run_batched(V, evs, tab, V.FP32_SPLIT, 0.0, max_steps=nsteps)
    errs = np.array([per_cell_errors(r32["B"].state(a, ("e", "u")), r64["B"].state(b, ("e", "u")))
                     for a, b in zip(r32["slots"], r64["slots"])])
    r32["B"].close()
    r64["B"].close()
    progress("plain per-event FP64 path")
    t_plain = sum(run_plain(V, ev, k, tab)[0] for ev, k in zip(evs, nsteps))
    lat = np.array(r64["stop_time"])
    cu = float(np.sum(cells * np.array(nsteps)))
    return {
        "workload": f"{n_events} boost-invariant MC-Glauber Pb+Pb events (1 fm nucleons), side U{{101..261}}, "
                    "dx U[0.08,0.12], dtau=0.2dx, b U[0,12] fm, tau0=0.5, eta/s=0.16, LQCD-like EOS table, "
                    "each to max T < 150 MeV, one batch",
        "precision": "fp64", "events": n_events, "frozen": frozen,
        "steps_min_max": [int(min(nsteps)), int(max(nsteps))],
        "events_per_s": n_events / r64["wall"], "wall_s": r64["wall"], "admission_s": r64["admit_s"],
        "cell_updates_per_s": cu / r64["wall"],
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "note": "all events admitted at once; admission to the poll that saw the event frozen"},
        "autotune": r64["autotune"],
        "steps_percentiles": {str(q): float(np.percentile(nsteps, q)) for q in (0, 50, 90, 99, 100)},
        "plain_fp64_events_per_s": n_events / t_plain,
        "speedup_vs_plain_fp64": t_plain / r64["wall"],
        "split_precision": {"events_per_s": n_events / r32t["wall"], "speedup_vs_plain_fp64": t_plain / r32t["wall"],
                            "per_cell_max_rel_error_vs_fp64": {"e": float(errs[:, 0].max()),
                                                               "u": float(errs[:, 1].max())},
                            "events_within_1e-4": int(np.sum(np.max(errs, axis=1) <= 1e-4))},
    }


def plain_service(V, r, steps, tab):
    """Service time of one request on the sequential plain FP64 path: its
    initial state generated on the host (vah_ic_mc_glauber), then one lattice
    context from set-up to teardown for the same number of steps."""
    t0 = time.perf_counter()
    e, _ = V.ic_mc_glauber(r.nx, r.ny, r.nz, r.dx, r.dz, r.glauber)
    t_ic = time.perf_counter() - t0
    return t_ic + run_plain(V, dict(nx=r.nx, ny=r.ny, nz=r.nz, dx=r.dx, dz=r.dz, dt=r.dt, t0=r.t0, etabar=r.etabar,
                                    e=e), steps, tab)[0]


def serving_sample(V, n_req, rate, tab, budget_gb):
    """The serving form on n_req requests of the serving mix (80 % boost-
    invariant L^2, 20 % 3+1D L^2 x nz, per-request MC-Glauber seeds) arriving
    as a Poisson stream: continuous batching over the FP64 batch engine,
    against the sequential plain FP64 path on the same arrivals."""
    serve = importlib.import_module("gpu-vah_amd.serve")
    reqs, base = serving_requests(V, n_req, seed=77, device_ic=True)
    arrivals = base / rate
    for r, a in zip(reqs, arrivals):
        r.arrival = float(a)
    B = V.Batch(budget_bytes=int(budget_gb * (1 << 30)), max_events=256, eos_table=tab, precision=V.FP64,
                freezeout_T_GeV=T_FREEZE_GEV, steps_per_graph=16)
    out = serve.serve(B, reqs, steps_per_poll=16)
    B.close()
    progress(f"served {out['events']} requests in {out['seconds']:.1f} s; sequential plain FP64 path")
    service = [plain_service(V, r, r.nstep, tab) for r in reqs]
    lat = serve.fifo_latencies(arrivals, service)
    finish = max(a + l for a, l in zip(sorted(arrivals), lat))
    return {
        "workload": f"{n_req} Poisson requests at {rate:g}/s, 80% boost-invariant L^2 (L U{{121..301}}), "
                    "20% 3+1D L^2 x nz (L U{121..201}, nz U{41..81}), dx=deta=0.1, dtau=0.02, MC-Glauber "
                    "(1 fm nucleons) per-request seeds generated on the device at admission, retire at "
                    f"max T < 150 MeV or after {SERVE_MAX_STEPS} steps; C++ request loop (vah_serve)",
        "precision": "fp64", "requests": n_req, "requests_3d": int(sum(r.nz > 1 for r in reqs)),
        "events_per_s": out["events_per_s"],
        "latency_s": {"p50": out["p50_latency_s"], "p99": out["p99_latency_s"],
                      "note": "arrival to retirement at freeze-out, queueing included"},
        "peak_resident": out["peak_resident"], "hbm_budget_gib": budget_gb,
        "sequential_plain_fp64": {"events_per_s": n_req / finish, "p50_latency_s": float(np.percentile(lat, 50)),
                                  "p99_latency_s": float(np.percentile(lat, 99))},
    }


def bench_batch(V, args):
    """The batched form alone on args.events events (one JSON line)."""
    tab = V.eos_lqcd_table(10000, 1e-8, 1e4)
    r = batch_sample(V, args.events, tab, autotune_steps=args.autotune_steps, max_steps=args.max_steps)
    line = {
        "metric": "events/s", "value": r["events_per_s"], "unit": "events/s", "n_gpus": 1,
        "steps": r["steps_min_max"][1], "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": r["workload"], "events": r["events"]},
        "batched": r,
        "roofline": {"bound": "hbm", "achieved": BYTES_PER_CELL_UPDATE * r["cell_updates_per_s"] / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": BYTES_PER_CELL_UPDATE * r["cell_updates_per_s"] / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "kernel": "batch stage kernels (whole-run average, 856 B per cell update)"},
        "cpu_baseline": None,
    }
    print(json.dumps(line))


def serving_requests(V, n, seed, device_ic=False):
    """Request mix: 80% boost-invariant L x L (L ~ U{121..301}), 20% 3+1D
    L x L x nz (L ~ U{121..201}, nz ~ U{41..81}); dx = deta = 0.1, dtau = 0.02,
    per-request MC-Glauber seeds.  device_ic: the requests carry only their
    Glauber geometry and seed and the engine generates each initial state on
    the device at admission (inside the latency); otherwise the states are
    built on the host beforehand.  Arrival times are set per rate."""
    serve = importlib.import_module("gpu-vah_amd.serve")
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        if rng.random() < 0.8:
            L, nz = int(rng.integers(121, 302)), 1
        else:
            L, nz = int(rng.integers(121, 202)), int(rng.integers(41, 82))
        specs.append((i, L, nz, float(rng.uniform(0.0, 12.0)), seed * 100003 + i))
    base = rng.exponential(1.0, size=n).cumsum()
    if device_ic:
        reqs = [serve.Request(i, 0.0, L, L, nz, 0.1, 0.1, 0.02, 0.5, 0.16, None, max_steps=SERVE_MAX_STEPS,
                              glauber=V.glauber(A=208, R=6.62, a=0.546, sigma_nn_mb=70.0, b=b, seed=sd,
                                                e_peak=60.0 / HBARC, nucleon_width=1.0, eta_flat=1.5, eta_sigma=1.4,
                                                e_floor=1e-3))
                for (i, L, nz, b, sd) in specs]
        return reqs, base
    es = parallel_map(lambda sp: mc_event(V, sp[1], sp[2], 0.1, 0.1, sp[3], sp[4]), specs)
    reqs = [serve.Request(i, 0.0, L, L, nz, 0.1, 0.1, 0.02, 0.5, 0.16, e, max_steps=SERVE_MAX_STEPS)
            for (i, L, nz, _, _), e in zip(specs, es)]
    return reqs, base


def bench_precision(V, args):
    """Split precision against FP64, both on the GPU: the 81x81x1 optical
    Glauber Au+Au event (b = 7 fm, conformal EOS, eta/s = 0.2, dtau = 0.01)
    with seeded +-5% lognormal noise on e, 450 steps to tau = 5 fm/c.  Per-field
    max and RMS relative errors (e; u^mu relative to u^tau; P_L; pi^munu
    relative to e) and the drift of the total energy tau * sum T^tautau."""
    nx = ny = 81
    n = nx * ny
    e = V.ic_optical_glauber(nx, ny, 1, 0.1, 0.1, V.glauber())
    e = e * np.exp(0.05 * np.random.default_rng(2024).standard_normal(n))
    steps = 450
    runs = {}
    for name, prec in (("split", V.FP32_SPLIT), ("fp64", V.FP64)):
        g = V.Vah(nx, ny, 1, 0.1, 0.1, 0.1, 0.01, 0.2, 0.5, precision=prec)
        g.set_primitives(e, np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n), pl=e / 3)
        g.synchronize()
        a = time.perf_counter()
        g.step(steps)
        g.synchronize()
        runs[name] = (g.state(), time.perf_counter() - a)
        g.close()
    a, b = runs["split"][0], runs["fp64"][0]
    rel = {"e": np.abs(a["e"] - b["e"]) / b["e"]}
    for d, k in enumerate(("u_tau", "u_x", "u_y", "u_eta")):
        rel[k] = np.abs(a["u"][d] - b["u"][d]) / b["u"][0]
    rel["P_L"] = np.abs(a["q"][4] - b["q"][4]) / b["q"][4]
    rel["pi"] = np.max(np.abs(a["q"][5:] - b["q"][5:]), axis=0) / b["e"]
    errs = {k: {"max": float(v.max()), "rms": float(np.sqrt(np.mean(v * v)))} for k, v in rel.items()}
    drift = float(abs(np.sum(a["q"][0]) / np.sum(b["q"][0]) - 1.0))
    worst = max(errs["e"]["max"], *(errs[k]["max"] for k in ("u_tau", "u_x", "u_y", "u_eta")))
    print(json.dumps({
        "metric": "max relative error of e and u^mu, split precision vs FP64", "value": worst, "unit": "relative",
        "n_gpus": 1, "steps": steps, "warmup": 0, "ms_per_step": runs["split"][1] / steps * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64", "data": "synthetic",
        "config": {"workload": "2+1D optical Glauber Au+Au b=7 fm with seeded +-5% lognormal noise on e, 81x81x1, "
                               "dx=0.1, tau0=0.5, dtau=0.01, 450 steps to tau=5 fm/c, conformal EOS, eta/s=0.2",
                   "fp64_ms_per_step": runs["fp64"][1] / steps * 1e3},
        "errors": errs, "total_energy_drift": drift,
        "roofline": None, "cpu_baseline": None}))


def bench_serving(V, args):
    serve = importlib.import_module("gpu-vah_amd.serve")
    tab = V.eos_lqcd_table(10000, 1e-8, 1e4)
    reqs, base = serving_requests(V, args.events, seed=77, device_ic=True)
    rates = [float(r) for r in args.rates.split(",")]
    prec = V.FP32_SPLIT if args.precision == "fp32_split" else V.FP64
    sweeps = []
    for rate in rates:
        for r, a in zip(reqs, base / rate):
            r.arrival, r.admitted, r.finished, r.state = float(a), None, None, -1
        B = V.Batch(budget_bytes=int(args.budget_gb * (1 << 30)), max_events=256, eos_table=tab,
                    precision=prec, freezeout_T_GeV=T_FREEZE_GEV, steps_per_graph=16)
        out = serve.serve(B, reqs, steps_per_poll=16)
        B.close()
        out["rate"] = rate
        out["steps"] = [r.nstep for r in reqs]
        sweeps.append(out)
        print(f"rate {rate}/s: {out['events_per_s']:.2f} events/s, p50 {out['p50_latency_s']:.2f} s, "
              f"p99 {out['p99_latency_s']:.2f} s, peak resident {out['peak_resident']}", file=sys.stderr, flush=True)
    # the sequential plain FP64 path: measured service time of every request
    nsteps = sweeps[-1]["steps"]
    service = [plain_service(V, r, k, tab) for r, k in zip(reqs, nsteps)]
    for sw in sweeps:
        lat = serve.fifo_latencies(base / sw["rate"], service)
        finish = max(a + l for a, l in zip(sorted(base / sw["rate"]), lat))
        sw["sequential_fp64"] = {"events_per_s": len(reqs) / finish, "p50_latency_s": float(np.percentile(lat, 50)),
                                 "p99_latency_s": float(np.percentile(lat, 99))}
        del sw["steps"]
    top = max(sweeps, key=lambda s: s["events_per_s"])
    line = {
        "metric": "events/s",
        "value": top["events_per_s"],
        "unit": "events/s",
        "n_gpus": 1,
        "steps": None,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32/f64",
        "data": "synthetic",
        "config": {"workload": f"serving: {len(reqs)} Poisson requests, 80% boost-invariant L^2 (L U{{121..301}}), "
                               "20% 3+1D L^2 x nz (L U{121..201}, nz U{41..81}), dx=deta=0.1, dtau=0.02, "
                               "MC-Glauber (1 fm nucleon width) per-request seeds generated on the device at "
                               f"admission, retire at max T < 150 MeV or after {SERVE_MAX_STEPS} steps, {args.precision}, C++ request loop, "
                               f"HBM budget {args.budget_gb} GiB",
                   "rates_per_s": rates},
        "latency_s": {"p50": top["p50_latency_s"], "p99": top["p99_latency_s"]},
        "sweep": sweeps,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line))


def measured_traffic(workload):
    """HBM bytes per launch of each stage kernel, as measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (scaled by the copy-kernel factor) and
    recorded in the committed traffic summary; None where no summary exists
    for this workload."""
    if workload != "mc3d" or not os.path.exists(TRAFFIC_SUMMARY):
        return None
    out = {}
    with open(TRAFFIC_SUMMARY) as f:
        for line in f:
            m = re.match(r"^traffic_bytes_per_launch\s+(stage[12])\s+([0-9.eE+]+)", line.strip())
            if m:
                out[m.group(1)] = float(m.group(2))
    return out if len(out) == 2 else None


def calibrate_copy(mib):
    """A plain device-to-device copy of a known size: the counter reading of
    FETCH_SIZE / WRITE_SIZE on it gives the scale of those counters."""
    import torch
    n = mib << 20
    a = torch.full((n,), 7, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    ok = bool(torch.equal(a, b))
    print(json.dumps({"copy_bytes_each_way": n, "copies": 3, "equal": ok}))
    if not ok:
        sys.exit(1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100,
                    help="timed steps per repetition.  With its 0.5 fm nucleon hot spots the 201x201x101 event "
                         "loses the Newton recovery at the dilute edge from step ~80 on (tau ~ 2.1 fm/c) at "
                         "dtau = 0.02 and blows up near step 100, on the CPU oracle as on the GPU, so the "
                         "event restarts from its initial state every 80 steps, outside the timed region")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps before the timed ones")
    ap.add_argument("--reps", type=int, default=5, help="repetitions of --steps timed steps; the median is reported")
    ap.add_argument("--batch-events", type=int, default=64,
                    help="default workload: events of the batched-form sample in the same line (0: none)")
    ap.add_argument("--serve-requests", type=int, default=48,
                    help="default workload: requests of the serving-form sample in the same line (0: none)")
    ap.add_argument("--serve-rate", type=float, default=16.0, help="arrival rate [1/s] of the serving sample")
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["batch", "serving", "precision"], default="mc3d",
                    help="mc3d (default): one 201x201x101 MC-Glauber event; optical2d: 81x81x1 optical Glauber; "
                         "batch: 256 boost-invariant events in one batch; serving: Poisson request stream; "
                         "precision: split precision against FP64 on the noisy 81x81x1 event to tau = 5 fm/c")
    ap.add_argument("--events", type=int, default=256, help="batch / serving: number of events / requests")
    ap.add_argument("--autotune-steps", type=int, default=4, help="batch: steps per variant trial (0: no autotune)")
    ap.add_argument("--max-steps", type=int, default=1000,
                    help="batch: step limit per event (events of this workload freeze out in 160-700 steps; "
                         "one still hot at the limit has an unstable edge)")
    ap.add_argument("--rates", default="2,4,8,16", help="serving: Poisson arrival rates [1/s] to sweep")
    ap.add_argument("--budget-gb", type=float, default=32.0, help="serving: HBM admission budget [GiB]")
    ap.add_argument("--flux-mode", choices=["shared", "per_cell"], default="shared")
    ap.add_argument("--precision", choices=["fp64", "fp32_split"], default="fp64",
                    help="fp32_split: stencil, fluxes and sources in float, Newton recovery in double")
    ap.add_argument("--eta-chunk", type=int, default=0, help="eta planes per workgroup (0: library default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--calibrate-copy", type=int, metavar="MIB", default=0,
                    help="only run a plain copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the step runs on one GPU; --gpus {args.gpus} is not supported "
              "(several GPUs would shard by event, or a single event by eta slabs with a halo exchange per stage)",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    V = importlib.import_module("gpu-vah_amd")
    if args.workload == "batch":
        bench_batch(V, args)
        return
    if args.workload == "serving":
        bench_serving(V, args)
        return
    if args.workload == "precision":
        bench_precision(V, args)
        return
    w = WORKLOADS[args.workload]
    e = initial_state(V, w)
    n = e.size
    eos_table = V.eos_lqcd_table(10000, 1e-8, 1e4) if w["eos"] == "table" else None
    mode = V.FLUX_SHARED if args.flux_mode == "shared" else V.FLUX_PER_CELL
    g = V.Vah(w["nx"], w["ny"], w["nz"], w["dx"], w["dx"], w["dz"], w["dt"], w["etabar"], w["t0"],
              eos_table=eos_table, flux_mode=mode,
              precision=V.FP32_SPLIT if args.precision == "fp32_split" else V.FP64, eta_chunk=args.eta_chunk)
    # The 201x201x101 event's 0.5 fm hot spots break the scheme down at the
    # dilute edge from step ~80 on (the CPU oracle does the same), so that event
    # is restarted from its initial state every reset_steps steps, outside the
    # timed region: runs of any length time steps of the healthy evolution.
    # Other workloads run as one evolution.
    reset_steps = w["restart"] or (args.warmup + args.steps * args.reps + 1)
    age = [reset_steps]

    def run(k):
        """k steps in chunks that never pass reset_steps since the last restart;
        returns the time of the steps alone, the device idle at both ends of
        each chunk."""
        t = 0.0
        while k > 0:
            if age[0] >= reset_steps:
                g.set_primitives(e, np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n), pl=e / 3)
                g.synchronize()
                age[0] = 0
            m = min(k, reset_steps - age[0])
            t0 = time.perf_counter()
            g.step(m)
            g.synchronize()
            t += time.perf_counter() - t0
            age[0] += m
            k -= m
        return t

    progress(f"{w['name']}: {n} cells, warmup {args.warmup} steps")
    run(args.warmup)
    # timed region: reps x K steps, the median repetition is the value
    g.profile(True)
    rep_ms = [run(args.steps) / args.steps * 1e3 for _ in range(args.reps)]
    prof = g.profile_read()
    g.profile(False)
    ms = float(np.median(rep_ms))
    progress(f"timed {args.reps} x {args.steps} steps: median {ms:.4f} ms/step")
    # one evolution from the initial state, timed in 10-step blocks: the cost
    # of the step as the flow develops (tau0 .. tau0 + reset_steps dtau)
    evolution = None
    if w["restart"]:
        age[0] = reset_steps
        blocks = [run(10) / 10 * 1e3 for _ in range(reset_steps // 10)]
        st = g.stats()
        evolution = {"ms_per_step_by_10_steps": [round(b, 4) for b in blocks],
                     "tau_range": [w["t0"], w["t0"] + 10 * len(blocks) * w["dt"]],
                     "recovery_fallbacks": st.recovery_fallbacks, "nan_cells": st.nan_cells,
                     "max_vT_end": st.max_vT}
    st = g.stats()
    # the validity maps (checkValidity, which the reference's GPU step runs after
    # every step) as the on-demand device pass, without host copies
    g.validity(copy=False)
    t0v = time.perf_counter()
    for _ in range(5):
        g.validity(copy=False)
    validity_ms = (time.perf_counter() - t0v) / 5 * 1e3
    g.close()

    value = n / (ms * 1e-3)
    # dominant kernel: the stage-2 launch (convex combination and the larger
    # read set); algorithmic bytes per launch over its mean launch duration
    stage_gbs = [BYTES_PER_CELL_STAGE[i] * n / (prof["avg_ms"][i] * 1e-3) / 1e9 for i in range(2)]
    dom = 1 if prof["avg_ms"][1] >= prof["avg_ms"][0] else 0
    achieved = stage_gbs[dom]
    traffic = measured_traffic(args.workload) if (args.flux_mode, args.precision) == ("shared", "fp64") else None
    line = {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32/f64",
        "data": "synthetic",
        "config": {"workload": w["name"], "cells": n, "flux_mode": args.flux_mode,
                   "precision": args.precision, "restart_every_steps": w["restart"],
                   "reps": args.reps, "rep_ms_per_step": [round(x, 4) for x in rep_ms],
                   "stage_kernel_ms": prof["avg_ms"], "stage_kernel_launches": prof["launches"],
                   "recovery_fallbacks": st.recovery_fallbacks, "nan_cells": st.nan_cells,
                   "evolution": evolution,
                   "validity_pass_ms": round(validity_ms, 4)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if traffic is None else traffic[f"stage{dom + 1}"],
                     "traffic_source": None if traffic is None else
                     "rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel (copy-calibrated), read from "
                     "profiles/stage_kernel_hbm_traffic_201x201x101_fp64.md; not measured in this run",
                     "kernel": f"fused KT stage-{dom + 1} kernel",
                     "algorithmic_bytes_per_launch": BYTES_PER_CELL_STAGE[dom] * n,
                     "avg_launch_ms": prof["avg_ms"][dom],
                     "per_stage_GBs": stage_gbs,
                     "step_GBs": BYTES_PER_CELL_UPDATE * value / 1e9,
                     "step_frac": BYTES_PER_CELL_UPDATE * value / 1e9 / HBM_PEAK_GBS,
                     "traffic_per_stage": traffic},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        progress("CPU baseline")
        line["cpu_baseline"] = cpu_baseline(w, e, eos_table, args.cpu_seconds)
    if args.workload == "mc3d":
        tab = eos_table if eos_table is not None else V.eos_lqcd_table(10000, 1e-8, 1e4)
        if args.batch_events > 0:
            progress(f"batched sample: {args.batch_events} events")
            line["batched"] = batch_sample(V, args.batch_events, tab)
        if args.serve_requests > 0:
            progress(f"serving sample: {args.serve_requests} requests at {args.serve_rate:g}/s")
            line["serving"] = serving_sample(V, args.serve_requests, args.serve_rate, tab, args.budget_gb)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
