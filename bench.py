#!/usr/bin/env python3
"""Throughput of the FP64 VAH Kurganov-Tadmor Heun-RK2 step on one MI355X.

Workload (default): one 3+1D event at real shape -- Monte-Carlo Glauber Pb+Pb
at b = 6 fm (Woods-Saxon R = 6.62 fm, a = 0.546 fm, sigma_NN = 70 mb, wounded
nucleons as Gaussians of width 0.5 fm, seed 42, peak 60 GeV/fm^3), eta plateau
half-width 1.5 with Gaussian tails 1.4, lattice 201 x 201 x 101,
dx = deta = 0.1, tau0 = 0.5 fm/c, dtau = 0.02 fm/c, lattice-QCD-like EOS
tabulated at 10^4 nodes, eta/s = 0.16.  A step is one Heun-RK2 step of the
whole lattice (two fused stage launches); the state is resident in HBM.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
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
    "c2": dict(nx=201, ny=201, nz=101, dx=0.1, dz=0.1, dt=0.02, t0=0.5, etabar=0.16, eos="table", ic="mc",
               name="3+1D MC-Glauber Pb+Pb b=6 fm, 201x201x101, dx=deta=0.1, tau0=0.5, dtau=0.02, "
                    "LQCD-like EOS table (1e4 nodes), eta/s=0.16"),
    "c1": dict(nx=81, ny=81, nz=1, dx=0.1, dz=0.1, dt=0.01, t0=0.5, etabar=0.2, eos="conformal", ic="optical",
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


def cpu_baseline(w, e, eos_table, seconds=20.0):
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
    o.close()
    return {"value": n * steps / t_used, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": f"{steps} full-lattice steps of the same workload on the CPU oracle ({t_used:.1f} s)"}


def measured_traffic(workload):
    """HBM bytes per launch of each stage kernel, as measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (scaled by the copy-kernel factor) and
    recorded in the committed traffic summary; None where no summary exists
    for this workload."""
    if workload != "c2" or not os.path.exists(TRAFFIC_SUMMARY):
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
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c2")
    ap.add_argument("--flux-mode", choices=["shared", "per_cell"], default="shared")
    ap.add_argument("--precision", choices=["fp64", "fp32_split"], default="fp64",
                    help="fp32_split: stencil, fluxes and sources in float, Newton recovery in double")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--calibrate-copy", type=int, metavar="MIB", default=0,
                    help="only run a plain copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a multi-GPU run would shard events or eta slabs, see the design notes)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    V = importlib.import_module("gpu-vah_amd")
    w = WORKLOADS[args.workload]
    e = initial_state(V, w)
    n = e.size
    eos_table = V.eos_lqcd_table(10000, 1e-8, 1e4) if w["eos"] == "table" else None
    mode = V.FLUX_SHARED if args.flux_mode == "shared" else V.FLUX_PER_CELL
    g = V.Vah(w["nx"], w["ny"], w["nz"], w["dx"], w["dx"], w["dz"], w["dt"], w["etabar"], w["t0"],
              eos_table=eos_table, flux_mode=mode,
              precision=V.FP32_SPLIT if args.precision == "fp32_split" else V.FP64)
    g.set_primitives(e, np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n), pl=e / 3)
    g.step(args.warmup)
    g.synchronize()

    # timed region: K steps, device idle on both sides
    g.profile(True)
    t0 = time.perf_counter()
    g.step(args.steps)
    g.synchronize()
    elapsed = time.perf_counter() - t0
    prof = g.profile_read()
    g.profile(False)
    st = g.stats()
    g.close()

    value = n * args.steps / elapsed
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
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32/f64",
        "data": "synthetic",
        "config": {"workload": w["name"], "cells": n, "flux_mode": args.flux_mode,
                   "precision": args.precision,
                   "stage_kernel_ms": prof["avg_ms"], "recovery_fallbacks": st.recovery_fallbacks,
                   "nan_cells": st.nan_cells},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if traffic is None else traffic[f"stage{dom + 1}"],
                     "kernel": f"fused KT stage-{dom + 1} kernel",
                     "algorithmic_bytes_per_launch": BYTES_PER_CELL_STAGE[dom] * n,
                     "avg_launch_ms": prof["avg_ms"][dom],
                     "per_stage_GBs": stage_gbs,
                     "step_GBs": BYTES_PER_CELL_UPDATE * value / 1e9,
                     "traffic_per_stage": traffic},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(w, e, eos_table, args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
