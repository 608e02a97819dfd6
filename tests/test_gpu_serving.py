"""GPU test of the serving loop (gpu-vah_amd/serve.py) over the real batch
engine: a burst of boost-invariant requests into an HBM budget that holds
only two of the largest at a time.  Every request is admitted first come,
first served, retired at its freeze-out step (hottest cell below 150 MeV),
and holds the fields the single-lattice path computes for that many steps."""
import importlib

import numpy as np
import pytest

from tests.parity import field_errors, gaussian_ic, rest_frame, vah

pytestmark = pytest.mark.gpu

HBARC = 0.197326938
T_F = 0.150  # GeV


def test_serving_loop_retires_every_request_at_its_freeze_out():
    V = vah()
    S = importlib.import_module("gpu-vah_amd.serve")
    tab = V.eos_lqcd_table(10000, 1e-8, 1e4)
    e_peak, _ = V.eos_lqcd(0.175 / HBARC)
    sizes = [(40, 36), (52, 48), (44, 44), (36, 50), (48, 40), (40, 40)]
    arrivals = S.poisson_arrivals(len(sizes), 1.0e4, np.random.default_rng(5))  # a burst: ~0.6 ms
    reqs = [S.Request(rid=i, arrival=float(a), nx=nx, ny=ny, nz=1, dx=0.1, dz=0.1, dt=0.02, t0=0.5, etabar=0.16,
                      e=gaussian_ic(nx, ny, seed=10 + i, peak=e_peak, floor=1e-4), max_steps=3000)
            for i, ((nx, ny), a) in enumerate(zip(sizes, arrivals))]
    biggest = max(V.batch_event_bytes(nx, ny, 1) for nx, ny in sizes)
    B = V.Batch(budget_bytes=2 * biggest, max_events=4, eos_table=tab, freezeout_T_GeV=T_F, steps_per_graph=8)
    out = S.serve(B, reqs, steps_per_poll=8, keep_fields=True)
    B.close()
    print("serving:", out)
    assert out["events"] == len(reqs) and out["rejected"] == 0
    assert out["peak_resident"] >= 2  # the burst shared the GPU
    admitted = [r.admitted for r in sorted(reqs, key=lambda r: r.arrival)]
    assert admitted == sorted(admitted)  # first come, first served
    for r in reqs:
        assert r.state == V.EVENT_FROZEN and 1 < r.nstep < r.max_steps
        g = V.Vah(r.nx, r.ny, 1, r.dx, r.dx, r.dz, r.dt, r.etabar, r.t0, eos_table=tab)
        g.set_primitives(r.e, *rest_frame(r.cells), pl=r.e / 3)
        g.step(r.nstep - 1)
        assert g.stats().max_T * HBARC >= T_F  # still hot one step earlier
        g.step(1)
        assert g.stats().max_T * HBARC < T_F
        err = max(field_errors(r.fields, g.state(), r.tau).values())
        g.close()
        print(f"request {r.rid}: {r.nx}x{r.ny}, frozen after {r.nstep} steps, latency {r.latency * 1e3:.1f} ms, "
              f"worst vs single-lattice path {err:.1e}")
        assert err <= 1e-12
