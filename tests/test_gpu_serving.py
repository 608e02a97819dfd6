"""GPU tests of the serving form: the C++ request loop (include/vah_serve.h,
through gpu-vah_amd/serve.py) over the real batch engine, and the
Monte-Carlo Glauber initial states the engine generates on the device at
admission (vah_batch_admit_glauber).

- a burst of boost-invariant requests into an HBM budget that holds only two
  of the largest at a time: every request is admitted first come, first
  served, retired at its freeze-out step (hottest cell below 150 MeV), and
  holds the fields the single-lattice path computes for that many steps;
- device-generated initial states equal the host generator's
  (vah_ic_mc_glauber) for the same seed, in 2+1D and 3+1D;
- requests that carry only a seed (one of them 3+1D) are served, each
  result equal to the single-lattice path started from the host-generated
  state."""
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
        g.set_primitives(r.e, *rest_frame(r.cells), pl=r.pl)
        g.step(r.nstep - 1)
        assert g.stats().max_T * HBARC >= T_F  # still hot one step earlier
        g.step(1)
        assert g.stats().max_T * HBARC < T_F
        err = max(field_errors(r.fields, g.state(), r.tau).values())
        g.close()
        print(f"request {r.rid}: {r.nx}x{r.ny}, frozen after {r.nstep} steps, latency {r.latency * 1e3:.1f} ms, "
              f"worst vs single-lattice path {err:.1e}")
        assert err <= 1e-12


def _glauber(V, seed, b, width=1.0):
    return V.glauber(A=208, R=6.62, a=0.546, sigma_nn_mb=70.0, b=b, seed=seed, e_peak=60.0 / HBARC,
                     nucleon_width=width, eta_flat=1.5, eta_sigma=1.4, e_floor=1e-3)


@pytest.mark.parametrize("dims", [(121, 121, 1, 0.1, 0.1), (61, 57, 15, 0.2, 0.25)], ids=["2+1D", "3+1D"])
def test_device_generated_glauber_state_equals_the_host_generator(dims):
    V = vah()
    nx, ny, nz, dx, dz = dims
    g = _glauber(V, seed=31, b=5.0)
    e_host, nw = V.ic_mc_glauber(nx, ny, nz, dx, dz, g)
    assert nw > 50
    B = V.Batch(budget_bytes=V.batch_event_bytes(nx, ny, nz) + (8 << 20), max_events=2, freezeout_T_GeV=0.0,
                steps_per_graph=0)
    slot = B.admit_glauber(nx, ny, nz, dx, dx, dz, 0.02, 0.5, 0.16, g, pl_ratio=1.0 / 3.0)
    st = B.state(slot)
    B.close()
    rel = np.abs(st["e"] - e_host) / e_host
    print(f"{nx}x{ny}x{nz}: {nw} wounded, max relative difference {rel.max():.1e}")
    assert rel.max() < 1e-13
    assert np.array_equal(st["q"][4], st["e"] * (1.0 / 3.0))  # P_L = e/3 of the device's e
    assert np.all(st["u"][0] == 1.0) and np.all(st["u"][1:] == 0.0)


def test_seed_only_requests_are_served_like_the_single_lattice_path():
    V = vah()
    S = importlib.import_module("gpu-vah_amd.serve")
    tab = V.eos_lqcd_table(10000, 1e-8, 1e4)
    shapes = [(61, 61, 1, 0.2, 0.1), (41, 41, 9, 0.3, 0.4), (71, 67, 1, 0.2, 0.1)]
    reqs = []
    for i, (nx, ny, nz, dx, dz) in enumerate(shapes):
        reqs.append(S.Request(rid=i, arrival=0.0005 * i, nx=nx, ny=ny, nz=nz, dx=dx, dz=dz, dt=0.2 * dx, t0=0.5,
                              etabar=0.16, e=None, glauber=_glauber(V, seed=100 + i, b=3.0 + i), max_steps=3000))
    budget = sum(V.batch_event_bytes(r.nx, r.ny, r.nz) for r in reqs)
    B = V.Batch(budget_bytes=budget, max_events=4, eos_table=tab, freezeout_T_GeV=T_F, steps_per_graph=8)
    out = S.serve(B, reqs, steps_per_poll=8, keep_fields=True)
    B.close()
    assert out["events"] == len(reqs) and out["rejected"] == 0
    for r in reqs:
        assert r.state == V.EVENT_FROZEN and 1 < r.nstep < r.max_steps
        e, _ = V.ic_mc_glauber(r.nx, r.ny, r.nz, r.dx, r.dz, r.glauber)
        g = V.Vah(r.nx, r.ny, r.nz, r.dx, r.dx, r.dz, r.dt, r.etabar, r.t0, eos_table=tab)
        g.set_primitives(e, *rest_frame(r.cells), pl=e * (1.0 / 3.0))
        g.step(r.nstep)
        g.synchronize()
        err = max(field_errors(r.fields, g.state(), r.tau).values())
        g.close()
        print(f"request {r.rid}: {r.nx}x{r.ny}x{r.nz} from seed {r.glauber.seed}, frozen after {r.nstep} steps, "
              f"latency {r.latency * 1e3:.1f} ms, worst vs single lattice {err:.1e}")
        assert err <= 1e-10


def test_serving_loop_retires_a_request_at_its_step_cap():
    """A request still hot at its max_steps retires there (state
    EVENT_STEP_LIMIT, exactly max_steps steps) with the single-lattice path's
    fields for that many steps, while the others freeze out as usual: the cap
    that bounds an event whose dilute edge turns unstable."""
    V = vah()
    S = importlib.import_module("gpu-vah_amd.serve")
    tab = V.eos_lqcd_table(10000, 1e-8, 1e4)
    e_peak, _ = V.eos_lqcd(0.175 / HBARC)
    caps = [4, 3000, 6]
    reqs = [S.Request(rid=i, arrival=0.0, nx=36, ny=40, nz=1, dx=0.1, dz=0.1, dt=0.02, t0=0.5, etabar=0.16,
                      e=gaussian_ic(36, 40, seed=30 + i, peak=e_peak, floor=1e-4), max_steps=cap)
            for i, cap in enumerate(caps)]
    B = V.Batch(budget_bytes=1 << 30, max_events=4, eos_table=tab, freezeout_T_GeV=T_F, steps_per_graph=8)
    out = S.serve(B, reqs, steps_per_poll=8, keep_fields=True)
    B.close()
    assert out["events"] == len(reqs) and out["rejected"] == 0
    for r in reqs:
        if r.max_steps < 100:
            assert r.state == V.EVENT_STEP_LIMIT and r.nstep == r.max_steps, (r.rid, r.state, r.nstep)
        else:
            assert r.state == V.EVENT_FROZEN and 1 < r.nstep < r.max_steps
        g = V.Vah(r.nx, r.ny, 1, r.dx, r.dx, r.dz, r.dt, r.etabar, r.t0, eos_table=tab)
        g.set_primitives(r.e, *rest_frame(r.cells), pl=r.pl)
        g.step(r.nstep)
        err = max(field_errors(r.fields, g.state(), r.tau).values())
        g.close()
        print(f"request {r.rid}: cap {r.max_steps}, state {r.state} after {r.nstep} steps, worst {err:.1e}")
        assert err <= 1e-12
