"""GPU parity on the real workloads of the benchmark configurations, not
stand-ins: the HIP path (single lattice, batch engine and serving loop, all
FP64) against the CPU oracle cell by cell at 1e-10 (metric in tests/parity.py).

- the 2+1D optical-Glauber Au+Au event at b = 7 fm (81x81x1, conformal EOS,
  eta/s = 0.2, 200 steps), in both flux modes;
- a lumpy Monte-Carlo Glauber Pb+Pb event of the batched workload (101x101,
  dx = 0.1, 1 fm nucleons, table EOS) run to its freeze-out (hottest cell below
  150 MeV), alone and in the batch engine, which must stop it at the same step;
- four events drawn from the batched workload's own distribution (side
  U{101..261}, dx U[0.08, 0.12], b U[0, 12]) in one batch to freeze-out, each
  against the single-lattice path and one against the oracle;
- a serving mix with a 3+1D request next to boost-invariant ones, every
  retired result against the oracle at its own size and step count;
- the conformal Bjorken solution (python/conformal_bjorken_test.py:46-68)
  reproduced by the HIP path on uniform 2+1D and 3+1D lattices."""
import importlib

import numpy as np
import pytest

from tests.parity import assert_parity, field_errors, gaussian_ic, rest_frame, table_T, vah

pytestmark = pytest.mark.gpu

HBARC = 0.197326938
T_F = 0.150  # GeV


def _table():
    return vah().eos_lqcd_table(10000, 1e-8, 1e4)


def _oracle_run(oracle_mod, nx, ny, nz, dx, dz, dt, etabar, t0, e, steps, tab=None, threads=None):
    n = nx * ny * nz
    o = oracle_mod.Oracle(nx, ny, nz, dx, dx, dz, dt, etabar, t0, eos_table=tab, threads=threads)
    o.set_primitives(e, *rest_frame(n), pl=e / 3)
    o.step(steps)
    st, tau, fails = o.state(), o.time, o.newton_fail
    o.close()
    return st, tau, fails


def _gpu_run(nx, ny, nz, dx, dz, dt, etabar, t0, e, steps, tab=None, flux_mode=0):
    V = vah()
    n = nx * ny * nz
    g = V.Vah(nx, ny, nz, dx, dx, dz, dt, etabar, t0, eos_table=tab, flux_mode=flux_mode)
    g.set_primitives(e, *rest_frame(n), pl=e / 3)
    g.step(steps)
    g.synchronize()
    st, tau = g.state(), g.time
    g.close()
    return st, tau


@pytest.mark.parametrize("flux_mode", [0, 1], ids=["shared", "per_cell"])
def test_optical_glauber_auau_81x81_200_steps(oracle_mod, flux_mode):
    """Au+Au optical Glauber (Woods-Saxon R 6.38 fm, a 0.535 fm, 42 mb, b 7 fm,
    e ~ 0.5 N_part + 0.5 N_coll scaled to 30 GeV/fm^3), P_L = P_T = e/3,
    81x81x1, dx 0.1, tau0 0.5, dtau 0.01, 200 steps, conformal, eta/s 0.2."""
    V = vah()
    e = V.ic_optical_glauber(81, 81, 1, 0.1, 0.1, V.glauber())
    assert e.max() == pytest.approx(30.0 / HBARC + 1e-3, rel=1e-3)
    a, tau_g = _gpu_run(81, 81, 1, 0.1, 0.1, 0.01, 0.2, 0.5, e, 200, flux_mode=flux_mode)
    b, tau, fails = _oracle_run(oracle_mod, 81, 81, 1, 0.1, 0.1, 0.01, 0.2, 0.5, e, 200)
    assert tau_g == tau and fails == 0
    worst = assert_parity(a, b, tau, label="optical Glauber 81x81x1")
    print("worst field", worst)
    assert np.max(np.abs(b["u"][1])) > 0.1  # developed transverse flow


def _lumpy_event(L, dx, b, seed):
    bench = importlib.import_module("bench")
    return bench.mc_event(vah(), L, 1, dx, 0.1, b, seed)


def _freeze_step_batch(nx, ny, dx, dt, e, tab, precision=0):
    """Steps the batch engine takes before it stops the event at freeze-out."""
    V = vah()
    B = V.Batch(budget_bytes=V.batch_event_bytes(nx, ny, 1) + (8 << 20), max_events=2, eos_table=tab,
                precision=precision, freezeout_T_GeV=T_F, steps_per_graph=8)
    slot = B.admit(nx, ny, 1, dx, dx, 0.1, dt, 0.5, 0.16, e, pl=e / 3, max_steps=5000)
    while True:
        B.step(16)
        st = {s["slot"]: s for s in B.poll()}[slot]
        if st["state"] != V.EVENT_LIVE:
            break
    fields = B.state(slot)
    B.close()
    assert st["state"] == V.EVENT_FROZEN
    return st["nstep"], st["tau"], fields


def test_lumpy_mc_glauber_event_to_freeze_out(oracle_mod):
    """One lumpy MC-Glauber Pb+Pb event (b = 4 fm, 1 fm nucleons, peak
    60 GeV/fm^3) on 101x101 at dx = 0.1, dtau = 0.02, table EOS, eta/s 0.16,
    evolved until its hottest cell falls below 150 MeV: the batch engine's
    freeze-out step is the oracle's, and the single-lattice and batch fields at
    that step equal the oracle's within 1e-10 per cell."""
    V = vah()
    tab = _table()
    L, dx = 101, 0.1
    e = _lumpy_event(L, dx, 4.0, 7)
    nstep, tau_b, fb = _freeze_step_batch(L, L, dx, 0.2 * dx, e, tab)
    assert 200 < nstep < 600
    b, tau, fails = _oracle_run(oracle_mod, L, L, 1, dx, 0.1, 0.2 * dx, 0.16, 0.5, e, nstep, tab)
    assert fails == 0 and tau_b == tau
    # the oracle crosses T_f at the same step
    Tmax = max(table_T(tab, b["e"].max()) * HBARC, 0.0)
    assert Tmax < T_F
    b1, _, _ = _oracle_run(oracle_mod, L, L, 1, dx, 0.1, 0.2 * dx, 0.16, 0.5, e, nstep - 1, tab)
    assert table_T(tab, b1["e"].max()) * HBARC >= T_F
    a, tau_g = _gpu_run(L, L, 1, dx, 0.1, 0.2 * dx, 0.16, 0.5, e, nstep, tab)
    w1 = assert_parity(a, b, tau, label="lumpy event, single lattice, freeze-out")
    w2 = assert_parity(fb, b, tau, label="lumpy event, batch engine, freeze-out")
    print(f"freeze-out after {nstep} steps (tau {tau:.2f}): worst single {w1}, batch {w2}")


def _oracle_sensitivity(oracle_mod, ev, steps, tab, eps=1e-15):
    """The oracle against itself: field errors (tests/parity.py metric) between
    runs from the event's initial state and from that state times (1 + eps g),
    g standard normal.  Where the reference's algorithm takes discontinuous
    decisions in dilute cells -- chiefly the Newton stop rule
    |du_T| <= 1e-7 |u_T| where the iteration converges slowly -- a rounding-level
    difference flips one and moves that cell by up to ~1e-8; no implementation
    that rounds differently from the oracle can agree with it more closely there."""
    n = ev["e"].size
    e2 = ev["e"] * (1 + eps * np.random.default_rng(3).standard_normal(n))
    a, tau, _ = _oracle_run(oracle_mod, ev["nx"], ev["ny"], 1, ev["dx"], 0.1, ev["dt"], ev["etabar"], ev["t0"],
                            ev["e"], steps, tab)
    b, _, _ = _oracle_run(oracle_mod, ev["nx"], ev["ny"], 1, ev["dx"], 0.1, ev["dt"], ev["etabar"], ev["t0"], e2,
                          steps, tab)
    return a, tau, max(field_errors(b, a, tau).values())


def test_batched_workload_draw_to_freeze_out(oracle_mod):
    """Four events of the batched workload's distribution in one FP64 batch with
    freeze-out at 150 MeV.  At each event's stop step the batch's fields equal
    the single-lattice path's and the oracle's within max(1e-10, 10 s), where s
    is the oracle's own spread under a 1e-15 relative perturbation of the
    initial state (_oracle_sensitivity): 1e-10 per cell where the oracle is
    well conditioned, its own reproducibility where it is not."""
    V = vah()
    bench = importlib.import_module("bench")
    tab = _table()
    evs = bench.batch_events(V, 4, seed=77)
    budget = sum(V.batch_event_bytes(ev["nx"], ev["ny"], 1) for ev in evs) + (16 << 20)
    B = V.Batch(budget_bytes=budget, max_events=4, eos_table=tab, freezeout_T_GeV=T_F, steps_per_graph=8)
    slots = [bench.admit_event(B, ev, max_steps=5000, tag=i) for i, ev in enumerate(evs)]
    done = {}
    while len(done) < len(evs):
        B.step(16)
        for s in B.poll():
            if s["state"] != V.EVENT_LIVE and s["slot"] not in done:
                done[s["slot"]] = s
    for i, (ev, slot) in enumerate(zip(evs, slots)):
        s = done[slot]
        assert s["state"] == V.EVENT_FROZEN, (i, s)
        fb = B.state(slot)
        a, tau_g = _gpu_run(ev["nx"], ev["ny"], 1, ev["dx"], 0.1, ev["dt"], ev["etabar"], ev["t0"], ev["e"],
                            s["nstep"], tab)
        assert tau_g == s["tau"]
        b, tau, sens = _oracle_sensitivity(oracle_mod, ev, s["nstep"], tab)
        assert tau == tau_g
        bar = max(1e-10, 10 * sens)
        e_bs = max(field_errors(fb, a, tau_g).values())
        e_bo = max(field_errors(fb, b, tau).values())
        e_so = max(field_errors(a, b, tau).values())
        print(f"event {i}: L={ev['nx']} dx={ev['dx']:.3f} b={ev['b']:.1f}: frozen after {s['nstep']} steps; "
              f"oracle self-spread {sens:.1e}; batch vs single {e_bs:.1e}, batch vs oracle {e_bo:.1e}, "
              f"single vs oracle {e_so:.1e} (bar {bar:.1e})")
        assert e_bs <= bar and e_bo <= bar and e_so <= bar
    B.close()


def test_serving_mix_with_a_3d_request_against_the_oracle(oracle_mod):
    """A burst of four requests, one of them 3+1D, through the serving loop over
    the FP64 batch engine into a budget that holds three of them: each retired
    result equals the oracle run of that request for the same number of
    steps, and the loop stopped it at its freeze-out step."""
    V = vah()
    S = importlib.import_module("gpu-vah_amd.serve")
    tab = _table()
    e_peak, _ = V.eos_lqcd(0.175 / HBARC)
    shapes = [(40, 36, 1, 0.1), (36, 32, 9, 0.3), (52, 48, 1, 0.1), (44, 44, 1, 0.1)]
    arrivals = S.poisson_arrivals(len(shapes), 1.0e4, np.random.default_rng(11))
    reqs = []
    for i, ((nx, ny, nz, dz), arr) in enumerate(zip(shapes, arrivals)):
        e = gaussian_ic(nx, ny, nz, dz=dz, seed=20 + i, peak=e_peak, floor=1e-4)
        reqs.append(S.Request(rid=i, arrival=float(arr), nx=nx, ny=ny, nz=nz, dx=0.1, dz=dz, dt=0.02, t0=0.5,
                              etabar=0.16, e=e, max_steps=3000))
    sizes = sorted(V.batch_event_bytes(r.nx, r.ny, r.nz) for r in reqs)
    B = V.Batch(budget_bytes=sum(sizes[-3:]), max_events=4, eos_table=tab, freezeout_T_GeV=T_F, steps_per_graph=8)
    out = S.serve(B, reqs, steps_per_poll=8, keep_fields=True)
    B.close()
    assert out["events"] == len(reqs) and out["rejected"] == 0
    assert out["peak_resident"] >= 2
    for r in reqs:
        assert r.state == V.EVENT_FROZEN
        b, tau, fails = _oracle_run(oracle_mod, r.nx, r.ny, r.nz, r.dx, r.dz, r.dt, r.etabar, r.t0, r.e, r.nstep,
                                    tab)
        assert tau == r.tau
        assert table_T(tab, b["e"].max()) * HBARC < T_F
        worst = assert_parity(r.fields, b, tau, label=f"request {r.rid} ({r.nx}x{r.ny}x{r.nz})")
        print(f"request {r.rid}: {r.nx}x{r.ny}x{r.nz} frozen after {r.nstep} steps, worst vs oracle {worst}")


@pytest.mark.parametrize("nz", [1, 4], ids=["2+1D", "3+1D"])
def test_conformal_bjorken_on_the_gpu(nz):
    """Uniform lattice at rest, e0 = e(T0 = 3.05 fm^-1), P_L = e/3, conformal,
    eta/s 0.2, tau 0.5 -> 2.5 fm/c: the HIP path is Heun's method on the
    Bjorken ODE (1e-12) and within 1e-4 of its exact solution; no transverse or
    longitudinal flow and no transverse shear appear."""
    from scipy.integrate import solve_ivp
    from tests.test_oracle_invariants import E0, T_single, heun, rhs
    nx, ny, steps, dt, t0 = 6, 5, 200, 0.01, 0.5
    n = nx * ny * nz
    st, tau = _gpu_run(nx, ny, nz, 0.1, 0.1, dt, 0.2, t0, np.full(n, E0), steps)
    assert tau == pytest.approx(2.5, abs=1e-12)
    e, pl = st["e"], st["q"][4]
    assert np.all(e == e[0]) and np.all(pl == pl[0])
    y = heun([E0, E0 / 3], t0, dt, steps)
    assert e[0] == pytest.approx(y[0], rel=1e-12)
    assert pl[0] == pytest.approx(y[1], rel=1e-12)
    sol = solve_ivp(lambda t, yy: rhs(t, yy, lambda x: (x / 13.8997) ** 0.25), (t0, tau), [E0, E0 / 3],
                    method="DOP853", rtol=1e-13, atol=1e-13)
    assert abs(e[0] / sol.y[0, -1] - 1) < 1e-4
    assert abs(pl[0] / sol.y[1, -1] - 1) < 1e-4
    assert np.max(np.abs(st["u"][1:])) == 0.0
    assert np.max(np.abs(st["q"][5:])) <= 1e-12 * e[0]


@pytest.mark.parametrize("nz", [1, 3], ids=["2+1D", "3+1D"])
def test_uniform_longitudinal_flow_against_the_oracle(oracle_mod, nz):
    """Uniform lattice with uniform flow relative to Bjorken,
    u = (cosh 0.5, 0, 0, sinh 0.5 / tau0): the time derivative of the flow
    drives P_L (the oracle's pairing of velocities in it is pinned against the
    exact solution in tests/test_oracle_pins.py); the HIP path follows the
    oracle within 1e-10 for 200 steps."""
    from tests.test_oracle_invariants import E0
    nx, ny, dt, t0, xi0 = 4, 3, 0.005, 0.5, 0.5
    n = nx * ny * nz
    e = np.full(n, E0)
    u = (np.full(n, np.cosh(xi0)), np.zeros(n), np.zeros(n), np.full(n, np.sinh(xi0) / t0))
    V = vah()
    g = V.Vah(nx, ny, nz, 0.1, 0.1, 0.1, dt, 0.2, t0)
    o = oracle_mod.Oracle(nx, ny, nz, 0.1, 0.1, 0.1, dt, 0.2, t0, threads=1)
    g.set_primitives(e, *u, pl=e / 3)
    o.set_primitives(e, *u, pl=e / 3)
    g.step(200)
    g.synchronize()
    o.step(200)
    a, b = g.state(), o.state()
    assert abs(np.arcsinh(o.time * b["u"][3][0]) - xi0) > 0.1  # the flow relaxed
    print("worst", assert_parity(a, b, o.time, label="uniform longitudinal flow"))
    g.close()
    o.close()
