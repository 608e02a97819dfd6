"""GPU tests of the batch engine (include/vah_batch.h): each resident event
evolves exactly as it would alone -- against the single-lattice path, the
CPU oracle, other kernel variants, graph replays and other admission
orders -- and the engine's admission, freeze-out and retirement rules."""
import numpy as np
import pytest

from tests.parity import TOL_FP64, assert_parity, field_errors, gaussian_ic, rest_frame, vah

pytestmark = pytest.mark.gpu

HBARC = 0.197326938


def table():
    return vah().eos_lqcd_table(10000, 1e-8, 1e4)


def peak_e_for_T(T_GeV):
    e, _ = vah().eos_lqcd(T_GeV / HBARC)
    return e


def events_mixed():
    """Two 2+1D events of different size, spacing and step, one 3+1D."""
    return [
        dict(nx=50, ny=45, nz=1, dx=0.1, dy=0.1, dz=0.1, dt=0.02, t0=0.5, etabar=0.16,
             e=gaussian_ic(50, 45, seed=1, peak=60.0)),
        dict(nx=37, ny=61, nz=1, dx=0.12, dy=0.12, dz=0.1, dt=0.024, t0=0.6, etabar=0.2,
             e=gaussian_ic(37, 61, dx=0.12, seed=2, peak=40.0, tilt=0.2)),
        dict(nx=24, ny=20, nz=12, dx=0.1, dy=0.1, dz=0.3, dt=0.02, t0=0.5, etabar=0.16,
             e=gaussian_ic(24, 20, 12, dz=0.3, seed=3, peak=60.0)),
    ]


def admit(B, ev, **kw):
    return B.admit(ev["nx"], ev["ny"], ev["nz"], ev["dx"], ev["dy"], ev["dz"], ev["dt"], ev["t0"], ev["etabar"],
                   ev["e"], pl=ev["e"] / 3, **kw)


def single(ev, steps, eos_table, precision=0):
    V = vah()
    g = V.Vah(ev["nx"], ev["ny"], ev["nz"], ev["dx"], ev["dy"], ev["dz"], ev["dt"], ev["etabar"], ev["t0"],
              eos_table=eos_table, precision=precision)
    n = ev["nx"] * ev["ny"] * ev["nz"]
    g.set_primitives(ev["e"], *rest_frame(n), pl=ev["e"] / 3)
    g.step(steps)
    g.synchronize()
    st = g.state()
    g.close()
    return st


def same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in ("q", "e", "p", "u", "up"))


def worst(a, b, tau):
    """Worst per-cell error over all fields (the parity metric of tests/parity.py)."""
    return max(field_errors(a, b, tau).values())


def run_until_stopped(B, slots, chunk=8, limit=10000):
    done = {}
    for _ in range(limit // chunk):
        B.step(chunk)
        for s in B.poll():
            if s["slot"] in slots and s["state"] != vah().EVENT_LIVE and s["slot"] not in done:
                done[s["slot"]] = s
        if len(done) == len(slots):
            return done
    raise AssertionError("events did not stop")


def test_batch_matches_single_lattice_path_and_oracle(oracle_mod):
    """FP64: every event of a mixed batch equals the single-lattice run of the
    same event (same per-cell arithmetic), and the oracle within 1e-10."""
    V = vah()
    tab = table()
    evs = events_mixed()
    steps = 20
    B = V.Batch(budget_bytes=1 << 30, max_events=8, eos_table=tab, freezeout_T_GeV=0.0, steps_per_graph=0)
    slots = [admit(B, ev, max_steps=steps, tag=i) for i, ev in enumerate(evs)]
    done = run_until_stopped(B, slots, chunk=5)
    for i, (ev, s) in enumerate(zip(evs, slots)):
        assert done[s]["state"] == V.EVENT_STEP_LIMIT and done[s]["nstep"] == steps and done[s]["tag"] == i
        a = B.state(s)
        b = single(ev, steps, tab)
        err = worst(a, b, done[s]["tau"])
        print(f"event {i}: batch vs single-lattice path worst {err:.3e}, identical={same(a, b)}")
        assert err <= 1e-12
    # the first event against the CPU oracle
    ev = evs[0]
    n = ev["nx"] * ev["ny"]
    o = oracle_mod.Oracle(ev["nx"], ev["ny"], 1, ev["dx"], ev["dy"], ev["dz"], ev["dt"], ev["etabar"], ev["t0"],
                          eos_table=tab)
    o.set_primitives(ev["e"], *rest_frame(n), pl=ev["e"] / 3)
    o.step(steps)
    tau = o.time
    assert done[slots[0]]["tau"] == tau
    assert_parity(B.state(slots[0]), o.state(), tau, TOL_FP64, label="batch vs oracle")
    o.close()
    B.close()


def test_batch_with_the_reference_eos_fits():
    """The Wuppertal-Budapest fits of the reference's non-conformal EOS
    (EquationOfState.cpp:31-142) in the batch engine: a 2+1D and a 3+1D event
    equal their single-lattice runs with the same EOS to 1e-12 per cell."""
    V = vah()
    evs = [dict(nx=30, ny=26, nz=1, dx=0.1, dy=0.1, dz=0.1, dt=0.02, t0=0.5, etabar=0.16,
                e=gaussian_ic(30, 26, seed=21, peak=60.0, sx=1.2, sy=1.5)),
           dict(nx=18, ny=16, nz=7, dx=0.1, dy=0.1, dz=0.3, dt=0.02, t0=0.5, etabar=0.16,
                e=gaussian_ic(18, 16, 7, dz=0.3, seed=22, peak=60.0, sx=0.8, sy=0.9))]
    steps = 15
    B = V.Batch(budget_bytes=1 << 30, max_events=4, freezeout_T_GeV=0.0, steps_per_graph=4, eos_kind=V.EOS_WB)
    slots = [admit(B, ev, max_steps=steps, tag=i) for i, ev in enumerate(evs)]
    done = run_until_stopped(B, slots, chunk=5)
    for ev, s in zip(evs, slots):
        g = V.Vah(ev["nx"], ev["ny"], ev["nz"], ev["dx"], ev["dy"], ev["dz"], ev["dt"], ev["etabar"], ev["t0"],
                  eos_kind=V.EOS_WB)
        n = ev["nx"] * ev["ny"] * ev["nz"]
        g.set_primitives(ev["e"], *rest_frame(n), pl=ev["e"] / 3)
        g.step(steps)
        g.synchronize()
        b = g.state()
        g.close()
        err = worst(B.state(s), b, done[s]["tau"])
        print(f"WB EOS event {ev['nx']}x{ev['ny']}x{ev['nz']}: batch vs single {err:.2e}")
        assert err <= 1e-12
    B.close()


def test_degenerate_events_in_a_batch(oracle_mod):
    """Events one or two cells wide, down to a single cell, resident beside a
    normal event: each within 1e-10 of the oracle run of the same event (the
    batch's tile lists and persistent launches hold tiles with a handful of
    live lanes)."""
    V = vah()
    shapes = [(1, 1, 1), (3, 4, 1), (5, 1, 3), (2, 2, 2), (40, 36, 1)]
    evs = [dict(nx=nx, ny=ny, nz=nz, dx=0.1, dy=0.1, dz=0.1, dt=0.01, t0=0.5, etabar=0.2,
                e=gaussian_ic(nx, ny, nz, sx=0.3 if nx * ny < 100 else 1.5, sy=0.4 if nx * ny < 100 else 1.8,
                              seed=9 + i))
           for i, (nx, ny, nz) in enumerate(shapes)]
    steps = 20
    B = V.Batch(budget_bytes=1 << 30, max_events=8, freezeout_T_GeV=0.0, steps_per_graph=4)
    slots = [admit(B, ev, max_steps=steps, tag=i) for i, ev in enumerate(evs)]
    done = run_until_stopped(B, slots, chunk=5)
    for i, (ev, s) in enumerate(zip(evs, slots)):
        assert done[s]["nstep"] == steps
        n = ev["nx"] * ev["ny"] * ev["nz"]
        o = oracle_mod.Oracle(ev["nx"], ev["ny"], ev["nz"], ev["dx"], ev["dy"], ev["dz"], ev["dt"], ev["etabar"],
                              ev["t0"])
        o.set_primitives(ev["e"], *rest_frame(n), pl=ev["e"] / 3)
        o.step(steps)
        assert done[s]["tau"] == o.time
        assert_parity(B.state(s), o.state(), o.time, TOL_FP64, label=f"batch event {shapes[i]}")
        o.close()
    B.close()


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32_split"])
def test_event_alone_equals_event_in_the_mix(precision):
    """One event run alone and run among others that are admitted and retired
    around it (continuous batching) ends bit-identical."""
    V = vah()
    tab = table()
    evs = events_mixed()
    target = dict(nx=60, ny=52, nz=1, dx=0.1, dy=0.1, dz=0.1, dt=0.02, t0=0.5, etabar=0.16,
                  e=gaussian_ic(60, 52, seed=11, peak=50.0))
    steps = 45
    kw = dict(budget_bytes=1 << 30, max_events=8, eos_table=tab, precision=precision, freezeout_T_GeV=0.0)
    A = V.Batch(steps_per_graph=4, **kw)
    sa = admit(A, target, max_steps=steps)
    run_until_stopped(A, [sa], chunk=6)
    alone = A.state(sa)
    A.close()

    M = V.Batch(steps_per_graph=4, **kw)
    s0 = admit(M, evs[0], max_steps=1000)
    s2 = admit(M, evs[2], max_steps=1000)
    M.step(7)
    st = admit(M, target, max_steps=steps)
    M.step(13)
    M.retire(s0)
    s1 = admit(M, evs[1], max_steps=1000)
    M.step(5)
    assert {s["slot"] for s in M.poll()} == {s1, s2, st}
    done = run_until_stopped(M, [st], chunk=6)
    assert done[st]["nstep"] == steps
    mixed = M.state(st)
    print("alone vs mix worst", worst(mixed, alone, done[st]["tau"]))
    assert same(mixed, alone)
    M.close()


def test_kernel_variants_compute_the_same_fields():
    """Every compiled tile shape / occupancy / eta-chunk variant computes the
    same per-cell arithmetic (up to the compiler's contraction choices in each
    instantiation), so the autotuner can time real steps."""
    V = vah()
    tab = table()
    evs = events_mixed()
    ref = None
    for v in range(len(V.batch_variants())):
        B = V.Batch(budget_bytes=1 << 30, max_events=4, eos_table=tab, freezeout_T_GeV=0.0, steps_per_graph=0,
                    variant=v)
        slots = [admit(B, ev) for ev in evs]
        B.step(6)
        states = [B.state(s) for s in slots]
        taus = [st["tau"] for st in B.poll()]
        B.close()
        if ref is None:
            ref = states
            continue
        for a, b, tau in zip(states, ref, taus):
            err = worst(a, b, tau)
            print(f"variant {v}: worst {err:.3e}, identical={same(a, b)}")
            assert err <= 1e-12, f"variant {v} differs"


def test_graph_replays_match_plain_launches():
    """Steps captured in a hipGraph, replayed several times with host work and
    synchronization in between, equal the same steps launched one by one."""
    V = vah()
    tab = table()
    evs = events_mixed()
    kw = dict(budget_bytes=1 << 30, max_events=4, eos_table=tab, freezeout_T_GeV=0.0)
    G = V.Batch(steps_per_graph=3, **kw)
    E = V.Batch(steps_per_graph=0, **kw)
    sg = [admit(G, ev) for ev in evs]
    se = [admit(E, ev) for ev in evs]
    for rep in range(4):
        G.step(3)
        E.step(3)
        pg, pe = G.poll(), E.poll()
        assert [s["nstep"] for s in pg] == [s["nstep"] for s in pe] == [3 * (rep + 1)] * 3
        for a, b in zip(sg, se):
            assert same(G.state(a), E.state(b)), f"replay {rep}"
    G.close()
    E.close()


def test_freezeout_stops_the_event_at_the_first_cool_step():
    """An event stops after the first step whose hottest cell is below
    T_f = 150 MeV, and holds that step's fields."""
    V = vah()
    tab = table()
    ev = dict(nx=40, ny=40, nz=1, dx=0.1, dy=0.1, dz=0.1, dt=0.02, t0=0.5, etabar=0.16,
              e=gaussian_ic(40, 40, peak=peak_e_for_T(0.170), floor=1e-4))
    B = V.Batch(budget_bytes=1 << 28, max_events=2, eos_table=tab, freezeout_T_GeV=0.150, steps_per_graph=8)
    s = admit(B, ev, max_steps=5000)
    done = run_until_stopped(B, [s], chunk=8, limit=5000)
    st = done[s]
    nf = st["nstep"]
    assert st["state"] == V.EVENT_FROZEN and 1 < nf < 5000
    assert st["max_T"] * HBARC < 0.150
    frozen = B.state(s)
    B.step(16)  # a stopped event does not move
    assert same(B.state(s), frozen) and B.poll()[0]["nstep"] == nf
    # the single-lattice path: hotter than T_f after nf - 1 steps, cooler after nf
    g = V.Vah(ev["nx"], ev["ny"], 1, ev["dx"], ev["dy"], ev["dz"], ev["dt"], ev["etabar"], ev["t0"], eos_table=tab)
    g.set_primitives(ev["e"], *rest_frame(ev["nx"] * ev["ny"]), pl=ev["e"] / 3)
    g.step(nf - 1)
    assert g.stats().max_T * HBARC >= 0.150
    g.step(1)
    assert g.stats().max_T * HBARC < 0.150
    assert worst(frozen, g.state(), st["tau"]) <= 1e-12
    print("frozen after", nf, "steps at tau", st["tau"])
    g.close()
    B.close()


def test_admission_budget_slots_and_retire():
    V = vah()
    one = V.batch_event_bytes(60, 60, 1)
    B = V.Batch(budget_bytes=2 * one + 4096, max_events=3, freezeout_T_GeV=0.0)
    ev = dict(nx=60, ny=60, nz=1, dx=0.1, dy=0.1, dz=0.1, dt=0.01, t0=0.5, etabar=0.2,
              e=gaussian_ic(60, 60, peak=30.0))
    a = admit(B, ev)
    b = admit(B, ev)
    assert B.memory()[0] == 2 * one
    assert admit(B, ev, raise_on_full=False) is None  # budget
    with pytest.raises(V.VahError) as exc:
        admit(B, ev)
    assert exc.value.code == V.VAH_ERR_BUDGET
    B.step(2)
    B.retire(a)
    assert B.memory()[0] == one
    c = admit(B, ev)  # reuses the freed region
    small = dict(ev, nx=10, ny=10, e=gaussian_ic(10, 10, peak=30.0))
    assert admit(B, small, raise_on_full=False) is None  # 4 KB left
    B.step(2)
    st = {s["slot"]: s["nstep"] for s in B.poll()}
    assert st[b] == 4 and st[c] == 2
    B.close()
    F = V.Batch(budget_bytes=1 << 26, max_events=2, freezeout_T_GeV=0.0)
    admit(F, small)
    admit(F, small)
    with pytest.raises(V.VahError) as exc:
        admit(F, small)
    assert exc.value.code == V.VAH_ERR_FULL
    F.close()
