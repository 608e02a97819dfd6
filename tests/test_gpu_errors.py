"""GPU tests of the C ABI's error paths (include/vah.h, include/vah_batch.h):
a lattice over one context's 4 GB field-set limit is refused before anything
is allocated, calls out of order return VAH_ERR_STATE, and a NaN cell -- where
the reference prints and exit(-1)s (EnergyMomentumTensor.cpp:193) -- comes back
as VAH_ERR_NAN from vah_synchronize, or stops only its own event in a batch.
The NaN is planted in the initial state from the host; the kernels' Newton
loops end at a NaN at once and their table lookups send it to the
below-table branch, so every wave exits normally."""
import numpy as np
import pytest

from tests.parity import field_errors, gaussian_ic, rest_frame, vah

pytestmark = pytest.mark.gpu


def test_lattice_over_the_field_set_limit_is_refused():
    V = vah()
    with pytest.raises(V.VahError) as exc:
        V.Vah(700, 700, 700, 0.1, 0.1, 0.1, 0.01, 0.2, 0.5)  # 45 GB per 15-field set
    assert exc.value.code == V.VAH_ERR_ARG and "4 GB" in str(exc.value)


def test_calls_out_of_order():
    V = vah()
    nx, ny = 20, 18
    g = V.Vah(nx, ny, 1, 0.1, 0.1, 0.1, 0.01, 0.2, 0.5)
    with pytest.raises(V.VahError) as exc:
        g.step(1)  # no initial state yet
    assert exc.value.code == V.VAH_ERR_STATE
    e = gaussian_ic(nx, ny, peak=30.0)
    g.set_primitives(e, *rest_frame(nx * ny), pl=e / 3)
    with pytest.raises(V.VahError) as exc:
        g.validity()  # needs a completed step
    assert exc.value.code == V.VAH_ERR_STATE
    g.step(1)
    g.synchronize()
    assert np.all(np.isfinite(g.validity()["kn"]))
    # a state uploaded at a later step has no step since: its q^n is unknown
    st = g.state()
    g.set_state(5, st["q"], st["e"], st["p"], st["u"], st["up"])
    with pytest.raises(V.VahError) as exc:
        g.validity()
    assert exc.value.code == V.VAH_ERR_STATE
    g.step(1)
    g.synchronize()
    assert np.all(np.isfinite(g.validity()["kn"]))
    g.close()


def test_nan_cell_is_reported_not_fatal():
    V = vah()
    nx, ny = 24, 20
    e = gaussian_ic(nx, ny, peak=40.0)
    e[(ny // 2) * nx + nx // 2] = np.nan
    g = V.Vah(nx, ny, 1, 0.1, 0.1, 0.1, 0.01, 0.2, 0.5)
    g.set_primitives(e, *rest_frame(nx * ny), pl=e / 3)
    g.step(2)
    with pytest.raises(V.VahError) as exc:
        g.synchronize()
    assert exc.value.code == V.VAH_ERR_NAN
    assert g.stats().nan_cells > 0
    g.synchronize()  # reported once: no new NaN since
    g.close()


def test_nan_event_stops_alone_in_a_batch():
    V = vah()
    ev = dict(nx=30, ny=28, dx=0.1, dt=0.01, t0=0.5, etabar=0.2)
    e_ok = gaussian_ic(30, 28, peak=40.0)
    e_bad = e_ok.copy()
    e_bad[14 * 30 + 15] = np.nan
    B = V.Batch(budget_bytes=1 << 28, max_events=4, freezeout_T_GeV=0.0, steps_per_graph=0)
    args = (ev["nx"], ev["ny"], 1, ev["dx"], ev["dx"], 0.1, ev["dt"], ev["t0"], ev["etabar"])
    bad = B.admit(*args, e_bad, pl=e_bad / 3, max_steps=50)
    ok = B.admit(*args, e_ok, pl=e_ok / 3, max_steps=50)
    B.step(10)
    st = {s["slot"]: s for s in B.poll()}
    assert st[bad]["state"] == V.EVENT_NAN and st[bad]["nstep"] < 10
    assert st[ok]["state"] == V.EVENT_LIVE and st[ok]["nstep"] == 10
    a = B.state(ok)
    B.close()
    g = V.Vah(ev["nx"], ev["ny"], 1, ev["dx"], ev["dx"], 0.1, ev["dt"], ev["etabar"], ev["t0"])
    g.set_primitives(e_ok, *rest_frame(30 * 28), pl=e_ok / 3)
    g.step(10)
    g.synchronize()
    err = max(field_errors(a, g.state(), g.time).values())
    g.close()
    assert err <= 1e-12  # the healthy event is untouched by its neighbour's NaN
