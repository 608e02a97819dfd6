"""GPU parity at the full 201x201x101 size of the timed 3+1D workload
(MC-Glauber Pb+Pb b = 6 fm, 0.5 fm nucleons, dx = deta = 0.1, tau0 = 0.5,
dtau = 0.02, eta/s = 0.16, the 10^4-node LQCD-like EOS table).

- A few steps against the CPU oracle on the same input, cell by cell: FP64
  within 1e-10 (metric in tests/parity.py), split precision within 1e-4.
- Boost invariance at that size, a property that needs no oracle: an
  eta-independent initial state stays eta-independent, every eta plane bit for
  bit (each plane runs the same arithmetic, whichever eta chunk and ring slot
  it falls in), and each plane equals the 201x201x1 run of the same
  transverse profile (the 3+1D kernels add eta fluxes and eta derivatives that
  are exactly zero here).
"""
import numpy as np
import pytest

from tests.parity import assert_parity, vah

pytestmark = pytest.mark.gpu

NX, NY, NZ = 201, 201, 101
DX, DZ, DT, T0, ETABAR = 0.1, 0.1, 0.02, 0.5, 0.16


def mc_glauber(V, nz):
    g = V.glauber(A=208, R=6.62, a=0.546, sigma_nn_mb=70.0, b=6.0, seed=42, e_peak=60.0 / V.HBARC,
                  nucleon_width=0.5, eta_flat=1.5, eta_sigma=1.4, e_floor=1e-3)
    e, _ = V.ic_mc_glauber(NX, NY, nz, DX, DZ, g)
    return e


def rest(n):
    return np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n)


def run_gpu(V, nz, e, steps, table, precision=0):
    n = e.size
    g = V.Vah(NX, NY, nz, DX, DX, DZ, DT, ETABAR, T0, eos_table=table, precision=precision)
    g.set_primitives(e, *rest(n), pl=e / 3)
    g.step(steps)
    g.synchronize()
    a, tau = g.state(), g.time
    g.close()
    return a, tau


def test_201x201x101_against_oracle(oracle_mod):
    """3 steps of the full-size event: FP64 within 1e-10 of the oracle, split
    precision within 1e-4 (e relative, u^mu relative to u^tau)."""
    V = vah()
    steps = 3
    table = V.eos_lqcd_table(10000, 1e-8, 1e4)
    e = mc_glauber(V, NZ)
    n = e.size
    print(f"201x201x101 MC-Glauber, table EOS, {steps} steps", flush=True)
    o = oracle_mod.Oracle(NX, NY, NZ, DX, DX, DZ, DT, ETABAR, T0, eos_table=table)
    o.set_primitives(e, *rest(n), pl=e / 3)
    o.step(steps)
    b, tau = o.state(), o.time
    assert o.newton_fail == 0
    o.close()
    a, tau_g = run_gpu(V, NZ, e, steps, table)
    assert tau_g == tau
    worst = assert_parity(a, b, tau, label="201x201x101 FP64")
    print("fp64 worst field", worst, flush=True)
    del a
    s, _ = run_gpu(V, NZ, e, steps, table, precision=V.FP32_SPLIT)
    err_e = float(np.max(np.abs(s["e"] - b["e"]) / b["e"]))
    err_u = max(float(np.max(np.abs(s["u"][d] - b["u"][d]) / b["u"][0])) for d in range(4))
    print("split max rel error e", err_e, "u", err_u, flush=True)
    assert err_e < 1e-4 and err_u < 1e-4


def test_201x201x101_boost_invariance():
    """20 steps of an eta-independent 201x201x101 state: all eta planes equal
    bit for bit, u^eta stays 0, and every plane matches the 201x201x1 run
    within 1e-12."""
    V = vah()
    steps = 20
    table = V.eos_lqcd_table(10000, 1e-8, 1e4)
    e2 = mc_glauber(V, 1)
    e3 = np.tile(e2, NZ)
    a3, tau3 = run_gpu(V, NZ, e3, steps, table)
    a2, tau2 = run_gpu(V, 1, e2, steps, table)
    assert tau3 == tau2
    n2 = e2.size
    q3 = a3["q"].reshape(-1, NZ, n2)
    u3 = a3["u"].reshape(4, NZ, n2)
    e3 = a3["e"].reshape(NZ, n2)
    assert np.all(e3 == e3[0])
    assert np.all(q3 == q3[:, :1])
    assert np.all(u3 == u3[:, :1])
    assert np.max(np.abs(u3[3])) == 0.0
    mid = NZ // 2
    plane = {"e": e3[mid], "q": q3[:, mid], "u": u3[:, mid]}
    worst = assert_parity(plane, a2, tau2, tol=1e-12, label="eta plane vs 2+1D")
    print("plane vs 2+1D worst field", worst, flush=True)
    assert np.max(np.abs(a2["u"][1])) > 1e-3  # the transverse flow has developed
