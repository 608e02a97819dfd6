"""GPU parity at the full 201x201x101 size of the timed 3+1D workload
(MC-Glauber Pb+Pb b = 6 fm, 0.5 fm nucleons, dx = deta = 0.1, tau0 = 0.5,
dtau = 0.02, eta/s = 0.16, the 10^4-node LQCD-like EOS table).

- A few steps against the CPU oracle on the same input, cell by cell: FP64
  within 1e-10 (metric in tests/parity.py), split precision within 1e-4; and
  a few steps from the developed flow of step 60, handed from the GPU to the
  oracle.
- Boost invariance at that size, a property that needs no oracle: an
  eta-independent initial state stays eta-independent, every eta plane bit for
  bit (each plane runs the same arithmetic, whichever eta chunk and ring slot
  it falls in), and each plane equals the 201x201x1 run of the same
  transverse profile (the 3+1D kernels add eta fluxes and eta derivatives that
  are exactly zero here).
"""
import numpy as np
import pytest

from tests.parity import assert_parity, field_errors, vah

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


def test_201x201x101_developed_flow_against_oracle(oracle_mod):
    """The full-size event after 60 GPU steps (tau = 1.7 fm/c, transverse
    velocity up to 0.8 at the fireball edge), handed to the oracle with its
    previous velocity: 3 more steps on both agree per cell within
    max(1e-10, 10 s), s the oracle's own spread under a 1e-15 relative
    perturbation of that state."""
    V = vah()
    table = V.eos_lqcd_table(10000, 1e-8, 1e4)
    e = mc_glauber(V, NZ)
    n = e.size
    print("201x201x101 MC-Glauber, 60 GPU steps, then 3 steps on both", flush=True)
    g = V.Vah(NX, NY, NZ, DX, DX, DZ, DT, ETABAR, T0, eos_table=table)
    g.set_primitives(e, *rest(n), pl=e / 3)
    g.step(60)
    g.synchronize()
    st = g.state()
    vT = float(np.max(np.hypot(st["u"][1], st["u"][2]) / st["u"][0]))
    print(f"max transverse velocity at step 60: {vT:.3f}", flush=True)
    assert vT > 0.5  # developed transverse flow (0.81 at step 60)
    g.step(3)
    g.synchronize()
    a, tau_g = g.state(), g.time
    g.close()

    def oracle_from(q):
        o = oracle_mod.Oracle(NX, NY, NZ, DX, DX, DZ, DT, ETABAR, T0, eos_table=table)
        o.set_state(60, q, st["e"], st["p"], st["u"], st["up"])
        o.step(3)
        b, tau, fails = o.state(), o.time, o.newton_fail
        o.close()
        return b, tau, fails

    b, tau, fails = oracle_from(st["q"])
    assert tau == tau_g and fails == 0
    rng = np.random.default_rng(5)
    b2, _, _ = oracle_from(st["q"] * (1.0 + 1e-15 * rng.standard_normal(st["q"].shape)))
    spread = max(field_errors(b2, b, tau).values())
    bar = max(1e-10, 10 * spread)
    err = field_errors(a, b, tau)
    worst = max(err.items(), key=lambda kv: kv[1])
    print(f"oracle self-spread {spread:.1e}; GPU vs oracle worst {worst[0]} {worst[1]:.1e} (bar {bar:.1e})", flush=True)
    assert worst[1] <= bar


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
