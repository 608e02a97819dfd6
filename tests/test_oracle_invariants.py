"""Whole-step pins of the oracle: the conformal Bjorken solution and lattice
symmetries.  They check the composition of the stage kernels
(FullyDiscreteKurganovTadmorScheme.cpp:738-792) that no reference build can:
a swapped direction, a wrong sign in one axis or a broken eta path breaks one
of them."""
import ctypes
import ctypes.util

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from tests.parity import gaussian_ic, rest_frame

libm = ctypes.CDLL(ctypes.util.find_library("m"))
libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
libm.powf.restype = ctypes.c_float

E0 = 93.21038818333332   # python/conformal_bjorken_test.py:96 (T0 = 3.05 fm^-1)
ETABAR = 0.2


def rtilde(a):
    # R-tilde(a) of the conformal Bjorken AH equations (python/conformal_bjorken_test.py:18-37)
    num = [-6.674731906076046e-6, 0.004617789933500251, 0.7207562721999754, 9.097427250602184,
           -4.475814747302824, -36.37501529319408, 46.868405146729316, -15.833867583743228]
    den = [0.06856675185266, 2.9181587012768597, 11.951184087839218, -29.708257843442173,
           -2.618233802059826, 34.646239784689065, -19.62596366454439, 2.374808442453899]
    return np.polyval(num[::-1], a) / np.polyval(den[::-1], a)


def rhs(t, y, T_of):
    """python/conformal_bjorken_test.py:46-68: d e/dt = -(e + P_L)/t,
    d P_L/dt = -(P_L - e/3) T/(5 eta/s) - (3 P_L - R-tilde(P_L/e) e)/t."""
    e, pl = y
    T = T_of(e)
    return np.array([-(e + pl) / t, -(pl - e / 3) * T / 5 / ETABAR - (3 * pl - rtilde(pl / e) * e) / t])


def T_single(e):
    return float(libm.powf(ctypes.c_float(e / 13.8997), ctypes.c_float(0.25)))


def heun(y0, t0, dt, n):
    y, t = np.array(y0, dtype=float), t0
    for _ in range(n):
        y1 = y + dt * rhs(t, y, T_single)
        y2 = y1 + dt * rhs(t + dt, y1, T_single)
        y = (y2 + y) / 2
        t = t0 + (_ + 1) * dt
    return y


@pytest.fixture(scope="module")
def bjorken_run(oracle_mod):
    nx, ny, steps, dt, t0 = 6, 5, 200, 0.01, 0.5
    n = nx * ny
    e = np.full(n, E0)
    o = oracle_mod.Oracle(nx, ny, 1, 0.1, 0.1, 0.1, dt, ETABAR, t0)
    o.set_primitives(e, *rest_frame(n), pl=e / 3)
    o.step(steps)
    st = o.state()
    return st, o.time, steps, dt, t0


def test_bjorken_uniform_lattice_is_heun_on_the_ode(bjorken_run):
    st, tau, steps, dt, t0 = bjorken_run
    assert tau == pytest.approx(2.5, abs=1e-12)
    e, pl = st["e"], st["q"][4]
    assert np.all(e == e[0]) and np.all(pl == pl[0])  # uniform stays uniform
    y = heun([E0, E0 / 3], t0, dt, steps)
    assert e[0] == pytest.approx(y[0], rel=1e-12)
    assert pl[0] == pytest.approx(y[1], rel=1e-12)
    # no transverse flow or shear appears
    assert np.max(np.abs(st["u"][1:])) == 0.0
    assert np.max(np.abs(st["q"][5:])) <= 1e-12 * e[0]


def test_bjorken_converges_to_the_exact_solution(bjorken_run):
    st, tau, steps, dt, t0 = bjorken_run
    sol = solve_ivp(lambda t, y: rhs(t, y, lambda e: (e / 13.8997) ** 0.25), (t0, tau), [E0, E0 / 3],
                    method="DOP853", rtol=1e-13, atol=1e-13)
    e_ex, pl_ex = sol.y[:, -1]
    # second-order scheme, dtau = 0.01: relative error O(1e-5)
    assert abs(st["e"][0] / e_ex - 1) < 1e-4
    assert abs(st["q"][4][0] / pl_ex - 1) < 1e-4


def test_reflection_symmetry(oracle_mod):
    """x -> -x is an exact symmetry of the KT fluxes, sources and recovery,
    but not of the reference's regulator: it takes the *signed* maximum of
    u.pi and z.pi ratios (FullyDiscreteKurganovTadmorScheme.cpp:631-650), and
    u_mu pi^{mu x} flips sign under the mirror.  Before that branch decides
    anything the evolution is mirror-exact; afterwards the mirror images drift
    apart only at the regulator's (1 - tanh(rho)/rho ~ rho^2/3) scale."""
    nx, ny = 25, 19
    e = gaussian_ic(nx, ny, sx=1.3, sy=1.8)
    o = oracle_mod.Oracle(nx, ny, 1, 0.1, 0.1, 0.1, 0.01, 0.2, 0.5)
    o.set_primitives(e, *rest_frame(nx * ny), pl=e / 3)

    def views(st):
        return (st["e"].reshape(ny, nx), st["u"][1].reshape(ny, nx), st["u"][2].reshape(ny, nx))

    o.step(12)
    E, UX, UY = views(o.state())
    assert np.max(np.abs(UX)) > 1e-3
    np.testing.assert_array_equal(E, E[:, ::-1])
    np.testing.assert_array_equal(E, E[::-1, :])
    np.testing.assert_array_equal(UX, -UX[:, ::-1])
    np.testing.assert_array_equal(UY, -UY[::-1, :])
    o.step(28)
    E, UX, UY = views(o.state())
    np.testing.assert_allclose(E, E[:, ::-1], rtol=1e-5)
    np.testing.assert_allclose(UX, -UX[:, ::-1], rtol=1e-4, atol=1e-9)


def test_xy_exchange_symmetry(oracle_mod):
    n1 = 21
    e = gaussian_ic(n1, n1, sx=1.5, sy=1.5, seed=None)
    # break reflection symmetry but keep x <-> y: add a lump on the diagonal
    x = (np.arange(n1) - 10) * 0.1
    X, Y = np.meshgrid(x, x, indexing="xy")
    e = e + (20.0 * np.exp(-((X - 0.4) ** 2 + (Y - 0.4) ** 2) / 0.3)).ravel()
    o = oracle_mod.Oracle(n1, n1, 1, 0.1, 0.1, 0.1, 0.01, 0.2, 0.5)
    o.set_primitives(e, *rest_frame(n1 * n1), pl=e / 3)
    o.step(30)
    st = o.state()
    E = st["e"].reshape(n1, n1)
    np.testing.assert_allclose(E, E.T, rtol=1e-12)
    UX = st["u"][1].reshape(n1, n1)
    UY = st["u"][2].reshape(n1, n1)
    np.testing.assert_allclose(UX, UY.T, rtol=1e-10, atol=1e-14)
    PIXX = st["q"][9].reshape(n1, n1)
    PIYY = st["q"][12].reshape(n1, n1)
    np.testing.assert_allclose(PIXX, PIYY.T, rtol=1e-9, atol=1e-12 * E.max())


def test_boost_invariance(oracle_mod):
    """An eta-independent state on an nz = 6 lattice evolves like the nz = 1 one."""
    nx, ny = 14, 12
    e2 = gaussian_ic(nx, ny, sx=0.6, sy=0.5)
    o2 = oracle_mod.Oracle(nx, ny, 1, 0.1, 0.1, 0.1, 0.01, 0.2, 0.5)
    o2.set_primitives(e2, *rest_frame(nx * ny), pl=e2 / 3)
    o2.step(15)
    e3 = np.tile(e2, 6)
    o3 = oracle_mod.Oracle(nx, ny, 6, 0.1, 0.1, 0.2, 0.01, 0.2, 0.5)
    o3.set_primitives(e3, *rest_frame(nx * ny * 6), pl=e3 / 3)
    o3.step(15)
    a, b = o2.state(), o3.state()
    for k in range(6):
        sl = slice(k * nx * ny, (k + 1) * nx * ny)
        np.testing.assert_array_equal(b["e"][sl], a["e"])
        np.testing.assert_array_equal(b["q"][:, sl], a["q"])


def test_validity_diagnostics_in_bjorken_flow(oracle_mod):
    """Bjorken flow (uniform lattice at rest): theta = 1/tau, the transverse
    expansion theta_T = 2 theta/3 + A vanishes (A = tau^2 sigma^etaeta =
    -2/(3 tau)) and zDzu = 1/tau, so Kn_T = 0 and Kn_L = Kn = 5 (eta/s)/(tau T).
    The diagnostics use the time after the last step, t_n + dt (rungeKutta2
    advances t before its checkValidity call), and T of the new e.  pi_perp
    stays 0, so R_pi = 0."""
    nx, ny, dt, t0 = 5, 4, 0.01, 0.5
    n = nx * ny
    e = np.full(n, E0)
    o = oracle_mod.Oracle(nx, ny, 1, 0.1, 0.1, 0.1, dt, ETABAR, t0)
    o.set_primitives(e, *rest_frame(n), pl=e / 3)
    with pytest.raises(RuntimeError):
        o.validity()
    o.step(7)
    v = o.validity()
    tn = (t0 + 6 * dt) + dt
    T = T_single(o.state()["e"][0])
    assert np.all(np.abs(v["kn_T"]) < 1e-12)
    assert np.allclose(v["kn_L"], 5 * ETABAR / (tn * T), rtol=1e-12)
    assert np.allclose(v["kn"], 5 * ETABAR / (tn * T), rtol=1e-12)
    assert np.all(v["r_pi"] < 1e-12)
    o.close()
