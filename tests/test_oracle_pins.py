"""Pins of the oracle parts that no reference build reaches, by what the
mathematics fixes (no formula of the oracle is retyped here):

1. The regulator's signed ratios (FullyDiscreteKurganovTadmorScheme.cpp:597-665):
   hand-built pi^{mu nu} whose rho follows in closed form -- a pure trace at
   rest (trace and u.pi ratios equal +1, so rho = 1 whatever the size), its
   mirror image (ratios -1: a signed maximum ignores them, an absolute value
   would not), an eta-eta component whose z.pi ratio is 1/tau, and a
   transverse traceless pi seen from a moving cell (u.pi ratio |u^x|/sqrt 2 on
   one side of the flow only).

2. The Heun composition with flow (rungeKutta2, :738-792): which velocity
   pair feeds the source terms' time derivative in each stage.  A uniform
   lattice with uniform longitudinal flow relative to Bjorken,
   u = (cosh xi, 0, 0, sinh xi / tau), has no fluxes; its exact evolution
   follows from energy-momentum conservation in Milne coordinates,
       d T^tautau/dtau = -(T^tautau + tau^2 T^etaeta)/tau,
       d T^taueta/dtau = -3 T^taueta / tau,
   and the anisotropic P_L relaxation equation
       u^tau dP_L/dtau = -(P_L - P)/tau_pi + (R~(P_L/e) e - 3 P_L) theta,
       theta = d u^tau/dtau + u^tau/tau
   (theta_T = 0, zDzu = theta, pi_perp = 0 for this flow).  Here theta
   contains the time derivative of the flow that the scheme differences from
   the stored velocities, so the pairing decides whether the lattice
   converges to this solution.  The test also plants the two plausible
   mistakes in the oracle and checks that they fail."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

from tests.test_oracle_invariants import E0, ETABAR, rtilde


def p(a):
    import ctypes
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ---------------------------------------------------------------- regulator
def regulate(oracle_mod, t, q, e, pl_p, u):
    """One cell through the oracle's regulator; returns the regulated q."""
    q = np.array(q, dtype=float)
    ee, pp = np.array([e]), np.array([pl_p])
    oracle_mod.lib().orc_regulate_cell(t, p(q), p(ee), p(pp), p(np.asarray(u, dtype=float)))
    return q


def magnitude_ratio(oracle_mod, e, pl, pipi_abs):
    """|pi| / sqrt(e^2 + P_L^2 + 2 P-hat_T^2) with P-hat_T = (e - P_L)/2 for e = 3p."""
    pt = (e - pl) / 2
    assert oracle_mod.lib().orc_pt_hat(e, e / 3, pl) == pytest.approx(pt, rel=1e-13)
    return np.sqrt(pipi_abs) / np.sqrt(e * e + pl * pl + 2 * pt * pt)


def fac(rho):
    return 1.0 if rho < 1e-7 else np.tanh(rho) / rho


E, PL = 10.0, 10.0 / 3


@pytest.mark.parametrize("c", [1e-6, 0.3, 5.0, 60.0])
def test_regulator_positive_trace_at_rest(oracle_mod, c):
    """pi^tautau = c > 0 alone, cell at rest: g.pi = c and (u.pi)^tau = c, each
    divided by |pi| = c gives +1, so rho = max(1, c/|T|) even for tiny c."""
    t = 0.7
    q = np.zeros(15)
    q[4], q[5] = PL, c
    out = regulate(oracle_mod, t, q, E, E / 3, [1, 0, 0, 0])
    rho = max(1.0, magnitude_ratio(oracle_mod, E, PL, c * c))
    assert out[5] == pytest.approx(c * fac(rho), rel=1e-14)
    if c < 1:
        assert out[5] == pytest.approx(c * np.tanh(1.0), rel=1e-14)


@pytest.mark.parametrize("c", [1e-6, 0.3, 5.0, 60.0])
def test_regulator_negative_trace_is_not_picked(oracle_mod, c):
    """The mirror image pi^tautau = -c: both ratios are -1, which the signed
    maximum passes over; rho is the magnitude ratio alone (an absolute value
    would give rho >= 1 and shrink small pi by tanh 1)."""
    t = 0.7
    q = np.zeros(15)
    q[4], q[5] = PL, -c
    out = regulate(oracle_mod, t, q, E, E / 3, [1, 0, 0, 0])
    rho = magnitude_ratio(oracle_mod, E, PL, c * c)
    assert out[5] == pytest.approx(-c * fac(rho), rel=1e-14)
    if c < 1e-3:
        assert out[5] == -c  # rho < 1e-7: untouched


@pytest.mark.parametrize("t", [0.5, 2.0])
@pytest.mark.parametrize("sign", [1, -1])
def test_regulator_eta_component_scales_with_tau(oracle_mod, t, sign):
    """pi^etaeta = -sign c / tau^2 at rest: |pi| = c, g.pi = sign c (ratio
    sign), (z.pi)^eta = z_eta pi^etaeta with z = (0,0,0,1/tau) gives
    sign/tau.  So rho = max(1, 1/t) for sign = +1 (the tau^2 of the metric and
    the 1/tau of z both enter), and the magnitude ratio for sign = -1."""
    c = 1e-3
    q = np.zeros(15)
    q[4], q[14] = PL, -sign * c / t ** 2
    out = regulate(oracle_mod, t, q, E, E / 3, [1, 0, 0, 0])
    rho = max(1.0, 1.0 / t) if sign > 0 else magnitude_ratio(oracle_mod, E, PL, c * c)
    assert out[14] == pytest.approx(q[14] * fac(rho), rel=1e-13)


@pytest.mark.parametrize("ux", [0.8, -0.8])
def test_regulator_u_ratio_with_flow(oracle_mod, ux):
    """Traceless pi^xx = -pi^yy = m in a cell moving along x: |pi| = sqrt 2 m and
    (u.pi)^x = u_x pi^xx = -u^x m, ratio -u^x/sqrt 2.  Flow towards -x makes
    that ratio positive (rho = |u^x|/sqrt 2 when it exceeds the magnitude
    ratio); flow towards +x leaves the magnitude ratio alone."""
    m = 1e-4
    ut = np.sqrt(1 + ux * ux)
    q = np.zeros(15)
    q[4], q[9], q[12] = PL, m, -m
    out = regulate(oracle_mod, 1.0, q, E, E / 3, [ut, ux, 0, 0])
    mag = magnitude_ratio(oracle_mod, E, PL, 2 * m * m)
    rho = max(mag, -ux / np.sqrt(2))
    assert out[9] == pytest.approx(m * fac(rho), rel=1e-13)
    assert out[12] == pytest.approx(-m * fac(rho), rel=1e-13)
    if ux < 0:
        assert out[9] == pytest.approx(m * np.tanh(0.8 / np.sqrt(2)) / (0.8 / np.sqrt(2)), rel=1e-13)


# ------------------------------------------------------- Heun composition
def longitudinal_flow_rhs(t, y):
    """Exact equations of uniform longitudinal flow in the conformal VAH
    theory (see the module docstring); y = (T^tautau, T^taueta, P_L)."""
    ttt, ttn, pl = y
    A = t * ttn / (ttt + pl)            # tanh xi
    xi = np.arctanh(A)
    ch, sh = np.cosh(xi), np.sinh(xi)
    e = (ttt + pl) / ch ** 2 - pl
    T = (e / 13.8997) ** 0.25
    dttt = -(ttt + e * sh * sh + pl * ch * ch) / t
    dttn = -3 * ttn / t
    # d xi/dtau = (al - be dP_L/dtau) / (1 - A^2), from A = tau T^taueta / (T^tautau + P_L)
    al = (ttn + t * dttn) / (ttt + pl) - t * ttn * dttt / (ttt + pl) ** 2
    be = t * ttn / (ttt + pl) ** 2
    X = -(pl - e / 3) * T / (5 * ETABAR)
    Y = rtilde(pl / e) * e - 3 * pl
    k = 1 / (1 - A * A)
    # u^tau dP_L/dtau = X + Y theta, theta = sinh xi xi' + cosh xi / tau: linear in dP_L/dtau
    dpl = (X + Y * (sh * al * k + ch / t)) / (ch + Y * sh * be * k)
    return [dttt, dttn, dpl]


def lattice_run(oracle_mod, dt, t1, xi0, mode=0):
    t0, n = 0.5, 4
    o = oracle_mod.Oracle(2, 2, 1, 0.1, 0.1, 0.1, dt, ETABAR, t0, threads=1)
    o.set_pairing(mode)
    e = np.full(n, E0)
    o.set_primitives(e, np.full(n, np.cosh(xi0)), np.zeros(n), np.zeros(n), np.full(n, np.sinh(xi0) / t0),
                     pl=e / 3)
    o.step(int(round((t1 - t0) / dt)))
    st, tau = o.state(), o.time
    o.close()
    assert np.all(st["e"] == st["e"][0])  # uniform stays uniform
    return st, tau


def exact(tau, xi0):
    t0 = 0.5
    ch, sh = np.cosh(xi0), np.sinh(xi0)
    y0 = [E0 * ch * ch + E0 / 3 * sh * sh, (E0 + E0 / 3) * ch * sh / t0, E0 / 3]
    sol = solve_ivp(longitudinal_flow_rhs, (t0, tau), y0, method="DOP853", rtol=1e-12, atol=1e-12)
    ttt, ttn, pl = sol.y[:, -1]
    xi = np.arctanh(tau * ttn / (ttt + pl))
    return (ttt + pl) / np.cosh(xi) ** 2 - pl, pl, xi


def errors(oracle_mod, dt, mode=0, xi0=0.5, t1=1.5):
    st, tau = lattice_run(oracle_mod, dt, t1, xi0, mode)
    e, pl, xi = exact(tau, xi0)
    return (abs(st["e"][0] / e - 1), abs(st["q"][4][0] / pl - 1), abs(np.arcsinh(tau * st["u"][3][0]) - xi), xi)


def test_uniform_longitudinal_flow_converges_to_the_exact_solution(oracle_mod):
    """The reference's pairing converges (first order: the scheme differences
    the flow's time derivative backward in each stage) to the exact
    solution, whose rapidity relaxes from 0.5 to ~0.24 by tau = 1.5 fm/c."""
    errs = [errors(oracle_mod, dt) for dt in (0.01, 0.005, 0.0025)]
    xi_end = errs[0][3]
    assert 0.2 < xi_end < 0.3
    for a, b in zip(errs, errs[1:]):
        for k in range(2):  # e, P_L
            assert 1.8 < a[k] / b[k] < 2.3
    e_err, pl_err, xi_err, _ = errs[-1]
    assert e_err < 3e-4 and pl_err < 2e-3 and xi_err < 2e-5


@pytest.mark.parametrize("mode", [1, 2], ids=["stage2_u_star_with_u_prev", "stage1_without_time_derivative"])
def test_planted_pairing_mistakes_fail_the_flow_pin(oracle_mod, mode):
    """Planted in the oracle: u* paired with u^{n-1} in stage 2 (twice the
    time derivative), or none in stage 1.  Both leave an O(1e-2) error in e
    and P_L that does not shrink with dtau, which the pin above rejects."""
    a = errors(oracle_mod, 0.005, mode)
    b = errors(oracle_mod, 0.0025, mode)
    assert a[0] > 1e-2 and b[0] > 1e-2
    assert a[0] / b[0] < 1.2  # no convergence


# ------------------------------------------------------ validity diagnostics
def test_validity_on_uniform_longitudinal_flow(oracle_mod):
    """checkValidity's Knudsen numbers (HydrodynamicValidity.cpp:195-247) with
    u^eta != 0: for uniform longitudinal flow theta = cosh xi / tau +
    sinh xi dxi/dtau, the transverse expansion theta_T vanishes and
    zDzu = theta.  The maps converge (first order, like the lattice's time
    derivative of u) to 5 (eta/s) theta / T from the exact solution, kn_T to 0."""
    rows = []
    for dt in (0.01, 0.005, 0.0025):
        t0, n = 0.5, 4
        o = oracle_mod.Oracle(2, 2, 1, 0.1, 0.1, 0.1, dt, ETABAR, t0, threads=1)
        e = np.full(n, E0)
        o.set_primitives(e, np.full(n, np.cosh(0.5)), np.zeros(n), np.zeros(n), np.full(n, np.sinh(0.5) / t0),
                         pl=e / 3)
        o.step(int(round(0.5 / dt)))
        v = o.validity()
        T = (o.state()["e"][0] / 13.8997) ** 0.25
        tau = o.time
        o.close()
        h = 1e-5
        _, _, xi = exact(tau, 0.5)
        dxi = (exact(tau + h, 0.5)[2] - exact(tau - h, 0.5)[2]) / (2 * h)
        kn_exact = 5 * ETABAR * (np.cosh(xi) / tau + np.sinh(xi) * dxi) / T
        rows.append((abs(v["kn"][0] / kn_exact - 1), v["kn_T"][0] / v["kn"][0], abs(v["kn_L"][0] / v["kn"][0] - 1)))
    for a, b in zip(rows, rows[1:]):
        for k in range(3):
            assert 1.8 < a[k] / b[k] < 2.3
    assert rows[-1][0] < 3e-4 and rows[-1][1] < 1e-3 and rows[-1][2] < 1e-3


def test_validity_longitudinal_rate_without_eta_flow(oracle_mod):
    """Boost-invariant transverse flow (u^eta = 0): the longitudinal expansion
    rate is exactly u^tau / tau (z = (0, 0, 0, 1/tau)), so
    kn_L = 5 (eta/s) u^tau / (tau T) cell by cell, and kn = kn_T + kn_L where
    theta_T >= 0 (an expanding fireball)."""
    from tests.parity import gaussian_ic, rest_frame
    nx, ny = 24, 20
    e = gaussian_ic(nx, ny, sx=1.0, sy=1.4, peak=60.0)
    o = oracle_mod.Oracle(nx, ny, 1, 0.1, 0.1, 0.1, 0.01, ETABAR, 0.5, threads=2)
    o.set_primitives(e, *rest_frame(nx * ny), pl=e / 3)
    o.step(40)
    v, st, tau = o.validity(), o.state(), o.time
    o.close()
    assert np.max(np.abs(st["u"][1])) > 0.05
    T = (st["e"] / 13.8997) ** 0.25
    np.testing.assert_allclose(v["kn_L"], 5 * ETABAR * st["u"][0] / (tau * T), rtol=1e-6)


def test_inverse_reynolds_number_by_tensor_contraction(oracle_mod):
    """R_pi = sqrt|pi^{mu nu} pi_{mu nu}| / |P-hat_T| on a flowing 3+1D state
    with pi^{mu nu} != 0 and u^eta != 0: the contraction is done here as a
    generic einsum with the Milne metric diag(1, -1, -1, -tau^2) at the time
    after the step, pi and P_L from the step's start (its conserved
    variables), P-hat_T = (e - P_L)/2 for the conformal EOS with the new e."""
    from tests.parity import gaussian_ic, rest_frame
    nx, ny, nz = 12, 10, 6
    e = gaussian_ic(nx, ny, nz, dz=0.4, sx=0.8, sy=1.1, tilt=0.3, peak=60.0, seed=3)
    n = nx * ny * nz
    o = oracle_mod.Oracle(nx, ny, nz, 0.1, 0.1, 0.4, 0.01, ETABAR, 0.5, threads=2)
    o.set_primitives(e, *rest_frame(n), pl=e / 3)
    o.step(15)
    q0 = o.state()["q"]
    o.step(1)
    v, st, tau = o.validity(), o.state(), o.time
    o.close()
    assert np.max(np.abs(st["u"][3])) > 1e-3 and np.max(np.abs(q0[8])) > 0  # u^eta and pi^taueta present
    pi = np.zeros((n, 4, 4))
    comp = {(0, 0): 5, (0, 1): 6, (0, 2): 7, (0, 3): 8, (1, 1): 9, (1, 2): 10, (1, 3): 11, (2, 2): 12, (2, 3): 13,
            (3, 3): 14}
    for (a, b), f in comp.items():
        pi[:, a, b] = pi[:, b, a] = q0[f]
    g = np.diag([1.0, -1.0, -1.0, -tau * tau])
    pipi = np.einsum("nab,ac,bd,ncd->n", pi, g, g, pi)
    r = np.sqrt(np.abs(pipi)) / np.abs((st["e"] - q0[4]) / 2)
    np.testing.assert_allclose(v["r_pi"], r, rtol=1e-12)


def test_validity_maps_by_tensor_calculus(oracle_mod):
    """Every validity map on a flowing 3+1D state with u^eta != 0 and pi != 0,
    against the covariant definitions evaluated in tests/milne.py (index
    contractions with the Milne metric, not the reference's component
    expansion): the Knudsen numbers
    5 (eta/s) |theta_T|, |zDzu|, |theta| / T, and the second-order inverse
    Reynolds number sqrt|J:J / sigma_T:sigma_T| / (2 etabar_TC) with
    J = delta_pipi theta_T pi + tau_pipi I4 - lambda_pipi zDzu pi,
    I4^{mu nu} = (pi^{mu a} sigma_T,a^nu + pi^{nu a} sigma_T,a^mu)/2 -
    Xi^{mu nu} (pi:sigma_T)/2 (HydrodynamicValidity.cpp:272-387; the
    reference's integer 1/2 makes tau_pipi = 4 delta_pipi / 3), pi and P_L
    from the step's start, T, e and u after it."""
    import ctypes
    import ctypes.util
    from tests.milne import contract, kinematics, pi_tensor
    from tests.parity import gaussian_ic, rest_frame
    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
    libm.powf.restype = ctypes.c_float
    nx, ny, nz, dx, dz, dt = 12, 10, 6, 0.1, 0.4, 0.01
    e0 = gaussian_ic(nx, ny, nz, dz=dz, sx=0.8, sy=1.1, tilt=0.3, peak=60.0, seed=3)
    n = nx * ny * nz
    o = oracle_mod.Oracle(nx, ny, nz, dx, dx, dz, dt, ETABAR, 0.5, threads=2)
    o.set_primitives(e0, *rest_frame(n), pl=e0 / 3)
    o.step(15)
    q0 = o.state()["q"]
    o.step(1)
    v, st, tau = o.validity(), o.state(), o.time
    o.close()
    assert np.max(np.abs(st["u"][3])) > 1e-3 and np.max(np.abs(q0[5:])) > 0
    theta, theta_T, zDzu, sT, Xi = kinematics(st["u"], st["up"], (nz, ny, nx), tau, dt, dx, dx, dz)
    T = np.array([libm.powf(ctypes.c_float(x / 13.8997), ctypes.c_float(0.25)) for x in st["e"]], dtype=float)
    kn = 5 * ETABAR / T
    for key, val in (("kn_T", theta_T), ("kn_L", zDzu), ("kn", theta)):
        np.testing.assert_allclose(v[key], kn * np.abs(val), rtol=1e-12, atol=1e-14 * np.max(kn * np.abs(theta)))
    # second order
    L = oracle_mod.lib()
    e, pl = st["e"], q0[4]
    Rt = np.array([L.orc_rtilde(a) for a in pl / e])
    Rg = np.array([L.orc_rgamma(a) for a in pl / e])
    pi = pi_tensor(q0)
    g = np.diag([1.0, -1.0, -1.0, -tau * tau])
    ps = contract(pi, sT, tau)
    I4 = np.einsum("nma,ab,nbv->nmv", pi, g, sT)
    I4 = 0.5 * (I4 + I4.transpose(0, 2, 1)) - 0.5 * Xi * ps[:, None, None]
    delta = (3 + Rg) / 2
    J = (delta * theta_T - (Rg - 1) * zDzu)[:, None, None] * pi + (4 * delta / 3)[:, None, None] * I4
    etaTC = ((3 - Rt) * e - 2 * pl) / 8
    r2 = np.sqrt(np.abs(contract(J, J, tau) / contract(sT, sT, tau))) / 2 / etaTC
    assert np.all(np.isfinite(r2)) and np.max(r2) > 1e-3
    np.testing.assert_allclose(v["r_pi2"], r2, rtol=1e-12)
