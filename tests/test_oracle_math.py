"""Pins of the oracle's functions that the reference build cannot reach
(EnergyMomentumTensor.cpp and FullyDiscreteKurganovTadmorScheme.cpp need the
absent libconfig.h): closed forms, round trips, special cases and invariants.
"""
import ctypes
import ctypes.util

import numpy as np
import pytest

_dp = ctypes.POINTER(ctypes.c_double)
libm = ctypes.CDLL(ctypes.util.find_library("m"))
libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
libm.powf.restype = ctypes.c_float


def p(a):
    return a.ctypes.data_as(_dp)


def conserved_from_primitives(t, e, pl, ut, ux, uy, un, pi):
    """T^{tau mu} of the anisotropic decomposition with P_T = (e - P_L)/2:
    T^{mu nu} = (e + P_T) u^mu u^nu - P_T g^{mu nu} + (P_L - P_T) z^mu z^nu + pi^{mu nu},
    z^mu = (t u^eta, 0, 0, u^tau / t) / sqrt(1 + u_T^2), metric diag(1, -1, -1, -1/t^2).
    Built from the tensor definition, independently of the component code."""
    PT = 0.5 * (e - pl)
    F = np.sqrt(1 + ux * ux + uy * uy)
    u = np.array([ut, ux, uy, un])
    z = np.array([t * un / F, 0.0, 0.0, ut / t / F])
    g = np.diag([1.0, -1.0, -1.0, -1.0 / t ** 2])
    T = (e + PT) * np.outer(u, u) - PT * g + (pl - PT) * np.outer(z, z)
    q = np.zeros(15)
    q[0:4] = T[0] + pi[0:4]
    q[4] = pl
    q[5:15] = pi
    return q


def random_primitives(rng, t):
    e = 10 ** rng.uniform(-2, 2.5)
    pl = e * rng.uniform(0.02, 0.9)
    ux, uy = rng.normal(0, 0.8, 2)
    un = rng.normal(0, 0.5) / t
    ut = np.sqrt(1 + ux * ux + uy * uy + t * t * un * un)
    # a pi orthogonal to u and z is not needed for the inversion: it is subtracted first
    pi = rng.normal(0, 0.05 * e, 10)
    return e, pl, ut, ux, uy, un, pi


def test_newton_round_trip(oracle_mod):
    L = oracle_mod.lib()
    rng = np.random.default_rng(21)
    out = np.zeros(6)
    worst = 0.0
    for _ in range(3000):
        t = rng.uniform(0.3, 5.0)
        e, pl, ut, ux, uy, un, pi = random_primitives(rng, t)
        q = conserved_from_primitives(t, e, pl, ut, ux, uy, un, pi)
        guess = np.hypot(ux, uy) * rng.uniform(0.8, 1.2)
        st = L.orc_recover(t, p(q), e, guess, p(out))
        assert st == 0
        rel = max(abs(out[0] - e) / e, abs(out[2] - ut) / ut, abs(out[3] - ux) / ut, abs(out[4] - uy) / ut,
                  abs(out[5] - un) * t / ut)
        worst = max(worst, rel)
        assert out[1] == out[0] / 3  # conformal p = e/3
    assert worst < 1e-11, worst


def test_newton_rest_frame_is_exact(oracle_mod):
    """u = (1,0,0,0): MBT hits the 1e-16 guard, F = 0, and e = T^tautau."""
    L = oracle_mod.lib()
    out = np.zeros(6)
    for e in (1e-3, 1.0, 93.21038818333332):
        q = np.zeros(15)
        q[0] = e
        q[4] = e / 3
        assert L.orc_recover(0.5, p(q), e, 0.0, p(out)) == 0
        assert out[0] == e and out[2] == 1.0 and out[3] == 0.0 and out[5] == 0.0


def face_root_closed_form(t, q):
    """|u_T| of a state from the quadratic the face solve's update zeroes
    (Pi = 0): with U = u/sqrt(1+u^2) and M = T^{tau mu} - pi^{tau mu},
    f(U) = -(b xM/2) U^2 + B U - xM, B = bL + b (c0 - P_L)/2; its smaller
    root 2 xM / (B + sqrt(B^2 - 2 b xM^2)).  The stage kernels take it as a
    face solve's result once the first Newton pass shows that the path ends
    there (gpu-vah_amd/csrc/vah_stage.h, face_root)."""
    M0, M1, M2, M3 = q[0] - q[5], q[1] - q[6], q[2] - q[7], q[3] - q[8]
    PL = q[4]
    Ft = t * M3 / (M0 + PL)
    x = np.sqrt(abs(1 - Ft * Ft))
    xM = x * np.hypot(M1, M2)
    bL = x * x * M0 - Ft * Ft * PL
    b = x * x + Ft * Ft
    c0 = M0 - t * Ft * M3
    B = bL + 0.5 * b * (c0 - PL)
    U = 2 * xM / (B + np.sqrt(B * B - 2 * b * xM * xM))
    return U / np.sqrt(1 - U * U)


def test_face_solve_root_in_closed_form(oracle_mod):
    """The closed-form root is the state's |u_T| (the primitives it was built
    from) and the oracle's Newton solution (EnergyMomentumTensor.cpp:30-72)
    from a guess 20 % off; a larger root (the other sign) would miss both."""
    L = oracle_mod.lib()
    rng = np.random.default_rng(33)
    out = np.zeros(6)
    worst_true = worst_newton = 0.0
    for _ in range(3000):
        t = rng.uniform(0.3, 5.0)
        e, pl, ut, ux, uy, un, pi = random_primitives(rng, t)
        q = conserved_from_primitives(t, e, pl, ut, ux, uy, un, pi)
        uT = np.hypot(ux, uy)
        u = face_root_closed_form(t, q)
        assert L.orc_recover(t, p(q), e, uT * rng.uniform(0.8, 1.2), p(out)) == 0
        worst_true = max(worst_true, abs(u - uT) / max(uT, 1e-3))
        worst_newton = max(worst_newton, abs(u - np.hypot(out[3], out[4])) / max(uT, 1e-3))
    assert worst_true < 1e-12, worst_true
    assert worst_newton < 1e-12, worst_newton


def face_root_jump(t, q, u0):
    """The stage kernels' test for a face solve started at u0 (face_root and
    face_speeds in gpu-vah_amd/csrc/vah_stage.h): None when the path goes on
    iterating, else the closed-form root the kernels take after the first
    pass.  The first pass is the reference's update
    (EnergyMomentumTensor.cpp:45-70); the path is then known to succeed when
    the start lies left of the quadratic's vertex, the update is clearly
    positive and e at the root is clearly positive."""
    M0, M1, M2, M3 = q[0] - q[5], q[1] - q[6], q[2] - q[7], q[3] - q[8]
    PL = q[4]
    M0PL = M0 + PL if M0 + PL != 0 else 1e-16
    Ft = t * M3 / M0PL
    x = np.sqrt(abs(1 - Ft * Ft))
    MBT = np.hypot(M1, M2) or 1e-16
    xM, bL, b, c0 = x * MBT, x * x * M0 - Ft * Ft * PL, x * x + Ft * Ft, M0 - t * Ft * M3
    B = bL + 0.5 * b * (c0 - PL)
    disc = B * B - 2 * b * xM * xM
    U = 2 * xM / (B + np.sqrt(max(disc, 0.0)))
    ok = disc >= 0 and B > 0 and 0 <= U < 1 and c0 - U * xM > 1e-12 * abs(c0)
    rs = 1 / np.sqrt(1 + u0 * u0)
    U0 = u0 * rs
    e = c0 - U0 * xM
    g = bL + b * 0.5 * (e - PL)
    fp = rs ** 3 * g - 0.5 * U0 * b * xM * rs ** 3
    u1 = u0 - (U0 * g - xM) / (fp if fp != 0 else 1e-16)
    if e < 0 or not (0 <= u1 <= 9.0072e15) or abs(u1 - u0) <= 1e-7 * abs(u1):
        return None
    if ok and U0 * b * xM <= B and u1 >= 1e-9 * u0:
        return U / np.sqrt(1 - U * U)
    return None


def test_face_root_jump_keeps_the_reference_path(oracle_mod):
    """Wherever the kernels' root jump fires, the oracle's Newton from the
    same start succeeds and ends on that root: physical states with near and
    far starts, near light-like flow, and random (often unphysical) conserved
    variables.  A jump that skipped a failure the full path would meet would
    show here as a failed orc_recover."""
    L = oracle_mod.lib()
    rng = np.random.default_rng(5)
    out = np.zeros(6)
    fired = 0
    worst = 0.0
    for k in range(20000):
        t = rng.uniform(0.3, 5.0)
        kind = k % 4
        if kind == 2:
            q = np.zeros(15)
            q[0] = 10 ** rng.uniform(-3, 2)
            q[1:4] = rng.normal(0, 1, 3) * q[0]
            q[3] /= t
            q[4] = q[0] * rng.uniform(-0.1, 1.0)
            q[5:15] = rng.normal(0, 0.1, 10) * q[0]
            u0 = 10 ** rng.uniform(-4, 1.5)
        else:
            e, pl, ut, ux, uy, un, pi = random_primitives(rng, t)
            if kind == 3:
                s = 10 ** rng.uniform(0.5, 1.5)
                ux, uy = ux * s, uy * s
                ut = np.sqrt(1 + ux * ux + uy * uy + t * t * un * un)
            q = conserved_from_primitives(t, e, pl, ut, ux, uy, un, pi)
            uT = np.hypot(ux, uy)
            u0 = uT * rng.uniform(0.5, 1.5) if kind != 1 else 10 ** rng.uniform(-6, 3)
        u = face_root_jump(t, q, u0)
        if u is None:
            continue
        fired += 1
        assert L.orc_recover(t, p(q), 1.0, u0, p(out)) == 0, (k, t, u0)
        worst = max(worst, abs(u - np.hypot(out[3], out[4])) / max(u, 1e-300))
    assert fired > 10000, fired
    assert worst < 1e-10, worst


def test_newton_fallback(oracle_mod):
    """Negative energy: the reference falls back to e = 0.999 e_prev, u = (1,0,0,0)
    (EnergyMomentumTensor.cpp:163-172)."""
    L = oracle_mod.lib()
    out = np.zeros(6)
    q = np.zeros(15)
    q[0] = -1.0
    q[1] = 0.3
    q[4] = 0.1
    assert L.orc_recover(1.0, p(q), 2.0, 0.1, p(out)) == -1
    assert out[0] == 2.0 * .999 and out[2] == 1.0 and out[3] == 0.0 and out[4] == 0.0 and out[5] == 0.0


def test_slope_closed_form(oracle_mod):
    L = oracle_mod.lib()
    rng = np.random.default_rng(4)
    for _ in range(20000):
        x, y, z = rng.normal(0, 1, 3)
        s = L.orc_slope(x, y, z)
        a, b, c = 1.1 * (y - x), (z - x) / 2, 1.1 * (z - y)
        if a > 0 and b > 0 and c > 0:
            assert s == min(a, b, c)
        elif a < 0 and b < 0 and c < 0:
            assert s == max(a, b, c)
        elif (a > 0 and c < 0) or (a < 0 and c > 0):  # local extremum
            assert s == 0.0


def _device_half_slope(x, y, z):
    """The stage kernels' double form of half the minmod slope (half_slope in
    gpu-vah_amd/csrc/vah_device.h), in numpy double arithmetic."""
    a, b = y - x, z - y
    m = np.minimum((1.1 * 0.5) * np.minimum(np.abs(a), np.abs(b)), 0.25 * np.abs(z - x))
    return np.where(a * b > 0, np.copysign(m, a), 0.0)


def _med3(u, v, w):
    return np.maximum(np.minimum(u, v), np.minimum(np.maximum(u, v), w))


def _device_half_slope_f32(x, y, z):
    """The float form: med3(l, med3(c, r, 0), 0) of the halved candidates."""
    h = np.float32(1.1) * np.float32(0.5)
    zero = np.float32(0.0)
    l, r, c = h * (y - x), h * (z - y), np.float32(0.25) * (z - x)
    return _med3(l, _med3(c, r, zero), zero)


def _scaled_triples(rng, n, lo, hi, dtype):
    scale = 10.0 ** rng.uniform(lo, hi, n)
    x, y, z = (rng.normal(0, 1, (3, n)) * scale).astype(dtype)
    # ties and zero differences
    y[:500] = x[:500]
    z[500:1000] = y[500:1000]
    z[1000:1500] = x[1000:1500]
    x[1500:2000] = 0.0
    return x, y, z


def test_device_half_slope_is_half_the_reference_slope(oracle_mod):
    """The kernels reconstruct faces with theta/2 * min(|a|, |b|), |z - x| / 4
    and a sign test a * b > 0 in place of the reference's sign()-based minmod
    chain (FluxLimiter.cpp:13-33, which the oracle reproduces bit for bit):
    equal to half the reference slope bit for bit over three hundred decades
    of scale, every sign pattern, zeros and ties; where a * b underflows the
    kernels' slope is 0 and the reference's below 2.4e-162."""
    L = oracle_mod.lib()
    rng = np.random.default_rng(12)
    x, y, z = _scaled_triples(rng, 20000, -150, 150, np.float64)
    got = _device_half_slope(x, y, z)
    want = np.array([0.5 * L.orc_slope(*v) for v in zip(x, y, z)])
    assert np.array_equal(got, want)
    # the underflow corner
    x, y, z = np.array([0.0, 0.0, 1e-170]), np.array([1e-170, -1e-170, 3e-170]), np.array([3e-170, -5e-170, 4e-170])
    want = np.array([0.5 * L.orc_slope(*v) for v in zip(x, y, z)])
    got = _device_half_slope(x, y, z)
    assert np.all(got == 0.0) and np.all(np.abs(want) < 2.4e-162) and np.all(want != 0.0)


def _device_slope4(x, y, z):
    """slope4 of gpu-vah_amd/csrc/vah_device.h: four times the half slope, with
    2.2 = 4 x 0.55 and |z - x| unscaled; the face value is fma(s4, 1/4, q)."""
    a, b = y - x, z - y
    m = np.minimum((1.1 * 2.0) * np.minimum(np.abs(a), np.abs(b)), np.abs(z - x))
    return np.where(a * b > 0, np.copysign(m, a), 0.0)


def test_device_slope4_is_four_half_slopes_to_the_bit():
    """The double kernels form the face values as fma(slope4, +-1/4, q): a
    quarter of slope4 must be the half slope bit for bit (scaling by a power of
    two commutes with rounding, and 2.2 as a double is 4 x 0.55 as a double), so
    that q +- slope4 / 4, rounded once by the fma, is q +- half_slope."""
    assert 1.1 * 2.0 == 4.0 * (1.1 * 0.5)
    rng = np.random.default_rng(13)
    x, y, z = _scaled_triples(rng, 20000, -150, 150, np.float64)
    s4 = _device_slope4(x, y, z)
    hs = _device_half_slope(x, y, z)
    assert np.array_equal(0.25 * s4, hs)
    assert np.array_equal(y + 0.25 * s4, y + hs) and np.array_equal(z - 0.25 * _device_slope4(y, z, x),
                                                                    z - _device_half_slope(y, z, x))


def test_float_half_slope_median_form():
    """The split-precision kernels' float slope, med3(l, med3(c, r, 0), 0),
    against the sign-test form min(theta/2 min(|a|, |b|), |z - x| / 4) with the
    sign of a when a * b > 0, in float32: equal wherever a * b does not
    underflow, and the median form is the minmod of its three candidates."""
    rng = np.random.default_rng(31)
    x, y, z = _scaled_triples(rng, 20000, -15, 15, np.float32)
    got = _device_half_slope_f32(x, y, z)
    a, b = y - x, z - y
    m = np.minimum(np.float32(0.55) * np.minimum(np.abs(a), np.abs(b)), np.float32(0.25) * np.abs(z - x))
    want = np.where(a * b > 0, np.copysign(m, a), np.float32(0.0))
    assert got.dtype == np.float32 and np.array_equal(got, want)
    # minmod of the three halved candidates, straight from its definition
    cand = np.stack([np.float32(0.55) * a, np.float32(0.25) * (z - x), np.float32(0.55) * b])
    same = np.all(cand > 0, axis=0) | np.all(cand < 0, axis=0)
    pick = np.take_along_axis(cand, np.argmin(np.abs(cand), axis=0)[None], axis=0)[0]
    assert np.array_equal(got, np.where(same, pick, np.float32(0.0)))


def test_conformal_temperature_uses_single_precision_powf(oracle_mod):
    L = oracle_mod.lib()
    rng = np.random.default_rng(8)
    for e in 10 ** rng.uniform(-7, 4, 5000):
        expect = float(libm.powf(ctypes.c_float(e / 13.8997), ctypes.c_float(0.25)))
        assert L.orc_temperature_conformal(e) == expect


def test_transverse_pressure_conformal_identity(oracle_mod):
    """For e = 3p the R-bar0 term of P-hat_T drops out: P-hat_T = (e - P_L)/2
    (AnisotropicDistributionFunctions.cpp:23)."""
    L = oracle_mod.lib()
    rng = np.random.default_rng(2)
    for e in 10 ** rng.uniform(-3, 3, 200):
        pl = e * rng.uniform(0.01, 0.99)
        assert L.orc_pt_hat(e, e / 3, pl) == pytest.approx((e - pl) / 2, rel=1e-13)


def test_regulator_closed_form(oracle_mod):
    """pi scaled by tanh(rho)/rho with rho = |pi|/sqrt(e^2 + P_L^2 + 2 P_T^2)
    when the magnitude ratio dominates (FullyDiscrete...:597-665)."""
    L = oracle_mod.lib()
    t = 1.0
    e, pl = 10.0, 10.0 / 3
    u = np.array([1.0, 0.0, 0.0, 0.0])
    for mag in (1e-9, 0.5, 4.0, 40.0):
        q = np.zeros(15)
        q[4] = pl
        q[9] = mag   # pixx
        q[12] = -mag  # piyy: traceless, orthogonal to u and z
        ee, pp = np.array([e]), np.array([e / 3])
        L.orc_regulate_cell(t, p(q), p(ee), p(pp), p(u))
        ptHat = L.orc_pt_hat(e, e / 3, pl)
        rho = np.sqrt(2 * mag * mag) / np.sqrt(e * e + pl * pl + 2 * ptHat * ptHat)
        fac = 1.0 if rho < 1e-7 else np.tanh(rho) / rho
        assert q[9] == pytest.approx(mag * fac, rel=1e-14)
        assert q[12] == pytest.approx(-mag * fac, rel=1e-14)


def test_regulator_floors(oracle_mod):
    L = oracle_mod.lib()
    q = np.zeros(15)
    q[4] = -1.0
    ee, pp = np.array([-2.0]), np.array([-1.0])
    L.orc_regulate_cell(1.0, p(q), p(ee), p(pp), p(np.array([1.0, 0, 0, 0])))
    assert ee[0] == 1e-7 and pp[0] == 1e-7 and q[4] == 1e-7


def test_wuppertal_budapest_fits_are_thermodynamically_consistent(oracle_mod):
    """The reference's non-conformal EOS is two independent rational fits,
    p(e) and T(e) (EquationOfState.cpp:31-76, :120-142).  With s = (e+p)/T,
    thermodynamics fixes T ds/de = 1; the fits satisfy it to the fits'
    accuracy (< 0.5 % over five decades of e), which a dropped term, a
    wrong sign or a mis-scaled coefficient in either fit would break.  The
    crossover (softest point p/e ~ 0.15 near T ~ 155 MeV) and the approach
    towards the conformal 1/3 at high e are checked as well."""
    L = oracle_mod.lib()
    P, T = L.orc_wb_pressure, L.orc_wb_temperature
    hbarc = 0.197326938
    s = lambda x: (x + P(x)) / T(x)
    for e in np.geomspace(0.05, 2000.0, 40):
        h = e * 1e-5
        ds = (s(e + h) - s(e - h)) / (2 * h)
        assert abs(T(e) * ds - 1.0) < 5e-3, e
    es = np.geomspace(0.05, 2000.0, 400)
    ratio = np.array([P(e) / e for e in es])
    soft = es[np.argmin(ratio)]
    assert 0.13 < ratio.min() < 0.16 and 140.0 < T(soft) * hbarc * 1e3 < 175.0
    assert 0.30 < P(2000.0) / 2000.0 < 1.0 / 3.0
    Ts = np.array([T(e) for e in es])
    assert np.all(np.diff(Ts) > 0)
