"""Host-side pieces of the library (no GPU calls): exported symbols, the
crossover EOS table, initial conditions and the reference's configuration
loader."""
import os
import re

import numpy as np
import pytest

from tests.parity import vah

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("vah.h", "vah_host.h", "vah_batch.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(vah_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_entry_point():
    L = vah().lib()
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_eos_table_thermodynamics():
    V = vah()
    lne0, h, p, T = V.eos_lqcd_table(10000, 1e-8, 1e4)
    e = np.exp(lne0 + h * np.arange(len(p)))
    assert np.all(np.diff(T) > 0) and np.all(np.diff(p) > 0)
    # nodes reproduce the analytic EOS: e(T_i) = e_i
    for i in (0, 1234, 5000, 9999):
        ei, pi = V.eos_lqcd(T[i])
        assert ei == pytest.approx(e[i], rel=1e-12) and pi == pytest.approx(p[i], rel=1e-12)
    # s = dp/dT = (e + p)/T along the table (finite differences)
    dpdT = np.gradient(p, T)
    s = (e + p) / T
    assert np.max(np.abs(dpdT[5:-5] / s[5:-5] - 1)) < 2e-4
    # high-T limit is the reference's conformal EOS e = 13.8997 T^4, p = e/3
    eh, ph = V.eos_lqcd(5.0)
    assert eh == pytest.approx(13.8997 * 5.0 ** 4, rel=1e-4) and ph == pytest.approx(eh / 3, rel=1e-4)
    # crossover: p/T^4 rises through T ~ 155 MeV
    Tc = 0.155 / V.HBARC
    lo, hi = V.eos_lqcd(0.8 * Tc), V.eos_lqcd(1.2 * Tc)
    assert lo[1] / (0.8 * Tc) ** 4 < 0.5 * hi[1] / (1.2 * Tc) ** 4


def test_optical_glauber_geometry():
    V = vah()
    g = V.glauber()
    e = V.ic_optical_glauber(81, 81, 1, 0.1, 0.1, g).reshape(81, 81)
    assert e.max() == pytest.approx(g.e_peak + g.e_floor, rel=1e-12)
    np.testing.assert_allclose(e, e[:, ::-1], rtol=1e-12)
    np.testing.assert_allclose(e, e[::-1, :], rtol=1e-12)
    # b = 7 along x: the overlap is longer in y than in x
    c = e[40, 40]
    half_x = np.argmax(e[40, 40:] < c / 2)
    half_y = np.argmax(e[40:, 40] < c / 2)
    assert half_y > half_x
    # central collision, tiny e_floor: rotationally symmetric profile
    g0 = V.glauber(b=0.0, e_floor=0.0)
    e0 = V.ic_optical_glauber(41, 41, 1, 0.2, 0.1, g0).reshape(41, 41)
    np.testing.assert_allclose(e0, e0.T, rtol=1e-12)


def test_mc_glauber_is_seeded_and_plausible():
    V = vah()
    g = V.glauber(A=208, R=6.62, a=0.546, sigma_nn_mb=70.0, b=6.0, seed=42, e_peak=60 / V.HBARC)
    e1, n1 = V.ic_mc_glauber(61, 61, 1, 0.2, 0.1, g)
    e2, n2 = V.ic_mc_glauber(61, 61, 1, 0.2, 0.1, g)
    assert n1 == n2 and np.array_equal(e1, e2)
    assert 200 < n1 < 360          # Pb+Pb at b = 6 fm
    g.seed = 43
    e3, _ = V.ic_mc_glauber(61, 61, 1, 0.2, 0.1, g)
    assert not np.array_equal(e1, e3)


def test_eta_profile():
    V = vah()
    L = V.lib()
    assert L.vah_ic_eta_profile(0.0, 1.5, 1.4) == 1.0
    assert L.vah_ic_eta_profile(1.5, 1.5, 1.4) == 1.0
    assert L.vah_ic_eta_profile(-2.9, 1.5, 1.4) == pytest.approx(np.exp(-1.4 ** 2 / (2 * 1.4 ** 2)))


def test_config_defaults_and_reference_files(tmp_path):
    V = vah()
    c = V.config_load(str(tmp_path / "missing"))
    # LatticeParameters.cpp:30-39, InitialConditionParameters.cpp:33-41, HydroParameters.cpp:25-29
    assert (c.numLatticePointsX, c.numLatticePointsY, c.numLatticePointsRapidity, c.numProperTimePoints) == \
        (128, 128, 64, 10)
    assert (c.latticeSpacingX, c.latticeSpacingRapidity, c.latticeSpacingProperTime) == (0.08, 0.3, 0.01)
    assert (c.initialConditionType, c.numberOfNucleonsPerNuclei) == (2, 208)
    assert (c.initialProperTimePoint, c.shearViscosityToEntropyDensity, c.freezeoutTemperatureGeV) == \
        (0.1, 0.0795775, 0.155)
    assert c.initializePimunuNavierStokes == 1
    conf = os.path.join(ROOT, "tests", "golden", "rhic-conf")
    c = V.config_load(conf)
    assert (c.numLatticePointsX, c.numLatticePointsY, c.numLatticePointsRapidity) == (201, 1, 1)
    assert c.numProperTimePoints == 10000 and c.latticeSpacingProperTime == 0.005
    assert c.initialConditionType == 4 and c.initialEnergyDensity == 15.6269
    assert c.initialProperTimePoint == 0.25 and c.freezeoutTemperatureGeV == 0.140
    assert c.initializePimunuNavierStokes == 0


def _printf_dat(cfg, var):
    """The reference's output() lines through the C library's snprintf."""
    import ctypes
    import ctypes.util
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    buf = ctypes.create_string_buffer(2048)
    nx, ny, nz = cfg.numLatticePointsX, cfg.numLatticePointsY, cfg.numLatticePointsRapidity
    out = []
    for k in range(nz):
        z = (k - (nz - 1) / 2.) * cfg.latticeSpacingRapidity
        for j in range(ny):
            y = (j - (ny - 1) / 2.) * cfg.latticeSpacingY
            for i in range(nx):
                x = (i - (nx - 1) / 2.) * cfg.latticeSpacingX
                v = var[(k * ny + j) * nx + i]
                libc.snprintf(buf, 2048, b"%.3f\t%.3f\t%.3f\t%.8f\n", ctypes.c_double(x), ctypes.c_double(y),
                              ctypes.c_double(z), ctypes.c_double(v))
                out.append(buf.value)
    return b"".join(out)


def test_dat_writer_is_printf_to_the_byte(tmp_path):
    """vah_write_dat (the .dat files of vah_run) formats on host threads with
    std::to_chars; its file equals the C library's printf of the reference's
    format line by line: random magnitudes over 40 decades, exact decimal
    ties of the last digit, signed zeros, values too long for the fast path,
    nan and inf."""
    V = vah()
    c = V.config_load(str(tmp_path / "missing"))
    c.numLatticePointsX, c.numLatticePointsY, c.numLatticePointsRapidity = 37, 23, 9
    c.latticeSpacingX, c.latticeSpacingY, c.latticeSpacingRapidity = 0.1, 0.13, 0.3
    n = 37 * 23 * 9
    rng = np.random.default_rng(4)
    var = rng.standard_normal(n) * 10.0 ** rng.uniform(-20, 20, n)
    special = [0.0, -0.0, 5e-9, -5e-9, 0.125e-7, 1.5e-8, 2.5e-8, 0.5, 1e300, -3e250, 1e-300, float("nan"),
               float("inf"), -float("inf"), 123456789.123456785, 2.0 ** -27, 2.0 ** -30]
    var[:len(special)] = special
    var[100:200] = np.round(var[100:200] * 1e8) / 1e8 + 5e-9  # near-ties of the 8th decimal
    V.write_dat(c, var, 1.25, str(tmp_path), "e")
    got = (tmp_path / "e_1.250.dat").read_bytes()
    want = _printf_dat(c, var)
    assert got == want


def test_dat_writer_argument_checks(tmp_path):
    """Bad arguments and an unwritable directory come back as VAH_ERR_ARG."""
    import ctypes
    V = vah()
    L = V.lib()
    c = V.config_load(str(tmp_path / "missing"))
    c.numLatticePointsX, c.numLatticePointsY, c.numLatticePointsRapidity = 3, 2, 1
    v = np.zeros(6)
    p = v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    d, nm = str(tmp_path).encode(), b"e"
    assert L.vah_write_dat(None, p, 0.5, d, nm) == V.VAH_ERR_ARG
    assert L.vah_write_dat(ctypes.byref(c), None, 0.5, d, nm) == V.VAH_ERR_ARG
    assert L.vah_write_dat(ctypes.byref(c), p, 0.5, None, nm) == V.VAH_ERR_ARG
    assert L.vah_write_dat(ctypes.byref(c), p, 0.5, str(tmp_path / "no" / "such").encode(), nm) == V.VAH_ERR_ARG
    c.numLatticePointsY = 0
    assert L.vah_write_dat(ctypes.byref(c), p, 0.5, d, nm) == V.VAH_ERR_ARG
    assert not list(tmp_path.glob("*.dat"))


def test_config_int_literal_for_float_keeps_default(tmp_path):
    """libconfig does not convert an integer literal for a float lookup, so
    the reference keeps the default (Properties.cpp:17-22)."""
    V = vah()
    (tmp_path / "lattice.properties").write_text("latticeSpacingX = 1\nlatticeSpacingY = 0.25;\n"
                                                 "numLatticePointsX = 33 // comment\n# numLatticePointsY = 7\n")
    c = V.config_load(str(tmp_path))
    assert c.latticeSpacingX == 0.08 and c.latticeSpacingY == 0.25
    assert c.numLatticePointsX == 33 and c.numLatticePointsY == 128


def test_reference_initial_conditions(tmp_path):
    V = vah()
    conf = os.path.join(ROOT, "tests", "golden", "rhic-conf")
    c = V.config_load(conf)
    c.numLatticePointsX, c.numLatticePointsY = 41, 41
    e, ut, ux, uy, un = V.ic_reference(c)
    assert np.all(ut == 1.0) and np.all(ux == 0.0)
    assert e.min() >= 1e-3 and e.max() > 1.0
    e2 = V.ic_reference(c)[0]
    assert np.array_equal(e, e2)  # srand(1328398221) makes it repeatable
    c.initialConditionType = 0
    e0 = V.ic_reference(c)[0]
    assert np.all(e0 == e0[0]) and e0[0] == pytest.approx(13.8997 * 3.05 ** 4, rel=1e-6)
    c.initialConditionType = 7
    with pytest.raises(V.VahError):
        V.ic_reference(c)


def test_batch_event_bytes_and_variants():
    """Arena accounting of the batch engine: 61 padded fields per event, one
    stored plane for boost-invariant lattices, 2 ghost planes per side in eta
    otherwise; every compiled variant's tile fits one 256-lane workgroup."""
    V = vah()
    def expect(nx, ny, nz):
        ncz = nz + 4 if nz > 1 else 1
        fs = -(-((nx + 4) * (ny + 4) * ncz) // 64) * 64
        return -(-(8 * fs * 61) // 256) * 256
    for dims in ((101, 101, 1), (261, 131, 1), (121, 121, 41), (1, 1, 1), (7, 3, 2)):
        assert V.batch_event_bytes(*dims) == expect(*dims), dims
    assert V.batch_event_bytes(0, 5, 5) == 0
    vs = V.batch_variants()
    assert len(vs) >= 4 and vs[0] == (16, 15, 2, 3)
    for tx, ty, b, kz in vs:
        assert tx * ty <= 256 and (tx + 1) * ty <= 256 and tx * (ty + 1) <= 256
        assert b in (1, 2, 3) and kz in (3, 6)


@pytest.mark.parametrize("kw, msg", [
    (dict(nx=0), "lattice sizes"),
    (dict(dx=-0.1), "lattice sizes"),
    (dict(t0=0.0), "t0 > 0"),
    (dict(etabar=0.0), "eta/s > 0"),
    (dict(precision=7), "unknown precision"),
    (dict(precision=1, flux_mode=1), "shared faces only"),
    (dict(eta_chunk=300), "eta_chunk"),
    (dict(flux_mode=5), "unknown flux mode"),
    (dict(eos_kind=9), "unknown EOS kind"),
    (dict(eos_table=(0.0, 0.0, [1.0, 2.0], [1.0, 2.0])), "EOS table"),
], ids=["nx", "dx", "t0", "etabar", "precision", "split-per-cell", "eta_chunk", "flux_mode", "eos_kind", "table_h"])
def test_vah_create_rejects_bad_arguments_before_touching_the_device(kw, msg):
    """vah_create validates its arguments before any HIP call: each bad value
    comes back as VAH_ERR_ARG with its message, here without a GPU."""
    V = vah()
    args = dict(nx=10, ny=10, nz=1, dx=0.1, dy=0.1, dz=0.1, dt=0.01, etabar=0.2, t0=0.5)
    args.update(kw)
    with pytest.raises(V.VahError) as exc:
        V.Vah(**args)
    assert exc.value.code == V.VAH_ERR_ARG and msg in str(exc.value)
