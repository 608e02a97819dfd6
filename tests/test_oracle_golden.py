"""The oracle against committed outputs of the reference's own functions
(tests/golden/ref_vectors.npz, written by tests/golden/make_ref_vectors.py from
the reference build of oracle/Makefile): loadSourceTerms and
loadSourceTermsX/Y/Z (SourceTerms.cpp), the conformal temperature
(EquationOfState.cpp:144), transversePressureHat, Rbar0 and Rbar0'
(rhic/src/AnisotropicDistributionFunctions.cpp:5-24), approximateDerivative
(FluxLimiter.cpp:13-33) and the half-site extrapolations
(HalfSiteExtrapolation.cpp:12-26), bit for bit.  Needs no reference tree:
these carry the pins of test_oracle_reference_build.py to any machine."""
import ctypes
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz")
_dp = ctypes.POINTER(ctypes.c_double)


def p(a):
    return np.ascontiguousarray(a).ctypes.data_as(_dp)


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_source_terms(oracle_mod, gold):
    L = oracle_mod.lib()
    for Q, f, u, sc, want in zip(gold["src_Q"], gold["src_fields"], gold["src_uprev"], gold["src_scalars"],
                                 gold["src_S"]):
        Q, f, u = (np.ascontiguousarray(x) for x in (Q, f, u))
        got = np.zeros(15)
        L.orc_source_block(p(Q), p(f), p(u), *map(float, sc), p(got))
        assert np.array_equal(got, want)


def test_axis_sources(oracle_mod, gold):
    L = oracle_mod.lib()
    for g, (ut, ui, t, d, direction), want in zip(gold["axis_gather"], gold["axis_args"], gold["axis_S"]):
        g = np.ascontiguousarray(g)
        got = np.zeros(15)
        L.orc_source_axis(p(g), ut, ui, d, int(direction), p(got))
        assert np.array_equal(got[:4], want)


def test_eos_and_fits(oracle_mod, gold):
    L = oracle_mod.lib()
    for e, pl, T, pth in zip(gold["eos_e"], gold["eos_pl"], gold["eos_T"], gold["eos_pt_hat"]):
        assert L.orc_temperature_conformal(e) == T
        assert L.orc_pt_hat(e, e / 3, pl) == pth
    for a, r0, r0p in zip(gold["fit_a"], gold["fit_rbar0"], gold["fit_rbar0p"]):
        assert L.orc_rbar0(a) == r0 and L.orc_rbar0p(a) == r0p


def test_limiter_and_halfsite(oracle_mod, gold):
    L = oracle_mod.lib()
    for v, slope, hs in zip(gold["lim_v"], gold["lim_slope"], gold["lim_halfsite"]):
        qmm, qm, q, qp, qpp = map(float, v)
        assert L.orc_slope(qmm, qm, q) == slope
        assert hs[0] == qp - L.orc_slope(q, qp, qpp) / 2
        assert hs[1] == q - L.orc_slope(qm, q, qp) / 2
        assert hs[2] == q + L.orc_slope(qm, q, qp) / 2
        assert hs[3] == qm + L.orc_slope(qmm, qm, q) / 2
