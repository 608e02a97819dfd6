"""Pins the oracle against the reference's own functions, compiled from the
reference sources (oracle/Makefile target ``ref``; SourceTerms.cpp,
EquationOfState.cpp, FluxLimiter.cpp, HalfSiteExtrapolation.cpp and
rhic/src/AnisotropicDistributionFunctions.cpp), bit for bit.

These run only where the reference tree is present; the committed fixtures in
tests/golden/ carry the same pins to machines without it
(tests/test_oracle_golden.py).
"""
import ctypes

import numpy as np
import pytest

from tests.blocks import random_axis_gather, random_source_block

_dp = ctypes.POINTER(ctypes.c_double)


def p(a):
    return a.ctypes.data_as(_dp)


def test_source_terms_bitexact(oracle_mod, ref_lib):
    L = oracle_mod.lib()
    rng = np.random.default_rng(1234)
    worst = 0
    for case in range(400):
        blk = random_source_block(rng, flow=case % 4 != 0, pi=case % 3 != 0, longitudinal=case % 5 != 0)
        S_ref = np.zeros(15)
        S_orc = np.zeros(15)
        args = [p(blk["Q"]), p(blk["fields"]), p(blk["uprev"]), blk["t"], blk["etabar"], blk["dt"], blk["dx"],
                blk["dy"], blk["dz"]]
        ref_lib.ref_source_block(*args, p(S_ref))
        L.orc_source_block(*args, p(S_orc))
        assert np.all(np.isfinite(S_ref))
        if not np.array_equal(S_ref, S_orc):
            diff = np.abs(S_ref - S_orc) / np.maximum(np.abs(S_ref), 1e-300)
            worst = max(worst, diff.max())
            raise AssertionError(f"case {case}: source terms differ, max rel {diff.max():.3e} at {diff.argmax()}")


@pytest.mark.parametrize("direction", [1, 2, 3])
def test_axis_sources_bitexact(oracle_mod, ref_lib, direction):
    L = oracle_mod.lib()
    rng = np.random.default_rng(7 + direction)
    for _ in range(200):
        g, ut, ui, t, d = random_axis_gather(rng)
        S_ref = np.zeros(15)
        S_orc = np.zeros(15)
        ref_lib.ref_source_axis(p(g), ut, ui, t, d, direction, p(S_ref))
        L.orc_source_axis(p(g), ut, ui, d, direction, p(S_orc))
        assert np.array_equal(S_ref[:4], S_orc[:4])


def test_eos_and_fits_bitexact(oracle_mod, ref_lib):
    L = oracle_mod.lib()
    rng = np.random.default_rng(3)
    es = np.concatenate([10.0 ** rng.uniform(-7, 3, 2000), [1e-3, 1.0, 13.8997, 93.21038818333332]])
    for e in es:
        assert L.orc_temperature_conformal(e) == ref_lib.ref_temperature(e)
        pl = e * rng.uniform(0.001, 1.0)
        assert L.orc_pt_hat(e, e / 3, pl) == ref_lib.ref_pt_hat(e, e / 3, pl)
    for a in np.concatenate([rng.uniform(0.0, 1.2, 3000), [0.0, 1 / 3, 1.0]]):
        assert L.orc_rbar0(a) == ref_lib.ref_rbar0(a)
        assert L.orc_rbar0p(a) == ref_lib.ref_rbar0p(a)


def test_limiter_and_halfsite_bitexact(oracle_mod, ref_lib):
    L = oracle_mod.lib()
    rng = np.random.default_rng(11)
    vals = rng.standard_normal((5000, 5))
    vals[::7, 2] = vals[::7, 1]  # flat pieces
    vals[::11, 3] = vals[::11, 2]
    for v in vals:
        x, y, z = v[:3]
        assert L.orc_slope(x, y, z) == ref_lib.ref_approx_derivative(x, y, z)
        qmm, qm, q, qp, qpp = v
        # HalfSiteExtrapolation.cpp:12-26 in terms of the slope
        assert ref_lib.ref_halfsite(0, *v) == qp - L.orc_slope(q, qp, qpp) / 2
        assert ref_lib.ref_halfsite(1, *v) == q - L.orc_slope(qm, q, qp) / 2
        assert ref_lib.ref_halfsite(2, *v) == q + L.orc_slope(qm, q, qp) / 2
        assert ref_lib.ref_halfsite(3, *v) == qm + L.orc_slope(qmm, qm, q) / 2
