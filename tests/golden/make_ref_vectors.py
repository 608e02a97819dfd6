"""Writes tests/golden/ref_vectors.npz: inputs and outputs of the reference's
own functions -- SourceTerms.cpp (loadSourceTerms, loadSourceTermsX/Y/Z),
EquationOfState.cpp, rhic/src/AnisotropicDistributionFunctions.cpp,
FluxLimiter.cpp and HalfSiteExtrapolation.cpp -- computed by the reference
build of oracle/Makefile (target ``ref``, compiled from the reference tree)
on seeded inputs.  Run from the repository root where /root/reference exists:

    python3 tests/golden/make_ref_vectors.py

tests/test_oracle_golden.py checks the oracle against these vectors bit for
bit on machines without the reference tree.  Arrays only (no pickles).
"""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402
from tests.blocks import random_axis_gather, random_source_block  # noqa: E402

_dp = ctypes.POINTER(ctypes.c_double)


def p(a):
    return a.ctypes.data_as(_dp)


def main():
    R = oracle.ref_lib()
    if R is None:
        raise SystemExit("the reference build needs /root/reference")
    rng = np.random.default_rng(2026)
    out = {}
    # loadSourceTerms on 3x3x3 blocks (flow, pi and longitudinal flow on and off)
    n = 80
    Q, F, U, S, scal = [], [], [], [], []
    for case in range(n):
        blk = random_source_block(rng, flow=case % 4 != 0, pi=case % 3 != 0, longitudinal=case % 5 != 0)
        s = np.zeros(15)
        sc = np.array([blk["t"], blk["etabar"], blk["dt"], blk["dx"], blk["dy"], blk["dz"]])
        R.ref_source_block(p(blk["Q"]), p(blk["fields"]), p(blk["uprev"]), *sc, p(s))
        Q.append(blk["Q"]); F.append(blk["fields"]); U.append(blk["uprev"]); scal.append(sc); S.append(s)
    out.update(src_Q=np.array(Q), src_fields=np.array(F), src_uprev=np.array(U), src_scalars=np.array(scal),
               src_S=np.array(S))
    # loadSourceTermsX/Y/Z from 5-point gathers
    G, A, SA = [], [], []
    for direction in (1, 2, 3):
        for _ in range(40):
            g, ut, ui, t, d = random_axis_gather(rng)
            s = np.zeros(15)
            R.ref_source_axis(p(g), ut, ui, t, d, direction, p(s))
            G.append(g); A.append([ut, ui, t, d, direction]); SA.append(s[:4])
    out.update(axis_gather=np.array(G), axis_args=np.array(A), axis_S=np.array(SA))
    # conformal temperature, transversePressureHat, Rbar0 and Rbar0' fits
    es = np.concatenate([10.0 ** rng.uniform(-7, 3, 500), [1e-3, 1.0, 13.8997, 93.21038818333332]])
    pls = es * rng.uniform(0.001, 1.0, es.size)
    out.update(eos_e=es, eos_pl=pls, eos_T=np.array([R.ref_temperature(e) for e in es]),
               eos_pt_hat=np.array([R.ref_pt_hat(e, e / 3, pl) for e, pl in zip(es, pls)]))
    a = np.concatenate([rng.uniform(0.0, 1.2, 500), [0.0, 1 / 3, 1.0]])
    out.update(fit_a=a, fit_rbar0=np.array([R.ref_rbar0(x) for x in a]),
               fit_rbar0p=np.array([R.ref_rbar0p(x) for x in a]))
    # approximateDerivative (minmod, theta = 1.1) and the four half-site extrapolations
    v = rng.standard_normal((800, 5)) * (10.0 ** rng.uniform(-6, 6, (800, 1)))
    v[::7, 2] = v[::7, 1]
    v[::11, 3] = v[::11, 2]
    out.update(lim_v=v, lim_slope=np.array([R.ref_approx_derivative(*r[:3]) for r in v]),
               lim_halfsite=np.array([[R.ref_halfsite(k, *r) for k in range(4)] for r in v]))
    path = os.path.join(ROOT, "tests", "golden", "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
