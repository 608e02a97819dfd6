"""Shared helpers for the parity tests: seeded initial states and the error
metric.

Error metric (per cell, the worst cell is reported):
  e, P_L, T^tautau, u^tau    |a - b| / |b|
  u^x, u^y, tau u^eta         |a - b| / u^tau
  T^taux, T^tauy, tau T^taueta   |a - b| / T^tautau
  pi^{mu nu} (eta legs scaled by tau)   |a - b| / e
The signed and transverse fields are measured against the cell's own
positive scale so that a component crossing zero is not divided by ~0.
"""
import importlib

import numpy as np

TOL_FP64 = 1e-10

# metric factor per conserved field: eta legs carry 1/tau each
ETA_LEGS = np.array([0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 1, 0, 1, 2])


def vah():
    return importlib.import_module("gpu-vah_amd")


def gaussian_ic(nx, ny, nz=1, dx=0.1, dz=0.1, peak=152.03, sx=2.0, sy=2.8, floor=1e-3, tilt=0.0, seed=None):
    """Elliptic Gaussian transverse profile x plateau-plus-Gaussian in eta.

    Interior arrays, x fastest.  ``seed`` adds +-5% lognormal lumps."""
    x = (np.arange(nx) - (nx - 1) / 2) * dx
    y = (np.arange(ny) - (ny - 1) / 2) * dx
    z = (np.arange(nz) - (nz - 1) / 2) * dz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    eT = np.exp(-((X - tilt * Y) ** 2 / (2 * sx ** 2) + Y ** 2 / (2 * sy ** 2)))
    a = np.maximum(np.abs(Z) - 1.5, 0.0)
    eL = np.exp(-a * a / (2 * 1.4 ** 2))
    e = peak * eT * eL
    if seed is not None:
        rng = np.random.default_rng(seed)
        e = e * np.exp(0.05 * rng.standard_normal(e.shape))
    return (e + floor).ravel()


def rest_frame(n):
    return np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n)


def field_errors(a, b, tau):
    """Worst-cell errors of state a against reference state b (dicts from .state())."""
    eb = b["e"]
    out = {"e": np.max(np.abs(a["e"] - eb) / np.abs(eb))}
    qa, qb = a["q"], b["q"]
    scale = tau ** ETA_LEGS
    for f, name in enumerate(vah().FIELDS):
        d = np.abs(qa[f] - qb[f]) * scale[f]
        if name in ("ttt", "pl"):
            out[name] = np.max(d / np.abs(qb[f]))
        elif f < 5:
            out[name] = np.max(d / np.abs(qb[0]))
        else:
            out[name] = np.max(d / eb)
    ua, ub = a["u"], b["u"]
    out["ut"] = np.max(np.abs(ua[0] - ub[0]) / ub[0])
    for d, name in ((1, "ux"), (2, "uy"), (3, "un")):
        s = tau if d == 3 else 1.0
        out[name] = np.max(np.abs(ua[d] - ub[d]) * s / ub[0])
    return out


def assert_parity(a, b, tau, tol=TOL_FP64, label=""):
    errs = field_errors(a, b, tau)
    bad = {k: v for k, v in errs.items() if not (v <= tol)}
    worst = max(errs.items(), key=lambda kv: kv[1])
    assert not bad, f"{label}: fields over {tol:g}: {bad}"
    return worst


def table_T(tab, e):
    """Temperature of the tabulated EOS (lne0, h, p, T) at energy density e:
    linear in ln e between nodes (the device's table lookup)."""
    lne0, h, _, T = tab
    f = (np.log(e) - lne0) / h
    i = int(np.clip(np.floor(f), 0, len(T) - 2))
    w = f - i
    return T[i] + w * (T[i + 1] - T[i])
