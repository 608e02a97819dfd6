"""Seeded random inputs of physical shape for the unit pins."""
import numpy as np


def random_source_block(rng, flow=True, pi=True, longitudinal=True):
    """Smooth fields on a 3x3x3 block (x fastest) around a random centre.

    Layout of ``fields``: ut, ux, uy, un, e, p, pl -- 27 values each, the
    argument order of orc_source_block / ref_source_block.
    """
    t = rng.uniform(0.5, 4.0)
    off = np.array([(i - 1, j - 1, k - 1) for k in range(3) for j in range(3) for i in range(3)], dtype=float)

    def smooth(c0, grad_scale, curv=0.0):
        g = rng.standard_normal(3) * grad_scale
        return c0 + off @ g + curv * rng.standard_normal(27)

    s = 1.0 if flow else 0.0
    ux = smooth(s * rng.normal(0, 0.6), s * 0.05, s * 0.005)
    uy = smooth(s * rng.normal(0, 0.6), s * 0.05, s * 0.005)
    sl = 1.0 if longitudinal else 0.0
    un = smooth(sl * rng.normal(0, 0.2) / t, sl * 0.02 / t, sl * 0.002 / t)
    ut = np.sqrt(1 + ux * ux + uy * uy + t * t * un * un)
    e = np.abs(smooth(10 ** rng.uniform(-1, 2), 0.0, 0.0)) * np.exp(smooth(0.0, 0.05, 0.01))
    p = e / 3
    pl = e * np.clip(smooth(rng.uniform(0.05, 0.6), 0.01, 0.002), 0.01, 0.99)
    Q = np.zeros(15)
    Q[0:4] = rng.uniform(0.5, 2.0, 4) * e[13]
    Q[4] = pl[13]
    if pi:
        Q[5:15] = rng.normal(0, 0.05, 10) * e[13]
        Q[8] /= t
        Q[11] /= t
        Q[13] /= t
        Q[14] /= t * t
    uprev = np.array([ut[13], ux[13], uy[13], un[13]]) + rng.normal(0, 1e-3, 4) * np.array([1, s, s, sl / t])
    fields = np.concatenate([ut, ux, uy, un, e, p, pl]).astype(np.float64)
    return {"Q": Q, "fields": fields, "uprev": uprev, "t": t, "etabar": rng.uniform(0.08, 0.3),
            "dt": rng.choice([0.005, 0.01, 0.02]), "dx": rng.choice([0.08, 0.1, 0.12]),
            "dy": rng.choice([0.08, 0.1, 0.12]), "dz": rng.choice([0.1, 0.2])}


def random_axis_gather(rng):
    """A 5 x 15 stencil gather (field-major, 5 per field) and local u."""
    g = rng.standard_normal(75) * 0.3 + 1.0
    ut = rng.uniform(1.0, 3.0)
    ui = rng.normal(0, 1.0)
    return g, ut, ui, rng.uniform(0.5, 4.0), rng.choice([0.08, 0.1, 0.12])
