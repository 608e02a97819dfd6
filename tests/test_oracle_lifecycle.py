"""The oracle's handle after close(): every method raises instead of passing a
freed lattice to the C library (the regression test for a use-after-close
that once crashed the process inside orc_time)."""
import numpy as np
import pytest


def test_every_method_raises_after_close(oracle_mod):
    nx, ny = 4, 3
    n = nx * ny
    o = oracle_mod.Oracle(nx, ny, 1, 0.1, 0.1, 0.1, 0.01, 0.2, 0.5, threads=1)
    e = np.full(n, 10.0)
    one, zero = np.ones(n), np.zeros(n)
    o.set_primitives(e, one, zero, zero, zero, pl=e / 3)
    o.step(1)
    st = o.state()
    o.close()
    o.close()  # a second close is harmless
    calls = {
        "step": lambda: o.step(1),
        "time": lambda: o.time,
        "newton_fail": lambda: o.newton_fail,
        "nan_hits": lambda: o.nan_hits,
        "state": lambda: o.state(),
        "validity": lambda: o.validity(),
        "set_state": lambda: o.set_state(1, st["q"], st["e"], st["p"], st["u"], st["up"]),
        "set_primitives": lambda: o.set_primitives(e, one, zero, zero, zero),
        "set_threads": lambda: o.set_threads(2),
    }
    for name, call in calls.items():
        with pytest.raises(RuntimeError, match="closed"):
            call()
