"""ctypes front end of the CPU oracle (oracle/vah_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
for the expected answer and by ``bench.py``'s CPU-baseline leg.  The product path never
imports it.

The oracle restates the reference's VAH Heun-RK2 step
(old_cpu-vah_code/src/FullyDiscreteKurganovTadmorScheme.cpp:738-792) on the
CPU in double precision; see the header of vah_oracle.c for the file list.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libvahref.so")

NQ = 15
FIELDS = ("ttt", "ttx", "tty", "ttn", "pl", "pitt", "pitx", "pity", "pitn",
          "pixx", "pixy", "pixn", "piyy", "piyn", "pinn")

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None
_ref = None


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and the reference build when ref=True)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 6 + [ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_eos_table.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_set_eos_table.restype = ctypes.c_int
        L.orc_set_primitives.argtypes = [ctypes.c_void_p] + [_dp] * 7
        L.orc_set_state.argtypes = [ctypes.c_void_p, ctypes.c_long] + [_dp] * 5
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_pairing.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_time.argtypes = [ctypes.c_void_p]
        L.orc_time.restype = ctypes.c_double
        L.orc_newton_fail.argtypes = [ctypes.c_void_p]
        L.orc_newton_fail.restype = ctypes.c_long
        L.orc_nan_hits.argtypes = [ctypes.c_void_p]
        L.orc_nan_hits.restype = ctypes.c_long
        L.orc_get_state.argtypes = [ctypes.c_void_p] + [_dp] * 5
        L.orc_validity.argtypes = [ctypes.c_void_p] + [_dp] * 5
        L.orc_validity.restype = ctypes.c_int
        L.orc_set_eos_wb.argtypes = [ctypes.c_void_p]
        L.orc_set_eos_wb.restype = None
        for name in ("orc_wb_pressure", "orc_wb_temperature"):
            getattr(L, name).argtypes = [ctypes.c_double]
            getattr(L, name).restype = ctypes.c_double
        for name in ("orc_temperature_conformal", "orc_rtilde", "orc_rbar0", "orc_rgamma", "orc_rbar0p"):
            getattr(L, name).argtypes = [ctypes.c_double]
            getattr(L, name).restype = ctypes.c_double
        L.orc_slope.argtypes = [ctypes.c_double] * 3
        L.orc_slope.restype = ctypes.c_double
        L.orc_pt_hat.argtypes = [ctypes.c_double] * 3
        L.orc_pt_hat.restype = ctypes.c_double
        L.orc_recover.argtypes = [ctypes.c_double, _dp, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_recover.restype = ctypes.c_int
        L.orc_source_block.argtypes = [_dp, _dp, _dp] + [ctypes.c_double] * 6 + [_dp]
        L.orc_source_axis.argtypes = [_dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_regulate_cell.argtypes = [ctypes.c_double, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def ref_lib():
    """The reference build (oracle/_ref), or None where it cannot be built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_LIB_PATH):
            if not os.path.isdir("/root/reference"):
                return None
            build(ref=True)
        R = ctypes.CDLL(REF_LIB_PATH)
        R.ref_source_block.argtypes = [_dp, _dp, _dp] + [ctypes.c_double] * 6 + [_dp]
        R.ref_source_axis.argtypes = [_dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, _dp]
        for name in ("ref_pressure", "ref_temperature", "ref_energy_density", "ref_rbar0", "ref_rbar0p"):
            getattr(R, name).argtypes = [ctypes.c_double]
            getattr(R, name).restype = ctypes.c_double
        R.ref_pt_hat.argtypes = [ctypes.c_double] * 3
        R.ref_pt_hat.restype = ctypes.c_double
        R.ref_approx_derivative.argtypes = [ctypes.c_double] * 3
        R.ref_approx_derivative.restype = ctypes.c_double
        R.ref_halfsite.argtypes = [ctypes.c_int] + [ctypes.c_double] * 5
        R.ref_halfsite.restype = ctypes.c_double
        _ref = R
    return _ref


class Oracle:
    """One lattice evolved by the CPU restatement.

    Arrays crossing this interface are interior-only, x fastest
    (index = i + nx*(j + ny*k)), exactly as the product's C-ABI takes them.
    """

    def __init__(self, nx, ny, nz, dx, dy, dz, dt, etabar, t0, threads=None, eos_table=None, eos_wb=False):
        L = lib()
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = nx * ny * nz
        if threads is None:
            threads = min(16, os.cpu_count() or 1)
        self.h = L.orc_create(nx, ny, nz, dx, dy, dz, dt, etabar, t0, threads)
        if not self.h:
            raise MemoryError("oracle allocation failed")
        if eos_table is not None:
            lne0, h, p, T = eos_table
            p = np.ascontiguousarray(p, dtype=np.float64)
            T = np.ascontiguousarray(T, dtype=np.float64)
            if L.orc_set_eos_table(self.h, len(p), lne0, h, _ptr(p), _ptr(T)) != 0:
                raise MemoryError("oracle EOS table allocation failed")
        if eos_wb:
            L.orc_set_eos_wb(self.h)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_threads(self, n):
        lib().orc_set_threads(self._handle(), n)

    def set_primitives(self, e, ut, ux, uy, un, pl=None, pi=None):
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()
        e, ut, ux, uy, un, pl, pi = map(c, (e, ut, ux, uy, un, pl, pi))
        lib().orc_set_primitives(self._handle(), _ptr(e), _ptr(pl), _ptr(ut), _ptr(ux), _ptr(uy), _ptr(un), _ptr(pi))

    def set_state(self, nstep, q, e, p, u, up):
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64).ravel()
        q, e, p, u, up = map(c, (q, e, p, u, up))
        lib().orc_set_state(self._handle(), nstep, _ptr(q), _ptr(e), _ptr(p), _ptr(u), _ptr(up))

    def _handle(self):
        if not self.h:
            raise RuntimeError("oracle lattice already closed")
        return self.h

    def set_pairing(self, mode):
        """Test-only: 0 the reference's velocity pairing in the sources' time
        derivative; 1 pairs u* with u^{n-1} in stage 2; 2 drops the stage-1
        time derivative.  Modes 1 and 2 are planted mistakes."""
        lib().orc_set_pairing(self._handle(), mode)

    def step(self, n=1):
        lib().orc_step(self._handle(), n)

    @property
    def time(self):
        return lib().orc_time(self._handle())

    @property
    def newton_fail(self):
        return lib().orc_newton_fail(self._handle())

    @property
    def nan_hits(self):
        return lib().orc_nan_hits(self._handle())

    def validity(self):
        """Knudsen numbers (theta_T, zDzu, theta), R_pi and the second-order
        R_pi2 after the last step."""
        n = self.n
        keys = ("kn_T", "kn_L", "kn", "r_pi", "r_pi2")
        out = {k: np.empty(n) for k in keys}
        if lib().orc_validity(self._handle(), *[_ptr(out[k]) for k in keys]) != 0:
            raise RuntimeError("validity needs a completed step")
        return out

    def state(self):
        n = self.n
        q = np.empty(NQ * n)
        e = np.empty(n)
        p = np.empty(n)
        u = np.empty(4 * n)
        up = np.empty(4 * n)
        lib().orc_get_state(self._handle(), _ptr(q), _ptr(e), _ptr(p), _ptr(u), _ptr(up))
        return {"q": q.reshape(NQ, n), "e": e, "p": p, "u": u.reshape(4, n), "up": up.reshape(4, n)}
