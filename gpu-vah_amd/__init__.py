"""Python front end of libvah.so (the C ABI in include/vah.h).

This is plumbing for tests and benchmarks: it allocates nothing on the device
itself and calls straight into the HIP library.  Import it with
``importlib.import_module("gpu-vah_amd")`` (the directory name is not a
Python identifier).  Loading fails loudly when the library has not been built:
there is no CPU fallback behind this interface.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("VAH_LIB") or os.path.join(HERE, "libvah.so")

NQ = 15
FIELDS = ("ttt", "ttx", "tty", "ttn", "pl", "pitt", "pitx", "pity", "pitn",
          "pixx", "pixy", "pixn", "piyy", "piyn", "pinn")

VAH_OK = 0
VAH_ERR_NAN = -4
EOS_CONFORMAL, EOS_TABLE = 0, 1
FP64, FP32_SPLIT = 0, 1
FLUX_SHARED, FLUX_PER_CELL = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)


class Lattice(ctypes.Structure):
    """rhic/include/LatticeParameters.h:17-31 (ghost cells implicit)."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("nt", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double)]


class Hydro(ctypes.Structure):
    """rhic/include/HydroParameters.h:15-21."""
    _fields_ = [("t0", ctypes.c_double), ("etabar", ctypes.c_double), ("freezeout_T_GeV", ctypes.c_double),
                ("init_pimunu_ns", ctypes.c_int)]


class EosTable(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("lne0", ctypes.c_double), ("h", ctypes.c_double), ("p", _dp), ("T", _dp)]


class Options(ctypes.Structure):
    _fields_ = [("eos", ctypes.c_int), ("table", EosTable), ("precision", ctypes.c_int), ("flux_mode", ctypes.c_int),
                ("device", ctypes.c_int), ("stream", ctypes.c_void_p)]


class Stats(ctypes.Structure):
    _fields_ = [("max_e", ctypes.c_double), ("max_T", ctypes.c_double), ("total_energy", ctypes.c_double),
                ("max_vT", ctypes.c_double), ("recovery_fallbacks", ctypes.c_longlong),
                ("nan_cells", ctypes.c_longlong)]


class Profile(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_int * 2), ("avg_ms", ctypes.c_double * 2)]


class VahError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vah error {code}: {msg}")
        self.code = code


_lib = None

class Glauber(ctypes.Structure):
    """vah_glauber of include/vah_host.h."""
    _fields_ = [("A", ctypes.c_int), ("R", ctypes.c_double), ("a", ctypes.c_double),
                ("sigma_nn_mb", ctypes.c_double), ("b", ctypes.c_double), ("alpha", ctypes.c_double),
                ("e_peak", ctypes.c_double), ("nucleon_width", ctypes.c_double), ("seed", ctypes.c_ulonglong),
                ("eta_flat", ctypes.c_double), ("eta_sigma", ctypes.c_double), ("e_floor", ctypes.c_double)]


class Config(ctypes.Structure):
    """vah_config of include/vah_host.h (the reference's three property files)."""
    _fields_ = [(n, ctypes.c_int) for n in ("numLatticePointsX", "numLatticePointsY", "numLatticePointsRapidity",
                                            "numProperTimePoints")] + \
               [(n, ctypes.c_double) for n in ("latticeSpacingX", "latticeSpacingY", "latticeSpacingRapidity",
                                               "latticeSpacingProperTime")] + \
               [(n, ctypes.c_int) for n in ("initialConditionType", "numberOfNucleonsPerNuclei")] + \
               [(n, ctypes.c_double) for n in ("initialEnergyDensity", "scatteringCrossSectionNN",
                                               "impactParameter", "fractionOfBinaryCollisions",
                                               "rapidityVariance", "rapidityMean", "initialProperTimePoint",
                                               "shearViscosityToEntropyDensity", "freezeoutTemperatureGeV")] + \
               [("initializePimunuNavierStokes", ctypes.c_int)]


HBARC = 0.197326938


def build():
    """Compile libvah.so for gfx950 (hipcc cross-compiles without a GPU)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        L.vah_default_options.argtypes = [ctypes.POINTER(Options)]
        L.vah_create.argtypes = [ctypes.POINTER(Lattice), ctypes.POINTER(Hydro), ctypes.POINTER(Options),
                                 ctypes.POINTER(ctypes.c_void_p)]
        L.vah_set_primitives.argtypes = [ctypes.c_void_p] + [_dp] * 7
        L.vah_set_state.argtypes = [ctypes.c_void_p, ctypes.c_longlong] + [_dp] * 5
        L.vah_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.vah_synchronize.argtypes = [ctypes.c_void_p]
        L.vah_get_state.argtypes = [ctypes.c_void_p] + [_dp] * 5
        L.vah_get_time.argtypes = [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_longlong)]
        L.vah_reduce.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats)]
        L.vah_profile_enable.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.vah_profile_read.argtypes = [ctypes.c_void_p, ctypes.POINTER(Profile)]
        L.vah_destroy.argtypes = [ctypes.c_void_p]
        L.vah_destroy.restype = None
        L.vah_last_error.restype = ctypes.c_char_p
        L.vah_build_info.restype = ctypes.c_char_p
        L.vah_eos_lqcd_table.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp]
        L.vah_eos_lqcd_eval.argtypes = [ctypes.c_double, _dp, _dp]
        L.vah_eos_lqcd_eval.restype = None
        L.vah_ic_optical_glauber.argtypes = [ctypes.POINTER(Lattice), ctypes.POINTER(Glauber), _dp]
        L.vah_ic_mc_glauber.argtypes = [ctypes.POINTER(Lattice), ctypes.POINTER(Glauber), _dp,
                                        ctypes.POINTER(ctypes.c_int)]
        L.vah_ic_eta_profile.argtypes = [ctypes.c_double] * 3
        L.vah_ic_eta_profile.restype = ctypes.c_double
        L.vah_config_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(Config)]
        L.vah_ic_reference.argtypes = [ctypes.POINTER(Config)] + [_dp] * 5
        L.vah_run.argtypes = [ctypes.POINTER(Config), ctypes.c_char_p, ctypes.c_int]
        _lib = L
    return _lib


def _check(rc):
    if rc != VAH_OK:
        raise VahError(rc, lib().vah_last_error().decode())


def _arr(a):
    if a is None:
        return None, None
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    return a, a.ctypes.data_as(_dp)


def eos_lqcd_table(n=10000, e_min=1e-8, e_max=1e4):
    """(lne0, h, p, T) of the crossover EOS tabulated at n nodes uniform in ln e."""
    p, T = np.empty(n), np.empty(n)
    lne0, h = ctypes.c_double(), ctypes.c_double()
    _check(lib().vah_eos_lqcd_table(n, e_min, e_max, p.ctypes.data_as(_dp), T.ctypes.data_as(_dp),
                                    ctypes.byref(lne0), ctypes.byref(h)))
    return lne0.value, h.value, p, T


def eos_lqcd(T):
    e, p = ctypes.c_double(), ctypes.c_double()
    lib().vah_eos_lqcd_eval(T, ctypes.byref(e), ctypes.byref(p))
    return e.value, p.value


def glauber(**kw):
    """A vah_glauber with the 200 GeV Au+Au optical defaults, overridden by kw."""
    g = Glauber(A=197, R=6.38, a=0.535, sigma_nn_mb=42.0, b=7.0, alpha=0.5, e_peak=30.0 / HBARC,
                nucleon_width=0.5, seed=1, eta_flat=1.5, eta_sigma=1.4, e_floor=1e-3)
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def ic_optical_glauber(nx, ny, nz, dx, dz, g):
    lat = Lattice(nx, ny, nz, 0, dx, dx, dz, 0.01)
    e = np.empty(nx * ny * nz)
    _check(lib().vah_ic_optical_glauber(ctypes.byref(lat), ctypes.byref(g), e.ctypes.data_as(_dp)))
    return e


def ic_mc_glauber(nx, ny, nz, dx, dz, g):
    lat = Lattice(nx, ny, nz, 0, dx, dx, dz, 0.01)
    e = np.empty(nx * ny * nz)
    nw = ctypes.c_int()
    _check(lib().vah_ic_mc_glauber(ctypes.byref(lat), ctypes.byref(g), e.ctypes.data_as(_dp), ctypes.byref(nw)))
    return e, nw.value


def config_load(directory):
    c = Config()
    _check(lib().vah_config_load(directory.encode(), ctypes.byref(c)))
    return c


def ic_reference(cfg):
    n = cfg.numLatticePointsX * cfg.numLatticePointsY * cfg.numLatticePointsRapidity
    arrs = [np.empty(n) for _ in range(5)]
    _check(lib().vah_ic_reference(ctypes.byref(cfg), *[a.ctypes.data_as(_dp) for a in arrs]))
    return arrs


class Vah:
    """One lattice evolved on the GPU through the C ABI."""

    def __init__(self, nx, ny, nz, dx, dy, dz, dt, etabar, t0, eos_table=None, flux_mode=FLUX_SHARED,
                 precision=FP64, device=0, nt=0, freezeout_T_GeV=0.155):
        L = lib()
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = nx * ny * nz
        lat = Lattice(nx, ny, nz, nt, dx, dy, dz, dt)
        hyd = Hydro(t0, etabar, freezeout_T_GeV, 0)
        opt = Options()
        L.vah_default_options(ctypes.byref(opt))
        opt.flux_mode = flux_mode
        opt.precision = precision
        opt.device = device
        self._keep = []
        if eos_table is not None:
            lne0, h, p, T = eos_table
            p, pp = _arr(p)
            T, pT = _arr(T)
            self._keep += [p, T]
            opt.eos = EOS_TABLE
            opt.table = EosTable(len(p), lne0, h, pp, pT)
        h = ctypes.c_void_p()
        _check(L.vah_create(ctypes.byref(lat), ctypes.byref(hyd), ctypes.byref(opt), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().vah_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_primitives(self, e, ut, ux, uy, un, pl=None, pi=None):
        args = [_arr(x) for x in (e, pl, ut, ux, uy, un, pi)]
        _check(lib().vah_set_primitives(self.h, *[a[1] for a in args]))

    def set_state(self, nstep, q, e, p, u, up):
        args = [_arr(x) for x in (q, e, p, u, up)]
        _check(lib().vah_set_state(self.h, nstep, *[a[1] for a in args]))

    def step(self, n=1):
        _check(lib().vah_step(self.h, n))

    def synchronize(self):
        _check(lib().vah_synchronize(self.h))

    def state(self):
        n = self.n
        q, e, p = np.empty(NQ * n), np.empty(n), np.empty(n)
        u, up = np.empty(4 * n), np.empty(4 * n)
        _check(lib().vah_get_state(self.h, *[x.ctypes.data_as(_dp) for x in (q, e, p, u, up)]))
        return {"q": q.reshape(NQ, n), "e": e, "p": p, "u": u.reshape(4, n), "up": up.reshape(4, n)}

    @property
    def time(self):
        t = ctypes.c_double()
        n = ctypes.c_longlong()
        _check(lib().vah_get_time(self.h, ctypes.byref(t), ctypes.byref(n)))
        return t.value

    def stats(self):
        s = Stats()
        _check(lib().vah_reduce(self.h, ctypes.byref(s)))
        return s

    def profile(self, on=True):
        _check(lib().vah_profile_enable(self.h, 1 if on else 0))

    def profile_read(self):
        pr = Profile()
        _check(lib().vah_profile_read(self.h, ctypes.byref(pr)))
        return {"launches": list(pr.launches), "avg_ms": list(pr.avg_ms)}
