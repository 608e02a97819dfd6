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
VAH_ERR_ARG, VAH_ERR_HIP, VAH_ERR_NOMEM, VAH_ERR_NAN, VAH_ERR_STATE, VAH_ERR_DEVICE = -1, -2, -3, -4, -5, -6
VAH_ERR_BUDGET = -7
VAH_ERR_FULL = -8
EVENT_LIVE, EVENT_FROZEN, EVENT_STEP_LIMIT, EVENT_NAN = 0, 1, 2, 3
EOS_CONFORMAL, EOS_TABLE, EOS_WB = 0, 1, 2
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
                ("device", ctypes.c_int), ("stream", ctypes.c_void_p), ("eta_chunk", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("max_e", ctypes.c_double), ("max_T", ctypes.c_double), ("total_energy", ctypes.c_double),
                ("max_vT", ctypes.c_double), ("recovery_fallbacks", ctypes.c_longlong),
                ("nan_cells", ctypes.c_longlong)]


class Profile(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_int * 2), ("avg_ms", ctypes.c_double * 2)]


class BatchOptions(ctypes.Structure):
    """vah_batch_options of include/vah_batch.h."""
    _fields_ = [("eos", ctypes.c_int), ("table", EosTable), ("precision", ctypes.c_int),
                ("budget_bytes", ctypes.c_size_t), ("max_events", ctypes.c_int),
                ("freezeout_T_GeV", ctypes.c_double), ("steps_per_graph", ctypes.c_int), ("variant", ctypes.c_int),
                ("device", ctypes.c_int), ("stream", ctypes.c_void_p)]


class BatchEvent(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
                ("t0", ctypes.c_double), ("etabar", ctypes.c_double), ("max_steps", ctypes.c_longlong),
                ("tag", ctypes.c_longlong)]


class BatchStatus(ctypes.Structure):
    _fields_ = [("slot", ctypes.c_int), ("state", ctypes.c_int), ("tag", ctypes.c_longlong),
                ("nstep", ctypes.c_longlong), ("tau", ctypes.c_double), ("max_T", ctypes.c_double),
                ("recovery_fallbacks", ctypes.c_longlong), ("nan_cells", ctypes.c_longlong)]


class ServeResult(ctypes.Structure):
    """vah_request_result of include/vah_serve.h."""
    _fields_ = [("tag", ctypes.c_longlong), ("slot", ctypes.c_int), ("state", ctypes.c_int),
                ("nstep", ctypes.c_longlong), ("tau", ctypes.c_double), ("admitted_s", ctypes.c_double),
                ("finished_s", ctypes.c_double), ("latency_s", ctypes.c_double)]


class ServeStats(ctypes.Structure):
    """vah_serve_stats of include/vah_serve.h."""
    _fields_ = [("served", ctypes.c_int), ("rejected", ctypes.c_int), ("peak_resident", ctypes.c_int),
                ("polls", ctypes.c_int), ("seconds", ctypes.c_double), ("events_per_s", ctypes.c_double),
                ("p50_latency_s", ctypes.c_double), ("p99_latency_s", ctypes.c_double)]


class BatchVariant(ctypes.Structure):
    _fields_ = [("tile_x", ctypes.c_int), ("tile_y", ctypes.c_int), ("blocks_per_cu", ctypes.c_int),
                ("kz", ctypes.c_int)]


class DeviceState(ctypes.Structure):
    """vah_device_state: caller-owned device buffers in the padded layout."""
    _fields_ = [(k, ctypes.c_void_p) for k in ("q", "q_next", "e", "p", "u", "u_prev")]


NOALLOC_SCRATCH_BYTES = 64 << 20


def field_stride(nx, ny, nz):
    """VAH_FIELD_STRIDE: doubles between two fields of the padded device layout."""
    return ((nx + 4) * (ny + 4) * (nz + 4 if nz > 1 else 1) + 63) // 64 * 64


def padded_index(nx, ny, nz):
    """Padded-layout offset of every interior cell (interior order, x fastest)."""
    i = np.arange(nx)
    j = np.arange(ny)
    k = np.arange(nz) + 2 if nz > 1 else np.zeros(1, dtype=np.int64)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    return ((K * (ny + 4) + J + 2) * (nx + 4) + I + 2).ravel()


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


class ServeRequest(ctypes.Structure):
    """vah_request of include/vah_serve.h (with its vah_glauber member)."""
    _fields_ = [("ev", BatchEvent), ("arrival_s", ctypes.c_double), ("e", ctypes.POINTER(ctypes.c_double)),
                ("glauber", Glauber), ("pl_ratio", ctypes.c_double)]


RETIRE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(ServeResult))
ADMIT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ServeRequest), ctypes.POINTER(ctypes.c_int))
STEP_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int)
POLL_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(BatchStatus), ctypes.c_int,
                           ctypes.POINTER(ctypes.c_int))
RETIRE_SLOT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int)
NOW_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p)
SLEEP_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_double)
ON_POLL_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int)


class ServeBackend(ctypes.Structure):
    """vah_serve_backend of include/vah_serve.h."""
    _fields_ = [("user", ctypes.c_void_p), ("admit", ADMIT_FN), ("step", STEP_FN), ("poll", POLL_FN),
                ("retire", RETIRE_SLOT_FN), ("now", NOW_FN), ("sleep", SLEEP_FN), ("on_poll", ON_POLL_FN)]


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
        L.vah_validity.argtypes = [ctypes.c_void_p] + [_dp] * 5
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
        L.vah_write_dat.argtypes = [ctypes.POINTER(Config), _dp, ctypes.c_double, ctypes.c_char_p, ctypes.c_char_p]
        L.vah_step_noalloc.argtypes = [ctypes.POINTER(Lattice), ctypes.POINTER(Hydro), ctypes.POINTER(Options),
                                       ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(DeviceState),
                                       ctypes.c_void_p]
        # batch engine (include/vah_batch.h)
        L.vah_batch_default_options.argtypes = [ctypes.POINTER(BatchOptions)]
        L.vah_batch_default_options.restype = None
        L.vah_batch_event_bytes.argtypes = [ctypes.c_int] * 3
        L.vah_batch_event_bytes.restype = ctypes.c_size_t
        L.vah_batch_create.argtypes = [ctypes.POINTER(BatchOptions), ctypes.POINTER(ctypes.c_void_p)]
        L.vah_batch_admit.argtypes = [ctypes.c_void_p, ctypes.POINTER(BatchEvent)] + [_dp] * 7 + \
                                     [ctypes.POINTER(ctypes.c_int)]
        L.vah_batch_admit_glauber.argtypes = [ctypes.c_void_p, ctypes.POINTER(BatchEvent), ctypes.POINTER(Glauber),
                                              ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
        L.vah_serve.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServeRequest), ctypes.c_int, ctypes.c_int, RETIRE_FN,
                                ctypes.c_void_p, ctypes.POINTER(ServeResult), ctypes.POINTER(ServeStats)]
        L.vah_serve_run.argtypes = [ctypes.POINTER(ServeBackend), ctypes.POINTER(ServeRequest), ctypes.c_int,
                                    ctypes.c_int, RETIRE_FN, ctypes.c_void_p, ctypes.POINTER(ServeResult),
                                    ctypes.POINTER(ServeStats)]
        L.vah_batch_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.vah_batch_synchronize.argtypes = [ctypes.c_void_p]
        L.vah_batch_poll.argtypes = [ctypes.c_void_p, ctypes.POINTER(BatchStatus), ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int)]
        L.vah_batch_get_state.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_dp] * 5
        L.vah_batch_retire.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.vah_batch_memory.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t),
                                       ctypes.POINTER(ctypes.c_size_t)]
        L.vah_batch_variant_count.argtypes = []
        L.vah_batch_variant_info.argtypes = [ctypes.c_int, ctypes.POINTER(BatchVariant)]
        L.vah_batch_set_variant.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.vah_batch_get_variant.argtypes = [ctypes.c_void_p]
        L.vah_batch_autotune.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp]
        L.vah_batch_destroy.argtypes = [ctypes.c_void_p]
        L.vah_batch_destroy.restype = None
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


def write_dat(cfg, var, tau, directory, name):
    """One interior field as <directory>/<name>_<tau>.dat in the reference's text format."""
    v = np.ascontiguousarray(var, dtype=np.float64).ravel()
    _check(lib().vah_write_dat(ctypes.byref(cfg), v.ctypes.data_as(_dp), tau, directory.encode(), name.encode()))


def ic_reference(cfg):
    n = cfg.numLatticePointsX * cfg.numLatticePointsY * cfg.numLatticePointsRapidity
    arrs = [np.empty(n) for _ in range(5)]
    _check(lib().vah_ic_reference(ctypes.byref(cfg), *[a.ctypes.data_as(_dp) for a in arrs]))
    return arrs


class Vah:
    """One lattice evolved on the GPU through the C ABI."""

    def __init__(self, nx, ny, nz, dx, dy, dz, dt, etabar, t0, eos_table=None, flux_mode=FLUX_SHARED,
                 precision=FP64, device=0, nt=0, freezeout_T_GeV=0.155, eos_kind=None, eta_chunk=0):
        """eos_table (lne0, h, p, T) selects the tabulated EOS; eos_kind=EOS_WB the
        reference's Wuppertal-Budapest fits; neither: the conformal EOS."""
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
        opt.eta_chunk = eta_chunk
        self._keep = []
        if eos_table is not None:
            lne0, h, p, T = eos_table
            p, pp = _arr(p)
            T, pT = _arr(T)
            self._keep += [p, T]
            opt.eos = EOS_TABLE
            opt.table = EosTable(len(p), lne0, h, pp, pT)
        if eos_kind is not None:
            opt.eos = eos_kind
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

    def validity(self, copy=True):
        """Knudsen numbers (theta_T, zDzu, theta), R_pi and R_pi2 after the last step.
        copy=False runs the device pass alone (no host copies) and returns None."""
        keys = ("kn_T", "kn_L", "kn", "r_pi", "r_pi2")
        if not copy:
            _check(lib().vah_validity(self.h, *[None] * len(keys)))
            return None
        out = {k: np.empty(self.n) for k in keys}
        _check(lib().vah_validity(self.h, *[out[k].ctypes.data_as(_dp) for k in keys]))
        return out

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


def batch_event_bytes(nx, ny, nz):
    """Arena bytes one resident event of nx x ny x nz interior cells takes."""
    return lib().vah_batch_event_bytes(nx, ny, nz)


def batch_variants():
    """The compiled kernel variants: (tile_x, tile_y, blocks_per_cu, kz)."""
    out = []
    for i in range(lib().vah_batch_variant_count()):
        v = BatchVariant()
        _check(lib().vah_batch_variant_info(i, ctypes.byref(v)))
        out.append((v.tile_x, v.tile_y, v.blocks_per_cu, v.kz))
    return out


class Batch:
    """Many resident events advanced together (include/vah_batch.h)."""

    def __init__(self, budget_bytes=8 << 30, max_events=1024, eos_table=None, precision=FP64,
                 freezeout_T_GeV=0.150, steps_per_graph=16, variant=-1, device=0, eos_kind=None):
        L = lib()
        opt = BatchOptions()
        L.vah_batch_default_options(ctypes.byref(opt))
        opt.budget_bytes = int(budget_bytes)
        opt.max_events = max_events
        opt.precision = precision
        opt.freezeout_T_GeV = freezeout_T_GeV
        opt.steps_per_graph = steps_per_graph
        opt.variant = variant
        opt.device = device
        self._keep = []
        if eos_table is not None:
            lne0, h, p, T = eos_table
            p, pp = _arr(p)
            T, pT = _arr(T)
            self._keep += [p, T]
            opt.eos = EOS_TABLE
            opt.table = EosTable(len(p), lne0, h, pp, pT)
        if eos_kind is not None:
            opt.eos = eos_kind
        h = ctypes.c_void_p()
        _check(L.vah_batch_create(ctypes.byref(opt), ctypes.byref(h)))
        self.h = h
        self.max_events = max_events
        self.dims = {}

    def close(self):
        if getattr(self, "h", None):
            lib().vah_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, nx, ny, nz, dx, dy, dz, dt, t0, etabar, e, ut=None, ux=None, uy=None, un=None, pl=None, pi=None,
              max_steps=0, tag=0, raise_on_full=True):
        """Admit one event; returns its slot, or None when the arena or the
        slot table is full and raise_on_full is False."""
        # ut = ux = uy = un = None: the rest frame (set on the device)
        ev = BatchEvent(nx, ny, nz, dx, dy, dz, dt, t0, etabar, max_steps, tag)
        args = [_arr(x) for x in (e, pl, ut, ux, uy, un, pi)]
        slot = ctypes.c_int(-1)
        rc = lib().vah_batch_admit(self.h, ctypes.byref(ev), *[a[1] for a in args], ctypes.byref(slot))
        if rc in (VAH_ERR_BUDGET, VAH_ERR_FULL) and not raise_on_full:
            return None
        _check(rc)
        self.dims[slot.value] = (nx, ny, nz)
        return slot.value

    def admit_glauber(self, nx, ny, nz, dx, dy, dz, dt, t0, etabar, g, pl_ratio=1.0 / 3.0, max_steps=0, tag=0,
                      raise_on_full=True):
        """Admit one event whose Monte-Carlo Glauber initial state (vah_glauber g)
        is generated on the device; P_L = pl_ratio e.  Slot, or None when full
        and raise_on_full is False."""
        ev = BatchEvent(nx, ny, nz, dx, dy, dz, dt, t0, etabar, max_steps, tag)
        slot = ctypes.c_int(-1)
        rc = lib().vah_batch_admit_glauber(self.h, ctypes.byref(ev), ctypes.byref(g), pl_ratio, ctypes.byref(slot))
        if rc in (VAH_ERR_BUDGET, VAH_ERR_FULL) and not raise_on_full:
            return None
        _check(rc)
        self.dims[slot.value] = (nx, ny, nz)
        return slot.value

    def step(self, n=1):
        _check(lib().vah_batch_step(self.h, n))

    def synchronize(self):
        _check(lib().vah_batch_synchronize(self.h))

    def poll(self):
        """Status of every resident event (dicts, slot order)."""
        buf = (BatchStatus * self.max_events)()
        n = ctypes.c_int()
        _check(lib().vah_batch_poll(self.h, buf, self.max_events, ctypes.byref(n)))
        return [{k: getattr(buf[i], k) for k, _ in BatchStatus._fields_} for i in range(n.value)]

    def state(self, slot, fields=("q", "e", "p", "u", "up")):
        """Fields of a resident event at its current step (only those named)."""
        nx, ny, nz = self.dims[slot]
        n = nx * ny * nz
        shape = {"q": (NQ, n), "e": (n,), "p": (n,), "u": (4, n), "up": (4, n)}
        out = {k: np.empty(shape[k]) for k in fields}
        ptrs = [out[k].ctypes.data_as(_dp) if k in out else None for k in ("q", "e", "p", "u", "up")]
        _check(lib().vah_batch_get_state(self.h, slot, *ptrs))
        return out

    def retire(self, slot):
        _check(lib().vah_batch_retire(self.h, slot))
        self.dims.pop(slot, None)

    def memory(self):
        used, budget = ctypes.c_size_t(), ctypes.c_size_t()
        _check(lib().vah_batch_memory(self.h, ctypes.byref(used), ctypes.byref(budget)))
        return used.value, budget.value

    @property
    def variant(self):
        return lib().vah_batch_get_variant(self.h)

    @variant.setter
    def variant(self, v):
        _check(lib().vah_batch_set_variant(self.h, v))

    def autotune(self, steps_per_trial=4):
        """Time every kernel variant on the resident events (real steps);
        returns (best variant, ms per step of each)."""
        nv = lib().vah_batch_variant_count()
        ms = np.zeros(nv)
        best = ctypes.c_int()
        _check(lib().vah_batch_autotune(self.h, steps_per_trial, ctypes.byref(best), ms.ctypes.data_as(_dp)))
        return best.value, ms
