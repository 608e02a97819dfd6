"""vah_step_noalloc (include/vah.h), the entry point that allocates nothing:
its intermediate stage lives in the library's static device array.  Checked
through the C ABI against vah_step (a context) on the same initial state,
bit for bit, on two calls in a row on one GPU -- a 3+1D lattice with the
tabulated EOS (table arrays in device memory) in two chunks of steps, then a
smaller boost-invariant lattice with the conformal EOS -- so the second call
reuses scratch the first one left behind.  The test is the caller: it owns
its device buffers, allocated with the HIP runtime libvah.so itself links
(libamdhip64.so.7), and passes them in the library's padded layout."""
import ctypes

import numpy as np
import pytest

from tests.parity import gaussian_ic, rest_frame, vah

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceBuffers:
    """hipMalloc'd buffers of the caller, freed on close()."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.shapes = {}

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        self.shapes[p.value] = a.shape
        return p.value

    def download(self, ptr):
        out = np.empty(self.shapes[ptr])
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.shapes:
            self.hip.hipFree(p)
        self.shapes = {}


def _padded(V, nx, ny, nz, nfields, arr):
    fs = V.field_stride(nx, ny, nz)
    idx = V.padded_index(nx, ny, nz)
    out = np.zeros((nfields, fs))
    out[:, idx] = np.asarray(arr).reshape(nfields, -1)
    return out


def _unpadded(V, nx, ny, nz, nfields, host):
    idx = V.padded_index(nx, ny, nz)
    return host.reshape(nfields, -1)[:, idx]


def _run_pair(nx, ny, nz, dz, dt, steps_chunks, tab):
    """The same initial state through a context and through vah_step_noalloc;
    returns both final states (interior arrays)."""
    V = vah()
    L = V.lib()
    e = gaussian_ic(nx, ny, nz, dz=dz, seed=nx + nz, tilt=0.3, peak=60.0 if tab else 152.0)
    n = nx * ny * nz
    g = V.Vah(nx, ny, nz, 0.1, 0.1, dz, dt, 0.16, 0.5, eos_table=tab)
    g.set_primitives(e, *rest_frame(n), pl=e / 3)
    s0 = g.state()
    g.step(sum(steps_chunks))
    g.synchronize()
    want = g.state()
    g.close()

    D = DeviceBuffers()
    try:
        lat = V.Lattice(nx, ny, nz, 0, 0.1, 0.1, dz, dt)
        hyd = V.Hydro(0.5, 0.16, 0.155, 0)
        opt = V.Options()
        L.vah_default_options(ctypes.byref(opt))
        if tab is not None:
            lne0, h, p, T = tab
            opt.eos = V.EOS_TABLE
            opt.table = V.EosTable(len(p), lne0, h, ctypes.cast(D.upload(p), V._dp), ctypes.cast(D.upload(T), V._dp))
        fs = V.field_stride(nx, ny, nz)
        st = V.DeviceState(D.upload(_padded(V, nx, ny, nz, 15, s0["q"])), D.upload(np.zeros((15, fs))),
                           D.upload(_padded(V, nx, ny, nz, 1, s0["e"])), D.upload(_padded(V, nx, ny, nz, 1, s0["p"])),
                           D.upload(_padded(V, nx, ny, nz, 4, s0["u"])), D.upload(_padded(V, nx, ny, nz, 4, s0["up"])))
        nstep = 0
        for k in steps_chunks:
            rc = L.vah_step_noalloc(ctypes.byref(lat), ctypes.byref(hyd), ctypes.byref(opt), nstep, k,
                                    ctypes.byref(st), None)
            assert rc == V.VAH_OK, L.vah_last_error().decode()
            nstep += k
        got = {
            "q": _unpadded(V, nx, ny, nz, 15, D.download(st.q)),
            "e": _unpadded(V, nx, ny, nz, 1, D.download(st.e))[0],
            "p": _unpadded(V, nx, ny, nz, 1, D.download(st.p))[0],
            "u": _unpadded(V, nx, ny, nz, 4, D.download(st.u)),
            "up": _unpadded(V, nx, ny, nz, 4, D.download(st.u_prev)),
        }
    finally:
        D.close()
    return got, want


def _assert_same(got, want, label):
    for k in ("q", "e", "p", "u", "up"):
        assert np.array_equal(got[k], want[k]), f"{label}: field set {k} differs, max " \
            f"{np.max(np.abs(got[k] - want[k]))}"


def test_noalloc_step_matches_the_context_path_on_two_calls():
    V = vah()
    tab = V.eos_lqcd_table(10000, 1e-8, 1e4)
    # first call: 3+1D, table EOS, 10 steps in chunks of 4 and 6
    got, want = _run_pair(40, 36, 12, 0.3, 0.02, [4, 6], tab)
    _assert_same(got, want, "3+1D 40x36x12")
    # second call, smaller lattice: boost-invariant, conformal EOS, 12 steps
    got, want = _run_pair(30, 26, 1, 0.1, 0.01, [12], None)
    _assert_same(got, want, "2+1D 30x26")


def test_noalloc_rejects_a_lattice_over_the_scratch_limit():
    V = vah()
    L = V.lib()
    lat = V.Lattice(700, 700, 1, 0, 0.1, 0.1, 0.1, 0.01)  # 704^2 padded cells > 399,454
    hyd = V.Hydro(0.5, 0.16, 0.155, 0)
    opt = V.Options()
    L.vah_default_options(ctypes.byref(opt))
    st = V.DeviceState(1, 1, 1, 1, 1, 1)  # never dereferenced: the size check comes first
    rc = L.vah_step_noalloc(ctypes.byref(lat), ctypes.byref(hyd), ctypes.byref(opt), 0, 1, ctypes.byref(st), None)
    assert rc == V.VAH_ERR_ARG and "scratch limit" in L.vah_last_error().decode()
