// vah_host_common.h -- host-side pieces shared by the single-lattice context
// (vah_capi.hip) and the batch engine (vah_batch.hip): the initial-state map
// from primitives to conserved variables, the padded device layout, and the
// device copy of the equation of state.
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/vah.h"
#include "vah_device.h"

namespace vah_host {

// Records the text vah_last_error() returns on this thread; returns code.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess) return vah_host::fail(VAH_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

// glibc powf(x, 0.25f) misrounding bitmap for the conformal temperature
// (built once per process from the host libm, 4 MB).
const std::vector<uint32_t> &powf_fix_table();

// Host view of the EOS, enough to evaluate p(e) for initial states with the
// device's arithmetic (linear interpolation in ln e, power-law outside).
struct HostEos {
    int kind = VAH_EOS_CONFORMAL;
    int n = 0;
    double lne0 = 0, h = 1;
    std::vector<double> p;
};

inline double host_eos_p(const HostEos &E, double e) {
    if (E.kind == VAH_EOS_CONFORMAL) return e / 3;
    if (E.kind == VAH_EOS_WB) return vah::wb_pressure(e);
    double f = (log(e) - E.lne0) * (1.0 / E.h);
    int last = E.n - 1;
    if (!(f >= 0.0)) return E.p[0] * (e / exp(E.lne0));
    int i = (int)f;
    if (i >= last) return E.p[last] * (e / exp(E.lne0 + last / (1.0 / E.h)));
    double w = f - (double)i;
    return E.p[i] + w * (E.p[i + 1] - E.p[i]);
}

// setConservedVariables (old_cpu-vah_code/src/DynamicalVariables.cpp:197-287):
// primitives of n interior cells -> 15 conserved fields (field-major), e, p
// and u^mu.  pl == NULL gives the reference's P_L = 0.0154483 e; pi == NULL
// gives pi^{mu nu} = 0.
inline void prims_to_conserved(double t, const HostEos &E, long n, const double *e, const double *pl,
                               const double *ut, const double *ux, const double *uy, const double *un,
                               const double *pi, double *q, double *ev, double *pv, double *u) {
    using namespace vah;
    for (long m = 0; m < n; ++m) {
        double eeq = e[m];
        double peq = host_eos_p(E, eeq);
        double vt = ut[m], vx = ux[m], vy = uy[m], vn = un[m];
        double pis[10];
        for (int f = 0; f < 10; ++f) pis[f] = pi ? pi[(size_t)f * n + m] : 0.0;
        double plv = pl ? pl[m] : 0.0154483 * eeq;
        double pt = 0.5 * (eeq - plv);
        double DP = plv - pt;
        double F = 1 + (vx * vx + vy * vy);
        double Ltt = DP * t * t * vn * vn / F;
        double Ltn = DP * vt * vn / F;
        q[(size_t)TTT * n + m] = (eeq + pt) * vt * vt - pt + Ltt + pis[0];
        q[(size_t)TTX * n + m] = (eeq + pt) * vt * vx + pis[1];
        q[(size_t)TTY * n + m] = (eeq + pt) * vt * vy + pis[2];
        q[(size_t)TTN * n + m] = (eeq + pt) * vt * vn + Ltn + pis[3];
        q[(size_t)PL * n + m] = plv;
        for (int f = 0; f < 10; ++f) q[(size_t)(PITT + f) * n + m] = pis[f];
        ev[m] = eeq;
        pv[m] = peq;
        u[m] = vt;
        u[(size_t)n + m] = vx;
        u[(size_t)2 * n + m] = vy;
        u[(size_t)3 * n + m] = vn;
    }
}

// Padded device layout of one lattice: (nx+4) x (ny+4) x (nz+4) cells, x
// fastest, 2 ghost cells per side, fields fs elements apart (fs rounded up to
// 64 doubles so every field starts 512-byte aligned).  A boost-invariant
// lattice (nz == 1) stores its one plane only: no stencil reads across eta
// there (the eta differences of such a lattice are identically 0).
struct Layout {
    int nx, ny, nz, ncx, ncy, ncz;
    long ncell, fs;
    Layout() = default;
    Layout(int nx_, int ny_, int nz_)
        : nx(nx_), ny(ny_), nz(nz_), ncx(nx_ + 4), ncy(ny_ + 4), ncz(nz_ > 1 ? nz_ + 4 : 1) {
        ncell = (long)ncx * ncy * ncz;
        fs = (ncell + 63) / 64 * 64;
    }
    long interior() const { return (long)nx * ny * nz; }
    long index(long m) const {
        int i = (int)(m % nx);
        int j = (int)((m / nx) % ny);
        int k = (int)(m / ((long)nx * ny));
        return (long)vah::plane_of(k, nz > 1) * ncx * ncy + (long)(j + 2) * ncx + (i + 2);
    }
    // interior field-major host arrays -> padded slab image (ghosts 0)
    void pad(int nfields, const double *src, std::vector<double> &dst) const {
        const long n = interior();
        dst.assign((size_t)nfields * fs, 0.0);
        for (int f = 0; f < nfields; ++f)
            for (long m = 0; m < n; ++m) dst[(size_t)f * fs + index(m)] = src[(size_t)f * n + m];
    }
    void unpad(int nfields, const double *src, double *dst) const {
        const long n = interior();
        for (int f = 0; f < nfields; ++f)
            for (long m = 0; m < n; ++m) dst[(size_t)f * n + m] = src[(size_t)f * fs + index(m)];
    }
};

// Device EOS: uploads the table (or the powf bitmap of the conformal EOS) on
// `stream` and fills *out.  Allocations are returned through fix_dev /
// tab_dev for the owner to free.  Returns VAH_OK or an error code with *err.
int make_device_eos(int kind, const vah_eos_table &T, hipStream_t stream, vah::Eos *out, uint32_t **fix_dev,
                    double **tab_dev, HostEos *host, std::string *err);

}  // namespace vah_host
