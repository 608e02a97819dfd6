// vah_glauber.h -- Monte-Carlo Glauber initial energy density on the device.
//
// The nucleon positions of the two nuclei arrive from the host (2A
// transverse positions sampled from the request's seed, the generator of
// vah_ic_mc_glauber); everything that scales with the lattice runs here:
//   wounded  a nucleon is wounded when a nucleon of the other nucleus is
//            closer than sqrt(sigma_NN / pi) (black disk,
//            rhic/src/MonteCarloGlauberModel.cpp:33-72);
//   deposit  each transverse cell sums a Gaussian of the nucleon width
//            around every wounded nucleon (:90-105), in the order the host
//            generator sums them (nucleus A's i-th, then B's i-th, i = 0..A-1),
//            deposits beyond 8 widths (< 1e-13 of a peak) left out;
//   extrude  scaled so the hottest transverse cell holds e_peak, times the
//            longitudinal plateau-plus-Gaussian factor of each eta plane,
//            plus the floor; P_L = pl_ratio e when pl_ratio > 0.
// Layout of `pos`: xa[A], ya[A], xb[A], yb[A].
#pragma once

#include "vah_device.h"

namespace vah {

// one lane per nucleon of both nuclei; lane 0 also clears the running maximum
static __global__ void glauber_wounded_kernel(const double *pos, int A, double d2, int *wounded,
                                              unsigned long long *emax) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i == 0) *emax = 0ull;
    if (i >= 2 * A) return;
    const double *xa = pos, *ya = pos + A, *xb = pos + 2 * A, *yb = pos + 3 * A;
    int w = 0;
    if (i < A) {
        for (int j = 0; j < A && !w; ++j) {
            const double dx = __dsub_rn(xa[i], xb[j]), dy = __dsub_rn(ya[i], yb[j]);
            w = __dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)) < d2;
        }
    } else {
        const int k = i - A;
        for (int j = 0; j < A && !w; ++j) {
            const double dx = __dsub_rn(xa[j], xb[k]), dy = __dsub_rn(ya[j], yb[k]);
            w = __dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)) < d2;
        }
    }
    wounded[i] = w;
}

// one lane per transverse cell
static __global__ void glauber_deposit_kernel(const double *pos, const int *wounded, int A, int nx, int ny, double dx,
                                              double dy, double inv2w2, double cut2, double *eT,
                                              unsigned long long *emax) {
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (m < (long)nx * ny) {
        const int i = (int)(m % nx), j = (int)(m / nx);
        const double x = __dmul_rn(i - (nx - 1) / 2., dx), y = __dmul_rn(j - (ny - 1) / 2., dy);
        const double *xa = pos, *ya = pos + A, *xb = pos + 2 * A, *yb = pos + 3 * A;
        auto add = [&](double px, double py) {
            const double ex = __dsub_rn(x, px), ey = __dsub_rn(y, py);
            const double r2 = __dadd_rn(__dmul_rn(ex, ex), __dmul_rn(ey, ey));
            if (r2 < cut2) s = __dadd_rn(s, exp(-__dmul_rn(r2, inv2w2)));
        };
        for (int n = 0; n < A; ++n) {
            if (wounded[n]) add(xa[n], ya[n]);
            if (wounded[A + n]) add(xb[n], yb[n]);
        }
        eT[m] = s;
    }
    // wave maximum, one atomic per wave (e >= 0: the bits order like the values)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s = fmax(s, __shfl_xor(s, off));
    if ((threadIdx.x & 63) == 0 && s > 0.0) atomicMax(emax, (unsigned long long)__double_as_longlong(s));
}

// one lane per cell of the event: out = [e | P_L] (P_L only when pl_ratio > 0)
static __global__ void glauber_extrude_kernel(const double *eT, const unsigned long long *emax, const double *fl,
                                              long nT, long n, double peak, double floor_, double pl_ratio,
                                              double *out) {
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const double mx = __longlong_as_double((long long)*emax);
    double v = eT[m % nT];
    if (mx > 0.0) v = __dmul_rn(v, __ddiv_rn(peak, mx));
    const double e = __dadd_rn(__dmul_rn(v, fl[m / nT]), floor_);
    out[m] = e;
    if (pl_ratio > 0.0) out[n + m] = __dmul_rn(pl_ratio, e);
}

}  // namespace vah
