// vah_sqrt.h -- square root from a reciprocal-square-root seed.
//
// The stage kernels take square roots of squares and sums of squares (the
// Newton inputs |1 - (tau F)^2| and |M_T|, the |u_T| guesses, the regulator's
// norms).  sqrt_refine(x, r) turns a seed r ~ 1/sqrt(x) (v_rsq_f64 on the
// device) into sqrt(x) with the device library's refinement:
//     g = x r, h = r / 2                      (g ~ sqrt x, h ~ 1 / (2 sqrt x))
//     e = 1/2 - h g;  g += g e;  h += h e     (one Goldschmidt step)
//     d = x - g^2;    g += d h                (twice: residual corrections)
// without the library's range scaling (which only acts below 2^-767) and its
// special-operand selects: x = 0 gives 0 (the select below), NaN gives NaN.
// The seed's relative error is squared by the Goldschmidt step and the
// corrections then round to the correctly rounded root; the host test feeds
// seeds off by up to 2^-20 and compares with the C library bit for bit.
#pragma once

namespace vah {

__host__ __device__ inline double sqrt_refine(double x, double r) {
    double g = x * r, h = 0.5 * r;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return x == 0.0 ? 0.0 : g;
}

}  // namespace vah
