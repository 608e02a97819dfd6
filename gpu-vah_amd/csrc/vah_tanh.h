// vah_tanh.h -- tanh(x) / x for the regulator's damping factor.
//
// tanh x = E / (E + 2) with E = e^{2x} - 1, which keeps full relative
// accuracy for small x (no cancellation).  E = expm1(y), y = 2x, by the
// textbook reduction y = k ln 2 + r, |r| <= ln2 / 2:
//     expm1(y) = 2^k expm1(r) + (2^k - 1),
//     expm1(r) = r + r^2 (1/2! + r/3! + ... + r^11/13!),
// whose truncation is below 2^-58 of expm1(r); 2^k - 1 is exact.  Above
// x = 20, tanh x rounds to 1 in double and the quotient is 1 / x.  The result
// is within a few units in the last place of tanh(x) / x from the host C
// library (checked on the CPU).  As for the EOS logarithm, the series
// coefficients sit in constant memory in device code, so that the compiler
// keeps them out of the vector registers of the stage kernels.
#pragma once

#include <math.h>

namespace vah {

// 1/13!, 1/12!, ..., 1/2!, highest first
#if defined(__HIP_DEVICE_COMPILE__)
static __constant__ double kExpm1Coef[12] = {
#else
static const double kExpm1Coef[12] = {
#endif
    1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0, 1.0 / 362880.0, 1.0 / 40320.0,
    1.0 / 5040.0,       1.0 / 720.0,       1.0 / 120.0,      1.0 / 24.0,      1.0 / 6.0,      1.0 / 2.0};

// e^y - 1 for 0 <= y <= 40 (and NaN)
__host__ __device__ inline double expm1_small(double y) {
    constexpr double inv_ln2 = 1.44269504088896338700e+00;
    constexpr double ln2_hi = 6.93147180369123816490e-01;  // low 21 bits zero: k ln2_hi exact
    constexpr double ln2_lo = 1.90821492927058770002e-10;
    const double kd = rint(y * inv_ln2);
    const double r = fma(-kd, ln2_lo, fma(-kd, ln2_hi, y));
    double p = kExpm1Coef[0];
    for (int j = 1; j < 12; ++j) p = fma(p, r, kExpm1Coef[j]);
    const double em = fma(r * r, p, r);  // expm1(r)
    const double t = ldexp(1.0, (int)kd);
    return kd == 0.0 ? em : fma(t, em, t - 1.0);
}

// tanh(x) / x for x > 0 (+inf -> 0, NaN -> NaN)
__host__ __device__ inline double tanh_over_x(double x) {
    if (x > 20.0) return 1.0 / x;
    const double E = expm1_small(2.0 * x);
    return E / ((E + 2.0) * x);
}

}  // namespace vah
