// vah_log.h -- natural logarithm for the tabulated equation of state.
//
// The table EOS is uniform in ln e (node i at ln e = lne0 + i / inv_h), so
// every lookup needs ln e.  This is the textbook reduction
//     x = 2^k m,  m in [sqrt(1/2), sqrt(2)),
//     ln x = k ln 2 + 2 atanh(s),  s = (m - 1) / (m + 1),
//     2 atanh(s) = 2 s (1 + s^2/3 + s^4/5 + ... ),
// with |s| <= 0.1716, so eleven odd terms leave a truncation below 2^-60 of
// the result; ln 2 is split so that k ln2_hi is exact.  It is within two
// units in the last place of the correctly rounded logarithm (checked
// against the host libm on the CPU), and, unlike the device library's
// logarithm, its few constants do not stay resident in vector registers
// across the stage kernels' plane loop.
// Special values: ln(+inf) = +inf; zero, negative and NaN arguments give
// -inf or NaN, which the table lookup sends to its below-table branch as it
// would the library's -inf or NaN.
#pragma once

#include <math.h>

namespace vah {

// 1/23, 1/21, ..., 1/3: the series coefficients, highest first.  In device
// code they are read from constant memory (scalar loads into scalar
// registers): as literals the compiler materialises them in vector
// registers once per kernel, where they compete with the stencil state.
#if defined(__HIP_DEVICE_COMPILE__)
static __constant__ double kEosLogCoef[11] = {
#else
static const double kEosLogCoef[11] = {
#endif
    1.0 / 23.0, 1.0 / 21.0, 1.0 / 19.0, 1.0 / 17.0, 1.0 / 15.0, 1.0 / 13.0,
    1.0 / 11.0, 1.0 / 9.0,  1.0 / 7.0,  1.0 / 5.0,  1.0 / 3.0};

__host__ __device__ inline double eos_log(double x) {
    if (!(x > 0.0)) return x == 0.0 ? -HUGE_VAL : __builtin_nan("");  // negative or NaN
    if (x == HUGE_VAL) return x;
    int k = 0;
    double m = frexp(x, &k);  // m in [1/2, 1)
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        k -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    double p = kEosLogCoef[0];
    for (int j = 1; j < 11; ++j) p = fma(p, z, kEosLogCoef[j]);
    const double s2 = 2.0 * s;
    const double lnm = fma(s2 * z, p, s2);  // 2 s + 2 s z P(z)
    // ln 2 = ln2_hi + ln2_lo, ln2_hi with its low 21 bits zero: k ln2_hi is exact
    constexpr double ln2_hi = 6.93147180369123816490e-01;
    constexpr double ln2_lo = 1.90821492927058770002e-10;
    const double kd = (double)k;
    return kd * ln2_hi + fma(kd, ln2_lo, lnm);
}

}  // namespace vah
