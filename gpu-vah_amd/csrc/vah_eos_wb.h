// vah_eos_wb.h -- the reference's non-conformal equation of state: the
// Wuppertal-Budapest rational fits p(e) and T(e) in e [fm^-4]
// (old_cpu-vah_code/src/EquationOfState.cpp:31-76 and :120-142, active
// there when CONFORMAL_EOS is not defined).  Shared by the device kernels
// and the host initial-state map.
#pragma once

#include <hip/hip_runtime.h>

namespace vah {

// p(e): numerator and denominator are nested fma chains over the explicit
// powers e, e^2 = e*e, ..., innermost term first; quotient in double.
__host__ __device__ inline double wb_pressure(double e) {
    constexpr double A[13] = {-0.25181736420168666, 9737.845799644809, 1.077580993288114e6,
                              3.1729694865420084e6, 1.6357487344679043e6, 334334.4309240126,
                              41913.439282708554, 6340.448389300905, 141.5073484468774,
                              0.7158279081255019, 0.0009417586777847889, 3.1188455176941583e-7,
                              1.9531729608963267e-11};
    constexpr double B[13] = {45829.44617893836, 4.0574329080826794e6, 2.0931169138134286e7,
                              1.3512402226067686e7, 1.7851642641834426e6, 278581.2989342773,
                              26452.34905933697, 499.04919730607065, 2.3405487982094204,
                              0.002962497695527404, 9.601103399348206e-7, 5.928138360995685e-11,
                              3.2581066229887368e-18};
    double pw = e, a = A[0], b = B[0];
#pragma unroll
    for (int k = 1; k <= 12; ++k) {
        a = __builtin_fma(A[k], pw, a);
        b = __builtin_fma(B[k], pw, b);
        pw = pw * e;
    }
    return a / b;
}

// T(e) [fm^-1]: degree 11 over degree 11, the terms added from the constant up.
__host__ __device__ inline double wb_temperature(double e) {
    constexpr double N[12] = {1.510073201405604e-29, 8.014062800678687e-18, 2.4954778310451065e-10,
                              0.000063810382643387, 0.4873490574161924, 207.48582344326206,
                              6686.07424325115, 14109.766109389702, 1471.6180520527757,
                              14.055788949565482, 0.015421252394182246, 1.5780479034557783e-6};
    constexpr double D[12] = {7.558667139355393e-28, 1.3686372302041508e-16, 2.998130743142826e-9,
                              0.0005036835870305458, 2.316902328874072, 578.0778724946719,
                              11179.193315394154, 17965.67607192861, 1051.0730543534657,
                              5.916312075925817, 0.003778342768228011, 1.8472801679382593e-7};
    double pw = e, n = N[0], d = D[0];
#pragma unroll
    for (int k = 1; k <= 11; ++k) {
        n = n + N[k] * pw;
        d = d + D[k] * pw;
        pw = pw * e;
    }
    return n / d;
}

}  // namespace vah
