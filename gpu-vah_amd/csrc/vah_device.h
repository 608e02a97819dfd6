// vah_device.h -- per-cell physics of the VAH Kurganov-Tadmor step, as
// device functions for gfx950.
//
// Reference behaviour restated here (paths relative to the reference repo):
//   Newton recovery      old_cpu-vah_code/src/EnergyMomentumTensor.cpp:30-197
//   KT face flux         old_cpu-vah_code/src/SemiDiscreteKurganovTadmorScheme.cpp:13-57
//   limiter / faces      old_cpu-vah_code/src/FluxLimiter.cpp:13-33, HalfSiteExtrapolation.cpp:12-26
//   sources              old_cpu-vah_code/src/SourceTerms.cpp:46-931 (PIMUNU; W and Pi off)
//   conformal EOS        old_cpu-vah_code/src/EquationOfState.cpp:78,144
//   regulator            old_cpu-vah_code/src/FullyDiscreteKurganovTadmorScheme.cpp:524-696
// The arithmetic is arranged for the GPU (reciprocals reused, F^1.5 = F*sqrt(F),
// contracted multiply-adds): results differ from the reference's by rounding
// only (~1e-15 relative after hundreds of steps).  Decisions the reference
// makes (Newton stop rule and fallback, regulator thresholds, P_L and e
// floors) are kept exactly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vah_eos_wb.h"
#include "vah_log.h"
#include "vah_sqrt.h"
#include "vah_tanh.h"

namespace vah {

constexpr int NQ = 15;
enum { TTT, TTX, TTY, TTN, PL, PITT, PITX, PITY, PITN, PIXX, PIXY, PIXN, PIYY, PIYN, PINN };

constexpr double kThird = 1.0 / 3.0;

// Storage plane of interior eta index k: 3+1D lattices carry 2 ghost planes
// on each side, a boost-invariant (one-plane) lattice stores its plane alone.
__host__ __device__ constexpr int plane_of(int k, bool d3) { return d3 ? k + 2 : 0; }

// ---------------------------------------------------------------------------
// Reciprocal, division and reciprocal square root from the hardware
// approximations (v_rcp_f64, v_rsq_f64) refined by two Newton-Raphson steps,
// a final residual correction for the quotient.  Accurate to ~1 ulp for the
// normal, finite operands of this solver -- about half the instructions of the
// correctly rounded IEEE sequences, whose special-case scaling is not needed
// here.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

__device__ __forceinline__ double fdiv(double a, double b) {
    double r = rcp(b);
    double q = a * r;
    double e = __builtin_fma(-b, q, a);
    return __builtin_fma(e, r, q);
}

// The step f / f' of a Newton iteration: the hardware reciprocal estimate
// (about 2^-23 relative) refined once (about 2^-46), times f.  The quotient
// only scales the correction, so its error changes each iterate by about
// 2^-46 of the step, far below the solvers' 1e-7 stopping tolerance: the
// iterates and the root agree with those of an exact division to rounding.
__device__ __forceinline__ double newton_quotient(double f, double fp) {
    double r = __builtin_amdgcn_rcp(fp);
    r = __builtin_fma(r, __builtin_fma(-fp, r, 1.0), r);
    return f * r;
}

// The failure test of a Newton update u (EnergyMomentumTensor.cpp:64):
// NaN, infinite, negative or above 9.0072e15.  One range test covers all
// four, since every comparison with NaN is false.
__device__ __forceinline__ bool newton_out_of_range(double u) { return !((u >= 0.0) & (u <= 9.0072e+15)); }

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }

__device__ __forceinline__ double rsqrt(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double h = 0.5 * x;
    double e = __builtin_fma(-h * y, y, 0.5);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-h * y, y, 0.5);
    return __builtin_fma(y, e, y);
}

// Square root of x >= 0 from the hardware reciprocal square root and the
// refinement of vah_sqrt.h: the device library's sequence without its range
// scaling, which only acts below 2^-767, so the result equals sqrt(x) for every
// x of at least that size (the solver's arguments are squares and sums of
// squares of physical quantities).  x = 0 gives 0, NaN gives NaN.
__device__ __forceinline__ double sqrt_pos(double x) { return sqrt_refine(x, __builtin_amdgcn_rsq(x)); }

// ---------------------------------------------------------------------------
// Equation of state
// ---------------------------------------------------------------------------
struct Eos {
    int kind;              // 0 conformal, 1 table, 2 Wuppertal-Budapest fits
    int n;                 // table nodes
    double lne0, inv_h;    // node i at ln e = lne0 + i / inv_h
    double e_first, e_last;
    const double *p, *T;   // node values
    const uint32_t *powf_fix;  // conformal: glibc powf misrounding bitmap (2^25 bits)
};

// T = powf((float)(e / 13.8997), 0.25f) with the host C library's rounding.
// x^(1/4) is formed in double (two correctly rounded square roots, < 1 ulp
// of double) and rounded to float, which is the correctly rounded result for
// every float x.  The host libm's powf is off by one float ulp on a fixed,
// 16-periodic subset of inputs (x and 16x behave alike); fix holds that subset
// for one period, biased exponents 124..127, built at start-up from the host
// libm.  The single exception to the periodicity sits at mantissa 0x2e5d6e of
// biased exponents = 0 mod 4 outside [64, 188].
__device__ __forceinline__ double conformal_T(double e, const uint32_t *__restrict__ fix) {
    float xf = (float)(e / 13.8997);
    double r = __builtin_sqrt(__builtin_sqrt((double)xf));
    float T = (float)r;
    uint32_t b = __float_as_uint(xf);
    uint32_t ex = b >> 23;
    if (fix != nullptr && ex >= 1u && ex <= 254u) {
        uint32_t off = (b - (124u << 23)) & ((1u << 25) - 1u);
        uint32_t flip = (__builtin_nontemporal_load(fix + (off >> 5)) >> (off & 31u)) & 1u;
        if (flip && (ex & 3u) == 0u && (b & 0x7fffffu) == 0x2e5d6eu && (ex < 64u || ex > 188u)) flip = 0u;
        if (flip) {
            uint32_t tb = __float_as_uint(T);
            T = __uint_as_float(r > (double)T ? tb + 1u : tb - 1u);
        }
    }
    return (double)T;
}

__device__ __forceinline__ void table_lookup(const Eos &E, double e, double *p, double *T) {
    double f = (eos_log(e) - E.lne0) * E.inv_h;
    int last = E.n - 1;
    if (!(f >= 0.0)) {
        double r = e / E.e_first;
        if (p) *p = E.p[0] * r;
        if (T) *T = E.T[0] * __builtin_sqrt(__builtin_sqrt(r));
        return;
    }
    int i = (int)f;
    if (i >= last) {
        double r = e / E.e_last;
        if (p) *p = E.p[last] * r;
        if (T) *T = E.T[last] * __builtin_sqrt(__builtin_sqrt(r));
        return;
    }
    double w = f - (double)i;
    if (p) *p = E.p[i] + w * (E.p[i + 1] - E.p[i]);
    if (T) *T = E.T[i] + w * (E.T[i + 1] - E.T[i]);
}

__device__ __forceinline__ double eos_p(const Eos &E, double e) {
    if (E.kind == 0) return e * kThird;
    if (E.kind == 2) return wb_pressure(e);
    double p;
    table_lookup(E, e, &p, nullptr);
    return p;
}

__device__ __forceinline__ double eos_T(const Eos &E, double e) {
    if (E.kind == 0) return conformal_T(e, E.powf_fix);
    if (E.kind == 2) return wb_temperature(e);
    double T;
    table_lookup(E, e, nullptr, &T);
    return T;
}

// ---------------------------------------------------------------------------
// Rational fits in a = P_L / e (SourceTerms.cpp:451-481,
// rhic/src/AnisotropicDistributionFunctions.cpp:18-21).  Coefficients are the
// reference's; each polynomial is summed over explicit powers of a in the
// reference's term order, so the rounding stays close to the reference's
// even where the large alternating coefficients cancel.
// ---------------------------------------------------------------------------
struct Powers {
    double a[16];
};

__device__ __forceinline__ Powers powers_of(double a) {
    Powers w;
    w.a[0] = 1.0;
    w.a[1] = a;
#pragma unroll
    for (int k = 2; k < 16; ++k) w.a[k] = w.a[k - 1] * a;
    return w;
}

// A coefficient table in constant memory, addressed so that the compiler
// reads it with scalar loads instead of folding its values into scalar moves
// (two per coefficient and evaluation): the empty asm hides which table the
// pointer holds.
typedef __attribute__((address_space(4))) const double ConstTable;

template <int N>
__device__ __forceinline__ ConstTable *scalar_table(const double (&c)[N]) {
    ConstTable *p = (ConstTable *)c;
    __asm__("" : "+s"(p));
    return p;
}

// sum c[i] * a^k[i], terms in the given order, constant term first
template <int N>
__device__ __forceinline__ double psum(const double (&tab)[N], const int (&k)[N], const Powers &w) {
    ConstTable *c = scalar_table(tab);
    double s = c[0];
#pragma unroll
    for (int i = 1; i < N; ++i) s = s + c[i] * w.a[k[i]];
    return s;
}

template <int N>
__device__ __forceinline__ double psum(const double (&tab)[N], const Powers &w) {
    ConstTable *c = scalar_table(tab);
    double s = c[0];
#pragma unroll
    for (int i = 1; i < N; ++i) s = s + c[i] * w.a[i];
    return s;
}

struct Fits {
    double Rtilde, Rbar0, Rgamma, Rbar0P;
};

// Coefficients of the fits (numerator N, denominator D, constant term first).
// They are read from constant memory, where each scalar load brings several
// of them; as literals every one costs two scalar moves per evaluation.
static __constant__ double kRbar0N[9] = {
    0.015267955823446243, 7.725572805021035, 421.0063884634789, 3422.877939650926, -5785.670846299543,
    -12261.66452089229, 31491.409484673808, -22737.05146992673, 5441.373392185447
};
static __constant__ double kRbar0D[9] = {
    0.05470696094814806, 14.505878005231883, 522.6643024173569, 2731.7776413939037, -6161.1991042880445,
    -3989.4375208972588, 15008.260526258282, -10243.036679405379, 2116.74060159494
};
static __constant__ double kRtildeN[8] = {
    -6.674731906076046e-6, 0.004617789933500251, 0.7207562721999754, 9.097427250602184, -4.475814747302824,
    -36.37501529319408, 46.868405146729316, -15.833867583743228
};
static __constant__ double kRtildeD[8] = {
    0.06856675185266, 2.9181587012768597, 11.951184087839218, -29.708257843442173, -2.618233802059826,
    34.646239784689065, -19.62596366454439, 2.374808442453899
};
static __constant__ double kRgammaN[8] = {
    0.0001373796340585521, -0.6149634004722385, 3.1968875683514253, 246.35973783799196, -764.319750320186,
    834.160165071214, -371.64955466673234, 52.87411921963021
};
static __constant__ double kRgammaD[8] = {
    -1.673322188488071, 13.343782941396997, 561.7566534476223, -1790.2296622275915, 1896.4688704912812,
    -658.7933063369629, -85.96181900698849, 65.09739194472589
};
static __constant__ double kRbar0PN[16] = {
    2.473173363908116e-10, -4.899839370307281e-6, 155055.91462124084, -275435.45350226434, 350689.68825705117,
    -299725.38986957155, 151477.08809203724, -33196.47417939176, -0.004301975027942015, -0.14858206981041563,
    6.249255189587875, -92.79641927240235, 807.175057749925, -4760.015905266286, 20324.533122685436,
    -64758.869552496515
};
static __constant__ double kRbar0PD[16] = {
    0.00008222793468208523, 0.03411917870833943, 4.895969276094396e6, -8.84162305829353e6,
    1.1445063656613324e7, -9.918442713390596e6, 5.065882388219598e6, -1.1181016364928822e6,
    0.23871740573818725, -23.50912574236691, 417.4953123877312, -4234.215775452717, 29824.022790048104,
    -157419.8447785501, 641300.6529027821, -2.0248032895288002e6
};

__device__ __forceinline__ double rbar0(const Powers &w) {
    return fdiv(psum(kRbar0N, w), psum(kRbar0D, w));
}

__device__ __forceinline__ Fits fits(const Powers &w) {
    // Rbar0P: the reference writes the terms in the order a^0, a^1, a^10..a^15, a^2..a^9
    constexpr int rp_k[16] = {0, 1, 10, 11, 12, 13, 14, 15, 2, 3, 4, 5, 6, 7, 8, 9};
    Fits f;
    f.Rtilde = fdiv(psum(kRtildeN, w), psum(kRtildeD, w));
    f.Rbar0 = rbar0(w);
    f.Rgamma = fdiv(psum(kRgammaN, w), psum(kRgammaD, w));
    f.Rbar0P = fdiv(psum(kRbar0PN, rp_k, w), psum(kRbar0PD, rp_k, w));
    return f;
}

// transversePressureHat (AnisotropicDistributionFunctions.cpp:5-24)
__device__ __forceinline__ double pt_hat(double e, double p, double pl) {
    return 0.5 * ((e - pl) - rbar0(powers_of(fdiv(pl, e))) * (e - 3.0 * p));
}

// ---------------------------------------------------------------------------
// Generalised minmod slope, theta = 1.1 (FluxLimiter.cpp:13-33):
// minmod(l, minmod(c, r)) of l = theta (y - x), c = (z - x) / 2,
// r = theta (z - y), where minmod of two values is sign*min(|x|,|y|) when the
// signs (with sign(0) = +1) agree, else 0.  The result is nonzero only when
// a = y - x and b = z - y are nonzero with one sign; z - x then has that sign
// too, and the result is the smallest of |l|, |c|, |r| with that sign.  In
// every other case -- opposite signs, or a zero, whose sign the reference
// counts as +1 -- one of the minima is 0.
// half_slope returns half of it, the step from the cell to its face: halving
// commutes with rounding, so (theta / 2) a, (theta / 2) b and (z - x) / 4 are
// the halves of the candidates exactly (short of subnormal results), and
// min(|l|, |r|) / 2 is (theta / 2) min(|a|, |b|) because rounding is monotonic.
// ---------------------------------------------------------------------------

// Double: "nonzero with one sign" is tested as a * b > 0.  That test differs
// from the reference's only when a * b underflows to 0; the slope is then 0
// instead of a value of at most 1.1 * sqrt(5e-324) = 2.4e-162 in magnitude.
// (Carrying the sign through min / max instead costs as many instructions:
// the compiler canonicalises a copysign result before v_min_f64.)
__device__ __forceinline__ double half_slope(double x, double y, double z) {
    constexpr double kHalfTheta = 1.1 * 0.5;
    const double a = y - x, b = z - y;
    const double m = fmin(kHalfTheta * fmin(fabs(a), fabs(b)), 0.25 * fabs(z - x));
    return a * b > 0.0 ? __builtin_copysign(m, a) : 0.0;
}

// Float, on the hardware median: the median of u, v and 0 is minmod(u, v)
// (0 when the signs differ, else the one nearer 0), so half the slope is
// med3(l, med3(c, r, 0), 0) with the halved candidates l = (theta / 2) a,
// c = (z - x) / 4, r = (theta / 2) b -- two v_med3_f32, and l of one face
// side is r of the other.  Equal to the double form's arithmetic done in float
// wherever a * b does not underflow (|a|, |b| below ~4e-23), where this one
// keeps the reference's tiny slope.
__device__ __forceinline__ float half_slope(float x, float y, float z) {
    constexpr float kHalfTheta = 1.1f * 0.5f;
    const float l = kHalfTheta * (y - x), r = kHalfTheta * (z - y), c = 0.25f * (z - x);
    return __builtin_amdgcn_fmed3f(l, __builtin_amdgcn_fmed3f(c, r, 0.0f), 0.0f);
}

// Four times half_slope, for a face value formed as fma(slope4, +-1/4, q):
// (theta / 2) m scaled by 4 is 2.2 m to the bit (2.2 as a double is 4 times
// 0.55 as a double, and scaling by a power of two commutes with rounding), and
// |z - x| / 4 scaled by 4 is |z - x|, so min(2.2 min(|a|, |b|), |z - x|) is 4 x
// the half slope exactly, and q + s4 / 4 rounds once, as q + half_slope does.
// One instruction fewer per slope: the quarter rides on the face value's add.
__device__ __forceinline__ double slope4(double x, double y, double z) {
    constexpr double kTwiceTheta = 1.1 * 2.0;
    const double a = y - x, b = z - y;
    const double m = fmin(kTwiceTheta * fmin(fabs(a), fabs(b)), fabs(z - x));
    return a * b > 0.0 ? __builtin_copysign(m, a) : 0.0;
}

// left and right face values of the stencil x0..x3 (faces of x1 | x2)
__device__ __forceinline__ void face_values(double q0, double q1, double q2, double q3, double &l, double &r) {
    l = __builtin_fma(slope4(q0, q1, q2), 0.25, q1);
    r = __builtin_fma(slope4(q1, q2, q3), -0.25, q2);
}
__device__ __forceinline__ void face_values(float q0, float q1, float q2, float q3, float &l, float &r) {
    l = q1 + half_slope(q0, q1, q2);
    r = q2 - half_slope(q1, q2, q3);
}

// ---------------------------------------------------------------------------
// Newton recovery of (e, p, u^mu) from T^tau-mu, P_L, pi^tau-mu.
// ---------------------------------------------------------------------------
struct Prim {
    double e, p, ut, ux, uy, un;
    int ok;       // 1 converged, 0 fallback
};

// Solve u_T/sqrt(1+u_T^2) (bL + b Pt) - bT = 0 (EnergyMomentumTensor.cpp:30-72).
// The loop body has no branches: each pass evaluates the update and then
// decides, as the reference does, whether it failed (e < 0 at the current
// u_T, where the reference stops before the update, or an update that is NaN,
// infinite, negative or above 9.0072e15), converged, or continues.
__device__ __forceinline__ bool newton_uT(double t, double uT0, double MB0, double MBT, double MB3, double PL,
                                          double Ft, double x, double &uT) {
    const double Ft2 = Ft * Ft;
    const double bT = x * MBT;
    const double bL = x * x * MB0 - Ft2 * PL;
    const double b = x * x + Ft2;
    const double c0 = MB0 - t * Ft * MB3;
    int st = 0;
    for (int it = 0; it < 10000 && st == 0; ++it) {
        double F = 1.0 + uT0 * uT0;
        double rs = rsqrt(F);                  // 1/sqrt(1+u^2)
        double U = uT0 * rs;
        double e = c0 - U * x * MBT;
        double Pt = 0.5 * (e - PL);
        double iP = rs * rs * rs;              // (1+u^2)^-1.5
        double g = bL + b * Pt;
        double dPtduT = -0.5 * x * MBT * iP;
        double f = U * g - bT;
        double fp = iP * g + U * b * dPtduT;
        fp = fabs(fp) == 0.0 ? 1.e-16 : fp;
        double u1 = uT0 - newton_quotient(f, fp);
        const bool bad = (e < 0.0) | newton_out_of_range(u1);
        const bool conv = fabs(u1 - uT0) <= 1.e-7 * fabs(u1);
        st = bad ? -1 : (int)conv;
        uT0 = u1;
    }
    uT = uT0;
    return st == 1;
}

// getInferredVariables with W = Pi = 0 (EnergyMomentumTensor.cpp:74-197):
// then F = A = M3/(M0+P_L), MB = M.
template <typename Q>
__device__ __forceinline__ Prim recover(const Eos &E, double t, const Q &q, double ePrev, double uT_guess) {
    double M0 = (double)q[TTT] - (double)q[PITT];
    double M1 = (double)q[TTX] - (double)q[PITX];
    double M2 = (double)q[TTY] - (double)q[PITY];
    double M3 = (double)q[TTN] - (double)q[PITN];
    double M0PL = M0 + (double)q[PL];
    if (M0PL == 0.0) M0PL = 1.e-16;
    double F = fdiv(M3, M0PL);
    double Ft = t * F;
    double x = sqrt_pos(fabs(1.0 - Ft * Ft));
    double MBT = sqrt_pos(M1 * M1 + M2 * M2);
    if (MBT == 0.0) MBT = 1.e-16;
    double uT = 0.0;
    Prim r;
    if (newton_uT(t, uT_guess, M0, MBT, M3, (double)q[PL], Ft, x, uT)) {
        double C2 = 1.0 + uT * uT;
        double iC = rsqrt(C2);
        double C = C2 * iC;
        double ix = rcp(x);
        double iMBT = rcp(MBT);
        r.ux = uT * M1 * iMBT;
        r.uy = uT * M2 * iMBT;
        r.un = F * C * ix;
        r.ut = C * ix;
        r.e = M0 - t * Ft * M3 - (uT * iC) * x * MBT;
        r.ok = 1;
    } else {
        r.e = ePrev * .999;
        r.ux = 0.0;
        r.uy = 0.0;
        r.un = 0.0;
        r.ut = 1.0;
        r.ok = 0;
    }
    r.p = eos_p(E, r.e);
    return r;
}

__device__ __forceinline__ bool prim_nan(const Prim &r) {
    return __builtin_isnan(r.e) || __builtin_isnan(r.ut) || __builtin_isnan(r.ux) || __builtin_isnan(r.uy) ||
           __builtin_isnan(r.un);
}

}  // namespace vah
