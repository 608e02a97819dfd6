// eos_table.cpp -- the lattice-QCD-like crossover equation of state used for
// the 3+1D workload, tabulated uniformly in ln e for the device lookup.
//
// Thermodynamics from one function: p(T) = T^4 f(T) with
//   f(T) = (pi^2/90) [g_h + (g_q - g_h) S(T)],  S(T) = (1 + tanh((T - Tc)/dT)) / 2,
// so s = dp/dT and e = T s - p = 3 p + T^5 f'(T) are consistent by
// construction.  g_q = 42.25 reproduces the conformal factor 13.8997 of the
// reference (old_cpu-vah_code/include/EquationOfState.h:17) at high T.
#include <math.h>

#include "../../include/vah_host.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kGq = 42.25;
constexpr double kGh = 5.0;
constexpr double kTc = 0.155 / VAH_HBARC;   // fm^-1
constexpr double kDT = 0.025 / VAH_HBARC;   // fm^-1

void eval(double T, double *e, double *p) {
    const double c = kPi * kPi / 90.0;
    const double x = (T - kTc) / kDT;
    const double S = 0.5 * (1.0 + tanh(x));
    const double ch = cosh(x);
    const double dS = 0.5 / (kDT * ch * ch);
    const double f = c * (kGh + (kGq - kGh) * S);
    const double fp = c * (kGq - kGh) * dS;
    const double T4 = T * T * T * T;
    *p = T4 * f;
    *e = 3.0 * T4 * f + T4 * T * fp;
}

// e(T) is strictly increasing: bracket and bisect to full double precision.
double temperature_of(double e) {
    double lo = 1e-6, hi = 1.0;
    double elo, ehi, p;
    eval(hi, &ehi, &p);
    while (ehi < e) {
        hi *= 2.0;
        eval(hi, &ehi, &p);
    }
    eval(lo, &elo, &p);
    while (elo > e) {
        lo *= 0.5;
        eval(lo, &elo, &p);
    }
    for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        double em;
        eval(mid, &em, &p);
        (em < e ? lo : hi) = mid;
    }
    return 0.5 * (lo + hi);
}

}  // namespace

extern "C" {

void vah_eos_lqcd_eval(double T, double *e, double *p) { eval(T, e, p); }

int vah_eos_lqcd_table(int n, double e_min, double e_max, double *p, double *T, double *lne0, double *h) {
    if (n < 2 || !(e_min > 0) || !(e_max > e_min) || !p || !T || !lne0 || !h) return VAH_ERR_ARG;
    const double a = log(e_min), b = log(e_max);
    const double step = (b - a) / (n - 1);
    for (int i = 0; i < n; ++i) {
        double e = exp(a + i * step);
        double Ti = temperature_of(e);
        double ei, pi;
        eval(Ti, &ei, &pi);
        T[i] = Ti;
        p[i] = pi * (e / ei);  // exact node energy density; ei == e to rounding
    }
    *lne0 = a;
    *h = step;
    return VAH_OK;
}

}  // extern "C"
