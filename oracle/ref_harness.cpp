// ref_harness.cpp -- extern "C" wrappers around reference functions that
// compile from their own sources with no stand-in headers:
//   old_cpu-vah_code/src/SourceTerms.cpp        loadSourceTerms, loadSourceTermsX/Y/Z
//   old_cpu-vah_code/src/EquationOfState.cpp    equilibriumPressure, effectiveTemperature
//   rhic/src/AnisotropicDistributionFunctions.cpp  transversePressureHat, Rbar0_fun, Rbar0P_fun
//   old_cpu-vah_code/src/FluxLimiter.cpp        approximateDerivative
//   old_cpu-vah_code/src/HalfSiteExtrapolation.cpp
//
// TEST INFRASTRUCTURE ONLY: built by oracle/Makefile (target `ref`) into
// oracle/_ref/ where /root/reference exists, and used by the CPU tests to pin
// the oracle's restatement.  Nothing here is shipped or run on the GPU.
#include "DynamicalVariables.h"
#include "EquationOfState.h"
#include "FluxLimiter.h"
#include "HalfSiteExtrapolation.h"
#include "SourceTerms.h"
#include "AnisotropicDistributionFunctions.h"

extern "C" {

// Source terms at the centre of a 3x3x3 block (x fastest, s = 13).
// fields: ut, ux, uy, un, e, p, pl -- 27 values each.
void ref_source_block(const double *Q, const double *fields, const double *uprev, double t, double etabar,
                      double dt, double dx, double dy, double dz, double *S) {
    FLUID_VELOCITY u;
    u.ut = const_cast<double *>(fields);
    u.ux = const_cast<double *>(fields + 27);
    u.uy = const_cast<double *>(fields + 54);
    u.un = const_cast<double *>(fields + 81);
    CONSERVED_VARIABLES cv = {};
    cv.pl = const_cast<double *>(fields + 162);
    loadSourceTerms(Q, S, &u, uprev[0], uprev[1], uprev[2], uprev[3], t, fields + 108, fields + 135, 13, 3, 3, 3,
                    etabar, dt, dx, dy, dz, 1, 1, 1, 0.0, 0.0, 0.0, &cv);
}

// Axis sources from a 5 x NUMBER_CONSERVED_VARIABLES gather.
void ref_source_axis(const double *g, double ut, double ui, double t, double d, int dir, double *S) {
    double vt[1] = {ut}, vi[1] = {ui}, zero[1] = {0.0};
    FLUID_VELOCITY u;
    u.ut = vt;
    u.ux = dir == 1 ? vi : zero;
    u.uy = dir == 2 ? vi : zero;
    u.un = dir == 3 ? vi : zero;
    if (dir == 1)
        loadSourceTermsX(g, S, &u, 0, d);
    else if (dir == 2)
        loadSourceTermsY(g, S, &u, 0, d);
    else
        loadSourceTermsZ(g, S, &u, 0, t, d);
}

double ref_pressure(double e) { return equilibriumPressure(e); }
double ref_temperature(double e) { return effectiveTemperature(e); }
double ref_energy_density(double T) { return equilibriumEnergyDensity(T); }
double ref_pt_hat(double e, double p, double pl) { return transversePressureHat(e, p, pl); }
double ref_rbar0(double a) { return Rbar0_fun(a); }
double ref_rbar0p(double a) { return Rbar0P_fun(a); }
double ref_approx_derivative(double x, double y, double z) { return approximateDerivative(x, y, z); }

// which: 0 right-forward, 1 right-backwards, 2 left-forward, 3 left-backwards
double ref_halfsite(int which, double qmm, double qm, double q, double qp, double qpp) {
    switch (which) {
    case 0: return rightHalfCellExtrapolationForward(qmm, qm, q, qp, qpp);
    case 1: return rightHalfCellExtrapolationBackwards(qmm, qm, q, qp, qpp);
    case 2: return leftHalfCellExtrapolationForward(qmm, qm, q, qp, qpp);
    default: return leftHalfCellExtrapolationBackwards(qmm, qm, q, qp, qpp);
    }
}
}
