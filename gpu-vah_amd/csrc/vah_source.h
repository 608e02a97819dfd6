// vah_source.h -- VAH source terms and regulator for one cell (device code).
//
// Restates old_cpu-vah_code/src/SourceTerms.cpp:46-931 (PIMUNU on, W_TZ_MU
// and PI off: every W and Pi term is an exact zero there and is left out) and
// the transverse-shear regulator of
// old_cpu-vah_code/src/FullyDiscreteKurganovTadmorScheme.cpp:524-696.
// Reference quirks kept: tau_pipi = 4 delta_pipi / 3 (integer 1/2 == 0,
// SourceTerms.cpp:634); dnLtn counted twice in S[3] (:686); pitn*u^x and
// pitn*u^y inside I1Ztx and I1Zty (:743-744); signed ratios and
// xi0 = rhomax = 1 in the regulator (FullyDiscrete...:556-650).
#pragma once

#include "vah_device.h"

namespace vah {

// Everything the source term reads around one cell.  Derivatives are the
// reference's centred differences: (f[s+1] - f[s-1]) * (1/d)/2.
struct SrcIn {
    double t, dt_inv, etabar_inv;
    double e, p, ut, ux, uy, un;      // at the cell
    double utp, uxp, uyp, unp;        // previous time level at the cell
    double dxut, dxux, dxuy, dxun, dyut, dyux, dyuy, dyun, dnut, dnux, dnuy, dnun;
    double dxp, dyp, dnp, dxe, dye, dne, dxpl, dypl, dnpl;
};

template <typename QA>
__device__ __forceinline__ void source_terms(const Eos &E, const SrcIn &in, const QA &Q, double (&S)[NQ],
                                             bool &nan) {
    const double t = in.t;
    const double ttt = Q[TTT], ttx = Q[TTX], tty = Q[TTY], ttn = Q[TTN], pl = Q[PL];
    const double pitt = Q[PITT], pitx = Q[PITX], pity = Q[PITY], pitn = Q[PITN], pixx = Q[PIXX],
                 pixy = Q[PIXY], pixn = Q[PIXN], piyy = Q[PIYY], piyn = Q[PIYN], pinn = Q[PINN];
    double e = in.e;
    const double p = in.p, ut = in.ut, ux = in.ux, uy = in.uy, un = in.un;
    const double dxut = in.dxut, dxux = in.dxux, dxuy = in.dxuy, dxun = in.dxun;
    const double dyut = in.dyut, dyux = in.dyux, dyuy = in.dyuy, dyun = in.dyun;
    const double dnut = in.dnut, dnux = in.dnux, dnuy = in.dnuy, dnun = in.dnun;

    const double iut = rcp(ut);
    const double it = rcp(t);
    const double vx = ux * iut, vy = uy * iut, vn = un * iut;
    const double dkvk = ((dxux - vx * dxut) + (dyuy - vy * dyut) + (dnun - vn * dnut)) * iut;

    const double t2 = t * t, it2 = it * it, un2 = un * un, ut2 = ut * ut;

    const double dtut = (ut - in.utp) * in.dt_inv, dtux = (ux - in.uxp) * in.dt_inv;
    const double dtuy = (uy - in.uyp) * in.dt_inv, dtun = (un - in.unp) * in.dt_inv;

    // (u.d) u^mu and covariant derivatives
    const double dut = ut * dtut + ux * dxut + uy * dyut + un * dnut;
    const double dux = ut * dtux + ux * dxux + uy * dyux + un * dnux;
    const double duy = ut * dtuy + ux * dxuy + uy * dyuy + un * dnuy;
    const double dun = ut * dtun + ux * dxun + uy * dyun + un * dnun;
    const double Dut = dut + t * un * un;
    const double Dux = -dux, Duy = -duy;
    const double Dun = -t2 * (dun + 2.0 * ut * un * it);

    const double theta = ut * it + dtut + dxux + dyuy + dnun;
    const double th3 = theta * kThird;

    // shear tensor sigma^{mu nu}
    const double stt = -t * ut * un2 + (dtut - ut * dut) + (ut2 - 1.0) * th3;
    const double stx = -0.5 * (t * un2 * ux) + 0.5 * (dtux - dxut) - 0.5 * (ux * dut + ut * dux) + ut * ux * th3;
    const double sty = -0.5 * (t * un2 * uy) + 0.5 * (dtuy - dyut) - 0.5 * (uy * dut + ut * duy) + ut * uy * th3;
    const double stn = -un * (2.0 * ut2 + t2 * un2) * 0.5 * it + 0.5 * (dtun - dnut * it2) - 0.5 * (un * dut + ut * dun) +
                       ut * un * th3;
    const double sxx = -(dxux + ux * dux) + (1.0 + ux * ux) * th3;
    const double sxy = -0.5 * (dxuy + dyux) - 0.5 * (uy * dux + ux * duy) + ux * uy * th3;
    const double sxn = -ut * ux * un * it - 0.5 * (dxun + dnux * it2) - 0.5 * (un * dux + ux * dun) + ux * un * th3;
    const double syy = -(dyuy + uy * duy) + (1.0 + uy * uy) * th3;
    const double syn = -ut * uy * un * it - 0.5 * (dyun + dnuy * it2) - 0.5 * (un * duy + uy * dun) + uy * un * th3;
    const double snn = -ut * (1.0 + 2.0 * t2 * un2) * it2 * it - dnun * it2 - un * dun + (it2 + un2) * th3;

    // longitudinal unit vector z^mu and the transverse / longitudinal split
    const double uTdtuT = ux * dtux + uy * dtuy;
    const double uTdxuT = ux * dxux + uy * dxuy;
    const double uTdyuT = ux * dyux + uy * dyuy;
    const double uTdnuT = ux * dnux + uy * dnuy;
    const double uTduT = ut * uTdtuT + ux * uTdxuT + uy * uTdyuT + un * uTdnuT;

    const double F = 1.0 + ux * ux + uy * uy;
    const double iF = rcp(F);
    const double iF2 = iF * iF;
    const double isF = rsqrt(F);

    const double z0 = t * un * isF;
    const double z3 = ut * it * isF;
    const double z02 = z0 * z0, z0z3 = z0 * z3, z32 = z3 * z3;

    const double dz0 = (ut * un + t * dun - t * un * iF * uTduT) * isF;
    const double dz3 = (-ut2 * it2 + dut * it - ut * it * iF * uTduT) * isF;
    const double Dz0 = dz0 + t * un * z3;
    const double Dz3 = -t2 * (dz3 + (ut * z3 + un * z0) * it);

    const double A = z02 * stt - 2.0 * t2 * z0z3 * stn + t2 * t2 * z32 * snn;
    const double B0 = z0 * stt - t2 * z3 * stn;
    const double B1 = z0 * stx - t2 * z3 * sxn;
    const double B2 = z0 * sty - t2 * z3 * syn;
    const double B3 = z0 * stn - t2 * z3 * snn;

    const double thetaT = 2.0 * th3 + A;
    const double zDzu = theta - thetaT;

    const double hA = 0.5 * A;
    const double sttT = stt + 2.0 * z0 * B0 + z02 * A - (1.0 - ut * ut + z02) * hA;
    const double stxT = stx + z0 * B1 + ut * ux * hA;
    const double styT = sty + z0 * B2 + ut * uy * hA;
    const double stnT = stn + z0 * B3 + z3 * B0 + z0z3 * A - (-ut * un + z0z3) * hA;
    const double sxxT = sxx + (1.0 + ux * ux) * hA;
    const double sxyT = sxy + ux * uy * hA;
    const double sxnT = sxn + z3 * B1 + ux * un * hA;
    const double syyT = syy + (1.0 + uy * uy) * hA;
    const double synT = syn + z3 * B2 + uy * un * hA;
    const double snnT = snn + 2.0 * z3 * B3 + z32 * A - (-it2 - un * un + z32) * hA;

    // anisotropic functions of a = P_L / e
    if (e == 0.0) e = 1.e-7;
    const double T = eos_T(E, e);
    const double ie = rcp(e);
    const double a = pl * ie;
    const Powers w = powers_of(a);
    const Fits fr = fits(w);

    const double ptHat = 0.5 * ((e - pl) - fr.Rbar0 * (e - 3.0 * p));
    const double pt = ptHat;
    const double DP = pl - pt;

    const double Ltt = DP * t2 * un2 * iF;
    const double Ltn = DP * ut * un * iF;
    const double Lnn = DP * ut2 * iF * it2;
    const double tnn = (e + pt) * un * un + pt * it2 + Lnn + pinn;

    const double dxpl = in.dxpl, dypl = in.dypl, dnpl = in.dnpl;
    const double dxe = in.dxe, dye = in.dye, dne = in.dne;
    const double dxRbar0 = fr.Rbar0P * (dxpl - a * dxe) * ie;
    const double dyRbar0 = fr.Rbar0P * (dypl - a * dye) * ie;
    const double dnRbar0 = fr.Rbar0P * (dnpl - a * dne) * ie;
    const double em3p = e - 3.0 * p;
    const double dxptHat = 0.5 * (dxe - dxpl - fr.Rbar0 * (dxe - 3.0 * in.dxp) - em3p * dxRbar0);
    const double dyptHat = 0.5 * (dye - dypl - fr.Rbar0 * (dye - 3.0 * in.dyp) - em3p * dyRbar0);
    const double dnptHat = 0.5 * (dne - dnpl - fr.Rbar0 * (dne - 3.0 * in.dnp) - em3p * dnRbar0);

    const double dxDP = dxpl - dxptHat, dyDP = dypl - dyptHat, dnDP = dnpl - dnptHat;

    const double DPiF = DP * iF, DPiF2 = DP * iF2;
    const double dxLtt = 2.0 * DPiF * dxun * t2 * un + dxDP * t2 * un2 * iF - 2.0 * DPiF2 * uTdxuT * t2 * un2;
    const double dyLtt = 2.0 * DPiF * dyun * t2 * un + dyDP * t2 * un2 * iF - 2.0 * DPiF2 * uTdyuT * t2 * un2;
    const double dnLtt = 2.0 * DPiF * dnun * t2 * un + dnDP * t2 * un2 * iF - 2.0 * DPiF2 * uTdnuT * t2 * un2;
    const double dxLtn = DPiF * dxut * un + DPiF * dxun * ut + dxDP * un * ut * iF - 2.0 * DPiF2 * uTdxuT * un * ut;
    const double dyLtn = DPiF * dyut * un + DPiF * dyun * ut + dyDP * un * ut * iF - 2.0 * DPiF2 * uTdyuT * un * ut;
    const double dnLtn = DPiF * dnut * un + DPiF * dnun * ut + dnDP * un * ut * iF - 2.0 * DPiF2 * uTdnuT * un * ut;
    const double dnLnn = (2.0 * DPiF * dnut * ut + dnDP * ut2 * iF - 2.0 * DPiF2 * uTdnuT * ut2) * it2;

    const double zeta_zz = fr.Rtilde * e - 3.0 * pl;
    const double zeta_zT = 0.5 * (fr.Rtilde * e + pl);

    const double lambda_lpi = fr.Rgamma;
    const double delta_pipi = 0.5 * (3.0 + fr.Rgamma);
    const double tau_pipi = 4.0 * delta_pipi * kThird;
    const double lambda_pipi = fr.Rgamma - 1.0;

    const double taupiInv = 0.2 * T * in.etabar_inv;

    const double psT = pitt * sttT - 2.0 * pitx * stxT - 2.0 * pity * styT + pixx * sxxT + 2.0 * pixy * sxyT +
                       piyy * syyT - 2.0 * pitn * stnT * t2 + 2.0 * pixn * sxnT * t2 + 2.0 * piyn * synT * t2 +
                       pinn * snnT * t2 * t2;

    const double dPL = -(pl - p) * taupiInv + zeta_zz * zDzu - zeta_zT * thetaT - lambda_lpi * psT;
    S[PL] = dPL * iut + dkvk * pl;

    S[TTT] = -(ttt * it + t * tnn) + dkvk * (Ltt + pitt - pt) - vx * dxptHat - vy * dyptHat - vn * dnptHat +
             vx * dxLtt + vy * dyLtt + vn * dnLtt - dnLtn;
    S[TTX] = -ttx * it - dxptHat + dkvk * pitx;
    S[TTY] = -tty * it - dyptHat + dkvk * pity;
    S[TTN] = -3.0 * ttn * it - dnptHat * it2 + dkvk * (Ltn + pitn) + vx * dxLtn + vy * dyLtn + 2.0 * vn * dnLtn -
             dnLnn;

    nan = nan || __builtin_isnan(S[PL]) || __builtin_isnan(S[TTT]);

    // transverse shear stress pi_perp^{mu nu}
    const double etaBar_TC = 0.125 * ((3.0 - fr.Rtilde) * e - 2.0 * pl);

    const double piDt = pitt * Dut + pitx * Dux + pity * Duy + pitn * Dun;
    const double I1tt = 2.0 * ut * piDt - 2.0 * z0 * (pitt * Dz0 + pitn * Dz3);
    const double I1tx = (pitt * ux + pitx * ut) * Dut + (pitx * ux + pixx * ut) * Dux + (pity * ux + pixy * ut) * Duy +
                        (pitn * ux + pixn * ut) * Dun - ((pitx * z0) * Dz0 + (pitn * ux + pixn * z0) * Dz3);
    const double I1ty = (pitt * uy + pity * ut) * Dut + (pitx * uy + pixy * ut) * Dux + (pity * uy + piyy * ut) * Duy +
                        (pitn * uy + piyn * ut) * Dun - ((pity * z0) * Dz0 + (pitn * uy + piyn * z0) * Dz3);
    const double I1tn = (pitt * un + pitn * ut) * Dut + (pitx * un + pixn * ut) * Dux + (pity * un + piyn * ut) * Duy +
                        (pitn * un + pinn * ut) * Dun -
                        ((pitt * z3 + pitn * z0) * Dz0 + (pitn * z3 + pinn * z0) * Dz3);
    const double I1xx = 2.0 * ux * (pitx * Dut + pixx * Dux + pixy * Duy + pixn * Dun);
    const double I1xy = (pitx * uy + pity * ux) * Dut + (pixx * uy + pixy * ux) * Dux + (pixy * uy + piyy * ux) * Duy +
                        (pixn * uy + piyn * ux) * Dun;
    const double I1xn = (pitx * un + pitn * ux) * Dut + (pixx * un + pixn * ux) * Dux + (pixy * un + piyn * ux) * Duy +
                        (pixn * un + pinn * ux) * Dun - ((pitx * z3) * Dz0 + (pixn * z3) * Dz3);
    const double I1yy = 2.0 * uy * (pity * Dut + pixy * Dux + piyy * Duy + piyn * Dun);
    const double I1yn = (pity * un + pitn * uy) * Dut + (pixy * un + pixn * uy) * Dux + (piyy * un + piyn * uy) * Duy +
                        (piyn * un + pinn * uy) * Dun - ((pity * z3) * Dz0 + (piyn * z3) * Dz3);
    const double I1nn = 2.0 * un * (pitn * Dut + pixn * Dux + piyn * Duy + pinn * Dun) -
                        2.0 * z3 * (pitn * Dz0 + pinn * Dz3);

    const double ps2 = 0.5 * psT;
    const double I4tt = (pitt * sttT - pitx * stxT - pity * styT - t2 * pitn * stnT) - (1.0 - ut2 + z02) * ps2;
    const double I4tx = 0.5 * ((pitt * stxT + pitx * sttT) - (pitx * sxxT + pixx * stxT) - (pity * sxyT + pixy * styT) -
                               t2 * (pitn * sxnT + pixn * stnT)) + (ut * ux) * ps2;
    const double I4ty = 0.5 * ((pitt * styT + pity * sttT) - (pitx * sxyT + pixy * stxT) - (pity * syyT + piyy * styT) -
                               t2 * (pitn * synT + piyn * stnT)) + (ut * uy) * ps2;
    const double I4tn = 0.5 * ((pitt * stnT + pitn * sttT) - (pitx * sxnT + pixn * stxT) - (pity * synT + piyn * styT) -
                               t2 * (pitn * snnT + pinn * stnT)) + (ut * un - z0z3) * ps2;
    const double I4xx = (pitx * stxT - pixx * sxxT - pixy * sxyT - t2 * pixn * sxnT) + (1.0 + ux * ux) * ps2;
    const double I4xy = 0.5 * ((pitx * styT + pity * stxT) - (pixx * sxyT + pixy * sxxT) - (pixy * syyT + piyy * sxyT) -
                               t2 * (pixn * synT + piyn * sxnT)) + (ux * uy) * ps2;
    const double I4xn = 0.5 * ((pitx * stnT + pitn * stxT) - (pixx * sxnT + pixn * sxxT) - (pixy * synT + piyn * sxyT) -
                               t2 * (pixn * snnT + pinn * sxnT)) + (ux * un) * ps2;
    const double I4yy = (pity * styT - pixy * sxyT - piyy * syyT - t2 * piyn * synT) + (1.0 + uy * uy) * ps2;
    const double I4yn = 0.5 * ((pity * stnT + pitn * styT) - (pixy * sxnT + pixn * sxyT) - (piyy * synT + piyn * syyT) -
                               t2 * (piyn * snnT + pinn * synT)) + (uy * un) * ps2;
    const double I4nn = (pitn * stnT - pixn * sxnT - piyn * synT - t2 * pinn * snnT) + (it2 + un2 - z32) * ps2;

    // I = I1 - delta_pipi I2 + tau_pipi I4 - lambda_pipi I5 with I2 = thetaT pi, I5 = zDzu pi
    const double c25 = delta_pipi * thetaT + lambda_pipi * zDzu;
    const double twoEta = 2.0 * etaBar_TC;

    const double dpitt = twoEta * sttT - pitt * taupiInv - (I1tt - c25 * pitt + tau_pipi * I4tt) - 2.0 * un * t * pitn;
    const double dpitx = twoEta * stxT - pitx * taupiInv - (I1tx - c25 * pitx + tau_pipi * I4tx) - un * t * pixn;
    const double dpity = twoEta * styT - pity * taupiInv - (I1ty - c25 * pity + tau_pipi * I4ty) - un * t * piyn;
    const double dpitn = twoEta * stnT - pitn * taupiInv - (I1tn - c25 * pitn + tau_pipi * I4tn) - un * t * pinn -
                         (ut * pitn + un * pitt) * it;
    const double dpixx = twoEta * sxxT - pixx * taupiInv - (I1xx - c25 * pixx + tau_pipi * I4xx);
    const double dpixy = twoEta * sxyT - pixy * taupiInv - (I1xy - c25 * pixy + tau_pipi * I4xy);
    const double dpixn = twoEta * sxnT - pixn * taupiInv - (I1xn - c25 * pixn + tau_pipi * I4xn) -
                         (ut * pixn + un * pitx) * it;
    const double dpiyy = twoEta * syyT - piyy * taupiInv - (I1yy - c25 * piyy + tau_pipi * I4yy);
    const double dpiyn = twoEta * synT - piyn * taupiInv - (I1yn - c25 * piyn + tau_pipi * I4yn) -
                         (ut * piyn + un * pity) * it;
    const double dpinn = twoEta * snnT - pinn * taupiInv - (I1nn - c25 * pinn + tau_pipi * I4nn) -
                         2.0 * (ut * pinn + un * pitn) * it;

    S[PITT] = dpitt * iut + pitt * dkvk;
    S[PITX] = dpitx * iut + pitx * dkvk;
    S[PITY] = dpity * iut + pity * dkvk;
    S[PITN] = dpitn * iut + pitn * dkvk;
    S[PIXX] = dpixx * iut + pixx * dkvk;
    S[PIXY] = dpixy * iut + pixy * dkvk;
    S[PIXN] = dpixn * iut + pixn * dkvk;
    S[PIYY] = dpiyy * iut + piyy * dkvk;
    S[PIYN] = dpiyn * iut + piyn * dkvk;
    S[PINN] = dpinn * iut + pinn * dkvk;
}

// Axis sources from the pi gradients (SourceTerms.cpp:46-176):
// d = (pi(+1) - pi(-1)) * (1/d)/2 of the listed components, v = u^i / u^tau.
struct AxisPi {
    double dtt, dtx, dty, dtn;  // d_i pi^{tau mu}
    double dix, diy, din;       // d_i pi^{i mu} for mu = x, y, eta (direction-dependent set)
};

template <int DIR>
__device__ __forceinline__ void axis_source(const AxisPi &d, double v, double (&S)[4]) {
    // DIR 1: pi^{x mu} = (pitx, pixx, pixy, pixn); DIR 2: (pity, pixy, piyy, piyn); DIR 3: (pitn, pixn, piyn, pinn)
    S[0] = d.dtt * v - (DIR == 1 ? d.dtx : DIR == 2 ? d.dty : d.dtn);
    S[1] = d.dtx * v - d.dix;
    S[2] = d.dty * v - d.diy;
    S[3] = d.dtn * v - d.din;
}

// Regulator (FullyDiscreteKurganovTadmorScheme.cpp:540-665), in place.
template <typename QA>
__device__ __forceinline__ void regulate(double t, QA &q, double &e, double &p, double ut, double ux, double uy,
                                         double un) {
    if (e < 0.0) {
        e = 1.e-7;
        p = 1.e-7;
    }
    if (q[PL] < 0.0) q[PL] = 1.e-7;
    const double pl = q[PL];
    const double ptHat = pt_hat(e, p, pl);
    const double pitt = q[PITT], pitx = q[PITX], pity = q[PITY], pitn = q[PITN], pixx = q[PIXX], pixy = q[PIXY],
                 pixn = q[PIXN], piyy = q[PIYY], piyn = q[PIYN], pinn = q[PINN];
    const double isF = 1.0 / __builtin_sqrt(1.0 + ux * ux + uy * uy);
    const double z0 = t * un * isF;
    const double z3 = ut / t * isF;
    const double t2 = t * t;
    const double sTrsTrs = __builtin_sqrt(fabs(e * e + pl * pl + 2.0 * ptHat * ptHat));
    const double pipi = pitt * pitt - 2.0 * pitx * pitx - 2.0 * pity * pity + pixx * pixx + 2.0 * pixy * pixy +
                        piyy * piyy - 2.0 * pitn * pitn * t2 + 2.0 * pixn * pixn * t2 + 2.0 * piyn * piyn * t2 +
                        pinn * pinn * t2 * t2;
    const double spipi = __builtin_sqrt(fabs(pipi));
    const double ar[9] = {pitt - pixx - piyy - pinn * t2,
                          -pitn * t2 * un + pitt * ut - pitx * ux - pity * uy,
                          -pixn * t2 * un + pitx * ut - pixx * ux - pixy * uy,
                          -piyn * t2 * un + pity * ut - pixy * ux - piyy * uy,
                          -pinn * t2 * un + pitn * ut - pixn * ux - piyn * uy,
                          z0 * pitt - t2 * z3 * pitn,
                          z0 * pitx - t2 * z3 * pixn,
                          z0 * pity - t2 * z3 * piyn,
                          z0 * pitn - t2 * z3 * pinn};
    // ratios to spipi; a zero spipi makes them 0/0 = NaN, which never wins
    double rho = spipi / sTrsTrs;
#pragma unroll
    for (int m = 0; m < 9; ++m) {
        double am = ar[m] / spipi;
        if (am > rho) rho = am;
    }
    double fac = tanh(rho) / rho;
    if (rho < 1.e-7) fac = 1.0;
#pragma unroll
    for (int n = PITT; n <= PINN; ++n) q[n] *= fac;
}

}  // namespace vah
