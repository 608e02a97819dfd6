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

// R is the arithmetic type of the source terms: double, or float for the
// split-precision form (the rational fits and the temperature stay double).
template <typename R, typename QA>
__device__ __forceinline__ void source_terms(const Eos &E, const SrcIn &in, const QA &Q, R (&S)[NQ], bool &nan) {
    const R t = in.t;
    const R ttt = Q[TTT], ttx = Q[TTX], tty = Q[TTY], ttn = Q[TTN], pl = Q[PL];
    const R pitt = Q[PITT], pitx = Q[PITX], pity = Q[PITY], pitn = Q[PITN], pixx = Q[PIXX],
                 pixy = Q[PIXY], pixn = Q[PIXN], piyy = Q[PIYY], piyn = Q[PIYN], pinn = Q[PINN];
    R e = (R)in.e;
    const R p = in.p, ut = in.ut, ux = in.ux, uy = in.uy, un = in.un;
    const R dxut = in.dxut, dxux = in.dxux, dxuy = in.dxuy, dxun = in.dxun;
    const R dyut = in.dyut, dyux = in.dyux, dyuy = in.dyuy, dyun = in.dyun;
    const R dnut = in.dnut, dnux = in.dnux, dnuy = in.dnuy, dnun = in.dnun;

    const R iut = rcp(ut);
    const R it = rcp(t);
    const R vx = ux * iut, vy = uy * iut, vn = un * iut;
    const R dkvk = ((dxux - vx * dxut) + (dyuy - vy * dyut) + (dnun - vn * dnut)) * iut;

    const R t2 = t * t, it2 = it * it, un2 = un * un, ut2 = ut * ut;

    const R dtut = (ut - in.utp) * in.dt_inv, dtux = (ux - in.uxp) * in.dt_inv;
    const R dtuy = (uy - in.uyp) * in.dt_inv, dtun = (un - in.unp) * in.dt_inv;

    // (u.d) u^mu and covariant derivatives
    const R dut = ut * dtut + ux * dxut + uy * dyut + un * dnut;
    const R dux = ut * dtux + ux * dxux + uy * dyux + un * dnux;
    const R duy = ut * dtuy + ux * dxuy + uy * dyuy + un * dnuy;
    const R dun = ut * dtun + ux * dxun + uy * dyun + un * dnun;
    const R Dut = dut + t * un * un;
    const R Dux = -dux, Duy = -duy;
    const R Dun = -t2 * (dun + R(2.0) * ut * un * it);

    const R theta = ut * it + dtut + dxux + dyuy + dnun;
    const R th3 = theta * R(kThird);

    // shear tensor sigma^{mu nu}
    const R stt = -t * ut * un2 + (dtut - ut * dut) + (ut2 - R(1.0)) * th3;
    const R stx = -R(0.5) * (t * un2 * ux) + R(0.5) * (dtux - dxut) - R(0.5) * (ux * dut + ut * dux) + ut * ux * th3;
    const R sty = -R(0.5) * (t * un2 * uy) + R(0.5) * (dtuy - dyut) - R(0.5) * (uy * dut + ut * duy) + ut * uy * th3;
    const R stn = -un * (R(2.0) * ut2 + t2 * un2) * R(0.5) * it + R(0.5) * (dtun - dnut * it2) - R(0.5) * (un * dut + ut * dun) +
                       ut * un * th3;
    const R sxx = -(dxux + ux * dux) + (R(1.0) + ux * ux) * th3;
    const R sxy = -R(0.5) * (dxuy + dyux) - R(0.5) * (uy * dux + ux * duy) + ux * uy * th3;
    const R sxn = -ut * ux * un * it - R(0.5) * (dxun + dnux * it2) - R(0.5) * (un * dux + ux * dun) + ux * un * th3;
    const R syy = -(dyuy + uy * duy) + (R(1.0) + uy * uy) * th3;
    const R syn = -ut * uy * un * it - R(0.5) * (dyun + dnuy * it2) - R(0.5) * (un * duy + uy * dun) + uy * un * th3;
    const R snn = -ut * (R(1.0) + R(2.0) * t2 * un2) * it2 * it - dnun * it2 - un * dun + (it2 + un2) * th3;

    // longitudinal unit vector z^mu and the transverse / longitudinal split
    const R uTdtuT = ux * dtux + uy * dtuy;
    const R uTdxuT = ux * dxux + uy * dxuy;
    const R uTdyuT = ux * dyux + uy * dyuy;
    const R uTdnuT = ux * dnux + uy * dnuy;
    const R uTduT = ut * uTdtuT + ux * uTdxuT + uy * uTdyuT + un * uTdnuT;

    const R F = R(1.0) + ux * ux + uy * uy;
    const R iF = rcp(F);
    const R iF2 = iF * iF;
    const R isF = rsqrt(F);

    const R z0 = t * un * isF;
    const R z3 = ut * it * isF;
    const R z02 = z0 * z0, z0z3 = z0 * z3, z32 = z3 * z3;

    const R dz0 = (ut * un + t * dun - t * un * iF * uTduT) * isF;
    const R dz3 = (-ut2 * it2 + dut * it - ut * it * iF * uTduT) * isF;
    const R Dz0 = dz0 + t * un * z3;
    const R Dz3 = -t2 * (dz3 + (ut * z3 + un * z0) * it);

    const R A = z02 * stt - R(2.0) * t2 * z0z3 * stn + t2 * t2 * z32 * snn;
    const R B0 = z0 * stt - t2 * z3 * stn;
    const R B1 = z0 * stx - t2 * z3 * sxn;
    const R B2 = z0 * sty - t2 * z3 * syn;
    const R B3 = z0 * stn - t2 * z3 * snn;

    const R thetaT = R(2.0) * th3 + A;
    const R zDzu = theta - thetaT;

    const R hA = R(0.5) * A;
    const R sttT = stt + R(2.0) * z0 * B0 + z02 * A - (R(1.0) - ut * ut + z02) * hA;
    const R stxT = stx + z0 * B1 + ut * ux * hA;
    const R styT = sty + z0 * B2 + ut * uy * hA;
    const R stnT = stn + z0 * B3 + z3 * B0 + z0z3 * A - (-ut * un + z0z3) * hA;
    const R sxxT = sxx + (R(1.0) + ux * ux) * hA;
    const R sxyT = sxy + ux * uy * hA;
    const R sxnT = sxn + z3 * B1 + ux * un * hA;
    const R syyT = syy + (R(1.0) + uy * uy) * hA;
    const R synT = syn + z3 * B2 + uy * un * hA;
    const R snnT = snn + R(2.0) * z3 * B3 + z32 * A - (-it2 - un * un + z32) * hA;

    // anisotropic functions of a = P_L / e
    if (e == R(0.0)) e = R(1.e-7);
    const R T = (R)eos_T(E, (double)e);
    const R ie = rcp(e);
    const R a = pl * ie;
    const Fits fr = fits(powers_of((double)a));
    const R Rtilde = (R)fr.Rtilde, Rbar0 = (R)fr.Rbar0, Rgamma = (R)fr.Rgamma, Rbar0P = (R)fr.Rbar0P;

    const R ptHat = R(0.5) * ((e - pl) - Rbar0 * (e - R(3.0) * p));
    const R pt = ptHat;
    const R DP = pl - pt;

    const R Ltt = DP * t2 * un2 * iF;
    const R Ltn = DP * ut * un * iF;
    const R Lnn = DP * ut2 * iF * it2;
    const R tnn = (e + pt) * un * un + pt * it2 + Lnn + pinn;

    const R dxpl = (R)in.dxpl, dypl = (R)in.dypl, dnpl = (R)in.dnpl;
    const R dxe = in.dxe, dye = in.dye, dne = in.dne;
    const R dxRbar0 = Rbar0P * (dxpl - a * dxe) * ie;
    const R dyRbar0 = Rbar0P * (dypl - a * dye) * ie;
    const R dnRbar0 = Rbar0P * (dnpl - a * dne) * ie;
    const R em3p = e - R(3.0) * p;
    const R dxptHat = R(0.5) * (dxe - dxpl - Rbar0 * (dxe - R(3.0) * (R)in.dxp) - em3p * dxRbar0);
    const R dyptHat = R(0.5) * (dye - dypl - Rbar0 * (dye - R(3.0) * (R)in.dyp) - em3p * dyRbar0);
    const R dnptHat = R(0.5) * (dne - dnpl - Rbar0 * (dne - R(3.0) * (R)in.dnp) - em3p * dnRbar0);

    const R dxDP = dxpl - dxptHat, dyDP = dypl - dyptHat, dnDP = dnpl - dnptHat;

    const R DPiF = DP * iF, DPiF2 = DP * iF2;
    const R dxLtt = R(2.0) * DPiF * dxun * t2 * un + dxDP * t2 * un2 * iF - R(2.0) * DPiF2 * uTdxuT * t2 * un2;
    const R dyLtt = R(2.0) * DPiF * dyun * t2 * un + dyDP * t2 * un2 * iF - R(2.0) * DPiF2 * uTdyuT * t2 * un2;
    const R dnLtt = R(2.0) * DPiF * dnun * t2 * un + dnDP * t2 * un2 * iF - R(2.0) * DPiF2 * uTdnuT * t2 * un2;
    const R dxLtn = DPiF * dxut * un + DPiF * dxun * ut + dxDP * un * ut * iF - R(2.0) * DPiF2 * uTdxuT * un * ut;
    const R dyLtn = DPiF * dyut * un + DPiF * dyun * ut + dyDP * un * ut * iF - R(2.0) * DPiF2 * uTdyuT * un * ut;
    const R dnLtn = DPiF * dnut * un + DPiF * dnun * ut + dnDP * un * ut * iF - R(2.0) * DPiF2 * uTdnuT * un * ut;
    const R dnLnn = (R(2.0) * DPiF * dnut * ut + dnDP * ut2 * iF - R(2.0) * DPiF2 * uTdnuT * ut2) * it2;

    const R zeta_zz = Rtilde * e - R(3.0) * pl;
    const R zeta_zT = R(0.5) * (Rtilde * e + pl);

    const R lambda_lpi = Rgamma;
    const R delta_pipi = R(0.5) * (R(3.0) + Rgamma);
    const R tau_pipi = R(4.0) * delta_pipi * R(kThird);
    const R lambda_pipi = Rgamma - R(1.0);

    const R taupiInv = R(0.2) * T * in.etabar_inv;

    const R psT = pitt * sttT - R(2.0) * pitx * stxT - R(2.0) * pity * styT + pixx * sxxT + R(2.0) * pixy * sxyT +
                       piyy * syyT - R(2.0) * pitn * stnT * t2 + R(2.0) * pixn * sxnT * t2 + R(2.0) * piyn * synT * t2 +
                       pinn * snnT * t2 * t2;

    const R dPL = -(pl - p) * taupiInv + zeta_zz * zDzu - zeta_zT * thetaT - lambda_lpi * psT;
    S[PL] = dPL * iut + dkvk * pl;

    S[TTT] = -(ttt * it + t * tnn) + dkvk * (Ltt + pitt - pt) - vx * dxptHat - vy * dyptHat - vn * dnptHat +
             vx * dxLtt + vy * dyLtt + vn * dnLtt - dnLtn;
    S[TTX] = -ttx * it - dxptHat + dkvk * pitx;
    S[TTY] = -tty * it - dyptHat + dkvk * pity;
    S[TTN] = -R(3.0) * ttn * it - dnptHat * it2 + dkvk * (Ltn + pitn) + vx * dxLtn + vy * dyLtn + R(2.0) * vn * dnLtn -
             dnLnn;

    nan = nan || __builtin_isnan(S[PL]) || __builtin_isnan(S[TTT]);

    // transverse shear stress pi_perp^{mu nu}
    const R etaBar_TC = R(0.125) * ((R(3.0) - Rtilde) * e - R(2.0) * pl);

    const R piDt = pitt * Dut + pitx * Dux + pity * Duy + pitn * Dun;
    const R I1tt = R(2.0) * ut * piDt - R(2.0) * z0 * (pitt * Dz0 + pitn * Dz3);
    const R I1tx = (pitt * ux + pitx * ut) * Dut + (pitx * ux + pixx * ut) * Dux + (pity * ux + pixy * ut) * Duy +
                        (pitn * ux + pixn * ut) * Dun - ((pitx * z0) * Dz0 + (pitn * ux + pixn * z0) * Dz3);
    const R I1ty = (pitt * uy + pity * ut) * Dut + (pitx * uy + pixy * ut) * Dux + (pity * uy + piyy * ut) * Duy +
                        (pitn * uy + piyn * ut) * Dun - ((pity * z0) * Dz0 + (pitn * uy + piyn * z0) * Dz3);
    const R I1tn = (pitt * un + pitn * ut) * Dut + (pitx * un + pixn * ut) * Dux + (pity * un + piyn * ut) * Duy +
                        (pitn * un + pinn * ut) * Dun -
                        ((pitt * z3 + pitn * z0) * Dz0 + (pitn * z3 + pinn * z0) * Dz3);
    const R I1xx = R(2.0) * ux * (pitx * Dut + pixx * Dux + pixy * Duy + pixn * Dun);
    const R I1xy = (pitx * uy + pity * ux) * Dut + (pixx * uy + pixy * ux) * Dux + (pixy * uy + piyy * ux) * Duy +
                        (pixn * uy + piyn * ux) * Dun;
    const R I1xn = (pitx * un + pitn * ux) * Dut + (pixx * un + pixn * ux) * Dux + (pixy * un + piyn * ux) * Duy +
                        (pixn * un + pinn * ux) * Dun - ((pitx * z3) * Dz0 + (pixn * z3) * Dz3);
    const R I1yy = R(2.0) * uy * (pity * Dut + pixy * Dux + piyy * Duy + piyn * Dun);
    const R I1yn = (pity * un + pitn * uy) * Dut + (pixy * un + pixn * uy) * Dux + (piyy * un + piyn * uy) * Duy +
                        (piyn * un + pinn * uy) * Dun - ((pity * z3) * Dz0 + (piyn * z3) * Dz3);
    const R I1nn = R(2.0) * un * (pitn * Dut + pixn * Dux + piyn * Duy + pinn * Dun) -
                        R(2.0) * z3 * (pitn * Dz0 + pinn * Dz3);

    const R ps2 = R(0.5) * psT;
    const R I4tt = (pitt * sttT - pitx * stxT - pity * styT - t2 * pitn * stnT) - (R(1.0) - ut2 + z02) * ps2;
    const R I4tx = R(0.5) * ((pitt * stxT + pitx * sttT) - (pitx * sxxT + pixx * stxT) - (pity * sxyT + pixy * styT) -
                               t2 * (pitn * sxnT + pixn * stnT)) + (ut * ux) * ps2;
    const R I4ty = R(0.5) * ((pitt * styT + pity * sttT) - (pitx * sxyT + pixy * stxT) - (pity * syyT + piyy * styT) -
                               t2 * (pitn * synT + piyn * stnT)) + (ut * uy) * ps2;
    const R I4tn = R(0.5) * ((pitt * stnT + pitn * sttT) - (pitx * sxnT + pixn * stxT) - (pity * synT + piyn * styT) -
                               t2 * (pitn * snnT + pinn * stnT)) + (ut * un - z0z3) * ps2;
    const R I4xx = (pitx * stxT - pixx * sxxT - pixy * sxyT - t2 * pixn * sxnT) + (R(1.0) + ux * ux) * ps2;
    const R I4xy = R(0.5) * ((pitx * styT + pity * stxT) - (pixx * sxyT + pixy * sxxT) - (pixy * syyT + piyy * sxyT) -
                               t2 * (pixn * synT + piyn * sxnT)) + (ux * uy) * ps2;
    const R I4xn = R(0.5) * ((pitx * stnT + pitn * stxT) - (pixx * sxnT + pixn * sxxT) - (pixy * synT + piyn * sxyT) -
                               t2 * (pixn * snnT + pinn * sxnT)) + (ux * un) * ps2;
    const R I4yy = (pity * styT - pixy * sxyT - piyy * syyT - t2 * piyn * synT) + (R(1.0) + uy * uy) * ps2;
    const R I4yn = R(0.5) * ((pity * stnT + pitn * styT) - (pixy * sxnT + pixn * sxyT) - (piyy * synT + piyn * syyT) -
                               t2 * (piyn * snnT + pinn * synT)) + (uy * un) * ps2;
    const R I4nn = (pitn * stnT - pixn * sxnT - piyn * synT - t2 * pinn * snnT) + (it2 + un2 - z32) * ps2;

    // I = I1 - delta_pipi I2 + tau_pipi I4 - lambda_pipi I5 with I2 = thetaT pi, I5 = zDzu pi
    const R c25 = delta_pipi * thetaT + lambda_pipi * zDzu;
    const R twoEta = R(2.0) * etaBar_TC;

    const R dpitt = twoEta * sttT - pitt * taupiInv - (I1tt - c25 * pitt + tau_pipi * I4tt) - R(2.0) * un * t * pitn;
    const R dpitx = twoEta * stxT - pitx * taupiInv - (I1tx - c25 * pitx + tau_pipi * I4tx) - un * t * pixn;
    const R dpity = twoEta * styT - pity * taupiInv - (I1ty - c25 * pity + tau_pipi * I4ty) - un * t * piyn;
    const R dpitn = twoEta * stnT - pitn * taupiInv - (I1tn - c25 * pitn + tau_pipi * I4tn) - un * t * pinn -
                         (ut * pitn + un * pitt) * it;
    const R dpixx = twoEta * sxxT - pixx * taupiInv - (I1xx - c25 * pixx + tau_pipi * I4xx);
    const R dpixy = twoEta * sxyT - pixy * taupiInv - (I1xy - c25 * pixy + tau_pipi * I4xy);
    const R dpixn = twoEta * sxnT - pixn * taupiInv - (I1xn - c25 * pixn + tau_pipi * I4xn) -
                         (ut * pixn + un * pitx) * it;
    const R dpiyy = twoEta * syyT - piyy * taupiInv - (I1yy - c25 * piyy + tau_pipi * I4yy);
    const R dpiyn = twoEta * synT - piyn * taupiInv - (I1yn - c25 * piyn + tau_pipi * I4yn) -
                         (ut * piyn + un * pity) * it;
    const R dpinn = twoEta * snnT - pinn * taupiInv - (I1nn - c25 * pinn + tau_pipi * I4nn) -
                         R(2.0) * (ut * pinn + un * pitn) * it;

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
template <typename R>
struct AxisPi {
    R dtt, dtx, dty, dtn;  // d_i pi^{tau mu}
    R dix, diy, din;       // d_i pi^{i mu} for mu = x, y, eta (direction-dependent set)
};

template <int DIR, typename R>
__device__ __forceinline__ void axis_source(const AxisPi<R> &d, R v, R (&S)[4]) {
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
    const double isF = rsqrt(1.0 + ux * ux + uy * uy);
    const double z0 = t * un * isF;
    const double z3 = ut / t * isF;
    const double t2 = t * t;
    const double sTrsTrs = sqrt_pos(fabs(e * e + pl * pl + 2.0 * ptHat * ptHat));
    const double pipi = pitt * pitt - 2.0 * pitx * pitx - 2.0 * pity * pity + pixx * pixx + 2.0 * pixy * pixy +
                        piyy * piyy - 2.0 * pitn * pitn * t2 + 2.0 * pixn * pixn * t2 + 2.0 * piyn * piyn * t2 +
                        pinn * pinn * t2 * t2;
    const double spipi = sqrt_pos(fabs(pipi));
    const double ar[9] = {pitt - pixx - piyy - pinn * t2,
                          -pitn * t2 * un + pitt * ut - pitx * ux - pity * uy,
                          -pixn * t2 * un + pitx * ut - pixx * ux - pixy * uy,
                          -piyn * t2 * un + pity * ut - pixy * ux - piyy * uy,
                          -pinn * t2 * un + pitn * ut - pixn * ux - piyn * uy,
                          z0 * pitt - t2 * z3 * pitn,
                          z0 * pitx - t2 * z3 * pixn,
                          z0 * pity - t2 * z3 * piyn,
                          z0 * pitn - t2 * z3 * pinn};
    // ratios to spipi; a zero spipi makes them 0 * inf = NaN (or +-inf for a
    // nonzero numerator), as the quotients would be
    double rho = spipi / sTrsTrs;
    const double ispipi = 1.0 / spipi;
#pragma unroll
    for (int m = 0; m < 9; ++m) {
        double am = ar[m] * ispipi;
        if (am > rho) rho = am;
    }
    double fac = tanh_over_x(rho);  // tanh(rho) / rho
    if (rho < 1.e-7) fac = 1.0;
#pragma unroll
    for (int n = PITT; n <= PINN; ++n) q[n] *= fac;
}

}  // namespace vah
