// vah_stage.h -- the fused Kurganov-Tadmor Heun stage kernel for gfx950.
//
// One launch performs one Euler stage of the reference's rungeKutta2
// (old_cpu-vah_code/src/FullyDiscreteKurganovTadmorScheme.cpp:761-789) for
// every interior cell:
//   out = in + dt*S  (+= x, y, eta flux differences and axis sources)
//   [stage 2: out = (out + q^n)/2]
//   Newton recovery of (e, p, u) at tau + dt, regulation, and the
//   zero-gradient ghost copies of all outputs (DynamicalVariables.cpp:349-489).
//
// Work decomposition.  A workgroup of 256 lanes (4 wave64) owns a 16 x 15
// tile of cells in (x, y) and marches through KZ consecutive eta planes.
// Face fluxes are the expensive part (two Newton solves each).  With
// SHARED, each x or y face of the tile is solved once by one lane, its
// 15-component flux parked in LDS and read by the two cells that share it;
// the eta face between planes k and k+1 is solved by the cell's own lane and
// carried to the next plane in an LDS slot only that lane touches (in
// registers it would push the kernel into spilling).  Without SHARED every cell solves
// both of its faces with its own Newton guess, the reference's work split.
// Workgroups are mapped to tiles so that the eight XCDs work through the eta
// chunks together (tile_of_block).
#pragma once

#include "vah_device.h"
#include "vah_source.h"

namespace vah {

// workgroups per CU the register allocation is sized for (occupancy target)
#ifndef VAH_BLOCKS_2D
#define VAH_BLOCKS_2D 2
#endif
#ifndef VAH_BLOCKS_3D
#define VAH_BLOCKS_3D 2
#endif

// default tile: 16 x 15 cells (x faces 17 x 15, y faces 16 x 16 fill 255 / 256 lanes)
constexpr int kTileX = 16;
constexpr int kTileY = 15;
constexpr int NTHREADS = 256;

// a tile shape fits one workgroup when its cells and both face sets do
template <int TX, int TY>
constexpr bool tile_fits() {
    return TX * TY <= NTHREADS && (TX + 1) * TY <= NTHREADS && TX * (TY + 1) <= NTHREADS;
}

struct StageParams {
    const double *qin;    // 15-field slab the stage reads
    double *qout;         // 15-field slab it writes
    const double *qn;     // q at tau_n (stage 2 convex combination)
    const double *e_in, *p_in;
    double *e_out, *p_out;
    const double *ucur;   // u used for gradients and face Newton guesses
    const double *uprev;  // u one level back (time derivative)
    double *uout;
    long fs;              // elements between fields of a slab
    int nx, ny, nz;       // interior cells
    int ncx, ncy;         // computational extents (interior + 4)
    long sy, sz;          // strides in y and eta
    int tiles_x, tiles_y, zchunks, kz;
    int nblocks;          // tiles of the lattice (tiles_x * tiles_y * zchunks)
    double t_src, t_rec;  // tau of the sources/fluxes, tau of the recovery
    double dt, dt_inv, etabar, etabar_inv;
    double inv_dx, inv_dy, inv_dz;  // 1/d
    double hx, hy, hz;              // (1/d)/2, the reference's centred-difference factor
    double dt_dx, dt_dy, dt_dz;     // dt * (1/d), the weight of a flux difference
    Eos eos;
    unsigned int *counters;  // [0] NaN cells, [1] recovery fallbacks
    unsigned long long *emax;  // running max of the new e (ordered bits), or null
};

// Workgroup -> tile.  The hardware deals workgroups round-robin over the
// eight XCDs (workgroup b runs on XCD b % 8).  All XCDs work through the eta
// chunks together, each on its own slice of a chunk's tiles: XCD x takes the
// tiles x * slice .. (x + 1) * slice - 1 of chunk 0, then of chunk 1, and so
// on.  The planes in flight are then those of a few neighbouring chunks, one
// band that the memory-side cache can hold, and a tile's eta neighbour in the
// next chunk runs on the same XCD shortly after it; giving each XCD a
// contiguous run of all tiles instead keeps eight bands in flight, one per
// XCD (7 % slower at 201x201x101).  The grid holds zchunks * 8 * slice
// workgroups; those past the end of a chunk's tiles get -1.
__device__ __forceinline__ int tile_of_block(const StageParams &P, int b) {
    const int tpl = P.tiles_x * P.tiles_y, slice = (tpl + 7) >> 3;
    const int l = b >> 3, c = l / slice, within = (b & 7) * slice + (l - c * slice);
    return within < tpl && c < P.zchunks ? c * tpl + within : -1;
}

// ---------------------------------------------------------------------------
// KT flux through one cell face (SemiDiscreteKurganovTadmorScheme.cpp:13-57)
// ---------------------------------------------------------------------------

// Newton inputs of one reconstructed face state (EnergyMomentumTensor.cpp:116-143
// with W = Pi = 0): M^mu = T^{tau mu} - pi^{tau mu}, F = M^3/(M^0 + P_L).
struct FaceState {
    double M0, M1, M2, M3, PL, F, Ft, x, MBT;
};

__device__ __forceinline__ FaceState face_state(double t, double ttt, double ttx, double tty, double ttn, double pl,
                                                double pitt, double pitx, double pity, double pitn) {
    FaceState f;
    f.M0 = ttt - pitt;
    f.M1 = ttx - pitx;
    f.M2 = tty - pity;
    f.M3 = ttn - pitn;
    f.PL = pl;
    double M0PL = f.M0 + pl;
    if (M0PL == 0.0) M0PL = 1.e-16;
    f.F = fdiv(f.M3, M0PL);
    f.Ft = t * f.F;
    f.x = sqrt_pos(fabs(1.0 - f.Ft * f.Ft));
    f.MBT = sqrt_pos(f.M1 * f.M1 + f.M2 * f.M2);
    if (f.MBT == 0.0) f.MBT = 1.e-16;
    return f;
}

// One Newton update of u_T for state S from u0 (EnergyMomentumTensor.cpp:45-70),
// evaluated without branches: *u1 receives the update; the status is 1
// converged, -1 failed (e < 0 at u0, or an update that is NaN, infinite,
// negative or above 9.0072e15), 0 continue.  The reference stops before the
// update when e < 0; the update computed here is then discarded, so the
// decisions and the values kept are the reference's.
__device__ __forceinline__ int newton_update(double t, const FaceState &S, double u0, double *u1) {
    const double Ft2 = S.Ft * S.Ft;
    const double xM = S.x * S.MBT;
    const double bL = S.x * S.x * S.M0 - Ft2 * S.PL;
    const double b = S.x * S.x + Ft2;
    const double c0 = S.M0 - t * S.Ft * S.M3;
    double rs = rsqrt(1.0 + u0 * u0);
    double U = u0 * rs;
    double e = c0 - U * xM;
    double iP = rs * rs * rs;
    double g = bL + b * 0.5 * (e - S.PL);
    double f = U * g - xM;
    double fp = iP * g - 0.5 * U * b * xM * iP;
    fp = fabs(fp) == 0.0 ? 1.e-16 : fp;
    double u = u0 - newton_quotient(f, fp);
    *u1 = u;
    // non-short-circuit operators: no branch splits the two sides' chains
    const bool bad = (e < 0.0) | newton_out_of_range(u);
    const bool conv = fabs(u - u0) <= 1.e-7 * fabs(u);
    return bad ? -1 : (int)conv;
}

// The root a face solve's Newton path ends on, and when the path is known to
// end there.  With Pi = 0 the function the update zeroes is a quadratic in
// U = u / sqrt(1 + u^2):
//     f = U (bL + b (e - PL) / 2) - xM,  e = c0 - U xM
//       = -(b xM / 2) U^2 + B U - xM,    B = bL + b (c0 - PL) / 2,
// whose smaller root U* = 2 xM / (B + sqrt(B^2 - 2 b xM^2)) is the physical
// one (the roots' product is 2 / b = 2 for Ft^2 <= 1, and U < 1 for every
// finite u).  On 0 <= U <= U_v = B / (b xM), left of the vertex, f is
// increasing and concave in u, so from a start there every Newton iterate
// after the first lies between the first update and the root: e, which falls
// as U grows, stays at or above e(U*) on all of them and u stays in [0, u*].
// A path whose start lies there, whose first update passed the tests and is
// clearly positive, and whose root has e(U*) clearly positive, can therefore
// neither fail nor end anywhere but on the root.
struct FaceRoot {
    double u;    // u* (when ok)
    double B;    // the vertex test: start U0 with U0 * bxM <= B
    double bxM;
    bool ok;     // a root in [0, 1) with e(U*) clearly positive
};

__device__ __forceinline__ FaceRoot face_root(double t, const FaceState &S) {
    const double Ft2 = S.Ft * S.Ft;
    const double xM = S.x * S.MBT;
    const double bL = S.x * S.x * S.M0 - Ft2 * S.PL;
    const double b = S.x * S.x + Ft2;
    const double c0 = S.M0 - t * S.Ft * S.M3;
    FaceRoot r;
    r.B = bL + 0.5 * b * (c0 - S.PL);
    r.bxM = b * xM;
    const double disc = r.B * r.B - 2.0 * b * xM * xM;
    const double U = fdiv(2.0 * xM, r.B + sqrt_pos(fmax(disc, 0.0)));
    r.ok = (disc >= 0.0) & (r.B > 0.0) & (U >= 0.0) & (U < 1.0) & (c0 - U * xM > 1.e-12 * fabs(c0));
    r.u = U * rsqrt(fmax(1.0 - U * U, 1.e-300));
    return r;
}

// Both sides of a face solved in one loop (independent iterations, each side
// stops at its own convergence or failure, at most 10^4 iterations each),
// from `guess`.  Every pass evaluates both updates and keeps each only while
// its side is still running, so the two dependency chains interleave and the
// loop body has no branches.  With ROOT_JUMP a side whose first pass shows
// that its path must end on the root (face_root) takes the root there and
// stops: two to three passes fewer, and every decision (which sides fail)
// is the one the full path makes.  Outputs the signed speed v = u^DIR / u^tau
// of each side; a failed recovery gives the reference's fallback
// u = (1,0,0,0), i.e. v = 0.
template <int DIR, bool ROOT_JUMP>
__device__ __forceinline__ void face_speeds(double t, double guess, const FaceState &A, const FaceState &B, double &vA,
                                            double &vB) {
    // the first pass, outside the loop: both sides start at `guess`
    double ua, ub;
    int sa = newton_update(t, A, guess, &ua);
    int sb = newton_update(t, B, guess, &ub);
    if (ROOT_JUMP) {
        const FaceRoot qa = face_root(t, A), qb = face_root(t, B);
        const double U0 = guess * rsqrt(1.0 + guess * guess);  // the first pass's U
        const bool ja = (sa == 0) & qa.ok & (U0 * qa.bxM <= qa.B) & (ua >= 1.e-9 * guess);
        const bool jb = (sb == 0) & qb.ok & (U0 * qb.bxM <= qb.B) & (ub >= 1.e-9 * guess);
        ua = ja ? qa.u : ua;
        sa = ja ? 1 : sa;
        ub = jb ? qb.u : ub;
        sb = jb ? 1 : sb;
    }
    for (int it = 1; it < 10000 && (sa == 0 || sb == 0); ++it) {
        double na, nb;
        const int ta = newton_update(t, A, ua, &na);
        const int tb = newton_update(t, B, ub, &nb);
        const bool ra = sa == 0, rb = sb == 0;
        ua = ra ? na : ua;
        sa = ra ? ta : sa;
        ub = rb ? nb : ub;
        sb = rb ? tb : sb;
    }
    auto speed = [](const FaceState &S, double u, int st) {
        if (st != 1) return 0.0;
        if (DIR == 3) return S.F;  // u^eta / u^tau = F C/x / (C/x)
        double iC = rsqrt(1.0 + u * u);
        double num = DIR == 1 ? S.M1 : S.M2;
        return u * num * S.x * iC * rcp(S.MBT);  // (u M/MBT) / (C/x)
    };
    vA = speed(A, ua, sa);
    vB = speed(B, ub, sb);
}

// ---------------------------------------------------------------------------
// Field sets seen through buffer resources.  A load or store takes the byte
// offset of the cell (per lane, 32-bit) and the byte offset of the field
// (uniform, an SGPR), so the 15-field stencils need no per-lane 64-bit address
// arithmetic.  An access outside the set's byte range reads 0 and drops the
// store instead of faulting.  A set must stay below 4 GB (checked on the host).
// ---------------------------------------------------------------------------
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));

constexpr int kBufWord3 = 0x00020000;  // raw dword access, gfx9 family

struct Buf {
    __amdgpu_buffer_rsrc_t r;
};

__device__ __forceinline__ Buf make_buf(const double *base, int nfields, long fs) {
    Buf b;
    b.r = __builtin_amdgcn_make_buffer_rsrc((void *)base, 0, (int)(uint32_t)((long)nfields * fs * 8), kBufWord3);
    return b;
}

__device__ __forceinline__ double ld(const Buf &b, uint32_t cell, uint32_t field) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(b.r, (int)cell, (int)field, 0));
}

__device__ __forceinline__ void st(const Buf &b, uint32_t cell, uint32_t field, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), b.r, (int)cell, (int)field, 0);
}

// the buffers of one stage, and the byte strides of the layout
struct Views {
    Buf qin, qout, qn, e_in, p_in, e_out, p_out, ucur, uprev, uout;
    uint32_t f8;        // field to field
    uint32_t sy8, sz8;  // row to row, plane to plane
};

__device__ __forceinline__ Views make_views(const StageParams &P) {
    Views v;
    v.qin = make_buf(P.qin, NQ, P.fs);
    v.qout = make_buf(P.qout, NQ, P.fs);
    v.qn = make_buf(P.qn, NQ, P.fs);
    v.e_in = make_buf(P.e_in, 1, P.fs);
    v.p_in = make_buf(P.p_in, 1, P.fs);
    v.e_out = make_buf(P.e_out, 1, P.fs);
    v.p_out = make_buf(P.p_out, 1, P.fs);
    v.ucur = make_buf(P.ucur, 4, P.fs);
    v.uprev = make_buf(P.uprev, 4, P.fs);
    v.uout = make_buf(P.uout, 4, P.fs);
    v.f8 = (uint32_t)(P.fs * 8);
    v.sy8 = (uint32_t)(P.sy * 8);
    v.sz8 = (uint32_t)(P.sz * 8);
    return v;
}

// Flux through the face between cells c-st and c (byte offsets).  The
// left/right states are the reference's half-cell extrapolations
// (HalfSiteExtrapolation.cpp:12-26); emit(n, H_n) receives the 15 components
// in order.  The nine components the Newton solve needs stay in registers,
// the other six are reconstructed after it, which keeps the live state small.
template <int DIR, typename R, bool ROOT_JUMP, typename Guess, typename Load, typename Emit>
__device__ __forceinline__ void face_flux_from(double t, Guess guess, Load load, Emit emit) {
    constexpr int NN = 9;  // TTT..PITN
    // all loads of the nine fields first, so that they are in flight together;
    // load(n, j) gives field n at stencil position j (0..3: c-2st .. c+st)
    double raw[NN][4];
#pragma unroll
    for (int n = 0; n < NN; ++n) {
#pragma unroll
        for (int j = 0; j < 4; ++j) raw[n][j] = load(n, j);
    }
    // the Newton guess |u_T| (loads of its own) after the stencil loads, so
    // that one wait covers both
    const double uT_guess = guess();
    R qL[NN], qR[NN];
#pragma unroll
    for (int n = 0; n < NN; ++n) {
        R q0 = (R)raw[n][0], q1 = (R)raw[n][1], q2 = (R)raw[n][2], q3 = (R)raw[n][3];
        face_values(q0, q1, q2, q3, qL[n], qR[n]);
    }
    // the Newton recovery of both sides runs in double whatever R is
    FaceState FR = face_state(t, qR[TTT], qR[TTX], qR[TTY], qR[TTN], qR[PL], qR[PITT], qR[PITX], qR[PITY], qR[PITN]);
    FaceState FL = face_state(t, qL[TTT], qL[TTX], qL[TTY], qL[TTN], qL[PL], qL[PITT], qL[PITX], qL[PITY], qL[PITN]);
    double vR, vL;
    face_speeds<DIR, ROOT_JUMP>(t, uT_guess, FR, FL, vR, vL);
    const double a = fmax(fabs(vL), fabs(vR));
    // (F(qR) + F(qL) - a (qR - qL)) / 2 with F(q) = v q
    const R cR = (R)(0.5 * (vR - a)), cL = (R)(0.5 * (vL + a));
#pragma unroll
    for (int n = 0; n < NN; ++n) emit(n, cR * qR[n] + cL * qL[n]);
    double rest[NQ - NN][4];
#pragma unroll
    for (int n = NN; n < NQ; ++n) {
#pragma unroll
        for (int j = 0; j < 4; ++j) rest[n - NN][j] = load(n, j);
    }
#pragma unroll
    for (int n = NN; n < NQ; ++n) {
        R q0 = (R)rest[n - NN][0], q1 = (R)rest[n - NN][1], q2 = (R)rest[n - NN][2], q3 = (R)rest[n - NN][3];
        R l, r;
        face_values(q0, q1, q2, q3, l, r);
        emit(n, cR * r + cL * l);
    }
}

// Flux through the face between cells c-st and c (byte offsets).  The
// left/right states are the reference's half-cell extrapolations
// (HalfSiteExtrapolation.cpp:12-26); emit(n, H_n) receives the 15 components
// in order.  The nine components the Newton solve needs stay in registers,
// the other six are loaded and reconstructed after it.
template <int DIR, typename R, bool ROOT_JUMP, typename Guess, typename Emit>
__device__ __forceinline__ void face_flux(double t, Guess guess, const Buf &q, uint32_t f8, uint32_t c,
                                          uint32_t stp, Emit emit) {
    face_flux_from<DIR, R, ROOT_JUMP>(
        t, guess, [&](int n, int j) { return ld(q, c + (uint32_t)j * stp - 2u * stp, n * f8); }, emit);
}

__device__ __forceinline__ double uT_of(const Buf &u, uint32_t f8, uint32_t c) {
    double ux = ld(u, c, f8), uy = ld(u, c, 2 * f8);
    return sqrt_pos(fabs(ux * ux + uy * uy));
}

// d_i of the pi components entering the axis source of direction DIR
// (all fourteen loads issued before the differences are formed)
template <int DIR, typename R>
__device__ __forceinline__ AxisPi<R> axis_pi(const Buf &q, uint32_t f8, uint32_t s, uint32_t stp, double h) {
    // pi^{tau mu} (4), then pi^{DIR mu} for mu = x, y, eta
    constexpr int F[7] = {PITT, PITX, PITY, PITN, DIR == 1 ? PIXX : DIR == 2 ? PIXY : PIXN,
                          DIR == 1 ? PIXY : DIR == 2 ? PIYY : PIYN, DIR == 1 ? PIXN : DIR == 2 ? PIYN : PINN};
    double up[7], dn[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        up[i] = ld(q, s + stp, F[i] * f8);
        dn[i] = ld(q, s - stp, F[i] * f8);
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the loads together (not sunk to their uses)
    auto d = [&](int i) { return (R)((up[i] - dn[i]) * h); };
    AxisPi<R> r;
    r.dtt = d(0);
    r.dtx = d(1);
    r.dty = d(2);
    r.dtn = d(3);
    r.dix = d(4);
    r.diy = d(5);
    r.din = d(6);
    return r;
}

// one component of the reference's per-direction update:
// acc += dt * [ (H- - H+)/d  (+ S_axis for the four T^{tau mu} components) ]
// (the difference in R, the accumulation in double)
template <typename R>
__device__ __forceinline__ void add_component(double &acc, int n, R Hm, R Hp, const R (&Sa)[4], R inv_d, double dt) {
    R r = (Hm - Hp) * inv_d;
    if (n < 4) r += Sa[n];
    acc += (double)r * dt;
}

// the flux part of one component: acc += dt * (H- - H+)/d, the axis sources
// being added separately.  In double one fused multiply-add with dt_d = dt/d
// (precomputed with the stage parameters); in float the difference is formed
// in float and accumulated in double.
__device__ __forceinline__ void add_flux(double &acc, double Hm, double Hp, double /*inv_d*/, double dt_d,
                                         double /*dt*/) {
    acc = __builtin_fma(Hm - Hp, dt_d, acc);
}
__device__ __forceinline__ void add_flux(double &acc, float Hm, float Hp, float inv_d, double /*dt_d*/, double dt) {
    acc += (double)((Hm - Hp) * inv_d) * dt;
}

// acc^{tau mu} += dt * S_axis of every direction, after the source terms.  The
// loads of all directions (fourteen each) are issued together, one wait for
// all; the velocities u^i come from the source terms' own loads.
template <bool D3, typename R>
__device__ __forceinline__ void add_axis_sources(const StageParams &P, const Views &V, uint32_t s, const SrcIn &in,
                                                 double iut, double (&acc)[NQ]) {
    const uint32_t f8 = V.f8;
    auto fields = [](int dir, int i) {
        constexpr int T[4] = {PITT, PITX, PITY, PITN};
        constexpr int X[3] = {PIXX, PIXY, PIXN}, Y[3] = {PIXY, PIYY, PIYN}, Z[3] = {PIXN, PIYN, PINN};
        return i < 4 ? T[i] : dir == 1 ? X[i - 4] : dir == 2 ? Y[i - 4] : Z[i - 4];
    };
    constexpr int ND = D3 ? 3 : 2;
    const uint32_t stp[3] = {8u, V.sy8, V.sz8};
    double up[ND][7], dn[ND][7];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
#pragma unroll
        for (int i = 0; i < 7; ++i) {
            up[d][i] = ld(V.qin, s + stp[d], fields(d + 1, i) * f8);
            dn[d][i] = ld(V.qin, s - stp[d], fields(d + 1, i) * f8);
        }
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the loads together (not sunk to their uses)
    const double h[3] = {P.hx, P.hy, P.hz};
    const double v[3] = {in.ux * iut, in.uy * iut, in.un * iut};
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        auto dd = [&](int i) { return (R)((up[d][i] - dn[d][i]) * h[d]); };
        AxisPi<R> a;
        a.dtt = dd(0);
        a.dtx = dd(1);
        a.dty = dd(2);
        a.dtn = dd(3);
        a.dix = dd(4);
        a.diy = dd(5);
        a.din = dd(6);
        R Sa[4];
        if (d == 0) axis_source<1>(a, (R)v[0], Sa);
        else if (d == 1) axis_source<2>(a, (R)v[1], Sa);
        else axis_source<3>(a, (R)v[2], Sa);
#pragma unroll
        for (int n = 0; n < 4; ++n) acc[n] += (double)Sa[n] * P.dt;
    }
}

template <bool D3>
__device__ __forceinline__ void load_src_in(const StageParams &P, const Views &V, uint32_t s, SrcIn &in) {
    const Buf &u = V.ucur;
    const uint32_t f8 = V.f8, sy8 = V.sy8, sz8 = V.sz8;
    in.t = P.t_src;
    in.dt_inv = P.dt_inv;
    in.etabar_inv = P.etabar_inv;
    in.e = ld(V.e_in, s, 0);
    in.p = ld(V.p_in, s, 0);
    in.ut = ld(u, s, 0);
    in.ux = ld(u, s, f8);
    in.uy = ld(u, s, 2 * f8);
    in.un = ld(u, s, 3 * f8);
    in.utp = ld(V.uprev, s, 0);
    in.uxp = ld(V.uprev, s, f8);
    in.uyp = ld(V.uprev, s, 2 * f8);
    in.unp = ld(V.uprev, s, 3 * f8);
    auto cd = [&](const Buf &b, uint32_t fo, uint32_t stp, double h) { return (ld(b, s + stp, fo) - ld(b, s - stp, fo)) * h; };
    in.dxut = cd(u, 0, 8, P.hx);
    in.dxux = cd(u, f8, 8, P.hx);
    in.dxuy = cd(u, 2 * f8, 8, P.hx);
    in.dxun = cd(u, 3 * f8, 8, P.hx);
    in.dyut = cd(u, 0, sy8, P.hy);
    in.dyux = cd(u, f8, sy8, P.hy);
    in.dyuy = cd(u, 2 * f8, sy8, P.hy);
    in.dyun = cd(u, 3 * f8, sy8, P.hy);
    const uint32_t plo = PL * f8;
    in.dxp = cd(V.p_in, 0, 8, P.hx);
    in.dyp = cd(V.p_in, 0, sy8, P.hy);
    in.dxe = cd(V.e_in, 0, 8, P.hx);
    in.dye = cd(V.e_in, 0, sy8, P.hy);
    in.dxpl = cd(V.qin, plo, 8, P.hx);
    in.dypl = cd(V.qin, plo, sy8, P.hy);
    if (D3) {
        in.dnut = cd(u, 0, sz8, P.hz);
        in.dnux = cd(u, f8, sz8, P.hz);
        in.dnuy = cd(u, 2 * f8, sz8, P.hz);
        in.dnun = cd(u, 3 * f8, sz8, P.hz);
        in.dnp = cd(V.p_in, 0, sz8, P.hz);
        in.dne = cd(V.e_in, 0, sz8, P.hz);
        in.dnpl = cd(V.qin, plo, sz8, P.hz);
    } else {
        // one eta cell: the ghost planes equal the cell, the differences are exactly 0
        in.dnut = in.dnux = in.dnuy = in.dnun = 0.0;
        in.dnp = in.dne = in.dnpl = 0.0;
    }
}

// Inputs of a cell's recovery besides its conserved variables: the e its
// fallback scales and the |u_T| its Newton solve starts from.
struct Recovery {
    double e_prev, uT_guess;
};

// Stage-2 convex combination, recovery, regulation, store with ghost copies.
template <bool D3, bool STAGE2>
__device__ __forceinline__ double finish_cell(const StageParams &P, const Views &V, uint32_t s, int ii, int jj, int kk,
                                              double (&acc)[NQ], const Recovery &rec, bool &nan, bool &fallback) {
    const uint32_t f8 = V.f8;
    if (STAGE2) {
        // all fifteen loads first, so that they are in flight together
        double qn[NQ];
#pragma unroll
        for (int n = 0; n < NQ; ++n) qn[n] = ld(V.qn, s, n * f8);
#pragma unroll
        for (int n = 0; n < NQ; ++n) acc[n] = 0.5 * (acc[n] + qn[n]);
    }
    Prim r = recover(P.eos, P.t_rec, acc, rec.e_prev, rec.uT_guess);
    nan = nan || prim_nan(r);
    fallback = !r.ok;
    double e = r.e, p = r.p;
    regulate(P.t_rec, acc, e, p, r.ut, r.ux, r.uy, r.un);
    double vals[NQ + 6];
#pragma unroll
    for (int n = 0; n < NQ; ++n) vals[n] = acc[n];
    vals[NQ] = e;
    vals[NQ + 1] = p;
    vals[NQ + 2] = r.ut;
    vals[NQ + 3] = r.ux;
    vals[NQ + 4] = r.uy;
    vals[NQ + 5] = r.un;
    auto store = [&](uint32_t d) {
#pragma unroll
        for (int n = 0; n < NQ; ++n) st(V.qout, d, n * f8, vals[n]);
        st(V.e_out, d, 0, vals[NQ]);
        st(V.p_out, d, 0, vals[NQ + 1]);
#pragma unroll
        for (int m = 0; m < 4; ++m) st(V.uout, d, m * f8, vals[NQ + 2 + m]);
    };
    store(s);
    // face ghosts: copies of the boundary cell (only one index outside)
    if (ii == 0) { store(s - 8); store(s - 16); }
    if (ii == P.nx - 1) { store(s + 8); store(s + 16); }
    if (jj == 0) { store(s - V.sy8); store(s - 2 * V.sy8); }
    if (jj == P.ny - 1) { store(s + V.sy8); store(s + 2 * V.sy8); }
    if (D3) {
        if (kk == 0) { store(s - V.sz8); store(s - 2 * V.sz8); }
        if (kk == P.nz - 1) { store(s + V.sz8); store(s + 2 * V.sz8); }
    }
    return e;
}

// NaN checks of the reference (EnergyMomentumTensor.cpp:178-194 and
// SourceTerms.cpp:658-719 exit there) become a counter the host reads.
// The wave's largest new e goes to P.emax (as ordered bits: e > 0) when set,
// which the batch uses for its freezeout test.
__device__ __forceinline__ void count_flags(const StageParams &P, bool nan, bool fallback, double e) {
    unsigned long long bn = __ballot(nan), bf = __ballot(fallback);
    if ((threadIdx.x & 63) == 0) {
        if (bn) atomicAdd(&P.counters[0], (unsigned)__popcll(bn));
        if (bf) atomicAdd(&P.counters[1], (unsigned)__popcll(bf));
    }
    if (P.emax) {
        double m = e > 0.0 ? e : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
        if ((threadIdx.x & 63) == 0 && m > 0.0) atomicMax(P.emax, (unsigned long long)__double_as_longlong(m));
    }
}

// One tile of one stage.  Hs holds the x / y face fluxes of the tile
// (SHARED) or this lane's forward face (reference split); Hz is the eta face
// carried from the plane below, one slot per lane that only its own lane
// touches.  R is the arithmetic type of the stencil, fluxes and sources; the
// tile is TX x TY cells (P.tiles_x / P.tiles_y must be counted with them).
template <bool D3, bool STAGE2, bool SHARED, typename R, int TX = kTileX, int TY = kTileY>
__device__ __forceinline__ void stage_tile(const StageParams &P, int tile, R (*Hs)[NTHREADS], R (*Hz)[NTHREADS],
                                           R *zring) {
    // split precision in 3+1D: the eta-face stencil of this lane's column is
    // kept in LDS for three planes (lane-private slots), so each plane's 15
    // fields are read from memory once by the eta faces instead of four times
    constexpr bool ZRING = D3 && SHARED && sizeof(R) == 4;
    auto zr = [&](int plane, int n) -> R & {  // ring slot of a plane (planes >= -1) and field
        return zring[(((plane + 3) % 3) * NQ + n) * NTHREADS + (int)threadIdx.x];
    };
    static_assert(tile_fits<TX, TY>(), "tile does not fit one workgroup");
    const int tpl = P.tiles_x * P.tiles_y;
    const int zc = tile / tpl;
    const int txy = tile - zc * tpl;
    const int ty = txy / P.tiles_x;
    const int tx = txy - ty * P.tiles_x;
    const int i0 = tx * TX, j0 = ty * TY;
    const int k0 = zc * P.kz;
    const int k1 = min(k0 + P.kz, P.nz);

    const Views V = make_views(P);
    const uint32_t f8 = V.f8, sy8 = V.sy8, sz8 = V.sz8;
    const int tid = (int)threadIdx.x;
    const uint32_t sy = (uint32_t)P.sy;
    const double t = P.t_src, dt = P.dt;
    const R inv_dx = (R)P.inv_dx, inv_dy = (R)P.inv_dy, inv_dz = (R)P.inv_dz;

    // this lane's cell
    const int ci = tid % TX, cj = tid / TX;
    const int ii = i0 + ci, jj = j0 + cj;
    const bool cell_ok = cj < TY && ii < P.nx && jj < P.ny;

    for (int kk = k0; kk < k1; ++kk) {
        const uint32_t plane = (uint32_t)plane_of(kk, D3) * (uint32_t)P.sz;
        const uint32_t s = (plane + (uint32_t)(jj + 2) * sy + (uint32_t)(ii + 2)) * 8u;
        bool nan = false, fallback = false;
        double acc[NQ];
        Recovery rec{0.0, 0.0};
        double iut = 0.0;

        // ---- sources: acc = q + dt S (eulerStepKernelSource)
        if (cell_ok) {
            double Q[NQ];
#pragma unroll
            for (int n = 0; n < NQ; ++n) Q[n] = ld(V.qin, s, n * f8);
            SrcIn in;
            load_src_in<D3>(P, V, s, in);
            R S[NQ];
            source_terms<R>(P.eos, in, Q, S, nan);
#pragma unroll
            for (int n = 0; n < NQ; ++n) acc[n] = Q[n] + dt * (double)S[n];
            iut = rcp(in.ut);
            if (SHARED) add_axis_sources<D3, R>(P, V, s, in, iut, acc);
            // the cell recovery's inputs come from the same loads: e at the
            // stage's input level (its fallback base) and |u_T| of the velocity
            // that seeds it, u_cur in stage 1 and u_prev in stage 2
            // (FullyDiscreteKurganovTadmorScheme.cpp:761-780)
            rec.e_prev = in.e;
            rec.uT_guess = STAGE2 ? sqrt_pos(fabs(in.uxp * in.uxp + in.uyp * in.uyp))
                                  : sqrt_pos(fabs(in.ux * in.ux + in.uy * in.uy));
        }

        if (SHARED) {
            // ---- x faces: (TX+1) x TY, face f between cells (fi-1, fj) and (fi, fj)
            {
                const int fi = tid % (TX + 1), fj = tid / (TX + 1);
                const int gi = i0 + fi, gj = j0 + fj;  // interior index of the right cell
                if (fj < TY && gi <= P.nx && gj < P.ny) {
                    const uint32_t c = (plane + (uint32_t)(gj + 2) * sy + (uint32_t)(gi + 2)) * 8u;
                    face_flux<1, R, true>(t, [&] { return uT_of(V.ucur, f8, c - 8); }, V.qin, f8, c, 8,
                                    [&](int n, R h) { Hs[n][tid] = h; });
                }
            }
            __syncthreads();
            if (cell_ok) {
                const int fm = cj * (TX + 1) + ci;
#pragma unroll
                for (int n = 0; n < NQ; ++n) add_flux(acc[n], Hs[n][fm], Hs[n][fm + 1], inv_dx, P.dt_dx, dt);
            }
            __syncthreads();
            // ---- y faces: TX x (TY+1), face f between cells (fi, fj-1) and (fi, fj)
            {
                const int fi = tid % TX, fj = tid / TX;
                const int gi = i0 + fi, gj = j0 + fj;
                if (fj <= TY && gi < P.nx && gj <= P.ny) {
                    const uint32_t c = (plane + (uint32_t)(gj + 2) * sy + (uint32_t)(gi + 2)) * 8u;
                    face_flux<2, R, true>(t, [&] { return uT_of(V.ucur, f8, c - sy8); }, V.qin, f8, c, sy8,
                                    [&](int n, R h) { Hs[n][tid] = h; });
                }
            }
            __syncthreads();
            if (cell_ok) {
                const int fm = cj * TX + ci;
#pragma unroll
                for (int n = 0; n < NQ; ++n) add_flux(acc[n], Hs[n][fm], Hs[n][fm + TX], inv_dy, P.dt_dy, dt);
            }
            __syncthreads();
            // ---- eta faces: solved by the cell's own lane, the upper one carried to the next plane
            if (D3 && cell_ok) {
                if (kk == k0) {
                    if constexpr (ZRING) {
                        // lower face: planes k0-1 .. k0+1 into the ring first (fifteen
                        // loads in flight at a time), plane k0-2 from memory
#pragma unroll
                        for (int j = 1; j < 4; ++j) {
                            double v[NQ];
#pragma unroll
                            for (int n = 0; n < NQ; ++n) v[n] = ld(V.qin, s + (uint32_t)j * sz8 - 2u * sz8, n * f8);
#pragma unroll
                            for (int n = 0; n < NQ; ++n) zr(kk + j - 2, n) = (R)v[n];
                        }
                        face_flux_from<3, R, true>(t, [&] { return uT_of(V.ucur, f8, s - sz8); },
                            [&](int n, int j) {
                                return j == 0 ? ld(V.qin, s - 2u * sz8, n * f8) : (double)zr(kk + j - 2, n);
                            },
                            [&](int n, R h) { Hz[n][tid] = h; });
                    } else {
                        face_flux<3, R, true>(t, [&] { return uT_of(V.ucur, f8, s - sz8); }, V.qin, f8, s, sz8,
                                        [&](int n, R h) { Hz[n][tid] = h; });
                    }
                }
                auto upper = [&](int n, R h) {
                    add_flux(acc[n], Hz[n][tid], h, inv_dz, P.dt_dz, dt);
                    Hz[n][tid] = h;
                };
                if constexpr (ZRING) {
                    // upper face: planes kk-1 .. kk+1 from the ring, kk+2 from memory (all 15
                    // loads issued together); plane kk+2 then takes the slot kk-1 leaves
                    double fresh[NQ];
#pragma unroll
                    for (int n = 0; n < NQ; ++n) fresh[n] = ld(V.qin, s + 2u * sz8, n * f8);
                    face_flux_from<3, R, true>(t, [&] { return uT_of(V.ucur, f8, s); },
                        [&](int n, int j) { return j < 3 ? (double)zr(kk - 1 + j, n) : fresh[n]; }, upper);
#pragma unroll
                    for (int n = 0; n < NQ; ++n) zr(kk + 2, n) = (R)fresh[n];
                } else {
                    face_flux<3, R, true>(t, [&] { return uT_of(V.ucur, f8, s); }, V.qin, f8, s + sz8, sz8, upper);
                }
            }
        } else if (cell_ok) {
            // reference work split: both faces of each direction, Newton guess from this cell
            const double uTs = uT_of(V.ucur, f8, s);
            {
                R Sa[4];
                axis_source<1>(axis_pi<1, R>(V.qin, f8, s, 8, P.hx), (R)(ld(V.ucur, s, f8) * iut), Sa);
                face_flux<1, R, false>(t, [&] { return uTs; }, V.qin, f8, s + 8, 8, [&](int n, R h) { Hs[n][tid] = h; });
                face_flux<1, R, false>(t, [&] { return uTs; }, V.qin, f8, s, 8,
                                [&](int n, R h) { add_component(acc[n], n, h, Hs[n][tid], Sa, inv_dx, dt); });
            }
            {
                R Sa[4];
                axis_source<2>(axis_pi<2, R>(V.qin, f8, s, sy8, P.hy), (R)(ld(V.ucur, s, 2 * f8) * iut), Sa);
                face_flux<2, R, false>(t, [&] { return uTs; }, V.qin, f8, s + sy8, sy8, [&](int n, R h) { Hs[n][tid] = h; });
                face_flux<2, R, false>(t, [&] { return uTs; }, V.qin, f8, s, sy8,
                                [&](int n, R h) { add_component(acc[n], n, h, Hs[n][tid], Sa, inv_dy, dt); });
            }
            if (D3) {
                R Sa[4];
                axis_source<3>(axis_pi<3, R>(V.qin, f8, s, sz8, P.hz), (R)(ld(V.ucur, s, 3 * f8) * iut), Sa);
                face_flux<3, R, false>(t, [&] { return uTs; }, V.qin, f8, s + sz8, sz8, [&](int n, R h) { Hs[n][tid] = h; });
                face_flux<3, R, false>(t, [&] { return uTs; }, V.qin, f8, s, sz8,
                                [&](int n, R h) { add_component(acc[n], n, h, Hs[n][tid], Sa, inv_dz, dt); });
            }
        }

        double e_new = 0.0;
        if (cell_ok) e_new = finish_cell<D3, STAGE2>(P, V, s, ii, jj, kk, acc, rec, nan, fallback);
        count_flags(P, nan, fallback, e_new);
    }
}

// The LDS of a stage workgroup, sized by the template arguments.
template <bool D3, bool SHARED, typename R>
struct StageLds {
    static constexpr bool kRing = D3 && SHARED && sizeof(R) == 4;
    R Hs[NQ][NTHREADS];
    R Hz[D3 && SHARED ? NQ : 1][NTHREADS];
    R Zr[kRing ? 3 : 1][kRing ? NQ : 1][NTHREADS];
};

// One event: workgroup -> tile (XCD-aware) -> stage_tile.
template <bool D3, bool STAGE2, bool SHARED, typename R>
__global__ __launch_bounds__(NTHREADS, D3 ? VAH_BLOCKS_3D : VAH_BLOCKS_2D) void stage_kernel(StageParams P_arg) {
    // The parameters are read where they are used, through the kernel-argument
    // segment (P_arg is its first and only argument, at offset 0), instead of
    // from the argument copy the compiler otherwise keeps in scalar registers
    // across the plane loop: some ninety uniform values there made it spill
    // about three times as many scalar registers and twice as many vector
    // registers as it does reading them in place.
    typedef __attribute__((address_space(4))) const StageParams KernargParams;
    const StageParams &P = *(const StageParams *)(KernargParams *)__builtin_amdgcn_kernarg_segment_ptr();
    (void)P_arg;
    __shared__ StageLds<D3, SHARED, R> lds;
    const int tile = tile_of_block(P, (int)blockIdx.x);
    if (tile < 0) return;  // padding workgroups: no barrier reached
    stage_tile<D3, STAGE2, SHARED, R>(P, tile, lds.Hs, lds.Hz, &lds.Zr[0][0][0]);
}

// Zero-gradient ghost copies of an initial state: each boundary interior cell
// writes its face ghosts (every read of the stencils touches only those).
static __global__ void ghost_fill_kernel(double *q, double *e, double *p, double *u, long fs, int nx, int ny, int nz,
                                  long sy, long sz, int d3) {
    long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)nx * ny * nz;
    if (m >= n) return;
    int ii = (int)(m % nx);
    int jj = (int)((m / nx) % ny);
    int kk = (int)(m / ((long)nx * ny));
    long s = (long)plane_of(kk, d3 != 0) * sz + (long)(jj + 2) * sy + (ii + 2);
    auto copy = [&](long d) {
        for (int f = 0; f < NQ; ++f) q[f * fs + d] = q[f * fs + s];
        e[d] = e[s];
        p[d] = p[s];
        for (int f = 0; f < 4; ++f) u[f * fs + d] = u[f * fs + s];
    };
    if (ii == 0) { copy(s - 1); copy(s - 2); }
    if (ii == nx - 1) { copy(s + 1); copy(s + 2); }
    if (jj == 0) { copy(s - sy); copy(s - 2 * sy); }
    if (jj == ny - 1) { copy(s + sy); copy(s + 2 * sy); }
    if (d3) {
        if (kk == 0) { copy(s - sz); copy(s - 2 * sz); }
        if (kk == nz - 1) { copy(s + sz); copy(s + 2 * sz); }
    }
}

// The launches of the 3+1D stage kernels (vah_stage3d.hip): a translation
// unit of their own, compiled with the scheduler settings that suit them.
// precision and flux_mode take the values of include/vah.h.
struct StageLaunch {
    int grid;
    hipStream_t stream;
    int precision, flux_mode;
};
void launch_stage_3d(const StageLaunch &L, int stage, const StageParams &P);

}  // namespace vah
