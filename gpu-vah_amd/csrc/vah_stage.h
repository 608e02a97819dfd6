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
// carried in registers to the next plane.  Without SHARED every cell solves
// both of its faces with its own Newton guess, the reference's work split.
// blockIdx is remapped so that consecutive tiles land on the same XCD (the
// hardware deals workgroups round-robin over the 8 XCDs), keeping tile halos
// in one L2.
#pragma once

#include "vah_device.h"
#include "vah_source.h"

namespace vah {

constexpr int TX = 16;
constexpr int TY = 15;
constexpr int NTHREADS = 256;

struct StageParams {
    const double *qin;    // 15-field slab the stage reads
    double *qout;         // 15-field slab it writes
    const double *qn;     // q at tau_n (stage 2 convex combination)
    const double *e_in, *p_in;
    double *e_out, *p_out;
    const double *ucur;   // u used for gradients and face Newton guesses
    const double *uprev;  // u one level back (time derivative)
    const double *uguess; // u whose |u_T| seeds the cell recovery
    double *uout;
    long fs;              // elements between fields of a slab
    int nx, ny, nz;       // interior cells
    int ncx, ncy;         // computational extents (interior + 4)
    long sy, sz;          // strides in y and eta
    int tiles_x, tiles_y, zchunks, kz;
    int nblocks;          // real workgroups (the grid is padded to a multiple of 8)
    double t_src, t_rec;  // tau of the sources/fluxes, tau of the recovery
    double dt, dt_inv, etabar, etabar_inv;
    double inv_dx, inv_dy, inv_dz;  // 1/d
    double hx, hy, hz;              // (1/d)/2, the reference's centred-difference factor
    Eos eos;
    unsigned int *counters;  // [0] NaN cells, [1] recovery fallbacks
};

__device__ __forceinline__ int remap_block(int b, int nb) {
    // workgroup b runs on XCD b % 8; give each XCD a contiguous run of tiles
    int per = nb >> 3;
    return (b & 7) * per + (b >> 3);
}

// Reconstruct the two sides of the face between cells c-st and c and
// solve the KT flux (SemiDiscreteKurganovTadmorScheme.cpp:13-57).
template <int DIR>
__device__ __forceinline__ void face_flux(const Eos &E, double t, double uT_guess, const double *__restrict__ q,
                                          long fs, long c, long st, double (&H)[NQ], bool &nan) {
    double qL[NQ], qR[NQ];
#pragma unroll
    for (int n = 0; n < NQ; ++n) {
        const double *a = q + n * fs;
        double q0 = a[c - 2 * st], q1 = a[c - st], q2 = a[c], q3 = a[c + st];
        qL[n] = q1 + 0.5 * slope(q0, q1, q2);
        qR[n] = q2 - 0.5 * slope(q1, q2, q3);
    }
    Prim R = recover(E, t, qR, 0.0, uT_guess);
    Prim L = recover(E, t, qL, 0.0, uT_guess);
    nan = nan || prim_nan(R) || prim_nan(L);
    double uiR = DIR == 1 ? R.ux : DIR == 2 ? R.uy : R.un;
    double uiL = DIR == 1 ? L.ux : DIR == 2 ? L.uy : L.un;
    double vR = uiR * rcp(R.ut), vL = uiL * rcp(L.ut);
    double a = fmax(fabs(vL), fabs(vR));
    // (F(qR) + F(qL) - a (qR - qL)) / 2 with F(q) = v q
    double cR = 0.5 * (vR - a), cL = 0.5 * (vL + a);
#pragma unroll
    for (int n = 0; n < NQ; ++n) H[n] = cR * qR[n] + cL * qL[n];
}

__device__ __forceinline__ double uT_of(const double *__restrict__ u, long fs, long s) {
    double ux = u[fs + s], uy = u[2 * fs + s];
    return __builtin_sqrt(fabs(ux * ux + uy * uy));
}

// d_i of the pi components entering the axis source of direction DIR
template <int DIR>
__device__ __forceinline__ AxisPi axis_pi(const double *__restrict__ q, long fs, long s, long st, double h) {
    auto d = [&](int n) { return (q[n * fs + s + st] - q[n * fs + s - st]) * h; };
    AxisPi r;
    r.dtt = d(PITT);
    r.dtx = d(PITX);
    r.dty = d(PITY);
    r.dtn = d(PITN);
    if (DIR == 1) {
        r.dix = d(PIXX);
        r.diy = d(PIXY);
        r.din = d(PIXN);
    } else if (DIR == 2) {
        r.dix = d(PIXY);
        r.diy = d(PIYY);
        r.din = d(PIYN);
    } else {
        r.dix = d(PIXN);
        r.diy = d(PIYN);
        r.din = d(PINN);
    }
    return r;
}

// acc += dt * [ (H- - H+)/d + S_axis ]   (the reference adds the axis source
// to the four T^{tau mu} components only)
template <int DIR>
__device__ __forceinline__ void add_direction(double (&acc)[NQ], const double (&Hm)[NQ], const double (&Hp)[NQ],
                                              const AxisPi &ap, double v, double inv_d, double dt) {
    double Sa[4];
    axis_source<DIR>(ap, v, Sa);
#pragma unroll
    for (int n = 0; n < NQ; ++n) {
        double r = (Hm[n] - Hp[n]) * inv_d;
        if (n < 4) r += Sa[n];
        acc[n] += r * dt;
    }
}

template <bool D3>
__device__ __forceinline__ void load_src_in(const StageParams &P, long s, SrcIn &in) {
    const double *u = P.ucur;
    const long fs = P.fs, sy = P.sy, sz = P.sz;
    in.t = P.t_src;
    in.dt_inv = P.dt_inv;
    in.etabar_inv = P.etabar_inv;
    in.e = P.e_in[s];
    in.p = P.p_in[s];
    in.ut = u[s];
    in.ux = u[fs + s];
    in.uy = u[2 * fs + s];
    in.un = u[3 * fs + s];
    in.utp = P.uprev[s];
    in.uxp = P.uprev[fs + s];
    in.uyp = P.uprev[2 * fs + s];
    in.unp = P.uprev[3 * fs + s];
    auto cd = [&](const double *f, long st, double h) { return (f[s + st] - f[s - st]) * h; };
    in.dxut = cd(u, 1, P.hx);
    in.dxux = cd(u + fs, 1, P.hx);
    in.dxuy = cd(u + 2 * fs, 1, P.hx);
    in.dxun = cd(u + 3 * fs, 1, P.hx);
    in.dyut = cd(u, sy, P.hy);
    in.dyux = cd(u + fs, sy, P.hy);
    in.dyuy = cd(u + 2 * fs, sy, P.hy);
    in.dyun = cd(u + 3 * fs, sy, P.hy);
    const double *pl = P.qin + PL * fs;
    in.dxp = cd(P.p_in, 1, P.hx);
    in.dyp = cd(P.p_in, sy, P.hy);
    in.dxe = cd(P.e_in, 1, P.hx);
    in.dye = cd(P.e_in, sy, P.hy);
    in.dxpl = cd(pl, 1, P.hx);
    in.dypl = cd(pl, sy, P.hy);
    if (D3) {
        in.dnut = cd(u, sz, P.hz);
        in.dnux = cd(u + fs, sz, P.hz);
        in.dnuy = cd(u + 2 * fs, sz, P.hz);
        in.dnun = cd(u + 3 * fs, sz, P.hz);
        in.dnp = cd(P.p_in, sz, P.hz);
        in.dne = cd(P.e_in, sz, P.hz);
        in.dnpl = cd(pl, sz, P.hz);
    } else {
        // one eta cell: the ghost planes equal the cell, the differences are exactly 0
        in.dnut = in.dnux = in.dnuy = in.dnun = 0.0;
        in.dnp = in.dne = in.dnpl = 0.0;
    }
}

// Stage-2 convex combination, recovery, regulation, store with ghost copies.
template <bool D3, bool STAGE2>
__device__ __forceinline__ void finish_cell(const StageParams &P, long s, int ii, int jj, int kk, double (&acc)[NQ],
                                            bool &nan, bool &fallback) {
    const long fs = P.fs;
    if (STAGE2) {
#pragma unroll
        for (int n = 0; n < NQ; ++n) acc[n] = 0.5 * (acc[n] + P.qn[n * fs + s]);
    }
    Prim r = recover(P.eos, P.t_rec, acc, P.e_in[s], uT_of(P.uguess, fs, s));
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
    auto store = [&](long d) {
#pragma unroll
        for (int n = 0; n < NQ; ++n) P.qout[n * fs + d] = vals[n];
        P.e_out[d] = vals[NQ];
        P.p_out[d] = vals[NQ + 1];
#pragma unroll
        for (int m = 0; m < 4; ++m) P.uout[m * fs + d] = vals[NQ + 2 + m];
    };
    store(s);
    // face ghosts: copies of the boundary cell (only one index outside)
    if (ii == 0) { store(s - 1); store(s - 2); }
    if (ii == P.nx - 1) { store(s + 1); store(s + 2); }
    if (jj == 0) { store(s - P.sy); store(s - 2 * P.sy); }
    if (jj == P.ny - 1) { store(s + P.sy); store(s + 2 * P.sy); }
    if (D3) {
        if (kk == 0) { store(s - P.sz); store(s - 2 * P.sz); }
        if (kk == P.nz - 1) { store(s + P.sz); store(s + 2 * P.sz); }
    }
}

__device__ __forceinline__ void count_flags(const StageParams &P, bool nan, bool fallback) {
    unsigned long long bn = __ballot(nan), bf = __ballot(fallback);
    if ((threadIdx.x & 63) == 0) {
        if (bn) atomicAdd(&P.counters[0], (unsigned)__popcll(bn));
        if (bf) atomicAdd(&P.counters[1], (unsigned)__popcll(bf));
    }
}

template <bool D3, bool STAGE2, bool SHARED>
__global__ __launch_bounds__(NTHREADS, 2) void stage_kernel(StageParams P) {
    __shared__ double Hs[SHARED ? NQ : 1][SHARED ? NTHREADS : 1];

    const int b = (int)blockIdx.x;
    const int tile = (gridDim.x & 7) == 0 ? remap_block(b, (int)gridDim.x) : b;
    if (tile >= P.nblocks) return;  // padding workgroups: no barrier reached
    const int tpl = P.tiles_x * P.tiles_y;
    const int zc = tile / tpl;
    const int txy = tile - zc * tpl;
    const int ty = txy / P.tiles_x;
    const int tx = txy - ty * P.tiles_x;
    const int i0 = tx * TX, j0 = ty * TY;
    const int k0 = zc * P.kz;
    const int k1 = min(k0 + P.kz, P.nz);

    const int tid = (int)threadIdx.x;
    const long fs = P.fs, sy = P.sy, sz = P.sz;
    const Eos &E = P.eos;

    // this lane's cell
    const int ci = tid % TX, cj = tid / TX;
    const int ii = i0 + ci, jj = j0 + cj;
    const bool cell_ok = cj < TY && ii < P.nx && jj < P.ny;

    double Hzm[NQ];  // eta face below the current plane (D3 && SHARED)

    for (int kk = k0; kk < k1; ++kk) {
        const long plane = (long)(kk + 2) * sz;
        const long s = plane + (long)(jj + 2) * sy + (ii + 2);
        bool nan = false, fallback = false;
        double acc[NQ];

        if (cell_ok) {
            double Q[NQ];
#pragma unroll
            for (int n = 0; n < NQ; ++n) Q[n] = P.qin[n * fs + s];
            SrcIn in;
            load_src_in<D3>(P, s, in);
            double S[NQ];
            source_terms(E, in, Q, S, nan);
#pragma unroll
            for (int n = 0; n < NQ; ++n) acc[n] = Q[n] + P.dt * S[n];
        }

        if (SHARED) {
            // ---- x faces: 17 x 15 faces, face f sits between cells (fi-1, fj) and (fi, fj)
            {
                const int fi = tid % (TX + 1), fj = tid / (TX + 1);
                const int gi = i0 + fi, gj = j0 + fj;  // interior index of the right cell
                if (fj < TY && gi <= P.nx && gj < P.ny) {
                    const long c = plane + (long)(gj + 2) * sy + (gi + 2);
                    double H[NQ];
                    face_flux<1>(E, P.t_src, uT_of(P.ucur, fs, c - 1), P.qin, fs, c, 1, H, nan);
#pragma unroll
                    for (int n = 0; n < NQ; ++n) Hs[n][tid] = H[n];
                }
            }
            __syncthreads();
            if (cell_ok) {
                double Hm[NQ], Hp[NQ];
                const int fm = cj * (TX + 1) + ci;
#pragma unroll
                for (int n = 0; n < NQ; ++n) {
                    Hm[n] = Hs[n][fm];
                    Hp[n] = Hs[n][fm + 1];
                }
                AxisPi ap = axis_pi<1>(P.qin, fs, s, 1, P.hx);
                add_direction<1>(acc, Hm, Hp, ap, P.ucur[fs + s] * rcp(P.ucur[s]), P.inv_dx, P.dt);
            }
            __syncthreads();
            // ---- y faces: 16 x 16 faces, face f sits between cells (fi, fj-1) and (fi, fj)
            {
                const int fi = tid % TX, fj = tid / TX;
                const int gi = i0 + fi, gj = j0 + fj;
                if (fj <= TY && gi < P.nx && gj <= P.ny) {
                    const long c = plane + (long)(gj + 2) * sy + (gi + 2);
                    double H[NQ];
                    face_flux<2>(E, P.t_src, uT_of(P.ucur, fs, c - sy), P.qin, fs, c, sy, H, nan);
#pragma unroll
                    for (int n = 0; n < NQ; ++n) Hs[n][tid] = H[n];
                }
            }
            __syncthreads();
            if (cell_ok) {
                double Hm[NQ], Hp[NQ];
                const int fm = cj * TX + ci;
#pragma unroll
                for (int n = 0; n < NQ; ++n) {
                    Hm[n] = Hs[n][fm];
                    Hp[n] = Hs[n][fm + TX];
                }
                AxisPi ap = axis_pi<2>(P.qin, fs, s, sy, P.hy);
                add_direction<2>(acc, Hm, Hp, ap, P.ucur[2 * fs + s] * rcp(P.ucur[s]), P.inv_dy, P.dt);
            }
            __syncthreads();
            if (D3 && cell_ok) {
                if (kk == k0) face_flux<3>(E, P.t_src, uT_of(P.ucur, fs, s - sz), P.qin, fs, s, sz, Hzm, nan);
                double Hzp[NQ];
                face_flux<3>(E, P.t_src, uT_of(P.ucur, fs, s), P.qin, fs, s + sz, sz, Hzp, nan);
                AxisPi ap = axis_pi<3>(P.qin, fs, s, sz, P.hz);
                add_direction<3>(acc, Hzm, Hzp, ap, P.ucur[3 * fs + s] * rcp(P.ucur[s]), P.inv_dz, P.dt);
#pragma unroll
                for (int n = 0; n < NQ; ++n) Hzm[n] = Hzp[n];
            }
        } else if (cell_ok) {
            // reference work split: both faces per direction, Newton guess from this cell
            const double uTs = uT_of(P.ucur, fs, s);
            {
                double Hm[NQ], Hp[NQ];
                face_flux<1>(E, P.t_src, uTs, P.qin, fs, s + 1, 1, Hp, nan);
                face_flux<1>(E, P.t_src, uTs, P.qin, fs, s, 1, Hm, nan);
                AxisPi ap = axis_pi<1>(P.qin, fs, s, 1, P.hx);
                add_direction<1>(acc, Hm, Hp, ap, P.ucur[fs + s] * rcp(P.ucur[s]), P.inv_dx, P.dt);
            }
            {
                double Hm[NQ], Hp[NQ];
                face_flux<2>(E, P.t_src, uTs, P.qin, fs, s + sy, sy, Hp, nan);
                face_flux<2>(E, P.t_src, uTs, P.qin, fs, s, sy, Hm, nan);
                AxisPi ap = axis_pi<2>(P.qin, fs, s, sy, P.hy);
                add_direction<2>(acc, Hm, Hp, ap, P.ucur[2 * fs + s] * rcp(P.ucur[s]), P.inv_dy, P.dt);
            }
            if (D3) {
                double Hm[NQ], Hp[NQ];
                face_flux<3>(E, P.t_src, uTs, P.qin, fs, s + sz, sz, Hp, nan);
                face_flux<3>(E, P.t_src, uTs, P.qin, fs, s, sz, Hm, nan);
                AxisPi ap = axis_pi<3>(P.qin, fs, s, sz, P.hz);
                add_direction<3>(acc, Hm, Hp, ap, P.ucur[3 * fs + s] * rcp(P.ucur[s]), P.inv_dz, P.dt);
            }
        }

        if (cell_ok) finish_cell<D3, STAGE2>(P, s, ii, jj, kk, acc, nan, fallback);
        count_flags(P, nan, fallback);
    }
}

// Zero-gradient ghost copies of an initial state: each boundary interior cell
// writes its face ghosts (every read of the stencils touches only those).
__global__ void ghost_fill_kernel(double *q, double *e, double *p, double *u, long fs, int nx, int ny, int nz,
                                  long sy, long sz, int d3) {
    long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)nx * ny * nz;
    if (m >= n) return;
    int ii = (int)(m % nx);
    int jj = (int)((m / nx) % ny);
    int kk = (int)(m / ((long)nx * ny));
    long s = (long)(kk + 2) * sz + (long)(jj + 2) * sy + (ii + 2);
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

}  // namespace vah
