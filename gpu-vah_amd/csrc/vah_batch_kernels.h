// vah_batch_kernels.h -- persistent stage kernels over many resident events.
//
// A batch holds events of different lattice sizes, spacings and time steps,
// each in its own region of one device arena.  One launch per Heun stage and
// dimensionality (2+1D / 3+1D lists) covers the tiles of every live event:
// the workgroups of an XCD take a contiguous run of the tile list (so the
// tiles of one event share that XCD's L2) and stride through it.  Each tile
// is the same fused stage as the single-lattice kernel (stage_tile), with
// the event's buffer roles picked from its own step counter, so a step needs
// no host work and a whole sequence of steps can be captured in a hipGraph.
// After stage 2 a one-lane-per-event kernel advances the step counters and
// stops events whose hottest cell fell below the freeze-out temperature.
#pragma once

#include "vah_stage.h"

namespace vah {

enum BatchEventState { EV_LIVE = 0, EV_FROZEN = 1, EV_STEP_LIMIT = 2, EV_NAN = 3 };

// Static description of one resident event, written by the host at
// admission.  base carries geometry, constants and the EOS; its buffer
// pointers and times are set per stage from the step counter.
struct BatchEvent {
    StageParams base;
    double *q[3];   // conserved slabs: q^n at [0] (even n) or [2] (odd n), q* at [1]
    double *u[3];   // velocity: u^n at [n & 1], u^{n-1} at [1 - (n & 1)], u* at [2]
    double *e[2], *p[2];  // [0] current level, [1] stage-1 output
    double t0, dt;
    double Tf;            // freeze-out temperature [fm^-1] (<= 0: never)
    long long max_steps;  // stop after this many steps (<= 0: no limit)
};

// Per-event state the device updates.
struct BatchDyn {
    long long nstep;
    unsigned long long emax;  // max e of the last stage-2 output (ordered bits)
    unsigned int counters[2]; // [0] NaN cells, [1] recovery fallbacks (cumulative)
    int live;                 // 1 while the event evolves
    int state;                // BatchEventState
    double maxT;              // T(max e) after the last step [fm^-1]
};

__device__ __forceinline__ void batch_roles(const BatchEvent &E, bool stage2, long long n, StageParams &P) {
    const int a = (n & 1) ? 2 : 0;  // q^n
    const int b = 2 - a;            // q^{n+1}
    const int uc = (int)(n & 1), up = 1 - uc;
    if (!stage2) {
        P.qin = E.q[a];
        P.qout = E.q[1];
        P.qn = E.q[a];
        P.ucur = E.u[uc];
        P.uprev = E.u[up];
        P.uout = E.u[2];
        P.e_in = E.e[0];
        P.p_in = E.p[0];
        P.e_out = E.e[1];
        P.p_out = E.p[1];
    } else {
        P.qin = E.q[1];
        P.qout = E.q[b];
        P.qn = E.q[a];
        P.ucur = E.u[2];
        P.uprev = E.u[uc];
        P.uout = E.u[up];
        P.e_in = E.e[1];
        P.p_in = E.p[1];
        P.e_out = E.e[0];
        P.p_out = E.p[0];
    }
}

// tau_n = t0 + n dt rounded as the host computes it (no contraction)
__device__ __forceinline__ double batch_tau(const BatchEvent &E, long long n) {
    return __dadd_rn(E.t0, __dmul_rn((double)n, E.dt));
}

// The resolved parameters of one event for one stage of its current step,
// written by batch_prepare_kernel at the start of the step, so that a tile
// needs one batch of uniform loads before its work starts.
struct BatchStageRec {
    StageParams P;
    int live;
    int pad_;
};

// Start of a step: one lane per resident event resolves the buffer roles and
// times of both stages from its step counter.  active[0 .. *nactive) lists
// resident slots; rec[2 slot + stage].
static __global__ void batch_prepare_kernel(const BatchEvent *__restrict__ ev, BatchDyn *dyn,
                                            const int *__restrict__ active, const int *__restrict__ nactive,
                                            BatchStageRec *rec) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= *nactive) return;
    const int slot = active[i];
    BatchDyn *d = dyn + slot;
    const BatchEvent &E = ev[slot];
    const long long nstep = d->nstep;
    const double t = batch_tau(E, nstep);
    for (int stage2 = 0; stage2 < 2; ++stage2) {
        BatchStageRec &r = rec[2 * slot + stage2];
        r.P = E.base;
        batch_roles(E, stage2 != 0, nstep, r.P);
        r.P.t_src = stage2 ? t + E.dt : t;
        r.P.t_rec = t + E.dt;
        r.P.counters = d->counters;
        r.P.emax = stage2 ? &d->emax : nullptr;
        r.live = d->live;
    }
}

// One stage over the tile list of one dimensionality.  tiles[w] = (slot,
// tile index within the event); ntiles[0] / ntiles[1] count the 2+1D / 3+1D
// lists.  gridDim.x is a multiple of 8 (one share per XCD).
template <bool D3, bool STAGE2, typename R, int TX, int TY, int BLOCKS>
__global__ __launch_bounds__(NTHREADS, BLOCKS) void batch_stage_kernel(const BatchStageRec *__restrict__ rec,
                                                                      const int2 *__restrict__ tiles,
                                                                      const int *__restrict__ ntiles) {
    __shared__ StageLds<D3, true, R> lds;
    const int n = ntiles[D3 ? 1 : 0];
    const int xcd = (int)(blockIdx.x & 7u), loc = (int)(blockIdx.x >> 3), nloc = (int)(gridDim.x >> 3);
    const int lo = (int)(((long long)n * xcd) >> 3), hi = (int)(((long long)n * (xcd + 1)) >> 3);
    int w = lo + loc;
    if (w >= hi) return;
    int2 tr = tiles[w];
    for (;;) {
        // the next list entry is fetched while this tile runs
        const int wn = w + nloc;
        const int2 next = wn < hi ? tiles[wn] : tr;
        const int slot = __builtin_amdgcn_readfirstlane(tr.x);
        const int tile = __builtin_amdgcn_readfirstlane(tr.y);
        const BatchStageRec &r = rec[2 * slot + (STAGE2 ? 1 : 0)];
        if (r.live != 0) {
            stage_tile<D3, STAGE2, true, R, TX, TY>(r.P, tile, lds.Hs, lds.Hz, &lds.Zr[0][0][0]);
            __syncthreads();  // the next tile rewrites the face buffers
        }
        if (wn >= hi) break;
        w = wn;
        tr = next;
    }
}

// After stage 2: one lane per resident event advances its step counter,
// evaluates T at the event's largest e and stops it at freeze-out, at its
// step limit or after a NaN.  active[0 .. *nactive) lists resident slots.
static __global__ void batch_advance_kernel(const BatchEvent *__restrict__ ev, BatchDyn *dyn, const int *__restrict__ active,
                                     const int *__restrict__ nactive) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= *nactive) return;
    const int slot = active[i];
    BatchDyn *d = dyn + slot;
    if (d->live == 0) return;
    const BatchEvent &E = ev[slot];
    const long long n = d->nstep + 1;
    const double em = __longlong_as_double((long long)d->emax);
    const double T = em > 0.0 ? eos_T(E.base.eos, em) : 0.0;
    int st = EV_LIVE;
    if (d->counters[0] != 0u)
        st = EV_NAN;
    else if (T < E.Tf)
        st = EV_FROZEN;
    else if (E.max_steps > 0 && n >= E.max_steps)
        st = EV_STEP_LIMIT;
    d->nstep = n;
    d->emax = 0ull;
    d->maxT = T;
    if (st != EV_LIVE) {
        d->state = st;
        d->live = 0;
    }
}

// Host-side launch of one variant's kernels (one translation unit per
// variant instantiates them).
struct BatchLaunch {
    const BatchStageRec *rec;
    const int2 *tiles2, *tiles3;  // 2+1D and 3+1D tile lists
    const int *counts;            // [0] 2+1D tiles, [1] 3+1D tiles, [2] active events
    int grid, grid3;              // persistent workgroups for 2+1D / 3+1D tiles (multiples of 8)
    hipStream_t stream;
};

typedef void (*BatchStageFn)(const BatchLaunch &L, bool d3, bool stage2, bool fp32);

}  // namespace vah
