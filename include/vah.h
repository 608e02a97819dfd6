/*
 * vah.h -- C ABI of the MI355X anisotropic viscous hydrodynamics (VAH) step.
 *
 * This is the drop-in boundary for the reference's hot path: the per-step
 * Kurganov-Tadmor Heun-RK2 update of GPU-VAH (reference paths relative to its
 * repository root).  Plain C types only; every entry point returns a status
 * code (VAH_OK == 0) instead of exiting, and vah_last_error() gives the text.
 *
 * Evolved fields (15, in this order): T^tautau, T^taux, T^tauy, T^taueta, P_L,
 * pi^tautau, pi^taux, pi^tauy, pi^taueta, pi^xx, pi^xy, pi^xeta, pi^yy,
 * pi^yeta, pi^etaeta  (old_cpu-vah_code/include/DynamicalVariables.h:47-75
 * built with PIMUNU).  Host arrays crossing this interface hold interior
 * cells only, x fastest: index = i + nx*(j + ny*k); multi-field arrays are
 * field-major (field f of cell m at [f*n + m], n = nx*ny*nz).
 */
#ifndef VAH_H_
#define VAH_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VAH_NQ 15

enum vah_status {
    VAH_OK = 0,
    VAH_ERR_ARG = -1,    /* invalid argument */
    VAH_ERR_HIP = -2,    /* HIP runtime failure */
    VAH_ERR_NOMEM = -3,  /* device or host allocation failed */
    VAH_ERR_NAN = -4,    /* a cell produced NaN (the reference exit(-1)s here) */
    VAH_ERR_STATE = -5,  /* call out of order (e.g. step before initial state) */
    VAH_ERR_DEVICE = -6  /* no usable gfx950 device / library built without HIP */
};

enum vah_field {
    VAH_TTT = 0, VAH_TTX, VAH_TTY, VAH_TTN, VAH_PL,
    VAH_PITT, VAH_PITX, VAH_PITY, VAH_PITN, VAH_PIXX, VAH_PIXY, VAH_PIXN, VAH_PIYY, VAH_PIYN, VAH_PINN
};

/* rhic/include/LatticeParameters.h:17-31 (ghost cells are implicit: 2 per side) */
typedef struct vah_lattice {
    int nx, ny, nz;            /* numLatticePointsX / Y / Rapidity */
    int nt;                    /* numProperTimePoints (the driver runs nt+1 steps) */
    double dx, dy, dz, dt;     /* latticeSpacingX / Y / Rapidity / ProperTime */
} vah_lattice;

/* rhic/include/HydroParameters.h:15-21 */
typedef struct vah_hydro {
    double t0;                 /* initialProperTimePoint [fm/c] */
    double etabar;             /* shearViscosityToEntropyDensity */
    double freezeout_T_GeV;    /* freezeoutTemperatureGeV */
    int init_pimunu_ns;        /* initializePimunuNavierStokes */
} vah_hydro;

enum vah_eos_kind {
    VAH_EOS_CONFORMAL = 0,     /* p = e/3, T = powf(e/13.8997, 1/4) as the reference */
    VAH_EOS_TABLE = 1,         /* tabulated e -> (p, T), linear in ln e */
    VAH_EOS_WB = 2             /* the reference's Wuppertal-Budapest fits p(e), T(e)
                                  (old_cpu-vah_code/src/EquationOfState.cpp:31-76,120-142) */
};

enum vah_precision {
    VAH_FP64 = 0,              /* everything in double precision */
    VAH_FP32_SPLIT = 1         /* stencil, fluxes and sources in float; Newton recovery,
                                  tau and the state in double */
};

enum vah_flux_mode {
    VAH_FLUX_SHARED = 0,       /* each cell face solved once, shared by its two cells */
    VAH_FLUX_PER_CELL = 1      /* each cell solves both of its faces (the reference's work split) */
};

/* Tabulated EOS: nodes at ln(e_i) = lne0 + i*h, i = 0..n-1. */
typedef struct vah_eos_table {
    int n;
    double lne0, h;
    const double *p;           /* pressure at the nodes [fm^-4] */
    const double *T;           /* temperature at the nodes [fm^-1] */
} vah_eos_table;

typedef struct vah_options {
    int eos;                   /* enum vah_eos_kind */
    vah_eos_table table;       /* read when eos == VAH_EOS_TABLE (copied) */
    int precision;             /* enum vah_precision */
    int flux_mode;             /* enum vah_flux_mode */
    int device;                /* HIP device ordinal */
    void *stream;              /* hipStream_t to launch on; NULL: a library-owned stream */
    int eta_chunk;             /* eta planes one workgroup marches through (3+1D); 0: 3 */
} vah_options;

typedef struct vah_stats {
    double max_e, max_T;       /* over interior cells [fm^-4], [fm^-1] */
    double total_energy;       /* tau * sum(T^tautau) dx dy deta  [fm^-3 * fm^3 ...] */
    double max_vT;             /* max transverse speed |v_T| */
    long long recovery_fallbacks; /* cell recoveries that fell back, since the last state upload */
    long long nan_cells;       /* cell updates that produced NaN, since the last state upload */
} vah_stats;

typedef struct vah_profile {
    int launches[2];           /* stage-1 / stage-2 kernel launches recorded */
    double avg_ms[2];          /* their average duration (HIP events on the launch stream) */
} vah_profile;

typedef struct vah_ctx vah_ctx;

/* Default options: conformal EOS, FP64, shared faces, library default eta
 * chunk, the context's own stream.  No counterpart: the reference fixes these
 * at compile time (CONFORMAL_EOS, PRECISION, CudaConfiguration.cu). */
void vah_default_options(vah_options *opt);

/* Allocates the device state for one lattice: 3 x 15 conserved fields, e and
 * p (double-buffered), u^mu at three time levels, all with 2 ghost cells per
 * side.  Replaces allocateDeviceMemory + initializeCUDALaunchParameters +
 * initializeCUDAConstantParameters (rhic/include/DynamicalVariables.cuh:145-181,
 * rhic/include/CudaConfiguration.cuh:76-77). */
int vah_create(const vah_lattice *lat, const vah_hydro *hydro, const vah_options *opt, vah_ctx **out);

/* Initial state from primitives at tau = t0 (setConservedVariables,
 * old_cpu-vah_code/src/DynamicalVariables.cpp:197-287, then setGhostCells).
 * pl == NULL: P_L = 0.0154483 e as the reference hardcodes; pi == NULL: pi = 0,
 * else 10 field-major arrays pitt..pinn.  p = p(e) from the EOS. */
int vah_set_primitives(vah_ctx *ctx, const double *e, const double *pl, const double *ut, const double *ux,
                       const double *uy, const double *un, const double *pi);

/* Exact state load: q (15 field-major), e, p, u and up (4 field-major each:
 * u^tau, u^x, u^y, u^eta); tau = t0 + nstep*dt.  Ghosts are filled on device.
 * No counterpart (the reference only starts from primitives): restarts from a
 * saved state and the parity tests' oracle states. */
int vah_set_state(vah_ctx *ctx, long long nstep, const double *q, const double *e, const double *p, const double *u,
                  const double *up);

/* Advance nsteps Heun-RK2 steps (twoStepRungeKutta + setCurrentConservedVariables,
 * rhic/src/FullyDiscreteKurganovTadmorScheme.cu:121-157, rhic/src/HydroPlugin.cpp:223-232).
 * Asynchronous on the context's stream. */
int vah_step(vah_ctx *ctx, int nsteps);

/* Blocks until the context's stream is idle; reports NaN cells as VAH_ERR_NAN
 * (the cudaDeviceSynchronize that ends each reference step,
 * rhic/src/FullyDiscreteKurganovTadmorScheme.cu:156; the CPU path exits on NaN
 * instead, old_cpu-vah_code/src/EnergyMomentumTensor.cpp:193). */
int vah_synchronize(vah_ctx *ctx);

/* Copies the current state to the host (copyDeviceToHostMemory).  Any pointer
 * may be NULL.  Same layouts as vah_set_state. */
int vah_get_state(vah_ctx *ctx, double *q, double *e, double *p, double *u, double *up);

/* tau = t0 + nstep * dt and nstep, the driver's clock (rhic/src/HydroPlugin.cpp:232). */
int vah_get_time(const vah_ctx *ctx, double *tau, long long *nstep);

/* Device-side reductions over interior cells (max e, max T, tau sum T^tautau,
 * max |v_T|, counters).  Replaces what run() reads from a full host copy of
 * the state: the centre-cell printout and the freeze-out test
 * (rhic/src/HydroPlugin.cpp:156-158,198-219). */
int vah_reduce(vah_ctx *ctx, vah_stats *out);

/* Validity diagnostics after a step, per interior cell: the Knudsen numbers
 * 5 (eta/s) |theta_T| / T, 5 (eta/s) |zDzu| / T, 5 (eta/s) |theta| / T, the
 * inverse Reynolds number sqrt|pi:pi| / |P_T hat| and the second-order one
 * sqrt|J:J / sigma_T:sigma_T| / (2 etabar_TC) (knudsenNumberTaupiT / TaupiL /
 * Taupi, Rpi and Rpi2 of old_cpu-vah_code/src/HydrodynamicValidity.cpp:195-387,
 * evaluated with the arguments of its call at
 * FullyDiscreteKurganovTadmorScheme.cpp:791: the time after the step,
 * t_n + dt, the conserved variables of the step's start, the new e, p, u).
 * Any output may be NULL.  Needs a step completed since the last
 * vah_set_state / vah_set_primitives (VAH_ERR_STATE otherwise). */
int vah_validity(vah_ctx *ctx, double *kn_T, double *kn_L, double *kn, double *r_pi, double *r_pi2);

/* ---- Device-resident state without a context --------------------------- */

/* Padded device layout of one field: (nx+4) x (ny+4) x (nz > 1 ? nz+4 : 1)
 * cells, x fastest, 2 ghost cells per side (a boost-invariant lattice stores
 * its one eta plane), rounded up to a multiple of 64 doubles; field f of a
 * multi-field buffer starts f * VAH_FIELD_STRIDE doubles in.  Interior cell
 * (i, j, k) sits at (k' * (ny+4) + j + 2) * (nx+4) + i + 2 with k' = k + 2 in
 * 3+1D and 0 in 2+1D (old_cpu-vah_code/src/DynamicalVariables.cpp:29-31). */
#define VAH_FIELD_STRIDE(nx, ny, nz)                                                            \
    (((long long)((nx) + 4) * ((ny) + 4) * ((nz) > 1 ? (nz) + 4 : 1) + 63) / 64 * 64)

/* Caller-owned device buffers of one lattice in that layout. */
typedef struct vah_device_state {
    double *q;        /* 15 fields: q^n in; after the call, the current level */
    double *q_next;   /* 15 fields written with q^{n+1} (roles swap every step) */
    double *e, *p;    /* level n in, the current level out (updated in place) */
    double *u;        /* 4 fields u^tau, u^x, u^y, u^eta of level n; current level out */
    double *u_prev;   /* 4 fields of level n-1 (the first call: equal to u); level n out */
} vah_device_state;

/* Bytes of the library's static device scratch array of vah_step_noalloc:
 * the intermediate stage of 21 padded fields (q*, e*, p*, u*) and 64 doubles
 * of counters must fit, i.e. VAH_FIELD_STRIDE(nx, ny, nz) <= 399,454 padded
 * cells (every boost-invariant lattice up to 628 x 628; 3+1D lattices such
 * as 101 x 101 x 30). */
#define VAH_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* nsteps Heun-RK2 steps of a state the caller keeps in device memory,
 * allocating nothing: no hipMalloc, hipFree or allocating library routine.
 * No counterpart in the reference; it is for callers that cannot allocate --
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * it.  Same arithmetic as vah_step (bit-identical fields).
 *  - lat, hydro, opt as for vah_create; opt->stream (NULL: the null stream)
 *    receives all work and the call returns without waiting.  With
 *    opt->eos == VAH_EOS_TABLE, opt->table.p and opt->table.T must be device
 *    pointers (read during the steps).
 *  - nstep: the step count of the input state (tau_n = t0 + nstep dt).
 *  - st: the state; the ghost cells of q, e, p, u are filled first (zero
 *    gradient, written in place), interior cells are only read.  On return
 *    the pointers are rotated as the reference swaps q <-> Q and u <-> up:
 *    st->q / st->u name the new level, st->q_next / st->u_prev the old one.
 *  - counters: NULL, or 2+ device unsigned ints to which the NaN and
 *    recovery-fallback counts are added (the caller clears them).  With
 *    NULL they go to the scratch array, cleared by a kernel at each call.
 * The intermediate stage lives in the library's static array of
 * VAH_NOALLOC_SCRATCH_BYTES per GPU (kept until the process ends; its device
 * address is looked up on the first call on a device, which also uploads the
 * conformal EOS's 4 MB powf bitmap to a second static array and waits for
 * it -- make that first call outside graph capture).  Every call on a GPU
 * shares that GPU's arrays: calls that overlap on one GPU (other streams,
 * other threads) race.  A lattice over the limit returns VAH_ERR_ARG. */
int vah_step_noalloc(const vah_lattice *lat, const vah_hydro *hydro, const vah_options *opt, long long nstep,
                     int nsteps, vah_device_state *st, unsigned int *counters);

/* Per-launch HIP-event timing of the two stage kernels (off by default); the
 * reference times each step with a host stopwatch (rhic/src/HydroPlugin.cpp:222-227,234). */
int vah_profile_enable(vah_ctx *ctx, int on);
int vah_profile_read(vah_ctx *ctx, vah_profile *out);

/* Frees the context: freeHostMemory + freeDeviceMemory (rhic/src/HydroPlugin.cpp:239-240,
 * rhic/include/DynamicalVariables.cuh:145-181). */
void vah_destroy(vah_ctx *ctx);

/* Text of the last error on this thread.  No counterpart: the reference prints
 * and exits where this library returns a status. */
const char *vah_last_error(void);

/* Build identification: "gfx950 <precision modes> <compiler>".  No counterpart. */
const char *vah_build_info(void);

#ifdef __cplusplus
}
#endif

#endif /* VAH_H_ */
