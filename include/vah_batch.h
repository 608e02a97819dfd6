/*
 * vah_batch.h -- C ABI of the batched / serving form of the VAH step.
 *
 * The reference evolves one event per process: run() in
 * rhic/src/HydroPlugin.cpp:130-242 allocates the lattice
 * (DynamicalVariables.cuh:145-181), loops twoStepRungeKutta
 * (FullyDiscreteKurganovTadmorScheme.cuh:18-20) and frees it.  This
 * interface has no single counterpart there: it keeps many events of
 * different lattice sizes resident in one device arena and advances all of
 * them with one persistent launch per Heun stage, so that small lattices,
 * which leave most of the GPU idle one at a time, are evolved together.
 * Each event's arithmetic is the one of vah_step (include/vah.h); what an
 * event computes does not depend on which other events share the batch.
 *
 * Life cycle:  vah_batch_create -> vah_batch_admit (any time, returns
 * VAH_ERR_BUDGET when the arena has no room) -> vah_batch_step (enqueues;
 * asynchronous) -> vah_batch_poll (waits, reports every resident event) ->
 * vah_batch_get_state / vah_batch_retire -> vah_batch_destroy.
 *
 * An event stops evolving at the end of the first step after which the
 * temperature of its hottest interior cell is below the freeze-out
 * temperature (or at its step limit, or when a cell produced NaN); its
 * fields then stay as they were at that step until it is retired.
 * Host arrays use the conventions of vah.h (interior cells, x fastest,
 * field-major).
 */
#ifndef VAH_BATCH_H_
#define VAH_BATCH_H_

#include "vah.h"
#include "vah_host.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VAH_ERR_BUDGET = -7,   /* the arena has no free region large enough (try after a retire) */
    VAH_ERR_FULL = -8      /* every event slot is taken */
};

enum vah_event_state {
    VAH_EVENT_LIVE = 0,
    VAH_EVENT_FROZEN = 1,      /* max T fell below the freeze-out temperature */
    VAH_EVENT_STEP_LIMIT = 2,  /* reached max_steps */
    VAH_EVENT_NAN = 3          /* a cell produced NaN; the event was stopped */
};

typedef struct vah_batch_options {
    int eos;                   /* enum vah_eos_kind, shared by all events */
    vah_eos_table table;       /* read when eos == VAH_EOS_TABLE (copied) */
    int precision;             /* VAH_FP64 or VAH_FP32_SPLIT */
    size_t budget_bytes;       /* device arena for event fields (HBM admission budget) */
    int max_events;            /* resident event slots */
    double freezeout_T_GeV;    /* stop an event when T(max e) falls below this; <= 0: never */
    int steps_per_graph;       /* steps captured per hipGraph launch; 0: plain launches */
    int variant;               /* kernel variant (vah_batch_variant_info); -1: default */
    int device;                /* HIP device ordinal */
    void *stream;              /* hipStream_t; NULL: a library-owned stream */
} vah_batch_options;

typedef struct vah_batch_event {
    int nx, ny, nz;            /* interior cells (nz == 1: boost-invariant 2+1D) */
    double dx, dy, dz, dt;     /* spacings [fm], [fm], [1], time step [fm/c] */
    double t0;                 /* initial proper time [fm/c] */
    double etabar;             /* eta/s */
    long long max_steps;       /* stop after this many steps; <= 0: no limit */
    long long tag;             /* caller's identifier, echoed by vah_batch_poll */
} vah_batch_event;

typedef struct vah_batch_status {
    int slot;
    int state;                 /* enum vah_event_state */
    long long tag;
    long long nstep;           /* steps taken */
    double tau;                /* t0 + nstep dt */
    double max_T;              /* T(max e) after the last step [fm^-1] */
    long long recovery_fallbacks;  /* cumulative Newton fallbacks */
    long long nan_cells;
} vah_batch_status;

typedef struct vah_batch_variant {
    int tile_x, tile_y;        /* cells per workgroup tile */
    int blocks_per_cu;         /* workgroups per CU the registers are sized for */
    int kz;                    /* eta planes per 3+1D tile */
} vah_batch_variant;

typedef struct vah_batch vah_batch;

void vah_batch_default_options(vah_batch_options *opt);

/* Device bytes one event of nx x ny x nz interior cells takes in the arena. */
size_t vah_batch_event_bytes(int nx, int ny, int nz);

/* Allocates the arena (budget_bytes, once) and the event tables. */
int vah_batch_create(const vah_batch_options *opt, vah_batch **out);

/* Admits one event from primitives (setConservedVariables semantics as in
 * vah_set_primitives: pl NULL -> 0.0154483 e, pi NULL -> 0; here also
 * ut = ux = uy = un = NULL -> the rest frame).  The inputs are copied before
 * the call returns; the upload and the initial-state map run on the device,
 * ordered after steps already enqueued.  *slot receives the event's slot. */
int vah_batch_admit(vah_batch *b, const vah_batch_event *ev, const double *e, const double *pl, const double *ut,
                    const double *ux, const double *uy, const double *un, const double *pi, int *slot);

/* Admits one event whose initial energy density is a Monte-Carlo Glauber
 * profile generated for it (rhic/src/MonteCarloGlauberModel.cpp:17-108, the
 * reference's IC type 4, with g's geometry): the 2A nucleon positions are
 * drawn on the host from g->seed (vah_ic_mc_glauber's generator and draw
 * order; microseconds), and the wounded-nucleon test, the Gaussian deposits of
 * the wounded nucleons (cells x nucleons), the scaling to g->e_peak, the eta
 * profile and the floor run on the device, ordered after steps already
 * enqueued.  The energy density equals vah_ic_mc_glauber's for the same
 * lattice and g up to the rounding of exp.  P_L = pl_ratio e (pl_ratio <= 0:
 * the reference's 0.0154483 e), u at rest, pi = 0.  No counterpart in the
 * reference, whose driver builds the initial state on the host once per
 * process; here a request carries only its seed. */
int vah_batch_admit_glauber(vah_batch *b, const vah_batch_event *ev, const vah_glauber *g, double pl_ratio,
                            int *slot);

/* Enqueues nsteps Heun steps of every live event; returns without waiting. */
int vah_batch_step(vah_batch *b, int nsteps);

/* Waits for the enqueued steps. */
int vah_batch_synchronize(vah_batch *b);

/* Waits, then writes the status of up to cap resident events to out (slot
 * order) and their number to *n. */
int vah_batch_poll(vah_batch *b, vah_batch_status *out, int cap, int *n);

/* Fields of a resident event at its current step (any pointer may be NULL):
 * q [15*n], e [n], p [n], u [4*n], up [4*n] as vah_get_state. */
int vah_batch_get_state(vah_batch *b, int slot, double *q, double *e, double *p, double *u, double *up);

/* Frees the event's slot and arena region.  Waits for steps in flight. */
int vah_batch_retire(vah_batch *b, int slot);

/* Arena bytes in use and the budget. */
int vah_batch_memory(const vah_batch *b, size_t *used, size_t *budget);

/* Kernel variants: tile shape, occupancy target and eta-chunk.  Every
 * variant computes the same per-cell arithmetic. */
int vah_batch_variant_count(void);
int vah_batch_variant_info(int index, vah_batch_variant *out);
int vah_batch_set_variant(vah_batch *b, int index);
int vah_batch_get_variant(const vah_batch *b);

/* Times steps_per_trial steps of the resident events with every variant
 * (the steps are real: events advance by count * steps_per_trial), selects
 * the fastest and returns it in *best; ms_per_step[count] (may be NULL)
 * receives each variant's time per step. */
int vah_batch_autotune(vah_batch *b, int steps_per_trial, int *best, double *ms_per_step);

void vah_batch_destroy(vah_batch *b);

#ifdef __cplusplus
}
#endif

#endif /* VAH_BATCH_H_ */
