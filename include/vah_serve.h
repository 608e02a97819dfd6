/*
 * vah_serve.h -- the serving form: a continuous-batching request loop over
 * the batch engine (include/vah_batch.h), in C++ behind a C ABI.
 *
 * No counterpart in the reference, which evolves one event per process
 * (rhic/src/HydroPlugin.cpp:130-242).  Requests -- events of different
 * lattice sizes -- arrive over time.  The loop admits waiting requests first
 * come, first served whenever the batch has a free slot and room in its HBM
 * budget (a request whose initial state is a Monte-Carlo Glauber seed is
 * generated on the device at admission), advances every resident event
 * together, and retires each one when the engine stops it (freeze-out, step
 * limit or NaN), recording its latency from arrival to retirement.  Requests
 * that arrive while others run join them at the next poll.  A request that
 * cannot fit even into an empty batch is rejected rather than waited on.
 */
#ifndef VAH_SERVE_H_
#define VAH_SERVE_H_

#include "vah_batch.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VAH_REQUEST_REJECTED (-2) /* vah_request_result.state of a rejected request */

typedef struct vah_request {
    vah_batch_event ev;        /* lattice, spacings, dt, t0, eta/s, step limit, tag */
    double arrival_s;          /* seconds after the start of the run */
    const double *e;           /* initial energy density (interior cells, x fastest),
                                  or NULL: generated on the device from `glauber` */
    vah_glauber glauber;       /* read when e == NULL */
    double pl_ratio;           /* P_L = pl_ratio e; <= 0: the reference's 0.0154483 e */
} vah_request;

typedef struct vah_request_result {
    long long tag;
    int slot;                  /* batch slot it ran in; -1 when rejected */
    int state;                 /* enum vah_event_state at retirement, or VAH_REQUEST_REJECTED */
    long long nstep;           /* steps taken */
    double tau;                /* proper time reached */
    double admitted_s, finished_s, latency_s;  /* latency = finished - arrival */
} vah_request_result;

typedef struct vah_serve_stats {
    int served, rejected, peak_resident, polls;
    double seconds;            /* wall time of the run */
    double events_per_s;       /* served / seconds */
    double p50_latency_s, p99_latency_s;  /* over served requests (linear interpolation) */
} vah_serve_stats;

/* Called once per event as it retires, before its slot is freed, so that its
 * fields can be read (vah_batch_get_state with the result's slot). */
typedef void (*vah_retire_fn)(void *user, const vah_request_result *r);

/* The operations the loop drives; vah_serve supplies the batch engine's.
 * admit returns VAH_OK with *slot, VAH_ERR_BUDGET or VAH_ERR_FULL when the
 * request does not fit now, any other code as an error that ends the run.
 * now / sleep may be NULL (the steady clock and a thread sleep); on_poll may
 * be NULL. */
typedef struct vah_serve_backend {
    void *user;
    int (*admit)(void *user, const vah_request *r, int *slot);
    int (*step)(void *user, int nsteps);
    int (*poll)(void *user, vah_batch_status *out, int cap, int *n);
    int (*retire)(void *user, int slot);
    double (*now)(void *user);
    void (*sleep)(void *user, double seconds);
    void (*on_poll)(void *user, double t, int resident, int waiting);
} vah_serve_backend;

/* Runs the loop until every request has been served or rejected.  results
 * (n entries, in request order) and stats may be NULL.  steps_per_poll >= 1:
 * steps enqueued between two polls. */
int vah_serve_run(const vah_serve_backend *be, const vah_request *reqs, int n, int steps_per_poll,
                  vah_retire_fn on_retire, void *user, vah_request_result *results, vah_serve_stats *stats);

/* The loop over a batch (vah_batch_admit / vah_batch_admit_glauber,
 * vah_batch_step, vah_batch_poll, vah_batch_retire) and the steady clock.
 * The batch should hold no events of its own: every resident event is
 * taken as one of these requests. */
int vah_serve(vah_batch *b, const vah_request *reqs, int n, int steps_per_poll, vah_retire_fn on_retire, void *user,
              vah_request_result *results, vah_serve_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* VAH_SERVE_H_ */
