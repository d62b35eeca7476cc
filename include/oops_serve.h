/*
 * oops_serve.h — decoherence-query serving loop on one GPU.
 *
 * No reference counterpart: the reference computes one CCE job per process
 * (source/app/cce.cpp:54-149, EnsNVnBath).  A request here is one CCE job:
 * a synthetic 13C bath (seed, radius), a field, a pulse sequence (Hahn echo
 * or CPMG-N, Pulse_Interval of misc.cpp:59-84), a CCE order and a time grid;
 * its result is the CCE product L(t) (cce.cpp:151-189) for that order.
 *
 * Continuous batching: requests are queued on submit; host worker threads
 * enumerate clusters and build every cluster's operators; on each poll the
 * server admits built requests while the device-memory budget and the number
 * of in-flight slots allow, uploads them and launches their size-bucketed
 * kernels asynchronously on per-request streams, so work of requests that
 * arrive while others run overlaps on the device; finished requests are
 * reduced (CCE product) and their memory returns to the budget.  Each
 * cluster is computed independently, so a request's result does not depend
 * on what else is in flight.
 */
#ifndef OOPS_SERVE_H
#define OOPS_SERVE_H

#include <stdint.h>

#include "oops_expv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oops_server oops_server;

typedef struct {
    uint64_t seed;      /* bath seed (oops_bath_diamond) */
    double radius;      /* bath sphere radius, A */
    double p;           /* 13C occupation probability */
    double B[3];        /* field, T */
    double cutoff;      /* pair cutoff, A */
    int32_t order;      /* CCE order (1..6) */
    int32_t npulse;     /* 1: Hahn echo, N >= 2: CPMG-N */
    int32_t nt;         /* output points on [0, t_end] */
    double t_end;       /* s */
} oops_request;

/* hbm_budget: device bytes the admitted requests may hold together;
 * max_inflight: requests on the device at once; m, tol, precision as for
 * oops_echo_batch_create; threads: host builder threads. */
int oops_server_create(oops_ctx* ctx, int64_t hbm_budget, int32_t max_inflight, int32_t m, double tol,
                       int32_t precision, int32_t threads, oops_server** out);
/* Queue a request (its arrival time is the call time); returns its id. */
int oops_server_submit(oops_server* s, const oops_request* r, int64_t* id);
/* Advance: admit and launch what fits, complete what finished.  Returns the
 * number of requests completed by this call in *completed. */
int oops_server_poll(oops_server* s, int32_t* completed);
/* State of a request: 0 queued, 1 building, 2 built, 3 on the device,
 * 4 done, negative on failure (its status code). */
int oops_server_state(oops_server* s, int64_t id, int32_t* state);
/* Result of a finished request: L (nt doubles, the CCE product for its order),
 * its latency (submit to completion, s), its cluster count and the device
 * bytes it held.  Any output may be NULL. */
int oops_server_result(oops_server* s, int64_t id, double* L, double* latency_s, int64_t* nclusters,
                       int64_t* footprint);
/* When the request finished building and was launched (s since submit). */
int oops_server_times(oops_server* s, int64_t id, double* built_s, double* launched_s);
/* Clusters of a completed request whose integration failed (step limit or
 * step-size collapse: failure flags in their counters).  A request with
 * failed clusters ends in state OOPS_ERR_NUMERIC and has no result. */
int oops_server_failed(oops_server* s, int64_t id, int64_t* failed);
/* Requests not yet done. */
int oops_server_pending(oops_server* s, int32_t* pending);
int oops_server_destroy(oops_server* s);

#ifdef __cplusplus
}
#endif

#endif
