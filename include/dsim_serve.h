/*
 * dsim serving: many simulation requests on one GPU at once (C ABI).
 *
 * No counterpart in the reference, which runs one simulation per process and
 * re-walks the spins for every gradient line (dSim3/src/diffusion.cpp:
 * 1068-1184, 564-566).  A server keeps substrates resident on the device,
 * takes requests (substrate, spin count, seed, physics, gradient lobes,
 * readout lines) as they arrive, admits them first-come first-served within
 * an HBM budget, and advances every admitted request together: one batched
 * step launch per round (a replayed HIP graph) carries all of them, each by
 * up to `round_steps` steps of its current gradient lobe.  A request's result
 * does not depend on what else is running: every spin draws from the Philox
 * stream of its request's seed, spin id and step.
 *
 * Requests use the physical mode (include/dsim.h: DSIM_MODE_PHYSICAL) with
 * the Philox stream and the shared-walk readout: the lobes accumulate
 * gradient moments, and the signal of every readout line is formed from them
 * at the end (dsim_signal_moment).
 */
#ifndef DSIM_SERVE_H
#define DSIM_SERVE_H

#include <stdint.h>

#include "dsim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dsim_server dsim_server;

typedef struct {
    uint32_t substrate;       /* id from dsim_server_add_substrate */
    uint32_t n_spins;
    uint64_t seed;            /* Philox key of this request */
    float dt;                 /* ms */
    float std_dev[4];         /* per compartment, normalised units */
    float t2[4];              /* ms, INFINITY = no decay */
    float permeability;       /* transmission probability per hit */
    uint32_t boundary;        /* DSIM_BOUNDARY_* */
    float start_box;          /* spins start uniform in [-start_box, start_box]^3 */
    float label_extent[3];    /* half-extent of the labelling rays (dsim_label_spins) */
    uint32_t n_lobes;         /* gradient lobes in time order */
    const float *lobe_weight; /* [n_lobes]: gradient as a multiple of each line's vector */
    const uint32_t *lobe_steps; /* [n_lobes] */
    uint32_t n_lines;         /* readout lines */
    const float *grads;       /* [n_lines*3] gradScaled (include/dsim.h units) */
    float phase_constant;     /* 2*pi*gamma*dt */
    float voxel_lo[3];        /* periodic voxel (dsim_params); when set, the spins start */
    float voxel_hi[3];        /* uniform in it and start_box is not used; all zero: cubic */
} dsim_request;

typedef struct {
    double submit_s, admit_s, done_s; /* server clock (dsim_server_clock) */
    uint64_t spin_steps;
    uint32_t rounds;                  /* batched rounds the request took part in */
    uint32_t ambiguous;               /* spins whose label stayed ambiguous */
} dsim_request_stats;

/* hbm_budget: device bytes the admitted requests may hold (the substrates are
 * not counted); round_steps: steps per request per round (0 = 256);
 * max_active_spins: spins in flight at most (0 = 64M). */
int dsim_server_create(uint64_t hbm_budget, uint32_t round_steps, uint32_t max_active_spins,
                       dsim_server **out);
/* Round policy (no counterpart in the reference, which runs one simulation per
 * process, diffusion.cpp:1068-1184).  small_first_blocks > 0: a round in which
 * some prepared request needs at most that many one-wave workgroups (64 spins
 * each) runs only such requests, and the others wait for a round without
 * one: small requests walk with the device to themselves, near their solo
 * latency, at the cost of the large requests' latency and of throughput
 * while they run.  0 (the default): every prepared request takes part in
 * every round.  Results do not depend on it. */
int dsim_server_set_policy(dsim_server *srv, uint32_t small_first_blocks);
/* Stops the background loop if running, waits for the device and frees every
 * request.  Substrates stay owned by the caller. */
int dsim_server_destroy(dsim_server *srv);
/* Registers a resident substrate (created on the device with an AABB or an
 * exact-membership grid). */
int dsim_server_add_substrate(dsim_server *srv, dsim_substrate *sub, uint32_t *id);
/* Queues a request (its arrays are copied).  Rejected with DSIM_ERR_ARG if it
 * alone exceeds the budget or the spin limit. */
int dsim_server_submit(dsim_server *srv, const dsim_request *req, uint64_t *ticket);
/* One scheduling round from the calling thread: admit what fits, advance all
 * admitted requests, read out and release the finished ones.  n_active and
 * n_queued (may be NULL) report the state after the round. */
int dsim_server_round(dsim_server *srv, uint32_t *n_active, uint32_t *n_queued);
/* Background loop: a server thread runs rounds while there is work. */
int dsim_server_start(dsim_server *srv);
int dsim_server_stop(dsim_server *srv);
/* Blocks until the request is done (DSIM_OK), failed (its error), or
 * timeout_s passes (DSIM_ERR_STATE; timeout_s < 0 waits without limit). */
int dsim_server_wait(dsim_server *srv, uint64_t ticket, double timeout_s);
/* Signals of a finished request: [n_lines*5] as dsim_signal_moment (total,
 * then compartments 0-3), and its timing.  DSIM_ERR_STATE if not finished. */
int dsim_server_result(dsim_server *srv, uint64_t ticket, double *signals,
                       dsim_request_stats *stats);
/* Seconds since the server was created (the clock of dsim_request_stats). */
double dsim_server_clock(const dsim_server *srv);

#ifdef __cplusplus
}
#endif

#endif
