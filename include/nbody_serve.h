/*
 * nbody_serve.h — a request server for simulations on one GPU.
 *
 * A request is a whole simulation: N bodies with seeded initial conditions,
 * `steps` steps of force + integrateGravitation (src/nbody.cu:158-236,293-470).
 * The server keeps every admitted request's positions and velocities resident
 * in HBM and runs continuous batching: each iteration advances every active
 * request by one step with the batched launches of nbody_engine.h (the
 * requests in two groups of about equal work, each group one persistent force
 * launch and one reduce + integrate launch on its own stream).
 * Waiting requests are admitted in arrival order while a slot is free and
 * their device memory fits the HBM budget; while a step runs they are
 * uploaded and join the next step. A request's slot frees when its last step
 * is launched, its memory when that step has finished. A request's result
 * depends only on the request, never on what else is in the batch.
 *
 * Functions return 0 or a nonzero code (nbody_hip_last_error()).
 */
#ifndef NBODY_SERVE_H
#define NBODY_SERVE_H

#include <stddef.h>
#include <stdint.h>

#include "nbody_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { NBODY_IC_UNIFORM_BALL = 0, NBODY_IC_PLUMMER = 1 };

typedef struct nbody_hip_request {
    uint64_t id;
    size_t N;               /* bodies */
    int steps;              /* steps to run */
    uint32_t seed;          /* initial-condition seed */
    int ic;                 /* NBODY_IC_UNIFORM_BALL (radius 1.54 (N/16384)^(1/3), unit masses,
                               velocities in a unit ball) or NBODY_IC_PLUMMER (a = 1, total mass 1,
                               velInvMass.w = 1) */
    float softeningSquared;
    double arrival_s;       /* arrival time, seconds after the run starts (trace replay) */
} nbody_hip_request;

typedef struct nbody_hip_result {
    uint64_t id;
    size_t N;
    int steps;
    int status;             /* 0 done; 1 rejected (cannot fit the budget); 2 failed */
    double arrival_s, admit_s, finish_s;  /* seconds after the run started */
} nbody_hip_result;

typedef struct nbody_hip_server nbody_hip_server;

/* budget_bytes: HBM the resident requests may hold (48 bytes per body plus
 * the batched force's block partials, as for nbody_hip_ensemble_create);
 * max_active: slots; dt, damping: the integrator's. keep_states != 0 keeps
 * each finished request's final posMass/velInvMass on the host. */
nbody_hip_server *nbody_hip_server_create(size_t budget_bytes, int max_active, float dt, float damping,
                                          int keep_states);
/* The same with options (NULL: as nbody_hip_server_create): opt->compensated
 * = 2 steps every request with FP64 accumulation, the batched FP64 mode of
 * nbody_hip_ensemble_create_opt; no other field is read. */
nbody_hip_server *nbody_hip_server_create_opt(size_t budget_bytes, int max_active, float dt, float damping,
                                              int keep_states, const nbody_hip_options *opt);

/* Device bytes a resident request of N bodies holds, the amount the budget
 * counts: 48 N (posMass, velInvMass, force) plus its block partials. */
size_t nbody_hip_request_bytes(size_t N);
/* The same for a server created with opt (FP64 partials with compensated = 2). */
size_t nbody_hip_request_bytes_opt(size_t N, const nbody_hip_options *opt);

/* Replays a trace: request k enters the queue when arrival_s seconds have
 * passed since the call started; returns when every request has finished or
 * been rejected. results[k] matches reqs[k]. *wall_s (may be NULL) gets the
 * run's wall time and *interactions the sum of N^2 * steps done. */
int nbody_hip_server_run_trace(nbody_hip_server *s, const nbody_hip_request *reqs, size_t count,
                               nbody_hip_result *results, double *wall_s, double *interactions);

/* Final state of a finished request (keep_states only): N float4 each. */
int nbody_hip_server_final_state(const nbody_hip_server *s, uint64_t id, float *posMass, float *velInvMass);

/* Largest number of requests that were active at once in the last run. */
int nbody_hip_server_peak_active(const nbody_hip_server *s);

void nbody_hip_server_destroy(nbody_hip_server *s);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_SERVE_H */
