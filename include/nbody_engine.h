/*
 * nbody_engine.h — device-resident stepping on top of the force entries of
 * nbody_hip.h: one system, or an ensemble of independent systems.
 *
 * The reference's step (src/nbody.cu:293-470, driven by src/nbody.cu:1015-1088)
 * copies positions to the device, runs the force entry, copies forces back and
 * integrates on the host (src/nbody.cu:352-370,420-441). Here positions,
 * velocities and forces stay in HBM; a step is the force evaluation followed by
 * integrateGravitation (src/nbody.cu:158-236), queued on the object's own
 * stream, optionally replayed from a captured HIP graph.
 *
 * All functions return 0 on success or a nonzero code with the message in
 * nbody_hip_last_error(); create functions return NULL on failure.
 * Buffers are host float4 arrays (posMass x,y,z,m; velInvMass vx,vy,vz,1/m).
 */
#ifndef NBODY_ENGINE_H
#define NBODY_ENGINE_H

#include <stddef.h>

#include "nbody_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- one system ------------------------------------------------------------ */

typedef struct nbody_hip_system nbody_hip_system;

/* Uploads N bodies and fixes the plan (opt as for nbody_hip_force; NULL: the
 * library's choice, including a plan nbody_hip_autotune found for this N). */
nbody_hip_system *nbody_hip_system_create(const float *posMass, const float *velInvMass, size_t N,
                                          float softeningSquared, float dt, float damping,
                                          const nbody_hip_options *opt);
/* Runs `steps` steps and waits for them; *ms (may be NULL) receives their
 * elapsed time measured with events on the system's stream. use_graph = 1
 * replays one captured graph per step (force kernel, slice/slot reduction,
 * integrator; captured on first use). */
int nbody_hip_system_run(nbody_hip_system *s, int steps, int use_graph, float *ms);
/* Copies the current state (and the last force) to the host; NULL skips one. */
int nbody_hip_system_read(const nbody_hip_system *s, float *posMass, float *velInvMass, float *force);
/* The plan in use. */
int nbody_hip_system_plan(const nbody_hip_system *s, nbody_hip_options *plan);
void nbody_hip_system_destroy(nbody_hip_system *s);

/* ---- ensembles of independent systems ------------------------------------ */

typedef struct nbody_hip_ensemble nbody_hip_ensemble;

enum {
    NBODY_HIP_BATCHED_GRAPH = 0, /* per step, for each of two groups of systems on its own stream: one
                                    persistent force launch + one slot-sum-and-integrate launch; from a graph */
    NBODY_HIP_BATCHED = 1,       /* the same launches per step, without a graph */
    NBODY_HIP_PER_SYSTEM = 2     /* every system on its own: its best single-system plan, one after another */
};

/* count systems of N[s] bodies, stored back to back in posMass/velInvMass
 * (system s starts at body sum_{r<s} N[r]); softeningSquared[s] per system;
 * dt and damping shared. The batched modes evaluate a system of N <= 2^18
 * bodies pairwise-symmetrically in blocks of 512, each pair once (block
 * pairs exactly as nbody_hip_force's symmetric plan with 8 bodies per lane
 * and blocks of 512; each diagonal block as one task, where that plan sums
 * it one-sidedly, so the two agree to rounding), keeping 12 * ceil(N/512)^2 * 512 bytes of
 * block partials in HBM; a larger one one-sidedly with compensated sums
 * (12 N bytes of partials). */
nbody_hip_ensemble *nbody_hip_ensemble_create(size_t count, const size_t *N, const float *posMass,
                                              const float *velInvMass, const float *softeningSquared, float dt,
                                              float damping);
/* The same with options (NULL: as nbody_hip_ensemble_create). opt->compensated
 * = 2 selects FP64 accumulation for every mode: the batched modes then run
 * the FP64 symmetric plan's pieces (4 bodies per lane, blocks of 256, FP64
 * block partials, 24 * ceil(N/256)^2 * 256 bytes per system of N <= 2^17;
 * a larger system one-sidedly with FP64 sums, 24 N bytes), so a system of
 * 1024 .. 2^17 bodies gets bit for bit the forces of nbody_hip_force with
 * compensated = 2; the per-system mode plans each system with opt. Other
 * option fields apply to the per-system mode only. No counterpart in the
 * reference, whose systems run one at a time (src/nbody.cu:1015-1088). */
nbody_hip_ensemble *nbody_hip_ensemble_create_opt(size_t count, const size_t *N, const float *posMass,
                                                  const float *velInvMass, const float *softeningSquared, float dt,
                                                  float damping, const nbody_hip_options *opt);
/* Runs `steps` steps of every system in `mode` and waits; *ms as above. In
 * the batched modes each body's force comes from the same arithmetic
 * whatever the other systems are, so a system gives bit-identical results
 * alone and in any ensemble. */
int nbody_hip_ensemble_run(nbody_hip_ensemble *e, int steps, int mode, float *ms);
int nbody_hip_ensemble_read(const nbody_hip_ensemble *e, float *posMass, float *velInvMass, float *force);
void nbody_hip_ensemble_destroy(nbody_hip_ensemble *e);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_ENGINE_H */
