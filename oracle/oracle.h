/*
 * Oracle: CPU restatement of tycho/nbody's O(N^2) softened-gravity force path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing under nbody_amd/ links, loads or calls
 * this code; only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py use it, as the reference the GPU path is compared against (and, in
 * bench.py, as the timed CPU baseline).
 *
 * Two translation units:
 *   nbody_oracle.c       the reference's FP32 CPU algorithms, compiled with the
 *                        reference's own CPU flags (-O3 -ffast-math -march
 *                        -fopenmp, meson.build:3-5,33-38) so 1.0f/sqrtf lowers
 *                        to the same hardware rsqrt estimate + Newton step.
 *   nbody_oracle_fp64.c  FP64 direct sum with Neumaier-compensated j-sums and
 *                        the FP64 energy/momentum diagnostic; strict IEEE
 *                        (no fast-math, no FMA contraction).
 *
 * Every function cites the reference file:line it restates.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FP32 restatements (nbody_oracle.c) ---------------------------------- */

/* src/bodybodyInteraction.cuh:40-65 with host rsqrtf = 1/sqrtf (src/chCUDA.h:143-146). */
void oracle_body_body(float *f3, float xi, float yi, float zi,
                      float xj, float yj, float zj, float mj, float eps2);

/* src/nbody_CPU_SOA.cpp:53-105 (the reference golden, src/nbody.cu:313-319). */
float oracle_cpu_soa(float **force, float **pos, float *mass, float eps2, size_t n);
/* src/nbody_CPU_AOS.cpp:51-98 */
float oracle_cpu_aos(float *force, float *posMass, float eps2, size_t n);
/* src/nbody_CPU_SOA_tiled.cpp:51-115 (returns 0 unless n % 4096 == 0; force +=) */
float oracle_cpu_soa_tiled(float **force, float **pos, float *mass, float eps2, size_t n);
/* src/nbody_CPU_AOS_tiled.cpp:51-110 (returns 0 unless n % 4096 == 0; force +=) */
float oracle_cpu_aos_tiled(float *force, float *posMass, float eps2, size_t n);
/* src/nbody_CPU_AVX.cpp:54-102 + src/bodybodyInteraction_AVX.h:39-110 (n % 8 == 0) */
float oracle_cpu_avx(float **force, float **pos, float *mass, float eps2, size_t n);

/* src/nbody.cu:158-236 (integrateOne + integrateGravitation_SOA). */
void oracle_integrate_soa(float **pos, float *invMass, float **vel, float **force,
                          float dt, float damping, size_t n);
/* src/nbody.cu:240-286 (integrateGravitation_AOS; same arithmetic on float4). */
void oracle_integrate_aos(float *posMass, float *velInvMass, const float *force,
                          float dt, float damping, size_t n);

/* src/nbody.cu:120-125 relError; src/nbody.cu:406-418 max over x,y,z of AOS float4. */
float oracle_rel_error(float a, float b);
float oracle_max_rel_error(const float *force, const float *golden, size_t n);

/* src/nbody_util.cpp:81-95 (LCG engine, seedRandom, nbodyRandom) and
 * src/nbody_util.cpp:102-144 (randomUnitBodies). Own engine state, so the
 * generator can be re-seeded independently of anything else in the process. */
void oracle_seed_random(uint32_t seed);
float oracle_nbody_random(float lo, float hi);
void oracle_random_unit_bodies(float *pos, float *vel, size_t n, float scale, float velscale);

/* ---- FP64 (nbody_oracle_fp64.c) ------------------------------------------- */

/* Exact-as-FP64 force on bodies idx[0..nidx) (all bodies when idx == NULL):
 * out[3k+c] = sum_j m_j d_c / (|d|^2 + eps2)^(3/2), d = x_j - x_i, with a
 * Neumaier-compensated j-sum. eps2 is taken as given (pass (double)eps2f to
 * match an FP32 kernel's softening exactly). */
void oracle_fp64_forces(double *out, const float *posMass, size_t n, double eps2,
                        const int64_t *idx, size_t nidx);

/* FP64 diagnostics: kinetic energy sum 1/2 m |v|^2, softened potential energy
 * -sum_{i<j} m_i m_j / sqrt(|d|^2 + eps2), and total momentum sum m v (3). */
void oracle_fp64_energy(double *kinetic, double *potential, double *momentum3,
                        const float *posMass, const float *velInvMass, size_t n, double eps2);

#ifdef __cplusplus
}
#endif
