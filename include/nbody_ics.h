/*
 * nbody_ics.h — seeded initial conditions of libnbody_hip.so (host code).
 * Layout: N float4 records each for posMass (x, y, z, m) and velInvMass
 * (vx, vy, vz, 1/m), as the reference's g_hostAOS_PosMass / g_hostAOS_VelInvMass
 * (src/nbody.cu:129-131). All return 0, or 1 on a bad argument.
 */
#ifndef NBODY_ICS_H
#define NBODY_ICS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces seedRandom + randomUnitBodies (src/nbody_util.cpp:83-86,102-144),
 * bit-for-bit: positions scaled by scale*N/1024, velocities by scale*velscale,
 * masses U(0.5, 1). The reference driver calls it with scale 0.9, velscale 8,
 * seed 7 (src/nbody.cu:81,155-156,996). */
int nbody_random_unit_bodies(float *posMass, float *velInvMass, size_t N, uint32_t seed, float scale,
                             float velscale);

/* Uniform ball of the given radius, velocities uniform in a ball of vradius,
 * equal masses; same engine and rejection draws as randomUnitBodies. */
int nbody_uniform_ball(float *posMass, float *velInvMass, size_t N, uint32_t seed, float radius, float vradius,
                       float mass);

/* Plummer sphere (Aarseth, Henon & Wielen 1974), G = 1, scale radius a,
 * total_mass split equally; velInvMass.w = 1 when unit_inv_mass != 0. */
int nbody_plummer(float *posMass, float *velInvMass, size_t N, uint32_t seed, float a, float total_mass,
                  int unit_inv_mass);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_ICS_H */
