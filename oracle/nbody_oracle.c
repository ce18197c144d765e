/*
 * Oracle, FP32 part: CPU restatement of tycho/nbody's force algorithms,
 * integrator, crosscheck metric and seeded initial conditions.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Built by oracle/Makefile with the
 * reference's CPU flags (-O3 -ffast-math -march=x86-64-v4 -fopenmp,
 * meson.build:3-5,33-38), so the host 1.0f/sqrtf (src/chCUDA.h:143-146)
 * lowers to the hardware reciprocal-sqrt estimate plus one Newton step, as in
 * the reference build, and the j-sums are vectorised with the same
 * `omp simd reduction`.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   oracle_cpu_*            against the reference compiled from its own sources
 *                           (oracle/_ref/libnbody_ref.so) on the same inputs,
 *                           against the FP64 oracle, closed-form two-body and
 *                           self/zero-mass cases, and the committed fixtures.
 *   oracle_integrate_*      against a one-step closed form evaluated in FP64.
 *   oracle_rel_error        hand-worked cases incl. the sign and zero quirks.
 *   oracle_random_*         bit-for-bit against the reference's
 *                           randomUnitBodies output (fixture + _ref).
 */
#include "oracle.h"

#include <immintrin.h>
#include <math.h>
#include <string.h>
#include <time.h>

#define ORACLE_ALIGN 64

static double wall_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* src/chCUDA.h:143-146: the host stand-in for rsqrtf is 1.0f / sqrtf. */
static inline float host_rsqrt(float v) { return 1.0f / sqrtf(v); }

/* src/bodybodyInteraction.cuh:40-65. The force on body 0 due to body 1,
 * without body 0's mass and without G: m1 * d / (|d|^2 + eps^2)^(3/2). */
static inline void pair_force(float *fx, float *fy, float *fz,
                              float x0, float y0, float z0,
                              float x1, float y1, float z1, float m1, float eps2)
{
    const float dx = x1 - x0;
    const float dy = y1 - y0;
    const float dz = z1 - z0;
    float d2 = dx * dx + dy * dy + dz * dz;
    d2 += eps2;
    const float inv = host_rsqrt(d2);
    const float inv3 = inv * inv * inv;
    const float s = m1 * inv3;
    *fx = dx * s;
    *fy = dy * s;
    *fz = dz * s;
}

void oracle_body_body(float *f3, float xi, float yi, float zi,
                      float xj, float yj, float zj, float mj, float eps2)
{
    pair_force(&f3[0], &f3[1], &f3[2], xi, yi, zi, xj, yj, zj, mj, eps2);
}

/* src/nbody_CPU_SOA.cpp:53-105: i over OpenMP threads (guided), j as an
 * `omp simd` sum reduction; the result overwrites force. */
float oracle_cpu_soa(float **force, float **pos, float *mass, float eps2, size_t n)
{
    const double t0 = wall_ms();
    const float *px = __builtin_assume_aligned(pos[0], ORACLE_ALIGN);
    const float *py = __builtin_assume_aligned(pos[1], ORACLE_ALIGN);
    const float *pz = __builtin_assume_aligned(pos[2], ORACLE_ALIGN);
    const float *pm = __builtin_assume_aligned(mass, ORACLE_ALIGN);
    float *fx = __builtin_assume_aligned(force[0], ORACLE_ALIGN);
    float *fy = __builtin_assume_aligned(force[1], ORACLE_ALIGN);
    float *fz = __builtin_assume_aligned(force[2], ORACLE_ALIGN);

#pragma omp parallel for schedule(guided)
    for (size_t i = 0; i < n; i++) {
        const float xi = px[i], yi = py[i], zi = pz[i];
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma omp simd reduction(+ : ax, ay, az)
        for (size_t j = 0; j < n; j++) {
            float gx, gy, gz;
            pair_force(&gx, &gy, &gz, xi, yi, zi, px[j], py[j], pz[j], pm[j], eps2);
            ax += gx;
            ay += gy;
            az += gz;
        }
        fx[i] = ax;
        fy[i] = ay;
        fz[i] = az;
    }
    return (float)(wall_ms() - t0);
}

/* src/nbody_CPU_AOS.cpp:51-98: same sum over float4 (x,y,z,m) records; x,y,z
 * of each force record are written, w is left alone. */
float oracle_cpu_aos(float *force, float *posMass, float eps2, size_t n)
{
    const double t0 = wall_ms();
    const float *p = __builtin_assume_aligned(posMass, ORACLE_ALIGN);
    float *f = __builtin_assume_aligned(force, ORACLE_ALIGN);

#pragma omp parallel for schedule(guided)
    for (size_t i = 0; i < n; i++) {
        const float xi = p[4 * i + 0], yi = p[4 * i + 1], zi = p[4 * i + 2];
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma omp simd reduction(+ : ax, ay, az)
        for (size_t j = 0; j < n; j++) {
            float gx, gy, gz;
            pair_force(&gx, &gy, &gz, xi, yi, zi,
                       p[4 * j + 0], p[4 * j + 1], p[4 * j + 2], p[4 * j + 3], eps2);
            ax += gx;
            ay += gy;
            az += gz;
        }
        f[4 * i + 0] = ax;
        f[4 * i + 1] = ay;
        f[4 * i + 2] = az;
    }
    return (float)(wall_ms() - t0);
}

#define TILE_BODIES 4096 /* src/nbody_CPU_SOA_tiled.cpp:51 */

/* src/nbody_CPU_SOA_tiled.cpp:51-115: j swept in 4096-body tiles inside one
 * parallel region; each tile's partial sum is added (+=) into force, which the
 * caller zeroes first (src/nbody.cu:328-331). Not applicable (returns 0)
 * unless n is a multiple of the tile. */
float oracle_cpu_soa_tiled(float **force, float **pos, float *mass, float eps2, size_t n)
{
    const double t0 = wall_ms();
    if (n % TILE_BODIES != 0)
        return 0.0f;
    const float *px = __builtin_assume_aligned(pos[0], ORACLE_ALIGN);
    const float *py = __builtin_assume_aligned(pos[1], ORACLE_ALIGN);
    const float *pz = __builtin_assume_aligned(pos[2], ORACLE_ALIGN);
    const float *pm = __builtin_assume_aligned(mass, ORACLE_ALIGN);
    float *fx = __builtin_assume_aligned(force[0], ORACLE_ALIGN);
    float *fy = __builtin_assume_aligned(force[1], ORACLE_ALIGN);
    float *fz = __builtin_assume_aligned(force[2], ORACLE_ALIGN);

#pragma omp parallel
    for (size_t t0j = 0; t0j < n; t0j += TILE_BODIES) {
        const size_t t1j = t0j + TILE_BODIES;
#pragma omp for schedule(guided)
        for (size_t i = 0; i < n; i++) {
            const float xi = px[i], yi = py[i], zi = pz[i];
            float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma omp simd reduction(+ : ax, ay, az)
            for (size_t j = t0j; j < t1j; j++) {
                float gx, gy, gz;
                pair_force(&gx, &gy, &gz, xi, yi, zi, px[j], py[j], pz[j], pm[j], eps2);
                ax += gx;
                ay += gy;
                az += gz;
            }
            fx[i] += ax;
            fy[i] += ay;
            fz[i] += az;
        }
    }
    return (float)(wall_ms() - t0);
}

/* src/nbody_CPU_AOS_tiled.cpp:51-110: the AOS twin of the tiled sweep. */
float oracle_cpu_aos_tiled(float *force, float *posMass, float eps2, size_t n)
{
    const double t0 = wall_ms();
    if (n % TILE_BODIES != 0)
        return 0.0f;
    const float *p = __builtin_assume_aligned(posMass, ORACLE_ALIGN);
    float *f = __builtin_assume_aligned(force, ORACLE_ALIGN);

#pragma omp parallel
    for (size_t t0j = 0; t0j < n; t0j += TILE_BODIES) {
        const size_t t1j = t0j + TILE_BODIES;
#pragma omp for schedule(guided)
        for (size_t i = 0; i < n; i++) {
            const float xi = p[4 * i + 0], yi = p[4 * i + 1], zi = p[4 * i + 2];
            float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma omp simd reduction(+ : ax, ay, az)
            for (size_t j = t0j; j < t1j; j++) {
                float gx, gy, gz;
                pair_force(&gx, &gy, &gz, xi, yi, zi,
                           p[4 * j + 0], p[4 * j + 1], p[4 * j + 2], p[4 * j + 3], eps2);
                ax += gx;
                ay += gy;
                az += gz;
            }
            f[4 * i + 0] += ax;
            f[4 * i + 1] += ay;
            f[4 * i + 2] += az;
        }
    }
    return (float)(wall_ms() - t0);
}

/* src/bodybodyInteraction_AVX.h:39-49: 8-wide hardware rsqrt estimate refined
 * by one Newton-Raphson step, y1 = (0.5 y0) * (3 - x y0^2). */
static inline __m256 avx_rsqrt_newton(__m256 x)
{
    const __m256 y0 = _mm256_rsqrt_ps(x);
    const __m256 xyy = _mm256_mul_ps(_mm256_mul_ps(y0, y0), x);
    const __m256 half_y = _mm256_mul_ps(_mm256_set1_ps(0.5f), y0);
    const __m256 corr = _mm256_sub_ps(_mm256_set1_ps(3.0f), xyy);
    return _mm256_mul_ps(half_y, corr);
}

/* src/bodybodyInteraction_AVX.h:51-59: hadd, hadd, then add the two halves. */
static inline float avx_hsum(__m256 a)
{
    const __m256 h1 = _mm256_hadd_ps(a, a);
    const __m256 h2 = _mm256_hadd_ps(h1, h1);
    const __m128 hi = _mm256_extractf128_ps(h2, 1);
    return _mm_cvtss_f32(_mm_add_ss(_mm256_castps256_ps128(h2), hi));
}

/* src/nbody_CPU_AVX.cpp:54-102 with the 8-wide interaction of
 * src/bodybodyInteraction_AVX.h:61-110 (products and sums kept in the
 * reference's order; the accumulators add d*s per lane). */
float oracle_cpu_avx(float **force, float **pos, float *mass, float eps2, size_t n)
{
    const double t0 = wall_ms();
    if (n % 8 != 0)
        return 0.0f;
    const __m256 e2 = _mm256_set1_ps(eps2);

#pragma omp parallel for schedule(guided)
    for (size_t i = 0; i < n; i++) {
        const __m256 x0 = _mm256_set1_ps(pos[0][i]);
        const __m256 y0 = _mm256_set1_ps(pos[1][i]);
        const __m256 z0 = _mm256_set1_ps(pos[2][i]);
        __m256 ax = _mm256_setzero_ps();
        __m256 ay = _mm256_setzero_ps();
        __m256 az = _mm256_setzero_ps();
        for (size_t j = 0; j < n; j += 8) {
            const __m256 dx = _mm256_sub_ps(_mm256_load_ps(&pos[0][j]), x0);
            const __m256 dy = _mm256_sub_ps(_mm256_load_ps(&pos[1][j]), y0);
            const __m256 dz = _mm256_sub_ps(_mm256_load_ps(&pos[2][j]), z0);
            __m256 d2 = _mm256_add_ps(_mm256_add_ps(_mm256_mul_ps(dx, dx), _mm256_mul_ps(dy, dy)),
                                      _mm256_mul_ps(dz, dz));
            d2 = _mm256_add_ps(d2, e2);
            const __m256 inv = avx_rsqrt_newton(d2);
            const __m256 inv3 = _mm256_mul_ps(inv, _mm256_mul_ps(inv, inv));
            const __m256 s = _mm256_mul_ps(_mm256_load_ps(&mass[j]), inv3);
            ax = _mm256_add_ps(ax, _mm256_mul_ps(dx, s));
            ay = _mm256_add_ps(ay, _mm256_mul_ps(dy, s));
            az = _mm256_add_ps(az, _mm256_mul_ps(dz, s));
        }
        force[0][i] = avx_hsum(ax);
        force[1][i] = avx_hsum(ay);
        force[2][i] = avx_hsum(az);
    }
    return (float)(wall_ms() - t0);
}

/* src/nbody.cu:158-182 integrateOne: kick with a = F * invMass, damp the
 * velocity, then drift with the new velocity. */
static inline void integrate_body(float *p, float *v, const float *f, float invMass,
                                  float dt, float damping)
{
    v[0] += (f[0] * invMass) * dt;
    v[1] += (f[1] * invMass) * dt;
    v[2] += (f[2] * invMass) * dt;
    v[0] *= damping;
    v[1] *= damping;
    v[2] *= damping;
    p[0] += v[0] * dt;
    p[1] += v[1] * dt;
    p[2] += v[2] * dt;
}

/* src/nbody.cu:186-236 integrateGravitation_SOA. */
void oracle_integrate_soa(float **pos, float *invMass, float **vel, float **force,
                          float dt, float damping, size_t n)
{
#pragma omp simd
    for (size_t i = 0; i < n; i++) {
        float p[3] = {pos[0][i], pos[1][i], pos[2][i]};
        float v[3] = {vel[0][i], vel[1][i], vel[2][i]};
        const float f[3] = {force[0][i], force[1][i], force[2][i]};
        integrate_body(p, v, f, invMass[i], dt, damping);
        pos[0][i] = p[0];
        pos[1][i] = p[1];
        pos[2][i] = p[2];
        vel[0][i] = v[0];
        vel[1][i] = v[1];
        vel[2][i] = v[2];
    }
}

/* src/nbody.cu:240-286 integrateGravitation_AOS: invMass is velInvMass.w. */
void oracle_integrate_aos(float *posMass, float *velInvMass, const float *force,
                          float dt, float damping, size_t n)
{
#pragma omp simd
    for (size_t i = 0; i < n; i++) {
        float p[3] = {posMass[4 * i + 0], posMass[4 * i + 1], posMass[4 * i + 2]};
        float v[3] = {velInvMass[4 * i + 0], velInvMass[4 * i + 1], velInvMass[4 * i + 2]};
        const float f[3] = {force[4 * i + 0], force[4 * i + 1], force[4 * i + 2]};
        integrate_body(p, v, f, velInvMass[4 * i + 3], dt, damping);
        posMass[4 * i + 0] = p[0];
        posMass[4 * i + 1] = p[1];
        posMass[4 * i + 2] = p[2];
        velInvMass[4 * i + 0] = v[0];
        velInvMass[4 * i + 1] = v[1];
        velInvMass[4 * i + 2] = v[2];
    }
}

/* src/nbody.cu:120-125. Signed on purpose: a negative golden component gives
 * a negative "error" that never raises the maximum; b == 0 != a gives inf. */
float oracle_rel_error(float a, float b)
{
    if (a == b)
        return 0.0f;
    return fabsf(a - b) / b;
}

/* src/nbody.cu:406-418: max relError over x,y,z (w skipped) of float4 records. */
float oracle_max_rel_error(const float *force, const float *golden, size_t n)
{
    float worst = 0.0f;
    for (size_t k = 0; k < 4 * n; k++) {
        if ((k + 1) % 4 == 0)
            continue;
        const float e = oracle_rel_error(force[k], golden[k]);
        if (e > worst)
            worst = e;
    }
    return worst;
}

/* src/nbody_util.cpp:81: std::linear_congruential_engine<uint32_t, 214013,
 * 2531011, 2^31>: x <- (214013 x + 2531011) mod 2^31, output = new state,
 * max() = 2^31 - 1, default seed 1. */
#define LCG_A 214013u
#define LCG_C 2531011u
#define LCG_M_MASK 0x7fffffffu
static uint32_t lcg_state = 1u;

/* src/nbody_util.cpp:83-86 via std::linear_congruential_engine::seed: the
 * state becomes seed mod 2^31 (c mod m != 0, so a zero state is allowed). */
void oracle_seed_random(uint32_t seed) { lcg_state = seed & LCG_M_MASK; }

static inline uint32_t lcg_next(void)
{
    lcg_state = (uint32_t)(((uint64_t)LCG_A * lcg_state + LCG_C) & LCG_M_MASK);
    return lcg_state;
}

/* src/nbody_util.cpp:88-95: u = (float)v / (float)max(), then the blend
 * (1-u)*lo + u*hi. */
float oracle_nbody_random(float lo, float hi)
{
    const uint32_t v = lcg_next();
    const float u = (float)v / (float)LCG_M_MASK;
    return (1.0f - u) * lo + u * hi;
}

static inline float dot3(const float *a, const float *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* src/nbody_util.cpp:102-144: position drawn in the cube, rejected outside
 * the unit ball; then velocity likewise; either rejection redraws both.
 * Position scale = scale * N / 1024 (linear in N), velocity scale =
 * scale * velscale, mass ~ U(0.5, 1), vel.w = 1 / mass. */
void oracle_random_unit_bodies(float *pos, float *vel, size_t n, float scale, float velscale)
{
    const float pscale = scale * (n / 1024.0f);
    const float vscale = scale * velscale;
    size_t i = 0;
    while (i < n) {
        float p[3], v[3];
        p[0] = oracle_nbody_random(-1.0f, 1.0f);
        p[1] = oracle_nbody_random(-1.0f, 1.0f);
        p[2] = oracle_nbody_random(-1.0f, 1.0f);
        if (dot3(p, p) > 1)
            continue;
        v[0] = oracle_nbody_random(-1.0f, 1.0f);
        v[1] = oracle_nbody_random(-1.0f, 1.0f);
        v[2] = oracle_nbody_random(-1.0f, 1.0f);
        if (dot3(v, v) > 1)
            continue;
        pos[4 * i + 0] = p[0] * pscale;
        pos[4 * i + 1] = p[1] * pscale;
        pos[4 * i + 2] = p[2] * pscale;
        pos[4 * i + 3] = oracle_nbody_random(0.5f, 1.0f);
        vel[4 * i + 0] = v[0] * vscale;
        vel[4 * i + 1] = v[1] * vscale;
        vel[4 * i + 2] = v[2] * vscale;
        vel[4 * i + 3] = 1.0f / pos[4 * i + 3];
        i++;
    }
}
