// Seeded initial conditions (host code, part of libnbody_hip.so).
//
// All generators draw from the reference's engine,
// std::linear_congruential_engine<uint32_t, 214013, 2531011, 2^31>
// (src/nbody_util.cpp:81), with u = (float)v / (float)max() (src/nbody_util.cpp:88-95).
//
// The reference builds src/nbody_util.cpp with -O3 -ffast-math, which fuses
// its float expressions; the operations that build executes are written out
// here explicitly (this file is compiled strict-IEEE, -ffp-contract=off):
//   (float)v / (float)max()        -> (float)v * 2^-31   (max() rounds to 2^31: exact)
//   (1-u)*lo + u*hi                -> fma(1-u, lo, u*hi)
//   x0*x0 + x1*x1 + x2*x2          -> fma(x2, x2, fma(x0, x0, x1*x1))
// so randomUnitBodies' roundings and rejection decisions come out bit for bit.
#include "nbody_ics.h"

#include <cmath>
#include <cstdint>

namespace {

struct Lcg {
    uint32_t x;
    explicit Lcg(uint32_t seed) : x(seed & 0x7fffffffu) {}
    uint32_t next()
    {
        x = (uint32_t)(((uint64_t)214013u * x + 2531011u) & 0x7fffffffu);
        return x;
    }
    // src/nbody_util.cpp:88-95: blend of the unit draw between lo and hi.
    float uniform(float lo, float hi)
    {
        const float u = (float)next() * 0x1p-31f;
        return std::fma(1.0f - u, lo, u * hi);
    }
};

// src/nbody_util.cpp:97-100 dot(v, v)
inline float norm2(const float* v) { return std::fma(v[2], v[2], std::fma(v[0], v[0], v[1] * v[1])); }

}  // namespace

extern "C" {

// src/nbody_util.cpp:102-144 randomUnitBodies, seeded per call
// (src/nbody.cu:919 seeds once before it; default seed 7 at src/nbody.cu:81).
int nbody_random_unit_bodies(float* pos, float* vel, size_t n, uint32_t seed, float scale, float velscale)
{
    if (!pos || !vel) return 1;
    Lcg rng(seed);
    const float pscale = scale * (n / 1024.0f);
    const float vscale = scale * velscale;
    size_t i = 0;
    while (i < n) {
        float p[3], v[3];
        p[0] = rng.uniform(-1.0f, 1.0f);
        p[1] = rng.uniform(-1.0f, 1.0f);
        p[2] = rng.uniform(-1.0f, 1.0f);
        if (norm2(p) > 1) continue;
        v[0] = rng.uniform(-1.0f, 1.0f);
        v[1] = rng.uniform(-1.0f, 1.0f);
        v[2] = rng.uniform(-1.0f, 1.0f);
        if (norm2(v) > 1) continue;
        pos[4 * i + 0] = p[0] * pscale;
        pos[4 * i + 1] = p[1] * pscale;
        pos[4 * i + 2] = p[2] * pscale;
        pos[4 * i + 3] = rng.uniform(0.5f, 1.0f);
        vel[4 * i + 0] = v[0] * vscale;
        vel[4 * i + 1] = v[1] * vscale;
        vel[4 * i + 2] = v[2] * vscale;
        vel[4 * i + 3] = 1.0f / pos[4 * i + 3];
        i++;
    }
    return 0;
}

// Positions uniform in a ball of `radius`, velocities uniform in a ball of
// `vradius`, every mass `mass` (velInvMass.w = 1/mass): the same
// cube-and-reject draws as randomUnitBodies, with the radius set directly.
int nbody_uniform_ball(float* pos, float* vel, size_t n, uint32_t seed, float radius, float vradius, float mass)
{
    if (!pos || !vel || !(mass > 0.0f)) return 1;
    Lcg rng(seed);
    size_t i = 0;
    while (i < n) {
        float p[3], v[3];
        p[0] = rng.uniform(-1.0f, 1.0f);
        p[1] = rng.uniform(-1.0f, 1.0f);
        p[2] = rng.uniform(-1.0f, 1.0f);
        if (norm2(p) > 1) continue;
        v[0] = rng.uniform(-1.0f, 1.0f);
        v[1] = rng.uniform(-1.0f, 1.0f);
        v[2] = rng.uniform(-1.0f, 1.0f);
        if (norm2(v) > 1) continue;
        pos[4 * i + 0] = p[0] * radius;
        pos[4 * i + 1] = p[1] * radius;
        pos[4 * i + 2] = p[2] * radius;
        pos[4 * i + 3] = mass;
        vel[4 * i + 0] = v[0] * vradius;
        vel[4 * i + 1] = v[1] * vradius;
        vel[4 * i + 2] = v[2] * vradius;
        vel[4 * i + 3] = 1.0f / mass;
        i++;
    }
    return 0;
}

// Plummer sphere (Aarseth, Henon & Wielen 1974, A&A 37, 183), G = 1, scale
// radius a, total mass M split equally. Radius from the inverted cumulative
// mass M(r)/M = r^3 / (r^2 + a^2)^(3/2) (draws giving r > 10a are redrawn);
// speed q * v_esc, v_esc = sqrt(2 M / a) (1 + r^2/a^2)^(-1/4), with q drawn by
// von Neumann rejection from g(q) = q^2 (1 - q^2)^(7/2) (bound 0.1); both
// directions isotropic. velInvMass.w is 1/m, or 1 when unit_inv_mass != 0:
// the force of src/bodybodyInteraction.cuh excludes m_i, so with
// invMass = 1 the reference integrator (src/nbody.cu:170-172) applies the
// physical acceleration.
int nbody_plummer(float* pos, float* vel, size_t n, uint32_t seed, float a, float total_mass, int unit_inv_mass)
{
    if (!pos || !vel || n == 0 || !(a > 0.0f) || !(total_mass > 0.0f)) return 1;
    Lcg rng(seed);
    const double m = (double)total_mass / (double)n;
    const double two_pi = 6.283185307179586;
    for (size_t i = 0; i < n; ++i) {
        double r;
        for (;;) {
            const double x1 = rng.uniform(0.0f, 1.0f);
            if (x1 <= 0.0) continue;
            r = a / std::sqrt(std::pow(x1, -2.0 / 3.0) - 1.0);
            if (std::isfinite(r) && r <= 10.0 * a) break;
        }
        const double z = (1.0 - 2.0 * rng.uniform(0.0f, 1.0f)) * r;
        const double rxy = std::sqrt(std::fmax(r * r - z * z, 0.0));
        const double phi = two_pi * rng.uniform(0.0f, 1.0f);
        double q;
        for (;;) {
            const double x4 = rng.uniform(0.0f, 1.0f);
            const double x5 = rng.uniform(0.0f, 1.0f);
            if (0.1 * x5 < x4 * x4 * std::pow(1.0 - x4 * x4, 3.5)) {
                q = x4;
                break;
            }
        }
        const double vesc = std::sqrt(2.0 * total_mass / a) * std::pow(1.0 + r * r / (a * a), -0.25);
        const double v = q * vesc;
        const double vz = (1.0 - 2.0 * rng.uniform(0.0f, 1.0f)) * v;
        const double vxy = std::sqrt(std::fmax(v * v - vz * vz, 0.0));
        const double psi = two_pi * rng.uniform(0.0f, 1.0f);
        pos[4 * i + 0] = (float)(rxy * std::cos(phi));
        pos[4 * i + 1] = (float)(rxy * std::sin(phi));
        pos[4 * i + 2] = (float)z;
        pos[4 * i + 3] = (float)m;
        vel[4 * i + 0] = (float)(vxy * std::cos(psi));
        vel[4 * i + 1] = (float)(vxy * std::sin(psi));
        vel[4 * i + 2] = (float)vz;
        vel[4 * i + 3] = unit_inv_mass ? 1.0f : (float)(1.0 / m);
    }
    return 0;
}

}  // extern "C"
