/*
 * Oracle, FP64 part: direct-sum forces and energy/momentum in double
 * precision. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Built strict-IEEE (-O2 -fno-fast-math -ffp-contract=off -fno-math-errno):
 * the Neumaier compensation below must not be reassociated away.
 *
 * The arithmetic follows the reference interaction (src/bodybodyInteraction.cuh:40-65)
 * term for term, in double: f_i = sum_j m_j d / (|d|^2 + eps2)^(3/2), d = x_j - x_i,
 * no m_i, no G; the self term is exactly zero because d = 0.
 *
 * Pinned by tests/test_oracle.py: the two-body closed form, exact zero for
 * self and zero-mass bodies, Newton's third law for equal masses, the
 * shell-theorem interior field of a uniform shell, and brute-force agreement
 * with a long-double sum on tiny inputs.
 */
#include "oracle.h"

#include <math.h>

#define LANES 8

/* Neumaier (improved Kahan-Babuska) step: s + c carries the running sum. */
static inline void neumaier_add(double *s, double *c, double x)
{
    const double t = *s + x;
    if (fabs(*s) >= fabs(x))
        *c += (*s - t) + x;
    else
        *c += (x - t) + *s;
    *s = t;
}

void oracle_fp64_forces(double *out, const float *posMass, size_t n, double eps2,
                        const int64_t *idx, size_t nidx)
{
    const size_t m = idx ? nidx : n;
    const size_t nblk = (m + LANES - 1) / LANES;

#pragma omp parallel for schedule(dynamic, 4)
    for (size_t b = 0; b < nblk; b++) {
        /* LANES bodies per block, each with its own independent sums, so the
         * lane loop vectorises without changing any body's arithmetic. */
        double xi[LANES], yi[LANES], zi[LANES];
        double sx[LANES], sy[LANES], sz[LANES], cx[LANES], cy[LANES], cz[LANES];
        size_t row[LANES];
        int live = 0;
        for (int l = 0; l < LANES; l++) {
            const size_t k = b * LANES + l;
            const size_t i = k < m ? (idx ? (size_t)idx[k] : k) : 0;
            row[l] = k;
            if (k < m)
                live++;
            xi[l] = posMass[4 * i + 0];
            yi[l] = posMass[4 * i + 1];
            zi[l] = posMass[4 * i + 2];
            sx[l] = sy[l] = sz[l] = cx[l] = cy[l] = cz[l] = 0.0;
        }
        for (size_t j = 0; j < n; j++) {
            const double xj = posMass[4 * j + 0], yj = posMass[4 * j + 1];
            const double zj = posMass[4 * j + 2], mj = posMass[4 * j + 3];
            for (int l = 0; l < LANES; l++) {
                const double dx = xj - xi[l], dy = yj - yi[l], dz = zj - zi[l];
                const double d2 = dx * dx + dy * dy + dz * dz + eps2;
                const double inv = 1.0 / sqrt(d2);
                const double s = mj * (inv * inv * inv);
                neumaier_add(&sx[l], &cx[l], dx * s);
                neumaier_add(&sy[l], &cy[l], dy * s);
                neumaier_add(&sz[l], &cz[l], dz * s);
            }
        }
        for (int l = 0; l < live; l++) {
            out[3 * row[l] + 0] = sx[l] + cx[l];
            out[3 * row[l] + 1] = sy[l] + cy[l];
            out[3 * row[l] + 2] = sz[l] + cz[l];
        }
    }
}

void oracle_fp64_energy(double *kinetic, double *potential, double *momentum3,
                        const float *posMass, const float *velInvMass, size_t n, double eps2)
{
    double ke = 0.0, kc = 0.0, px = 0.0, pxc = 0.0, py = 0.0, pyc = 0.0, pz = 0.0, pzc = 0.0;
    for (size_t i = 0; i < n; i++) {
        const double m = posMass[4 * i + 3];
        const double vx = velInvMass[4 * i + 0], vy = velInvMass[4 * i + 1], vz = velInvMass[4 * i + 2];
        neumaier_add(&ke, &kc, 0.5 * m * (vx * vx + vy * vy + vz * vz));
        neumaier_add(&px, &pxc, m * vx);
        neumaier_add(&py, &pyc, m * vy);
        neumaier_add(&pz, &pzc, m * vz);
    }
    double pe = 0.0, pc = 0.0;
#pragma omp parallel
    {
        double lpe = 0.0, lpc = 0.0;
#pragma omp for schedule(dynamic, 64)
        for (size_t i = 0; i < n; i++) {
            const double xi = posMass[4 * i + 0], yi = posMass[4 * i + 1], zi = posMass[4 * i + 2];
            const double mi = posMass[4 * i + 3];
            double s = 0.0, c = 0.0;
            for (size_t j = i + 1; j < n; j++) {
                const double dx = posMass[4 * j + 0] - xi, dy = posMass[4 * j + 1] - yi;
                const double dz = posMass[4 * j + 2] - zi;
                neumaier_add(&s, &c, posMass[4 * j + 3] / sqrt(dx * dx + dy * dy + dz * dz + eps2));
            }
            neumaier_add(&lpe, &lpc, -mi * (s + c));
        }
#pragma omp critical
        {
            neumaier_add(&pe, &pc, lpe + lpc);
        }
    }
    *kinetic = ke + kc;
    *potential = pe + pc;
    momentum3[0] = px + pxc;
    momentum3[1] = py + pyc;
    momentum3[2] = pz + pzc;
}
