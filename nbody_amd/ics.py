"""Seeded initial conditions (include/nbody_ics.h), returned as numpy float4
arrays: posMass (N, 4) = (x, y, z, m) and velInvMass (N, 4) = (vx, vy, vz, 1/m).
"""
import math

import numpy as np

from . import _native


def _call(name, n, *args):
    pm = np.zeros((n, 4), np.float32)
    vm = np.zeros((n, 4), np.float32)
    rc = getattr(_native.lib(), name)(pm.ctypes.data, vm.ctypes.data, n, *args)
    _native.check(rc, name)
    return pm, vm


def random_unit_bodies(n, seed=7, scale=0.9, velscale=8.0):
    """The reference driver's ICs: seedRandom(seed) + randomUnitBodies
    (src/nbody_util.cpp:83-86,102-144; defaults src/nbody.cu:81,155-156)."""
    return _call("nbody_random_unit_bodies", n, seed, scale, velscale)


def ball_radius(n, r16k=1.54):
    """Radius of the uniform-ball workloads: 1.54 at N = 16384, growing as
    N^(1/3) so the density is fixed."""
    return r16k * (n / 16384.0) ** (1.0 / 3.0)


def uniform_ball(n, seed=1, radius=None, vradius=1.0, mass=1.0):
    """Uniform ball (default radius ball_radius(n)), velocities uniform in a
    ball of vradius, equal masses."""
    r = ball_radius(n) if radius is None else radius
    return _call("nbody_uniform_ball", n, seed, r, vradius, mass)


def plummer(n, seed=1, a=1.0, total_mass=1.0, unit_inv_mass=True):
    """Plummer sphere (Aarseth, Henon & Wielen 1974), G = 1."""
    return _call("nbody_plummer", n, seed, a, total_mass, 1 if unit_inv_mass else 0)


def log_uniform_sizes(count, lo, hi, seed):
    """`count` integers log-uniform in [lo, hi] from a seeded numpy generator."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(math.log(lo), math.log(hi + 1), size=count)).astype(np.int64).clip(lo, hi)
