"""ctypes access to the oracle (oracle/liboracle.so) and, where it was built,
to the reference's own CPU path (oracle/_ref/libnbody_ref.so).

Test infrastructure only: the product package never imports this module.
Array arguments are numpy float32/float64 arrays; SOA entry points get
64-byte-aligned copies, as the reference's ASSERT_ALIGNED expects
(src/nbody_util.h:45,86-90).
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libnbody_ref.so")

_f = ctypes.c_float
_d = ctypes.c_double
_sz = ctypes.c_size_t
_p = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)


def aligned_empty(n, dtype=np.float32, align=64):
    """A 1-D array of n elements whose data pointer is `align`-byte aligned."""
    dt = np.dtype(dtype)
    buf = np.empty(n * dt.itemsize + align, dtype=np.uint8)
    off = (-buf.ctypes.data) % align
    return buf[off:off + n * dt.itemsize].view(dt)


def aligned_copy(a, align=64):
    out = aligned_empty(a.size, a.dtype, align)
    out[:] = a.ravel()
    return out


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _ptr3(arrs):
    return (ctypes.c_void_p * 3)(*[a.ctypes.data for a in arrs])


class _Lib:
    def __init__(self, path):
        self.path = path
        self.lib = ctypes.CDLL(path)


class Oracle(_Lib):
    """The C restatement (oracle/nbody_oracle.c, oracle/nbody_oracle_fp64.c)."""

    def __init__(self, path=ORACLE_SO):
        super().__init__(path)
        L = self.lib
        for name in ("oracle_cpu_soa", "oracle_cpu_soa_tiled", "oracle_cpu_avx"):
            fn = getattr(L, name)
            fn.restype = _f
            fn.argtypes = [_p, _p, _p, _f, _sz]
        for name in ("oracle_cpu_aos", "oracle_cpu_aos_tiled"):
            fn = getattr(L, name)
            fn.restype = _f
            fn.argtypes = [_p, _p, _f, _sz]
        L.oracle_body_body.restype = None
        L.oracle_body_body.argtypes = [_p] + [_f] * 8
        L.oracle_integrate_soa.restype = None
        L.oracle_integrate_soa.argtypes = [_p, _p, _p, _p, _f, _f, _sz]
        L.oracle_integrate_aos.restype = None
        L.oracle_integrate_aos.argtypes = [_p, _p, _p, _f, _f, _sz]
        L.oracle_rel_error.restype = _f
        L.oracle_rel_error.argtypes = [_f, _f]
        L.oracle_max_rel_error.restype = _f
        L.oracle_max_rel_error.argtypes = [_p, _p, _sz]
        L.oracle_seed_random.restype = None
        L.oracle_seed_random.argtypes = [ctypes.c_uint32]
        L.oracle_nbody_random.restype = _f
        L.oracle_nbody_random.argtypes = [_f, _f]
        L.oracle_random_unit_bodies.restype = None
        L.oracle_random_unit_bodies.argtypes = [_p, _p, _sz, _f, _f]
        L.oracle_fp64_forces.restype = None
        L.oracle_fp64_forces.argtypes = [_p, _p, _sz, _d, _p, _sz]
        L.oracle_fp64_energy.restype = None
        L.oracle_fp64_energy.argtypes = [_p, _p, _p, _p, _p, _sz, _d]

    # -- FP32 reference algorithms -------------------------------------------
    def body_body(self, xi, xj, mj, eps2):
        out = np.zeros(3, np.float32)
        self.lib.oracle_body_body(_ptr(out), *[float(v) for v in xi], *[float(v) for v in xj],
                                  float(mj), float(eps2))
        return out

    def _soa_call(self, fn, pos_mass, eps2, accumulate=None):
        n = pos_mass.shape[0]
        pos = [aligned_copy(np.ascontiguousarray(pos_mass[:, c], np.float32)) for c in range(3)]
        mass = aligned_copy(np.ascontiguousarray(pos_mass[:, 3], np.float32))
        force = [aligned_empty(n) for _ in range(3)]
        for c in range(3):
            force[c][:] = 0.0 if accumulate is None else accumulate[:, c]
        ms = fn(_ptr3(force), _ptr3(pos), _ptr(mass), float(eps2), n)
        out = np.zeros((n, 4), np.float32)
        for c in range(3):
            out[:, c] = force[c]
        return out, ms

    def _aos_call(self, fn, pos_mass, eps2):
        n = pos_mass.shape[0]
        pm = aligned_copy(np.ascontiguousarray(pos_mass, np.float32).ravel())
        force = aligned_empty(4 * n)
        force[:] = 0.0
        ms = fn(_ptr(force), _ptr(pm), float(eps2), n)
        return force.reshape(n, 4).copy(), ms

    def cpu_soa(self, pos_mass, eps2):
        return self._soa_call(self.lib.oracle_cpu_soa, pos_mass, eps2)

    def cpu_soa_tiled(self, pos_mass, eps2):
        return self._soa_call(self.lib.oracle_cpu_soa_tiled, pos_mass, eps2)

    def cpu_avx(self, pos_mass, eps2):
        return self._soa_call(self.lib.oracle_cpu_avx, pos_mass, eps2)

    def cpu_aos(self, pos_mass, eps2):
        return self._aos_call(self.lib.oracle_cpu_aos, pos_mass, eps2)

    def cpu_aos_tiled(self, pos_mass, eps2):
        return self._aos_call(self.lib.oracle_cpu_aos_tiled, pos_mass, eps2)

    # -- integrator, crosscheck, ICs ------------------------------------------
    def integrate_aos(self, pos_mass, vel_inv_mass, force, dt, damping):
        n = pos_mass.shape[0]
        pm = np.ascontiguousarray(pos_mass, np.float32).copy()
        vm = np.ascontiguousarray(vel_inv_mass, np.float32).copy()
        f = np.ascontiguousarray(force, np.float32)
        self.lib.oracle_integrate_aos(_ptr(pm), _ptr(vm), _ptr(f), float(dt), float(damping), n)
        return pm, vm

    def integrate_soa(self, pos_mass, vel_inv_mass, force, dt, damping):
        n = pos_mass.shape[0]
        pos = [aligned_copy(np.ascontiguousarray(pos_mass[:, c], np.float32)) for c in range(3)]
        vel = [aligned_copy(np.ascontiguousarray(vel_inv_mass[:, c], np.float32)) for c in range(3)]
        frc = [aligned_copy(np.ascontiguousarray(force[:, c], np.float32)) for c in range(3)]
        inv = aligned_copy(np.ascontiguousarray(vel_inv_mass[:, 3], np.float32))
        self.lib.oracle_integrate_soa(_ptr3(pos), _ptr(inv), _ptr3(vel), _ptr3(frc),
                                      float(dt), float(damping), n)
        pm = np.array(pos_mass, np.float32, copy=True)
        vm = np.array(vel_inv_mass, np.float32, copy=True)
        for c in range(3):
            pm[:, c] = pos[c]
            vm[:, c] = vel[c]
        return pm, vm

    def rel_error(self, a, b):
        return self.lib.oracle_rel_error(float(a), float(b))

    def max_rel_error(self, force, golden):
        f = np.ascontiguousarray(force, np.float32)
        g = np.ascontiguousarray(golden, np.float32)
        return self.lib.oracle_max_rel_error(_ptr(f), _ptr(g), f.shape[0])

    def random_unit_bodies(self, n, seed=7, scale=0.9, velscale=8.0):
        self.lib.oracle_seed_random(seed)
        pos = np.zeros((n, 4), np.float32)
        vel = np.zeros((n, 4), np.float32)
        self.lib.oracle_random_unit_bodies(_ptr(pos), _ptr(vel), n, scale, velscale)
        return pos, vel

    # -- FP64 -----------------------------------------------------------------
    def fp64_forces(self, pos_mass, eps2, idx=None):
        pm = np.ascontiguousarray(pos_mass, np.float32)
        n = pm.shape[0]
        if idx is None:
            out = np.zeros((n, 3), np.float64)
            self.lib.oracle_fp64_forces(_ptr(out), _ptr(pm), n, float(eps2), None, 0)
        else:
            ix = np.ascontiguousarray(idx, np.int64)
            out = np.zeros((ix.size, 3), np.float64)
            self.lib.oracle_fp64_forces(_ptr(out), _ptr(pm), n, float(eps2), _ptr(ix), ix.size)
        return out

    def fp64_energy(self, pos_mass, vel_inv_mass, eps2):
        pm = np.ascontiguousarray(pos_mass, np.float32)
        vm = np.ascontiguousarray(vel_inv_mass, np.float32)
        ke = ctypes.c_double()
        pe = ctypes.c_double()
        mom = np.zeros(3, np.float64)
        self.lib.oracle_fp64_energy(ctypes.byref(ke), ctypes.byref(pe), _ptr(mom), _ptr(pm), _ptr(vm),
                                    pm.shape[0], float(eps2))
        return ke.value, pe.value, mom


class RefCPU(_Lib):
    """The reference's own CPU path compiled from its sources (oracle/Makefile `ref`)."""

    def __init__(self, path=REF_SO):
        super().__init__(path)
        L = self.lib
        for name in ("ComputeGravitation_SOA", "ComputeGravitation_SOA_tiled", "ComputeGravitation_SIMD"):
            fn = getattr(L, name)
            fn.restype = _f
            fn.argtypes = [_p, _p, _p, _f, _sz]
        for name in ("ComputeGravitation_AOS", "ComputeGravitation_AOS_tiled"):
            fn = getattr(L, name)
            fn.restype = _f
            fn.argtypes = [_p, _p, _f, _sz]
        L.seedRandom.restype = None
        L.seedRandom.argtypes = [ctypes.c_uint]
        L.randomUnitBodies.restype = None
        L.randomUnitBodies.argtypes = [_p, _p, _sz, _f, _f]

    _soa_call = Oracle._soa_call
    _aos_call = Oracle._aos_call

    def cpu_soa(self, pos_mass, eps2):
        return self._soa_call(self.lib.ComputeGravitation_SOA, pos_mass, eps2)

    def cpu_soa_tiled(self, pos_mass, eps2):
        return self._soa_call(self.lib.ComputeGravitation_SOA_tiled, pos_mass, eps2)

    def cpu_avx(self, pos_mass, eps2):
        return self._soa_call(self.lib.ComputeGravitation_SIMD, pos_mass, eps2)

    def cpu_aos(self, pos_mass, eps2):
        return self._aos_call(self.lib.ComputeGravitation_AOS, pos_mass, eps2)

    def cpu_aos_tiled(self, pos_mass, eps2):
        return self._aos_call(self.lib.ComputeGravitation_AOS_tiled, pos_mass, eps2)

    def random_unit_bodies(self, n, seed=7, scale=0.9, velscale=8.0):
        self.lib.seedRandom(seed)
        pos = aligned_empty(4 * n)
        vel = aligned_empty(4 * n)
        self.lib.randomUnitBodies(_ptr(pos), _ptr(vel), n, scale, velscale)
        return pos.reshape(n, 4).copy(), vel.reshape(n, 4).copy()


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        _oracle = Oracle()
    return _oracle


def ref_cpu():
    """The compiled reference, or None where it was not built (e.g. on a host
    without the reference sources)."""
    if not os.path.exists(REF_SO):
        return None
    return RefCPU()


def force_errors(force, f64):
    """Per-body normwise relative error |F - F64| / |F64| of an (n,>=3) FP32
    force against FP64 forces (n,3): returns (max, rms)."""
    f = np.asarray(force, np.float64)[:, :3]
    d = np.linalg.norm(f - f64, axis=1)
    ref = np.linalg.norm(f64, axis=1)
    rel = d / ref
    return float(rel.max()), float(np.sqrt(np.mean(rel * rel)))
