"""Python driver of the device-resident engine (include/nbody_engine.h): one
system stepped in HBM, or an ensemble of independent systems stepped by the
batched launches. Arrays are numpy float32 (N, 4): posMass (x, y, z, m) and
velInvMass (vx, vy, vz, 1/m)."""
import ctypes

import numpy as np

from . import _native

BATCHED_GRAPH, BATCHED, PER_SYSTEM = _native.BATCHED_GRAPH, _native.BATCHED, _native.PER_SYSTEM


def _options(opt):
    return ctypes.byref(_native.Options.make(**opt)) if opt else None


class System:
    """nbody_hip_system_*: `opt` takes the nbody_hip_options fields
    (variant, i_per_lane, jslice, compensated, newton, waves_per_simd,
    slot_bytes); none given is the library's plan for N."""

    def __init__(self, pos_mass, vel_inv_mass, softening_squared, dt=0.016, damping=0.95, **opt):
        pm = np.ascontiguousarray(pos_mass, np.float32)
        vm = np.ascontiguousarray(vel_inv_mass, np.float32)
        self.n = pm.shape[0]
        self.h = _native.lib().nbody_hip_system_create(pm.ctypes.data, vm.ctypes.data, self.n, softening_squared,
                                                       dt, damping, _options(opt))
        if not self.h:
            raise _native.NativeError(_native.last_error())

    def run(self, steps, graph=True):
        """Runs `steps` steps (graph: one captured graph replay per step) and
        returns their milliseconds."""
        ms = ctypes.c_float()
        _native.check(_native.lib().nbody_hip_system_run(self.h, steps, 1 if graph else 0, ctypes.byref(ms)), "run")
        return ms.value

    def read(self):
        """(posMass, velInvMass, force of the last step)."""
        pm = np.zeros((self.n, 4), np.float32)
        vm = np.zeros((self.n, 4), np.float32)
        f = np.zeros((self.n, 4), np.float32)
        _native.check(_native.lib().nbody_hip_system_read(self.h, pm.ctypes.data, vm.ctypes.data, f.ctypes.data),
                      "read")
        return pm, vm, f

    def plan(self):
        out = _native.Options()
        _native.check(_native.lib().nbody_hip_system_plan(self.h, ctypes.byref(out)), "plan")
        return out

    def close(self):
        if self.h:
            _native.lib().nbody_hip_system_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Ensemble:
    """nbody_hip_ensemble_*: `systems` is a list of (posMass, velInvMass)
    pairs; `softening_squared` one value or one per system. `opt` as for
    System (compensated=2: FP64 accumulation in every mode)."""

    def __init__(self, systems, softening_squared, dt=0.016, damping=0.95, **opt):
        self.sizes = np.array([s[0].shape[0] for s in systems], np.uint64)
        pm = np.ascontiguousarray(np.concatenate([s[0] for s in systems]), np.float32)
        vm = np.ascontiguousarray(np.concatenate([s[1] for s in systems]), np.float32)
        e2 = np.ascontiguousarray(np.broadcast_to(np.asarray(softening_squared, np.float32), (len(systems),)))
        self.total = pm.shape[0]
        self.h = _native.lib().nbody_hip_ensemble_create_opt(len(systems), self.sizes.ctypes.data, pm.ctypes.data,
                                                             vm.ctypes.data, e2.ctypes.data, dt, damping,
                                                             _options(opt))
        if not self.h:
            raise _native.NativeError(_native.last_error())

    def run(self, steps, mode=BATCHED_GRAPH):
        """Runs `steps` steps of every system in `mode` and returns their
        milliseconds."""
        ms = ctypes.c_float()
        _native.check(_native.lib().nbody_hip_ensemble_run(self.h, steps, mode, ctypes.byref(ms)), "ensemble run")
        return ms.value

    def read(self):
        """(posMass, velInvMass, force) of all bodies, systems back to back."""
        pm = np.zeros((self.total, 4), np.float32)
        vm = np.zeros((self.total, 4), np.float32)
        f = np.zeros((self.total, 4), np.float32)
        _native.check(_native.lib().nbody_hip_ensemble_read(self.h, pm.ctypes.data, vm.ctypes.data, f.ctypes.data),
                      "read")
        return pm, vm, f

    def split(self, a):
        """An array over all bodies cut into one array per system."""
        return np.split(a, np.cumsum(self.sizes.astype(np.int64))[:-1])

    def close(self):
        if self.h:
            _native.lib().nbody_hip_ensemble_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
