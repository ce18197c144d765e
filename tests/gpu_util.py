"""Device-memory plumbing for the GPU tests: torch tensors hold the buffers,
the C ABI of libnbody_hip.so does the work."""
import ctypes

import numpy as np
import torch

from nbody_amd import _native


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def force_buffer(n, w=-7.0):
    """A device force buffer whose x,y,z are NaN (so unwritten entries show) and
    whose w holds a sentinel that the kernels must leave alone."""
    f = np.full((n, 4), np.nan, np.float32)
    f[:, 3] = w
    return dev(f)


def run_entry(symbol, pm_dev, eps2):
    n = pm_dev.shape[0]
    f = force_buffer(n)
    ms = getattr(_native.lib(), symbol)(ptr(f), ptr(pm_dev), eps2, n)
    _native.check_ms(ms, symbol)
    return host(f), ms


def run_force(pm_dev, eps2, base=0, count=None, **opt):
    """nbody_hip_force with options; returns the (count, 4) force array."""
    n = pm_dev.shape[0]
    count = n - base if count is None else count
    f = force_buffer(max(count, 1))
    o = _native.Options.make(**opt)
    rc = _native.lib().nbody_hip_force(ptr(f), ptr(pm_dev), eps2, n, base, count, ctypes.byref(o), None)
    _native.check(rc, "nbody_hip_force")
    return host(f)[:count]
