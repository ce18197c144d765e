"""nbody_hip_force_noalloc: the force entry that allocates nothing, its
partial sums in a fixed device array of NBODY_HIP_NOALLOC_SCRATCH_BYTES.

- With the plan fixed by the caller, two calls in a row (the second on a
  smaller system, so the array still holds the first call's partial sums)
  give nbody_hip_force's forces with the same plan bit for bit.
- With the library's choice, the N = 98304 workload (whose default symmetric
  plan keeps 226 MB of block slots, so it runs in column bands that fit the
  array) and then a small system, each within the path's tolerance of the
  FP64 oracle.
- The i-range form matches the whole evaluation, and a fixed shape whose
  partial sums exceed the array is refused before anything runs.
- Captured into a HIP graph (a caller that cannot allocate) and replayed
  with eager work between the replays: every replay equals the eager call.
"""
import ctypes

import numpy as np
import pytest

from conftest import EPS2
from gpu_util import dev, force_buffer, host, ptr, run_force
from oracle_lib import force_errors

from nbody_amd import _native, ics

pytestmark = pytest.mark.gpu

TOL_MAX, TOL_RMS = 1e-5, 1e-6


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def noalloc(pm_dev, eps2, base=0, count=None, stream=None, **opt):
    n = pm_dev.shape[0]
    count = n - base if count is None else count
    f = force_buffer(max(count, 1))
    o = _native.Options.make(**opt) if opt else None
    rc = _native.lib().nbody_hip_force_noalloc(ptr(f), ptr(pm_dev), eps2, n, base, count,
                                               ctypes.byref(o) if o else None, stream)
    _native.check(rc, "nbody_hip_force_noalloc")
    return host(f)[:count]


PLANS = [
    dict(variant=_native.SYMMETRIC, i_per_lane=8, jslice=512),
    dict(variant=_native.SYMMETRIC, i_per_lane=4, jslice=256),
    dict(variant=_native.LDS, i_per_lane=2, jslice=1024, compensated=1),
    dict(variant=_native.SGPR, i_per_lane=1, jslice=2048, compensated=0),
    dict(variant=_native.DPP, i_per_lane=4, jslice=4096, compensated=1),
]


@pytest.mark.parametrize("opt", PLANS, ids=lambda o: f"{_native.VARIANT_NAMES[o['variant']]}-{o['jslice']}")
def test_matches_existing_entry_twice(opt):
    for n, seed in ((20000, 3), (7001, 4)):
        pm, _ = ics.random_unit_bodies(n, seed=seed)
        d = dev(pm)
        a = noalloc(d, EPS2, **opt)
        b = run_force(d, EPS2, **opt)
        assert np.array_equal(bits(a), bits(b)), (opt, n)
        assert np.all(a[:, 3] == -7.0)


def test_library_choice_within_tolerance(oracle):
    for n, seed in ((98304, 1), (5000, 2)):
        pm, _ = ics.uniform_ball(n, seed=seed)
        f = noalloc(dev(pm), 0.01)
        r = np.linalg.norm(pm[:, :3], axis=1)
        rng = np.random.default_rng(n)
        idx = np.unique(np.concatenate([np.argsort(r)[:256], rng.choice(n, min(n, 1792), replace=False)]))
        mx, rms = force_errors(f[idx], oracle.fp64_forces(pm, 0.01, idx))
        assert mx <= TOL_MAX and rms <= TOL_RMS, (n, mx, rms)
        assert np.all(f[:, 3] == -7.0)


def test_i_range_and_refusal():
    pm, _ = ics.random_unit_bodies(9000, seed=5)
    d = dev(pm)
    opt = dict(variant=_native.LDS, i_per_lane=2, jslice=1024, compensated=1)
    full = noalloc(d, EPS2, **opt)
    part = noalloc(d, EPS2, base=2048, count=3000, **opt)
    assert np.array_equal(bits(part), bits(full[2048:5048]))
    # LDS with fixed 256-body compensated slices at N = 30000: 118 slices of
    # 32 bytes per body, 113 MB > 64 MiB
    n = 30000
    big = dev(np.zeros((n, 4), np.float32))
    f = force_buffer(n)
    o = _native.Options.make(variant=_native.LDS, i_per_lane=1, jslice=256, compensated=1)
    rc = _native.lib().nbody_hip_force_noalloc(ptr(f), ptr(big), EPS2, n, 0, n, ctypes.byref(o), None)
    assert rc != 0 and "fixed scratch array" in _native.last_error()
    assert np.all(np.isnan(host(f)[:, :3]))


def test_graph_capture_replays():
    import torch
    n = 12000
    pm, _ = ics.uniform_ball(n, seed=6)
    d = dev(pm)
    opt = _native.Options.make(variant=_native.SYMMETRIC, i_per_lane=8, jslice=512)
    L = _native.lib()
    eager = noalloc(d, 0.01, variant=_native.SYMMETRIC, i_per_lane=8, jslice=512)
    out = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            rc = L.nbody_hip_force_noalloc(ptr(out), ptr(d), 0.01, n, 0, n, ctypes.byref(opt),
                                           ctypes.c_void_p(s.cuda_stream))
    torch.cuda.current_stream().wait_stream(s)
    _native.check(rc, "captured nbody_hip_force_noalloc")
    for k in range(4):
        out.fill_(float("nan"))
        g.replay()
        got = host(out)
        assert np.array_equal(bits(got[:, :3]), bits(eager[:, :3])), k
        # eager work on the same array between the replays
        noalloc(d[: 3000 + 1000 * k].contiguous(), 0.01, variant=_native.SYMMETRIC, i_per_lane=8, jslice=512)
