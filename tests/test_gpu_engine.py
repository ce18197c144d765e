"""Device-resident stepping (include/nbody_engine.h) on the GPU.

- graph replays against eager launches, bit for bit, with eager work and
  synchronisations between the replays;
- the reference's per-step crosscheck (src/nbody.cu:313-325,406-418): at every
  step the GPU force on the GPU's current positions against the oracle's
  CPU_AOS on the same positions, and the GPU integrator against the oracle's
  integrator given the same force (bit-exact);
- ensembles: batched (graph and plain) against one system at a time, each
  system alone against the same system in a mix (bit-exact), forces against
  the FP64 oracle.
"""
import ctypes

import numpy as np
import pytest

from conftest import EPS2
from oracle_lib import force_errors

from nbody_amd import _native, ics
from nbody_amd.engine import Ensemble, System

pytestmark = pytest.mark.gpu
L = None


def lib():
    global L
    if L is None:
        L = _native.lib()
    return L


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("opt", [{}, {"variant": _native.LDS}, {"variant": _native.SGPR, "i_per_lane": 2}])
def test_graph_replays_match_eager(opt):
    pm, vm = ics.uniform_ball(6000, seed=3)
    a = System(pm, vm, 0.01, **opt)
    b = System(pm, vm, 0.01, **opt)
    try:
        for k in range(4):
            a.run(2, graph=True)         # replay 2 steps from the graph
            b.run(2, graph=False)        # the same two steps launched one by one
            sa, sb = a.read(), b.read()  # host synchronisation between replays
            for x, y in zip(sa, sb):
                assert np.array_equal(bits(x), bits(y)), (opt, k)
            _ = np.linalg.norm(sa[0])    # eager host work in between
    finally:
        a.close()
        b.close()


def test_step_crosscheck_every_step(oracle):
    """Ten steps of the N = 16384 uniform-ball workload, crosschecked at every
    step on identical positions against the oracle's CPU_AOS (the reference's
    crosscheck target) and CPU_SOA (the reference driver's golden,
    src/nbody.cu:313-325), and the integrator bit-exact against the oracle's."""
    n = 16384
    pm0, vm0 = ics.uniform_ball(n, seed=1)
    s = System(pm0, vm0, 0.01)
    prev_pm, prev_vm = pm0, vm0
    try:
        for step in range(10):
            s.run(1, graph=True)
            pm, vm, f = s.read()
            aos, _ = oracle.cpu_aos(prev_pm, np.float32(0.01))
            soa, _ = oracle.cpu_soa(prev_pm, np.float32(0.01))
            f64 = oracle.fp64_forces(prev_pm, np.float32(0.01))
            norm = np.linalg.norm(f64, axis=1)
            d = np.linalg.norm(f[:, :3].astype(np.float64) - aos[:, :3], axis=1) / norm
            d_soa = np.linalg.norm(f[:, :3].astype(np.float64) - soa[:, :3], axis=1) / norm
            mx, rms = force_errors(f, f64)
            assert d.max() <= 2e-5 and d_soa.max() <= 2e-5 and mx <= 1e-5 and rms <= 1e-6, (
                step, d.max(), d_soa.max(), mx, rms)
            p_ref, v_ref = oracle.integrate_aos(prev_pm, prev_vm, f, 0.016, 0.95)
            assert np.array_equal(bits(pm), bits(p_ref)) and np.array_equal(bits(vm), bits(v_ref)), step
            prev_pm, prev_vm = pm, vm
    finally:
        s.close()


def test_energy_and_momentum_conserved_without_damping():
    """Physics of the whole device-resident step (forces + integrator): a
    virialised Plummer sphere (N = 4096, eps^2 = 1e-2) stepped 200 times with
    dt = 2e-3 and no damping keeps its total energy to 1e-4 (an FP64
    restatement of the same symplectic-Euler steps drifts by 3e-6 here) and
    its total momentum to 1e-6 of sum m|v|, both measured with the FP64
    energy diagnostic."""
    from gpu_util import dev, ptr
    n, eps2 = 4096, 1e-2
    pm, vm = ics.plummer(n, seed=5)

    def energy(p, v):
        out = np.zeros(5, np.float64)
        d_p, d_v = dev(p), dev(v)  # both kept alive until the call returns
        _native.check(lib().nbody_hip_energy_fp64(out.ctypes.data, ptr(d_p), ptr(d_v), eps2, n), "energy")
        return out[0] + out[1], out[2:]

    e0, p0 = energy(pm, vm)
    s = System(pm, vm, eps2, dt=2e-3, damping=1.0)
    try:
        s.run(200, graph=True)
        pm1, vm1, _ = s.read()
    finally:
        s.close()
    e1, p1 = energy(pm1, vm1)
    scale = float((pm[:, 3].astype(np.float64) * np.linalg.norm(vm[:, :3].astype(np.float64), axis=1)).sum())
    assert abs(e1 - e0) / abs(e0) < 1e-4, (e0, e1)
    assert np.abs(p1 - p0).max() < 1e-6 * scale, (p0, p1, scale)


def make_ensemble_systems(count, lo, hi, seed):
    sizes = ics.log_uniform_sizes(count, lo, hi, seed)
    return [ics.uniform_ball(int(n), seed=seed * 1000 + k) for k, n in enumerate(sizes)]


def test_ensemble_modes_and_forces(oracle):
    systems = make_ensemble_systems(24, 17, 3000, seed=5)
    e_graph = Ensemble(systems, EPS2)
    e_plain = Ensemble(systems, EPS2)
    e_each = Ensemble(systems, EPS2)
    try:
        e_graph.run(1, _native.BATCHED_GRAPH)
        e_plain.run(1, _native.BATCHED)
        e_each.run(1, _native.PER_SYSTEM)
        g, p, q = e_graph.read(), e_plain.read(), e_each.read()
        for x, y in zip(g, p):
            assert np.array_equal(bits(x), bits(y))
        # forces of step 1 are the forces of the initial positions
        for (pm, _), fg, fq in zip(systems, e_graph.split(g[2]), e_each.split(q[2])):
            f64 = oracle.fp64_forces(pm, EPS2)
            if pm.shape[0] == 1:
                continue
            for f in (fg, fq):
                mx, rms = force_errors(f, f64)
                assert mx <= 1e-5 and rms <= 1e-6, (pm.shape[0], mx, rms)
        for k in range(3):  # several graph replays with host work between
            e_graph.run(2, _native.BATCHED_GRAPH)
            e_plain.run(2, _native.BATCHED)
            for x, y in zip(e_graph.read(), e_plain.read()):
                assert np.array_equal(bits(x), bits(y)), k
    finally:
        for e in (e_graph, e_plain, e_each):
            e.close()


def test_ensemble_alone_equals_in_mix():
    systems = make_ensemble_systems(40, 256, 5000, seed=6)
    mix = Ensemble(systems, EPS2)
    try:
        mix.run(3, _native.BATCHED_GRAPH)
        got = [mix.split(a) for a in mix.read()]
        for k in (0, 7, 39):
            alone = Ensemble([systems[k]], EPS2)
            try:
                alone.run(3, _native.BATCHED_GRAPH)
                for a, m in zip(alone.read(), got):
                    assert np.array_equal(bits(a), bits(m[k])), k
            finally:
                alone.close()
    finally:
        mix.close()


def test_ensemble_ragged_sizes_and_masses(oracle):
    """Batched systems of up to 2^18 bodies run as 512-body block pairs plus
    one task per diagonal block that takes each of its pairs once. For ragged
    sizes, a single block, a single body, unit masses (the equal-mass weights)
    and the reference's unequal masses U(0.5, 1): the forces of the first step
    against the FP64 oracle, and the positions after it against a System with
    the symmetric plan (8 bodies per lane, blocks of 512; its diagonal is
    summed one-sidedly, so the two agree to rounding, not bit for bit)."""
    sizes = [1, 63, 512, 513, 3000, 20000]
    systems = [ics.uniform_ball(n, seed=40 + k) for k, n in enumerate(sizes)]
    rng = np.random.default_rng(41)
    for pm, vm in systems[3:]:  # unequal masses U(0.5, 1), as the reference's own ICs
        pm[:, 3] = rng.uniform(0.5, 1.0, pm.shape[0]).astype(np.float32)
        vm[:, 3] = 1.0 / pm[:, 3]
    e = Ensemble(systems, EPS2)
    try:
        e.run(1, _native.BATCHED_GRAPH)
        got = [e.split(a) for a in e.read()]
    finally:
        e.close()
    for k, (pm, vm) in enumerate(systems):
        f = got[2][k]
        assert np.all(np.isfinite(f[:, :3])), sizes[k]
        if sizes[k] == 1:
            assert np.all(f[:, :3] == 0.0)
        else:
            mx, rms = force_errors(f, oracle.fp64_forces(pm, EPS2))
            assert mx <= 1e-5 and rms <= 1e-6, (sizes[k], mx, rms)
        s = System(pm, vm, EPS2, variant=_native.SYMMETRIC, i_per_lane=8, jslice=512)
        try:
            s.run(1, graph=False)
            pos = s.read()[0]
            assert np.allclose(pos[:, :3], got[0][k][:, :3], rtol=1e-6, atol=1e-6), sizes[k]
            assert np.array_equal(pos[:, 3], got[0][k][:, 3]), sizes[k]
        finally:
            s.close()


def test_ensemble_large_system_rows():
    """A system above 2^18 bodies is batched as one-sided rows with TwoSum-
    folded chunks: its forces against the FP64 direct sum (on the GPU, itself
    checked against the oracle), next to a small system whose forces stay
    bit-equal to its forces when batched alone."""
    import torch
    from gpu_util import dev, host, ptr
    big = ics.uniform_ball(2 ** 18 + 1000, seed=71)
    small = ics.uniform_ball(700, seed=72)
    e = Ensemble([big, small], EPS2)
    try:
        e.run(1, _native.BATCHED)
        f_big, f_small = e.split(e.read()[2])
    finally:
        e.close()
    n = big[0].shape[0]
    out = dev(np.zeros((n, 4), np.float64))
    d_pm = dev(big[0])
    assert _native.lib().nbody_hip_force_fp64(ptr(out), ptr(d_pm), float(EPS2), n) > 0
    f64 = host(out)[:, :3]
    del d_pm, out
    torch.cuda.empty_cache()
    mx, rms = force_errors(f_big, f64)
    assert mx <= 1e-5 and rms <= 1e-6, (mx, rms)
    alone = Ensemble([small], EPS2)
    try:
        alone.run(1, _native.BATCHED)
        assert np.array_equal(bits(alone.read()[2]), bits(f_small))
    finally:
        alone.close()


def test_ensemble_rejects_bad_input():
    n = np.array([0], np.uint64)
    pm = np.zeros((1, 4), np.float32)
    e2 = np.zeros(1, np.float32)
    h = lib().nbody_hip_ensemble_create(1, n.ctypes.data, pm.ctypes.data, pm.ctypes.data, e2.ctypes.data, 0.016,
                                        0.95)
    assert not h and "N = 0" in _native.last_error()
