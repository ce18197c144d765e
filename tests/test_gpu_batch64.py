"""FP64 accumulation in the batched ensemble and in the server
(compensated = 2, nbody_hip_ensemble_create_opt / nbody_hip_server_create_opt).

- The batched FP64 pieces give, bit for bit, the forces of the single-system
  FP64 symmetric plan (nbody_hip_force with compensated = 2: 4 bodies per
  lane, blocks of 256) for systems of 1024 .. 2^17 bodies, ragged last blocks
  included, and the per-system mode gives the same bits.
- Small systems (fewer bodies than a block) and a system above 2^17 (one-sided
  FP64 rows) against the FP64 oracle.
- Graph against eager, alone against the mix, the server against the
  ensemble: bit for bit.
- The 512-system ensemble (N log-uniform in [256, 16384], seed 4) after 55
  steps: every body within the bound, with FP32 block partials and with FP64
  accumulation.

Tolerances: per-body |F - F64| / |F64| max <= 1e-5 and RMS <= 1e-6.
"""
import ctypes

import numpy as np
import pytest

from gpu_util import dev, host, ptr, run_force
from nbody_amd import _native, ics, serve

pytestmark = pytest.mark.gpu

TOL_MAX, TOL_RMS = 1e-5, 1e-6
EPS2 = 0.01
DT, DAMPING = 0.016, 0.95


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def create(systems, compensated=2):
    L = _native.lib()
    sizes = np.array([s[0].shape[0] for s in systems], np.uint64)
    pm = np.ascontiguousarray(np.concatenate([s[0] for s in systems]))
    vm = np.ascontiguousarray(np.concatenate([s[1] for s in systems]))
    e2 = np.full(len(systems), EPS2, np.float32)
    opt = _native.Options.make(compensated=compensated)
    h = L.nbody_hip_ensemble_create_opt(len(systems), sizes.ctypes.data, pm.ctypes.data, vm.ctypes.data,
                                        e2.ctypes.data, DT, DAMPING, ctypes.byref(opt))
    assert h, _native.last_error()
    return h, sizes


def run(h, sizes, steps, mode):
    """Steps the ensemble, then reads (pm, vm, force) per system and destroys it."""
    L = _native.lib()
    total = int(sizes.sum())
    pm = np.zeros((total, 4), np.float32)
    vm = np.zeros((total, 4), np.float32)
    f = np.zeros((total, 4), np.float32)
    try:
        _native.check(L.nbody_hip_ensemble_run(h, steps, mode, None), "ensemble run")
        _native.check(L.nbody_hip_ensemble_read(h, pm.ctypes.data, vm.ctypes.data, f.ctypes.data), "read")
    finally:
        L.nbody_hip_ensemble_destroy(h)
    cut = np.cumsum(sizes.astype(np.int64))[:-1]
    return np.split(pm, cut), np.split(vm, cut), np.split(f, cut)


def balls(sizes, seed0):
    return [ics.uniform_ball(int(n), seed=seed0 + k) for k, n in enumerate(sizes)]


def test_batched64_equals_fp64_symmetric_plan():
    sizes = [1024, 1500, 4096, 9001, 70000]
    systems = balls(sizes, 300)
    h, sz = create(systems)
    _, _, f = run(h, sz, 1, _native.BATCHED)
    for k, (pm, _) in enumerate(systems):
        single = run_force(dev(pm), EPS2, variant=_native.SYMMETRIC, compensated=2)
        assert np.array_equal(bits(f[k][:, :3]), bits(single[:, :3])), sizes[k]


@pytest.mark.parametrize("shape", [(4, 256), (8, 512)])
def test_fp64_symmetric_equal_mass_blocks(oracle, shape):
    """Block pairs whose bodies share one mass take the mass-free FP64 terms
    (inv^3 converted once, no mass products, the masses do not rotate) and
    multiply the partials by the mass in FP64 at the store. One mass
    everywhere (not a power of two), the same with one heavier body (pairs
    with its block take the general terms), a ragged N (the last block holds
    zero-mass padding, so its pairs are general) against the oracle; a common
    mass of 1/4 scales the forces of unit masses exactly. (Unit-mass systems
    in the batched FP64 mode give the single-system plan's bits:
    test_batched64_equals_fp64_symmetric_plan.)"""
    a, b = shape
    for n in (4 * b, 4 * b + 37):
        pm, _ = ics.uniform_ball(n, seed=21)
        pm[:, 3] = 0.37
        mixed = pm.copy()
        mixed[2 * b + 5, 3] = 0.6
        for case, x in (("one mass", pm), ("one heavier body", mixed)):
            f = run_force(dev(x), EPS2, variant=_native.SYMMETRIC, i_per_lane=a, jslice=b, compensated=2)
            f64 = oracle.fp64_forces(x, EPS2)
            rel = np.linalg.norm(f[:, :3].astype(np.float64) - f64, axis=1) / np.linalg.norm(f64, axis=1)
            assert rel.max() <= TOL_MAX and np.sqrt(np.mean(rel ** 2)) <= TOL_RMS, (shape, n, case, rel.max())
            assert np.all(f[:, 3] == -7.0)
        unit = pm.copy()
        unit[:, 3] = 1.0
        quarter = pm.copy()
        quarter[:, 3] = 0.25
        f1 = run_force(dev(unit), EPS2, variant=_native.SYMMETRIC, i_per_lane=a, jslice=b, compensated=2)
        fq = run_force(dev(quarter), EPS2, variant=_native.SYMMETRIC, i_per_lane=a, jslice=b, compensated=2)
        assert np.array_equal(fq[:, :3], np.float32(0.25) * f1[:, :3]), (shape, n)


def test_batched64_small_and_row_systems_against_oracle(oracle):
    # fewer bodies than one block, a ragged pair of blocks, and one system
    # above 2^17 bodies, which runs as one-sided FP64 rows
    sizes = [1, 2, 17, 255, 300, 1023, 2 ** 17 + 100]
    systems = balls(sizes, 400)
    h, sz = create(systems)
    _, _, f = run(h, sz, 1, _native.BATCHED_GRAPH)
    assert np.array_equal(f[0][:, :3], np.zeros((1, 3), np.float32))  # a lone body feels nothing
    rng = np.random.default_rng(5)
    for k, (pm, _) in enumerate(systems[1:], start=1):
        n = pm.shape[0]
        r = np.linalg.norm(pm[:, :3], axis=1)
        idx = np.unique(np.concatenate([[np.argmin(r), np.argmax(r)], rng.choice(n, min(n, 254), replace=False)]))
        f64 = oracle.fp64_forces(pm, EPS2, idx)
        rel = np.linalg.norm(f[k][idx, :3].astype(np.float64) - f64, axis=1) / np.linalg.norm(f64, axis=1)
        assert rel.max() <= TOL_MAX and np.sqrt(np.mean(rel ** 2)) <= TOL_RMS, (n, rel.max())


def test_batched64_graph_eager_alone_and_per_system():
    sizes = [300, 2000, 5000]
    systems = balls(sizes, 500)
    g = run(*create(systems), 3, _native.BATCHED_GRAPH)
    e = run(*create(systems), 3, _native.BATCHED)
    for a, b in zip(g, e):
        for k in range(len(sizes)):
            assert np.array_equal(bits(a[k]), bits(b[k])), sizes[k]
    alone = run(*create([systems[1]]), 3, _native.BATCHED_GRAPH)
    for a, m in zip(alone, g):
        assert np.array_equal(bits(a[0]), bits(m[1]))
    # the per-system mode plans each system with the same options: the FP64
    # symmetric plan for N >= 1024 (the same bits), FP64 LDS below
    p = run(*create(systems), 1, _native.PER_SYSTEM)
    b = run(*create(systems), 1, _native.BATCHED)
    for k in (1, 2):
        assert np.array_equal(bits(p[2][k]), bits(b[2][k])), sizes[k]
    rel = np.linalg.norm(p[2][0][:, :3].astype(np.float64) - b[2][0][:, :3], axis=1) / np.linalg.norm(
        b[2][0][:, :3].astype(np.float64), axis=1)
    assert rel.max() <= 1e-6


def test_served64_equals_ensemble64():
    reqs = (_native.Request * 6)()
    spec = [(300, 0), (2000, 1), (4097, 0), (9000, 1), (513, 1), (20000, 0)]
    for k, (n, ic) in enumerate(spec):
        reqs[k] = _native.Request(k, n, 3, 700 + k, ic, EPS2, 0.0)
    srv = serve.Server(max_active=4, keep_states=True, compensated=2)
    try:
        res, _, _ = srv.run(reqs)
        assert all(r.status == 0 for r in res)
        assert srv.peak_active() >= 2
        final = [srv.final_state(k, spec[k][0]) for k in range(len(spec))]
    finally:
        srv.close()
    systems = [ics.plummer(n, seed=700 + k) if ic == _native.IC_PLUMMER else ics.uniform_ball(n, seed=700 + k)
               for k, (n, ic) in enumerate(spec)]
    pm, vm, _ = run(*create(systems), 3, _native.BATCHED_GRAPH)
    for k in range(len(spec)):
        assert np.array_equal(bits(final[k][0]), bits(pm[k])), spec[k]
        assert np.array_equal(bits(final[k][1]), bits(vm[k])), spec[k]
    n = 1000
    nb = -(-n // 256)
    assert serve.request_bytes(n, compensated=2) == 48 * n + 24 * nb * nb * 256


def fp64_errors(systems_pm, forces):
    """Per-body |F - F64|/|F64| over every body, F64 from the library's GPU
    FP64 diagnostic (pinned to the oracle by the force tests)."""
    L = _native.lib()
    out = []
    for pm, f in zip(systems_pm, forces):
        n = pm.shape[0]
        f64 = dev(np.zeros((n, 4), np.float64))
        _native.check_ms(L.nbody_hip_force_fp64(ptr(f64), ptr(dev(pm)), EPS2, n), "fp64")
        ref = host(f64)[:, :3]
        out.append(np.linalg.norm(f[:, :3].astype(np.float64) - ref, axis=1) / np.linalg.norm(ref, axis=1))
    return np.concatenate(out)


@pytest.mark.slow
def test_c4_evolved_state_holds_the_bound():
    """The 512-system ensemble stepped 55 times by the FP32 batched path (the
    state the ensemble workload ends in): the batched force on it, with FP32
    block partials and with FP64 accumulation, against the FP64 direct sum
    over every body."""
    sizes = ics.log_uniform_sizes(512, 256, 16384, seed=4)
    systems = balls(sizes, 100000)
    pm55, vm55, _ = run(*create(systems, compensated=-1), 55, _native.BATCHED_GRAPH)
    evolved = list(zip(pm55, vm55))
    _, _, f64acc = run(*create(evolved, compensated=2), 1, _native.BATCHED)
    _, _, f32 = run(*create(evolved, compensated=-1), 1, _native.BATCHED)
    rel = fp64_errors(pm55, f64acc)
    rel32 = fp64_errors(pm55, f32)
    mx, rms = float(rel.max()), float(np.sqrt(np.mean(rel ** 2)))
    mx32, rms32 = float(rel32.max()), float(np.sqrt(np.mean(rel32 ** 2)))
    print(f"512 systems after 55 steps: FP64 accumulation max {mx:.3e} rms {rms:.3e}; FP32 batched max "
          f"{mx32:.3e} rms {rms32:.3e}")
    assert mx <= TOL_MAX and rms <= TOL_RMS, (mx, rms)
    assert mx32 <= TOL_MAX and rms32 <= TOL_RMS, (mx32, rms32)


def test_options_the_fp64_batches_refuse():
    # FP64 accumulation is built for the LDS and symmetric plans: an ensemble
    # whose per-system plans would need another variant is refused at creation
    L = _native.lib()
    pm, vm = ics.uniform_ball(2000, seed=9)
    n = np.array([2000], np.uint64)
    e2 = np.array([EPS2], np.float32)
    opt = _native.Options.make(variant=_native.SGPR, compensated=2)
    h = L.nbody_hip_ensemble_create_opt(1, n.ctypes.data, pm.ctypes.data, vm.ctypes.data, e2.ctypes.data, DT,
                                        DAMPING, ctypes.byref(opt))
    assert not h and "FP64 accumulation" in _native.last_error()
    # NULL options behave as nbody_hip_ensemble_create
    h = L.nbody_hip_ensemble_create_opt(1, n.ctypes.data, pm.ctypes.data, vm.ctypes.data, e2.ctypes.data, DT,
                                        DAMPING, None)
    assert h, _native.last_error()
    L.nbody_hip_ensemble_destroy(h)


@pytest.mark.slow
@pytest.mark.parametrize("compensated", [-1, 2])
def test_batched_size_boundaries_unequal_masses(oracle, compensated):
    """Every size boundary of the batched layouts, with the reference's own
    ICs (randomUnitBodies: masses in U(0.5, 1), so the general pair weights):
    block edges of both block sizes, and the thresholds above which a system
    runs as one-sided rows (2^17 with FP64 accumulation, 2^18 without)."""
    sizes = [1, 2, 63, 64, 65, 255, 256, 257, 511, 512, 513, 1023, 1024, 1025,
             2 ** 17 - 1, 2 ** 17, 2 ** 17 + 1, 2 ** 18 - 1, 2 ** 18, 2 ** 18 + 1]
    systems = [ics.random_unit_bodies(n, seed=20 + k) for k, n in enumerate(sizes)]
    h, sz = create(systems, compensated=compensated)
    _, _, f = run(h, sz, 1, _native.BATCHED_GRAPH)
    assert np.array_equal(f[0][:, :3], np.zeros((1, 3), np.float32))
    rng = np.random.default_rng(11)
    worst = 0.0
    for k, (pm, _) in enumerate(systems[1:], start=1):
        n = pm.shape[0]
        # both ends of the system (first and last blocks) and random bodies
        idx = np.unique(np.concatenate([np.arange(min(n, 8)), np.arange(max(0, n - 8), n),
                                        rng.choice(n, min(n, 48), replace=False)]))
        f64 = oracle.fp64_forces(pm, EPS2, idx)
        rel = np.linalg.norm(f[k][idx, :3].astype(np.float64) - f64, axis=1) / np.linalg.norm(f64, axis=1)
        worst = max(worst, float(rel.max()))
        assert rel.max() <= TOL_MAX, (n, compensated, float(rel.max()))
    print(f"compensated={compensated}: worst {worst:.3e} over {len(sizes) - 1} systems")
