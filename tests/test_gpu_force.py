"""Parity of the HIP force path (through the C ABI) with the oracle.

Tolerances (FP32 kernels against the FP64 oracle, per-body normwise
|F - F64| / |F64|): max <= 1e-5 and RMS <= 1e-6 for compensated plans and for
plain sums of up to 4096 terms per slice; plain sums over longer slices get
the sqrt(terms) growth of a sequential FP32 sum (see tolerances()). Against
the reference's FP32 CPU_AOS crosscheck: that bound + 1e-5 (CPU_AOS's own
error). Integer-like properties (w untouched, determinism, i-range equality,
power-of-two mass scaling, the integrator) are bit-exact.
"""
import ctypes

import numpy as np
import pytest

from conftest import EPS2
from gpu_util import dev, force_buffer, host, ptr, run_entry, run_force
from oracle_lib import force_errors

from nbody_amd import _native, ics

pytestmark = pytest.mark.gpu

TOL_MAX, TOL_RMS, TOL_AOS = 1e-5, 1e-6, 2e-5


def tolerances(chain, compensated):
    """(max, rms) bound for a plan: the target 1e-5 / 1e-6 for compensated
    sums; a plain FP32 sum's error grows like sqrt(terms), so plain chains
    longer than 4096 terms get that growth on a 2e-5 / 2e-6 base."""
    if compensated:
        return TOL_MAX, TOL_RMS
    g = max(1.0, chain / 4096.0) ** 0.5
    return 2e-5 * g, 2e-6 * g


def entry_plan(entry, n):
    """(chain length, compensated) of the plan an entry uses for n bodies."""
    if entry.endswith("_AOS"):
        return n, False
    variant = {"LDS": _native.LDS, "DPP": _native.DPP, "SGPR": _native.SGPR,
               "Symmetric": _native.SYMMETRIC}.get(entry.rsplit("_", 1)[-1], _native.AUTO)
    o = _native.Options.make(variant=variant)
    plan = _native.Options()
    _native.check(_native.lib().nbody_hip_plan(n, n, ctypes.byref(o), ctypes.byref(plan)), "plan")
    return plan.jslice, bool(plan.compensated)


ENTRIES = ["ComputeGravitation_HIP_AOS", "ComputeGravitation_HIP_LDS", "ComputeGravitation_HIP_DPP",
           "ComputeGravitation_HIP_SGPR", "ComputeGravitation_HIP_Symmetric", "ComputeGravitation_HIP"]
SYM_SHAPES = [(2, 128), (2, 512), (4, 256), (4, 512), (4, 1024), (8, 512), (8, 1024), (8, 2048)]
SIZES = [1, 2, 3, 17, 255, 256, 257, 1000, 1024, 4097, 16384]


def bodies(n, seed=7):
    pm, vm = ics.random_unit_bodies(n, seed=seed)
    return pm, vm


def check_against_oracle(oracle, pm, f, what, tol=(TOL_MAX, TOL_RMS)):
    n = pm.shape[0]
    assert np.all(f[:, 3] == -7.0), f"{what}: force.w was written"
    assert np.all(np.isfinite(f[:, :3])), f"{what}: unwritten or non-finite force"
    if n == 1:
        assert np.all(f[:, :3] == 0.0)
        return
    f64 = oracle.fp64_forces(pm, EPS2)
    mx, rms = force_errors(f, f64)
    assert mx <= tol[0] and rms <= tol[1], (what, n, mx, rms)
    aos, _ = oracle.cpu_aos(pm, EPS2)
    d = np.linalg.norm(f[:, :3].astype(np.float64) - aos[:, :3], axis=1) / np.linalg.norm(f64, axis=1)
    assert d.max() <= tol[0] + 1e-5, (what, n, d.max())


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("entry", ENTRIES)
def test_entry_matches_oracle(oracle, entry, n):
    pm, _ = bodies(n)
    d_pm = dev(pm)
    f, ms = run_entry(entry, d_pm, EPS2)
    assert ms > 0
    check_against_oracle(oracle, pm, f, entry, tolerances(*entry_plan(entry, n)))


@pytest.mark.parametrize("variant", [_native.SGPR, _native.LDS, _native.DPP])
@pytest.mark.parametrize("ipl", [1, 2, 4])
@pytest.mark.parametrize("comp", [0, 1])
@pytest.mark.parametrize("newton", [0, 1])
def test_plan_options(oracle, variant, ipl, comp, newton):
    pm, _ = bodies(5003, seed=2)
    d = dev(pm)
    f64 = oracle.fp64_forces(pm, EPS2)
    for jslice in (256, 1000, 5003):
        f = run_force(d, EPS2, variant=variant, i_per_lane=ipl, jslice=jslice, compensated=comp, newton=newton)
        mx, rms = force_errors(f, f64)
        tmax, trms = tolerances(jslice, comp)
        assert mx <= tmax and rms <= trms, (variant, ipl, jslice, comp, newton, mx, rms)
        assert np.all(f[:, 3] == -7.0)


@pytest.mark.parametrize("comp", [0, 1])
@pytest.mark.parametrize("newton", [0, 1])
def test_dpp_eight_per_lane(oracle, comp, newton):
    """The DPP tile rotation with 8 i bodies per lane (its default shape at
    large N), on slices that end in a partial tile (1000, 5003) and on whole
    tiles (256); the variants whose j body costs no move per step are built
    for at most 4 and refuse 8."""
    pm, _ = bodies(5003, seed=2)
    d = dev(pm)
    f64 = oracle.fp64_forces(pm, EPS2)
    for jslice in (256, 1000, 5003):
        f = run_force(d, EPS2, variant=_native.DPP, i_per_lane=8, jslice=jslice, compensated=comp, newton=newton)
        mx, rms = force_errors(f, f64)
        tmax, trms = tolerances(jslice, comp)
        assert mx <= tmax and rms <= trms, (jslice, comp, newton, mx, rms)
        assert np.all(f[:, 3] == -7.0)
    plan = _native.Options()
    for variant in (_native.LDS, _native.SGPR):
        o = _native.Options.make(variant=variant, i_per_lane=8)
        assert _native.lib().nbody_hip_plan(5003, 5003, ctypes.byref(o), ctypes.byref(plan)) != 0
    o = _native.Options.make(variant=_native.DPP)
    _native.check(_native.lib().nbody_hip_plan(98304, 98304, ctypes.byref(o), ctypes.byref(plan)), "plan")
    assert plan.i_per_lane == 8


@pytest.mark.parametrize("variant", [_native.SGPR, _native.LDS, _native.DPP])
@pytest.mark.parametrize("comp", [0, 1])
@pytest.mark.parametrize("newton", [0, 1])
def test_one_sided_equal_mass(oracle, variant, comp, newton):
    """A j slice whose bodies share one mass takes the mass-free chain and
    multiplies its sums by that mass once. With unit masses that gives the
    bits of the general chain: the same system with one zero-mass body
    appended to the slice (which sends the slice down the general chain and
    adds exact zeros) must give the first n bodies the same forces. A mass of
    1/4 scales them exactly, and both stay within the oracle's bounds."""
    n = 4096
    pm, _ = ics.uniform_ball(n, seed=3, radius=ics.ball_radius(n))
    assert np.all(pm[:, 3] == 1.0)
    extra = np.concatenate([pm, np.array([[3.0, -2.0, 1.0, 0.0]], np.float32)])
    quarter = pm.copy()
    quarter[:, 3] = 0.25
    f64 = oracle.fp64_forces(pm, EPS2)
    d, d_extra, d_quarter = dev(pm), dev(extra), dev(quarter)
    for ipl in ((1, 2, 4, 8) if variant == _native.DPP else (1, 2, 4)):
        opt = dict(variant=variant, i_per_lane=ipl, jslice=8192, compensated=comp, newton=newton)
        a = run_force(d, EPS2, **opt)
        b = run_force(d_extra, EPS2, **opt)
        c = run_force(d_quarter, EPS2, **opt)
        assert np.array_equal(a[:, :3].view(np.uint32), b[:n, :3].view(np.uint32)), (ipl, "unit vs general")
        assert np.array_equal((a[:, :3] * np.float32(0.25)).view(np.uint32), c[:, :3].view(np.uint32)), (ipl, "1/4")
        mx, rms = force_errors(a, f64)
        tmax, trms = tolerances(n, comp)
        assert mx <= tmax and rms <= trms, (ipl, mx, rms)
        assert np.all(a[:, 3] == -7.0)


@pytest.mark.parametrize("shape", SYM_SHAPES)
@pytest.mark.parametrize("newton", [0, 1])
def test_symmetric_shapes(oracle, shape, newton):
    """Every compiled symmetric block shape on a ragged N (last block partly
    empty) with the reference's unequal masses U(0.5, 1): the reaction on j
    must carry m_i (the reference's tiled kernels use m_j, which is only right
    for equal masses)."""
    a, b = shape
    for n in (5003, 2 * b + 3):
        pm, _ = bodies(n, seed=2)
        d = dev(pm)
        f = run_force(d, EPS2, variant=_native.SYMMETRIC, i_per_lane=a, jslice=b, newton=newton)
        mx, rms = force_errors(f, oracle.fp64_forces(pm, EPS2))
        assert mx <= TOL_MAX and rms <= TOL_RMS, (shape, newton, n, mx, rms)
        assert np.all(f[:, 3] == -7.0)


@pytest.mark.parametrize("shape", [(8, 512), (4, 256), (2, 512), (2, 128)])
def test_symmetric_equal_mass_blocks(oracle, shape):
    """Block pairs whose bodies all share one mass take the mass-free pair
    weights and scale the partials once. Cases: one mass everywhere (not a
    power of two), the same with one heavier body in block 2 (pairs with block
    2 take the general weights, the rest the equal-mass ones), a ragged N
    (the last block holds zero-mass padding), and power-of-two scaling of the
    common mass, which must scale the forces exactly."""
    a, b = shape
    for n in (4 * b, 4 * b + 37):
        pm, _ = bodies(n, seed=9)
        pm[:, 3] = 0.37
        mixed = pm.copy()
        mixed[2 * b + 5, 3] = 0.6
        for case, x in (("one mass", pm), ("one heavier body", mixed)):
            f = run_force(dev(x), EPS2, variant=_native.SYMMETRIC, i_per_lane=a, jslice=b)
            mx, rms = force_errors(f, oracle.fp64_forces(x, EPS2))
            assert mx <= TOL_MAX and rms <= TOL_RMS, (shape, n, case, mx, rms)
            assert np.all(f[:, 3] == -7.0)
        unit = pm.copy()
        unit[:, 3] = 1.0
        quarter = pm.copy()
        quarter[:, 3] = 0.25
        f1 = run_force(dev(unit), EPS2, variant=_native.SYMMETRIC, i_per_lane=a, jslice=b)
        fq = run_force(dev(quarter), EPS2, variant=_native.SYMMETRIC, i_per_lane=a, jslice=b)
        assert np.array_equal(fq[:, :3], 0.25 * f1[:, :3]), (shape, n)


@pytest.mark.parametrize("shape", [(8, 512), (4, 256), (2, 128), (8, 2048)])
def test_symmetric_column_bands(oracle, shape):
    """A symmetric plan whose slots exceed slot_bytes runs in column bands of
    blocks, carrying each body's running (sum, error) pair from band to band.
    On a ragged N with unequal masses, for slot caps giving 2 bands up to one
    column per band: forces against the FP64 oracle, run-to-run bit-equality,
    and the cap reported by the plan."""
    a, b = shape
    n = 9 * b + 37
    nb = -(-n // b)
    pm, _ = bodies(n, seed=12)
    d = dev(pm)
    f64 = oracle.fp64_forces(pm, EPS2)
    whole = 12 * b * nb * nb
    for cap in (whole // 2, whole // 5, 12 * b * 3):
        f = run_force(d, EPS2, variant=_native.SYMMETRIC, i_per_lane=a, jslice=b, slot_bytes=cap)
        mx, rms = force_errors(f, f64)
        assert mx <= TOL_MAX and rms <= TOL_RMS, (shape, cap, mx, rms)
        assert np.all(f[:, 3] == -7.0)
        again = run_force(d, EPS2, variant=_native.SYMMETRIC, i_per_lane=a, jslice=b, slot_bytes=cap)
        assert np.array_equal(f.view(np.uint32), again.view(np.uint32)), (shape, cap)
    o = _native.Options.make(variant=_native.SYMMETRIC, i_per_lane=a, jslice=b, slot_bytes=12345678)
    plan = _native.Options()
    _native.check(_native.lib().nbody_hip_plan(n, n, ctypes.byref(o), ctypes.byref(plan)), "plan")
    assert plan.slot_bytes == 12345678


@pytest.mark.parametrize("opt", [dict(variant=_native.SYMMETRIC, i_per_lane=4),
                                 dict(variant=_native.SYMMETRIC, i_per_lane=8),
                                 dict(variant=_native.SYMMETRIC, i_per_lane=4, slot_bytes=24 * 256 * 4),
                                 dict(variant=_native.LDS, i_per_lane=1, jslice=1024),
                                 dict(variant=_native.LDS, i_per_lane=2, jslice=5003),
                                 dict(variant=_native.LDS, i_per_lane=4)],
                         ids=["sym4", "sym8", "sym4-bands", "lds1-slices", "lds2-one-slice", "lds4"])
def test_fp64_accumulation_plans(oracle, opt):
    """compensated = 2: every term added in FP64 and the partial sums (block
    slots, j slices) kept in FP64 until one rounding into the force. On a
    ragged N with the reference's unequal masses: within the tolerance of
    the FP64 oracle, deterministic, and reported by the plan."""
    n = 5003
    pm, _ = bodies(n, seed=13)
    d = dev(pm)
    f = run_force(d, EPS2, compensated=2, **opt)
    mx, rms = force_errors(f, oracle.fp64_forces(pm, EPS2))
    assert mx <= TOL_MAX and rms <= TOL_RMS, (opt, mx, rms)
    assert np.all(f[:, 3] == -7.0)
    again = run_force(d, EPS2, compensated=2, **opt)
    assert np.array_equal(f.view(np.uint32), again.view(np.uint32))
    o = _native.Options.make(compensated=2, **opt)
    plan = _native.Options()
    _native.check(_native.lib().nbody_hip_plan(n, n, ctypes.byref(o), ctypes.byref(plan)), "plan")
    assert plan.compensated == 2 and plan.variant == opt["variant"]


def test_fp64_accumulation_refusals():
    pm, _ = bodies(3000, seed=14)
    d = dev(pm)
    f = force_buffer(3000)
    for opt in (dict(variant=_native.SGPR), dict(variant=_native.DPP),
                dict(variant=_native.SYMMETRIC, i_per_lane=2), dict(variant=_native.SYMMETRIC, i_per_lane=4,
                                                                      jslice=512)):
        o = _native.Options.make(compensated=2, **opt)
        rc = _native.lib().nbody_hip_force(ptr(f), ptr(d), EPS2, 3000, 0, 3000, ctypes.byref(o), None)
        assert rc != 0 and ("FP64" in _native.last_error()), (opt, _native.last_error())
    assert np.all(np.isnan(host(f)[:, :3]))


def pair_shares(d_pm, n, shares, **opt):
    """Every share's FP64 partial forces and their sum."""
    L = _native.lib()
    o = _native.Options.make(**opt) if opt else None
    parts = []
    for k in range(shares):
        out = dev(np.full((n, 4), -9.0, np.float64))
        rc = L.nbody_hip_force_pair_share(ptr(out), ptr(d_pm), EPS2, n, k, shares, ctypes.byref(o) if o else None,
                                          None)
        _native.check(rc, "nbody_hip_force_pair_share")
        parts.append(host(out))
    return parts, sum(p[:, :3] for p in parts)


@pytest.mark.parametrize("shares", [1, 3, 8, 13])
@pytest.mark.parametrize("opt", [{}, dict(i_per_lane=4, jslice=256, slot_bytes=12 * 256 * 6),
                                 dict(compensated=2)], ids=["default", "bands", "fp64"])
def test_pair_shares_add_up(oracle, shares, opt):
    """The block pairs split over G GPUs by columns: the G shares' FP64
    partial forces add up to the force (within the path's tolerance of the
    FP64 oracle), each share leaves w alone, a share past the last column
    writes zeros, and one share is the whole symmetric evaluation bit for
    bit."""
    n = 9 * 512 + 37
    pm, _ = bodies(n, seed=20)
    d = dev(pm)
    parts, total = pair_shares(d, n, shares, **opt)
    for p in parts:
        assert np.all(p[:, 3] == -9.0)
    mx, rms = force_errors(total, oracle.fp64_forces(pm, EPS2))
    assert mx <= TOL_MAX and rms <= TOL_RMS, (shares, opt, mx, rms)
    if shares == 1:
        f = run_force(d, EPS2, variant=_native.SYMMETRIC, **opt)
        assert np.array_equal(total.astype(np.float32).view(np.uint32), f[:, :3].view(np.uint32))


def test_pair_share_refusals():
    pm, _ = bodies(3000, seed=21)
    d = dev(pm)
    out = dev(np.full((3000, 4), -9.0, np.float64))
    L = _native.lib()
    assert L.nbody_hip_force_pair_share(ptr(out), ptr(d), EPS2, 3000, 3, 3, None, None) != 0
    assert "share" in _native.last_error()
    o = _native.Options.make(variant=_native.LDS)
    assert L.nbody_hip_force_pair_share(ptr(out), ptr(d), EPS2, 3000, 0, 2, ctypes.byref(o), None) != 0
    assert "symmetric" in _native.last_error()
    assert np.all(host(out) == -9.0)


def test_symmetric_rejects_i_ranges():
    pm, _ = bodies(3000, seed=3)
    d = dev(pm)
    f = force_buffer(100)
    o = _native.Options.make(variant=_native.SYMMETRIC)
    rc = _native.lib().nbody_hip_force(ptr(f), ptr(d), EPS2, 3000, 100, 100, ctypes.byref(o), None)
    assert rc != 0 and "whole systems" in _native.last_error()
    o = _native.Options.make(variant=_native.SYMMETRIC, i_per_lane=4, jslice=300)
    rc = _native.lib().nbody_hip_force(ptr(f), ptr(d), EPS2, 3000, 0, 3000, ctypes.byref(o), None)
    assert rc != 0 and "block" in _native.last_error()


def test_symmetric_waves_per_simd_leaves_results_alone():
    """Capping the symmetric kernel's resident waves (extra LDS per workgroup)
    changes only the schedule: the forces stay bit-identical, the plan reports
    the cap, and caps outside 0..8 are rejected."""
    pm, _ = bodies(7000, seed=6)
    d = dev(pm)
    base = run_force(d, EPS2, variant=_native.SYMMETRIC, i_per_lane=8, jslice=512)
    for waves in (6, 4, 2, 1):
        f = run_force(d, EPS2, variant=_native.SYMMETRIC, i_per_lane=8, jslice=512, waves_per_simd=waves)
        assert np.array_equal(f.view(np.uint32), base.view(np.uint32)), waves
    o = _native.Options.make(variant=_native.SYMMETRIC, waves_per_simd=5)
    plan = _native.Options()
    _native.check(_native.lib().nbody_hip_plan(7000, 7000, ctypes.byref(o), ctypes.byref(plan)), "plan")
    assert plan.waves_per_simd == 5
    o = _native.Options.make(variant=_native.SYMMETRIC, waves_per_simd=9)
    f = force_buffer(7000)
    rc = _native.lib().nbody_hip_force(ptr(f), ptr(d), EPS2, 7000, 0, 7000, ctypes.byref(o), None)
    assert rc != 0 and "waves_per_simd" in _native.last_error()


@pytest.mark.parametrize("variant", [_native.SGPR, _native.LDS, _native.DPP, _native.SYMMETRIC])
def test_deterministic(variant):
    pm, _ = bodies(20000, seed=4)
    d = dev(pm)
    a = run_force(d, EPS2, variant=variant, compensated=1)
    b = run_force(d, EPS2, variant=variant, compensated=1)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("variant,ipl", [(_native.SGPR, 2), (_native.LDS, 2), (_native.DPP, 2), (_native.DPP, 4),
                                         (_native.DPP, 8)])
@pytest.mark.parametrize("masses", ["unequal", "one"])
def test_i_range_equals_full(variant, ipl, masses):
    """The base/n entry (the multi-GPU i-slice of src/nbody_multiGPU_shared.cuh)
    gives each body exactly the force of the full evaluation under one plan,
    with the reference's unequal masses and with one mass (the mass-free
    chain); the DPP tile rotation orders each body's terms by its own index,
    whatever the range's first body."""
    if masses == "one":
        pm, _ = ics.uniform_ball(9000, seed=5)
    else:
        pm, _ = bodies(9000, seed=5)
    d = dev(pm)
    opt = dict(variant=variant, i_per_lane=ipl, jslice=1024, compensated=1)
    full = run_force(d, EPS2, **opt)
    for base, count in [(0, 1), (4500, 1125), (8999, 1), (1, 8999), (2048, 2048), (37, 5000)]:
        part = run_force(d, EPS2, base=base, count=count, **opt)
        assert np.array_equal(part.view(np.uint32), full[base:base + count].view(np.uint32)), (base, count)


def test_mass_scaling_is_exact():
    pm, _ = bodies(4096, seed=6)
    d1 = dev(pm)
    a = run_force(d1, EPS2, variant=_native.SGPR)
    pm2 = pm.copy()
    pm2[:, 3] *= 2.0
    d2 = dev(pm2)
    b = run_force(d2, EPS2, variant=_native.SGPR)
    assert np.array_equal(b[:, :3], 2.0 * a[:, :3])


@pytest.mark.parametrize("n", [1, 3000, 70000, 2000])
def test_host_twin_matches_device_entry(n):
    """The host-pointer entry (page-locked staging, growing and shrinking
    across calls) gives the device entry's forces bit for bit and leaves the
    caller's w untouched."""
    pm, _ = bodies(n, seed=8)
    d_pm = dev(pm)
    f_dev, _ = run_entry("ComputeGravitation_HIP", d_pm, EPS2)
    fh = np.full((n, 4), np.nan, np.float32)
    fh[:, 3] = -7.0
    ms = _native.lib().ComputeGravitation_HIP_Host(fh.ctypes.data, np.ascontiguousarray(pm).ctypes.data, EPS2, n)
    assert ms > 0
    assert np.array_equal(fh.view(np.uint32), f_dev.view(np.uint32))


def test_reference_fixture_crosscheck(oracle):
    """seed-7 ICs of the reference driver against the forces its own CPU_SOA
    build produced (tests/golden), using the reference's crosscheck metric
    (src/nbody.cu:406-418) and the normwise one."""
    import os
    from conftest import GOLDEN
    g = np.load(os.path.join(GOLDEN, "ref_seed7_n4096.npz"))
    pm = g["pos_mass"]
    d_pm = dev(pm)
    f, _ = run_entry("ComputeGravitation_HIP", d_pm, EPS2)
    golden = np.zeros((4096, 4), np.float32)
    golden[:, :3] = g["force_cpu_soa"]
    rel = oracle.max_rel_error(f, golden)
    f64 = oracle.fp64_forces(pm, EPS2)
    d = np.linalg.norm(f[:, :3].astype(np.float64) - golden[:, :3], axis=1) / np.linalg.norm(f64, axis=1)
    print(f"reference-style max relError {rel:.3e}, normwise max {d.max():.3e}")
    assert d.max() <= TOL_AOS


def test_integrate_bitexact_vs_oracle(oracle):
    n = 5000
    pm, vm = bodies(n, seed=9)
    rng = np.random.default_rng(0)
    f = rng.normal(size=(n, 4)).astype(np.float32)
    dpm, dvm, df = dev(pm), dev(vm), dev(f)
    ms = _native.lib().integrateGravitation_HIP(ptr(dpm), ptr(dvm), ptr(df), 0.016, 0.95, n)
    assert ms > 0
    p1, v1 = oracle.integrate_aos(pm, vm, f, 0.016, 0.95)
    assert np.array_equal(host(dpm).view(np.uint32), p1.view(np.uint32))
    assert np.array_equal(host(dvm).view(np.uint32), v1.view(np.uint32))


def test_fp64_force_matches_oracle(oracle):
    pm, _ = bodies(3001, seed=10)
    out = dev(np.full((3001, 4), -9.0, np.float64))
    d_pm = dev(pm)
    ms = _native.lib().nbody_hip_force_fp64(ptr(out), ptr(d_pm), EPS2, 3001)
    assert ms > 0
    o = host(out)
    f64 = oracle.fp64_forces(pm, EPS2)
    assert np.all(o[:, 3] == -9.0)
    err = np.linalg.norm(o[:, :3] - f64, axis=1) / np.linalg.norm(f64, axis=1)
    assert err.max() < 1e-12


def test_energy_fp64_matches_oracle(oracle):
    pm, vm = bodies(2000, seed=11)
    out = np.zeros(5, np.float64)
    d_pm, d_vm = dev(pm), dev(vm)  # keep both alive: a freed temporary's block is reused by the next tensor
    rc = _native.lib().nbody_hip_energy_fp64(out.ctypes.data, ptr(d_pm), ptr(d_vm), EPS2, 2000)
    _native.check(rc, "energy")
    ke, pe, mom = oracle.fp64_energy(pm, vm, EPS2)
    assert out[0] == pytest.approx(ke, rel=1e-12)
    assert out[1] == pytest.approx(pe, rel=1e-10)
    assert out[2:] == pytest.approx(mom, rel=1e-10, abs=1e-10 * np.abs(mom).max())


def test_bad_arguments_fail_loudly():
    L = _native.lib()
    pm, _ = bodies(16)
    d = dev(pm)
    f = force_buffer(16)
    assert L.ComputeGravitation_HIP_SGPR(ptr(f), ptr(d), EPS2, 0) == 0.0
    assert "N = 0" in _native.last_error()
    assert L.ComputeGravitation_HIP_SGPR(None, ptr(d), EPS2, 16) == 0.0
    o = _native.Options.make(variant=_native.SGPR, i_per_lane=3)
    assert L.nbody_hip_force(ptr(f), ptr(d), EPS2, 16, 0, 16, ctypes.byref(o), None) != 0
    assert "i_per_lane" in _native.last_error()
    assert L.nbody_hip_force(ptr(f), ptr(d), EPS2, 16, 10, 7, None, None) != 0
    assert np.all(np.isnan(host(f)[:, :3]))  # nothing ran


def test_autotune_then_default(oracle):
    pm, _ = bodies(12000, seed=12)
    d = dev(pm)
    f = force_buffer(12000)
    best = _native.lib().nbody_hip_autotune(ptr(f), ptr(d), EPS2, 12000, _native.AUTO)
    _native.check_ms(best, "autotune")
    g, _ = run_entry("ComputeGravitation_HIP", d, EPS2)
    mx, rms = force_errors(g, oracle.fp64_forces(pm, EPS2))
    assert mx <= TOL_MAX and rms <= TOL_RMS


def test_autotune_without_a_fitting_plan():
    """No symmetric shape fits N = 2^26 (more than 16384 blocks each): the
    autotuner says so and launches nothing (the buffers are 16 bodies)."""
    pm, _ = bodies(16)
    d = dev(pm)
    f = force_buffer(16)
    assert _native.lib().nbody_hip_autotune(ptr(f), ptr(d), EPS2, 1 << 26, _native.SYMMETRIC) == 0.0
    assert "no plan" in _native.last_error()
    assert np.all(np.isnan(host(f)[:, :3]))


@pytest.mark.slow
def test_full_size_98304_sampled(oracle):
    """The single-system workload (N = 98304 uniform ball, eps^2 = 0.01): the
    default entry against FP64 on a sample that includes the 256 bodies
    nearest the centre (smallest |F|, largest relative error), and the
    size-independent equal-mass property sum_i F_i = 0."""
    n = 98304
    pm, _ = ics.uniform_ball(n, seed=1)
    d_pm = dev(pm)
    f, ms = run_entry("ComputeGravitation_HIP", d_pm, 0.01)
    r = np.linalg.norm(pm[:, :3], axis=1)
    idx = np.unique(np.concatenate([np.argsort(r)[:256], np.random.default_rng(0).choice(n, 1792, replace=False)]))
    f64 = oracle.fp64_forces(pm, 0.01, idx)
    mx, rms = force_errors(f[idx], f64)
    print(f"N=98304 HIP: {ms:.3f} ms, sampled max rel {mx:.3e} rms {rms:.3e}")
    assert mx <= TOL_MAX and rms <= TOL_RMS
    tot = f[:, :3].astype(np.float64).sum(0)
    assert np.abs(tot).max() <= 1e-6 * np.abs(f[:, :3]).sum()
