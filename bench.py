"""Benchmark of the MI355X force path: every part of the metric in one line.

    python bench.py [--steps K] [--warmup W] [--n N] [--variant NAME] [--single-only] [--no-cpu-baseline]

Headline workload (`value`): one system of N = 98304 bodies, positions
uniform in a ball of radius 1.54 (N/16384)^(1/3), unit masses, velocities
uniform in a unit ball (seeded), softeningSquared = 0.01, dt = 0.016, damping
0.95 (src/nbody.cu:152). A step is one force evaluation over all N^2 pairs
plus the integrator, with the state resident in HBM. Before the W warm-up steps, about 300 ms of
untimed force evaluations on the initial positions (the state does not move)
bring the GPU to its sustained speed: from idle, the force kernel's time falls
from about 1.9 to 1.59 ms over its first ~30 ms of sustained work. The timed
region is K steps between device synchronisations:

    value = N^2 K / time / 1e9   (10^9 interactions/s, the reference's figure,
                                  src/nbody.cu:1026-1027)

The same JSON line also carries (the reference prints every row it measures
in one run, src/nbody.cu:1024-1053):
  roofline      the force kernel's credited TFLOP/s (20 flops per interaction)
                against the FP32 vector peak; kernel time from events the
                library records around the force kernel on its stream over
                the timed region;
  accuracy      the force on the initial and on the final state against the
                FP64 direct sum, and the reference's crosscheck against
                CPU_AOS with its stated tolerance and a pass flag;
  general_mass  the reference's own ICs (randomUnitBodies, m in U(0.5, 1),
                src/nbody_util.cpp:137), which take the general pair weights;
                like the ensemble's timed steps, timed after the same untimed
                settle (force evaluations on a separate 98304-body ball, so
                the measured state is not touched);
  ensemble      512 independent systems, batched graph vs one launch per system,
                and the batched force on the initial and final states of every
                system against the FP64 direct sum;
  served        a Poisson stream of simulation requests: p50/p99 latency and
                interactions/s, and the force error on the final states of
                sampled requests (replayed untimed after the timed run); the
                same with FP64 accumulation;
  plummer       N = 2^20 Plummer sphere: default vs plain vs compensated sums;
  cpu_baseline  the reference's CPU_SOA_tiled (src/nbody_CPU_SOA_tiled.cpp:51-115)
                compiled from its sources, else the oracle's restatement, on
                this host's physical cores, at N = 98304 and 16384, plus one
                CPU step of every ensemble system and of sampled requests, and
                the N = 2^20 Plummer force extrapolated from 2^18 of its bodies.

Single workloads on their own (--workload): ensemble, served, plummer, share
(one GPU's share of the N = 98304 system split over --shares GPUs: its i
range against all N with the one-sided plans, and its share of the block
pairs with the symmetric plan).
"""
import argparse
import ctypes
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))

PEAK_FP32_TFLOPS = 157.3       # MI355X FP32 vector / dense matrix peak (published)
REF_BEST_GPU = 501.5           # 10^9 int/s, GPU_Shared clang build, RTX 2080 Ti, N=98304 (README.md:230-233)
METRIC = "10^9 interactions/s + % FP32 peak at N=96K; ensemble throughput + p99 latency"
TOL_FP64 = 1e-5                # per-body |F - F64| / |F64|, the force tolerance of the path


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--n", type=int, default=98304)
    ap.add_argument("--variant", default="auto", choices=["auto", "sgpr", "lds", "dpp", "aos", "symmetric"])
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--ipl", type=int, default=0, help="pin i bodies per lane (skips the autotuner)")
    ap.add_argument("--jslice", type=int, default=0, help="pin the j slice length (skips the autotuner)")
    ap.add_argument("--compensated", type=int, default=-1, choices=[-1, 0, 1, 2],
                    help="accumulation of the single system's plan (2: FP64); with --workload served, the "
                         "server's (2: FP64, else FP32 block partials)")
    ap.add_argument("--waves", type=int, default=0, help="pin the symmetric plan's waves per SIMD (skips the autotuner)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy 5 times (counter calibration for --pmc runs)")
    ap.add_argument("--workload", default="single", choices=["single", "ensemble", "served", "plummer", "share"],
                    help="single: one N-body system with every other part of the metric beside it (the "
                         "headline line); ensemble, served, plummer, share: that workload alone")
    ap.add_argument("--single-only", action="store_true",
                    help="the single-system line without the ensemble, served, plummer and general-mass parts")
    ap.add_argument("--shares", type=int, default=8, help="share: GPUs the single system would be split over")
    ap.add_argument("--systems", type=int, default=512)
    ap.add_argument("--ensemble-steps", type=int, default=50)
    ap.add_argument("--requests", type=int, default=400)
    ap.add_argument("--max-active", type=int, default=64)
    ap.add_argument("--rate", type=float, default=200.0, help="served: mean request arrivals per second")
    ap.add_argument("--plummer-steps", type=int, default=5)
    ap.add_argument("--settle-ms", type=float, default=300.0,
                    help="untimed force evaluations on the initial positions (results discarded) for about this "
                         "long before the warm-up steps: the kernel's time falls by about 20%% over its first "
                         "~30 ms of sustained work, which a short warm-up would leave inside the timed steps")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget-s", type=float, default=10.0)
    return ap.parse_args()


# ---- CPU baseline ------------------------------------------------------------

def physical_cores():
    """Physical cores this process may run on: distinct (core, socket) pairs
    of the CPUs in its affinity set, as the reference's benchmark script
    counts them (`lscpu -b -p=CORE,SOCKET`, scripts/compare-compilers:8-10)."""
    import subprocess
    allowed = os.sched_getaffinity(0)
    try:
        out = subprocess.run(["lscpu", "-b", "-p=CPU,CORE,SOCKET"], stdout=subprocess.PIPE, text=True,
                             check=True).stdout
        pairs = set()
        for line in out.splitlines():
            if line.startswith("#") or not line.strip():
                continue
            cpu, core, sock = line.split(",")[:3]
            if int(cpu) in allowed:
                pairs.add((core, sock))
        if pairs:
            return len(pairs), len(allowed)
    except (OSError, ValueError, subprocess.CalledProcessError):
        pass
    return len(allowed), len(allowed)


def cpu_model():
    """lscpu's model name, with family / model / stepping and the widest x86
    vector extension, since a virtual machine may report only a generic
    name."""
    import subprocess
    fields = {}
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True, check=True).stdout
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                fields[k.strip()] = v.strip()
    except (OSError, subprocess.CalledProcessError):
        pass
    if "Model name" not in fields:
        try:
            for line in open("/proc/cpuinfo"):
                if ":" in line:
                    k, v = line.split(":", 1)
                    k = {"model name": "Model name", "cpu family": "CPU family", "model": "Model",
                         "stepping": "Stepping", "flags": "Flags"}.get(k.strip(), k.strip())
                    fields.setdefault(k, v.strip())
        except OSError:
            pass
    name = fields.get("Model name", "unknown")
    flags = set(fields.get("Flags", "").split())
    isa = "AVX-512" if "avx512f" in flags else "AVX2" if "avx2" in flags else None
    extra = [f"{k.lower()} {fields[k]}" for k in ("CPU family", "Model", "Stepping") if k in fields]
    if isa:
        extra.append(isa)
    return f"{name} ({', '.join(extra)})" if extra else name


def _omp():
    """libgomp, which the CPU libraries' OpenMP regions run on."""
    g = ctypes.CDLL("libgomp.so.1")
    g.omp_get_max_threads.restype = ctypes.c_int
    g.omp_set_num_threads.argtypes = [ctypes.c_int]
    return g


def _cpu_force_fn(impl, n):
    """CPU_SOA_tiled where it applies (N % 4096 == 0, src/nbody_CPU_SOA_tiled.cpp:57-58), else CPU_SOA."""
    return (impl.cpu_soa_tiled, "CPU_SOA_tiled") if n % 4096 == 0 else (impl.cpu_soa, "CPU_SOA")


def _cpu_rate(impl, integ, pm, vm, eps2, dt, damping, budget_s, max_steps=20):
    """Median per-step force time of CPU force + host integrate steps (the
    force timed by the reference's own steady_clock, its return value)."""
    import numpy as np
    fn, algo = _cpu_force_fn(impl, pm.shape[0])
    fn(pm, eps2)  # warm-up (scripts/compare-compilers:116)
    times, spent = [], 0.0
    while len(times) < max_steps:
        t0 = time.perf_counter()
        f, ms = fn(pm, eps2)
        pm, vm = integ.integrate_aos(pm, vm, f, dt, damping)
        times.append(ms / 1e3)
        spent += time.perf_counter() - t0
        if len(times) >= 3 and spent + spent / len(times) > budget_s:
            break
    return float(np.median(times)), len(times), algo, pm


def cpu_baseline(pm_host, vm_host, eps2, dt, damping, budget_s, topology, ensemble_sizes=None, served_reqs=None,
                 plummer=False):
    """The reference's CPU path on this host (rank 0, one GPU only): the
    reference library compiled from its sources when present (kind
    "reference"), else the oracle's restatement (kind "port"). `topology` is
    physical_cores() taken at start-up: once an OpenMP region has run with
    OMP_PROC_BIND set, the main thread is bound to one CPU and the process's
    affinity shows only that one."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import numpy as np
    from nbody_amd import ics
    from oracle_lib import oracle, ref_cpu
    ref = ref_cpu()
    kind, impl = ("reference", ref) if ref is not None else ("port", oracle())
    integ = oracle()
    phys, logical = topology
    omp = _omp()
    env_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(phys, env_threads) if env_threads > 0 else phys
    omp.omp_set_num_threads(threads)
    n = pm_host.shape[0]
    med, steps, algo, pm_end = _cpu_rate(impl, integ, pm_host.copy(), vm_host.copy(), eps2, dt, damping, budget_s)
    out = {
        "value": round(n * n / med / 1e9, 3),
        "unit": "10^9 interactions/s",
        "cores": threads,
        "kind": kind,
        "sample": f"{steps} of the workload's steps ({algo} force, median of its own steady_clock ms, + host "
                  f"integrate), N={n}, {threads} OpenMP threads = physical cores of this process's CPU set, "
                  f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND', '')}",
        "ms_per_force": round(med * 1e3, 2),
        "cpu": cpu_model(),
        "physical_cores": phys,
        "logical_cpus": logical,
        "omp_threads": omp.omp_get_max_threads(),
        "build": "reference sources compiled with g++ -O3 -ffast-math -fopenmp -march=x86-64-v4 (AVX-512) in the "
                 "build container; -march=native there would target that container's CPU, not this host's"
                 if kind == "reference" else "oracle restatement, gcc -O3 -ffast-math -fopenmp -march=x86-64-v4",
    }
    # the reference's other CPU rows on the same positions, three calls each
    others = {}
    for name, fn in (("CPU_SOA", impl.cpu_soa), ("AVX", impl.cpu_avx), ("CPU_AOS", impl.cpu_aos)):
        t = sorted(fn(pm_end, eps2)[1] for _ in range(3))
        others[name] = round(n * n / (t[1] / 1e3) / 1e9, 3)
    out["other_cpu_rows"] = others
    # the 16384-body configuration
    pm1, vm1 = ics.uniform_ball(16384, seed=1)
    med1, steps1, algo1, _ = _cpu_rate(impl, integ, pm1, vm1, eps2, dt, damping, 2.0)
    out["n16384"] = {"value": round(16384.0 ** 2 / med1 / 1e9, 3), "ms_per_force": round(med1 * 1e3, 3),
                     "sample": f"{steps1} steps, {algo1}"}
    # with every logical CPU of the set, when that differs
    if logical != threads:
        omp.omp_set_num_threads(logical)
        medl, stepsl, _, _ = _cpu_rate(impl, integ, pm_host.copy(), vm_host.copy(), eps2, dt, damping,
                                       budget_s / 2, max_steps=10)
        out["all_logical_cpus"] = {"threads": logical, "value": round(n * n / medl / 1e9, 3), "steps": stepsl}
        omp.omp_set_num_threads(threads)
    if ensemble_sizes is not None:
        # one CPU step (force + integrate) of every ensemble system
        t_force = 0.0
        work = 0.0
        for k, m in enumerate(ensemble_sizes):
            p, v = ics.uniform_ball(int(m), seed=100000 + k)
            fn, _ = _cpu_force_fn(impl, int(m))
            f, ms = fn(p, eps2)
            integ.integrate_aos(p, v, f, dt, damping)
            t_force += ms / 1e3
            work += float(m) ** 2
        out["ensemble"] = {"value": round(work / t_force / 1e9, 3), "ms_per_step": round(t_force * 1e3, 2),
                           "sample": f"one step of each of the {len(ensemble_sizes)} systems, one after another "
                                     "(CPU_SOA, CPU_SOA_tiled where N % 4096 == 0)"}
    if plummer:
        # the N = 2^20 Plummer sphere extrapolated from one force on its first
        # 2^18 bodies (a quarter of the sphere's sample, 1/16 of the work):
        # the rate holds at this size, and the full force would take ~12 s
        p, _ = ics.plummer(1 << 20, seed=1)
        sub = np.ascontiguousarray(p[: 1 << 18])
        ms = impl.cpu_soa_tiled(sub, np.float32(1e-4))[1]
        rate = float(1 << 18) ** 2 / (ms / 1e3)
        out["plummer"] = {"value": round(rate / 1e9, 3),
                          "ms_per_force_extrapolated": round(float(1 << 20) ** 2 / rate * 1e3, 1),
                          "sample": "one CPU_SOA_tiled force on the first 2^18 bodies of the N = 2^20 Plummer sphere "
                                    "(eps^2 = 1e-4), extrapolated to N^2 = 2^40 interactions"}
    if served_reqs is not None:
        # one CPU step of every 20th request of the served trace
        t_force, work, picked, total = 0.0, 0.0, 0, 0.0
        for k in range(0, len(served_reqs), 20):
            r = served_reqs[k]
            p, _ = (ics.plummer(int(r.N), seed=int(r.seed)) if r.ic == 1
                    else ics.uniform_ball(int(r.N), seed=int(r.seed)))
            fn, _ = _cpu_force_fn(impl, int(r.N))
            t_force += fn(p, eps2)[1] / 1e3
            work += float(r.N) ** 2
            picked += 1
        total = sum(float(r.N) ** 2 * r.steps for r in served_reqs)
        rate = work / t_force
        out["served"] = {"value": round(rate / 1e9, 3),
                         "trace_cpu_time_s": round(total / rate, 1),
                         "sample": f"one force step of {picked} of the trace's {len(served_reqs)} requests "
                                   "(every 20th); trace_cpu_time_s = the trace's interactions at that rate"}
    return out


def crosscheck_cpu_aos(pm_host, force, f64, eps2):
    """The reference's crosscheck (src/nbody.cu:313-325,406-418) of the GPU
    force against CPU_AOS (its own build, or the oracle's restatement) on the
    same positions. Stated tolerance, per body: |F - F_AOS| / |F64| <= 1e-5 +
    CPU_AOS's own |F_AOS - F64| / |F64| (the triangle inequality through the
    FP64 sum). Also the reference's componentwise relError maximum, which has
    no bound (it divides by near-zero components)."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import numpy as np
    from oracle_lib import oracle, ref_cpu
    impl = ref_cpu() or oracle()
    aos, _ = impl.cpu_aos(np.ascontiguousarray(pm_host, np.float32), eps2)
    norm = np.linalg.norm(f64[:, :3], axis=1)
    d = np.linalg.norm(force[:, :3].astype(np.float64) - aos[:, :3], axis=1) / norm
    aos_err = np.linalg.norm(aos[:, :3].astype(np.float64) - f64[:, :3], axis=1) / norm
    bound = TOL_FP64 + aos_err
    return {"normwise_max": float(d.max()), "cpu_aos_vs_fp64_max": float(aos_err.max()),
            "tolerance": "per body |F-F_AOS|/|F64| <= 1e-5 + |F_AOS-F64|/|F64|",
            "worst_margin": float((d - bound).max()), "pass": bool(np.all(d <= bound)),
            "rel_error_max": float(oracle().max_rel_error(force, aos)),
            "golden": "reference CPU_AOS" if ref_cpu() is not None else "oracle CPU_AOS"}


KERNEL_NAMES = {0: "k_force_aos", 1: "k_force_lds", 2: "k_force_dpp", 3: "k_force_sgpr", 4: "k_force_sym"}


def pair_evaluations(plan, n):
    """Interaction evaluations one launch executes. The symmetric kernel
    evaluates each off-diagonal block pair once for both bodies and the
    diagonal blocks in full; the others evaluate all N^2. The roofline is
    credited with N^2 (20 flops each) either way."""
    if plan.variant != 4:
        return n * n
    b = plan.jslice
    nb = -(-n // b)
    return nb * (nb - 1) // 2 * b * b + nb * b * b


def traffic_from_profiles(workload_key):
    """HBM bytes per force launch measured by rocprofv3 counters (FETCH_SIZE,
    WRITE_SIZE, calibrated on a copy kernel), committed under profiles/."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    try:
        data = json.load(open(path))
        return data[workload_key]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


# ---- workloads ---------------------------------------------------------------

def ensemble_systems(count):
    from nbody_amd import ics
    sizes = ics.log_uniform_sizes(count, 256, 16384, seed=4)
    return sizes, [ics.uniform_ball(int(n), seed=100000 + k) for k, n in enumerate(sizes)]


def batched_force_errors(sizes, pm, vm, eps2, compensated=-1):
    """Per-body |F - F64|/|F64| of the batched force path (FP64 accumulation
    with compensated = 2) on the given states: the systems (bodies back to
    back, `sizes[k]` each) form one ensemble, one batched step computes each
    system's force on these positions, and
    the FP64 direct sum (GPU diagnostic) of every system is the reference.
    Returns (max, rms) over all bodies of all systems."""
    import numpy as np
    import torch
    from nbody_amd import _native
    L = _native.lib()
    nsz = np.ascontiguousarray(np.asarray(sizes, np.uint64))
    e2 = np.full(len(nsz), eps2, np.float32)
    pm = np.ascontiguousarray(pm, np.float32)
    vm = np.ascontiguousarray(vm, np.float32)
    total = int(nsz.sum())
    opt = _native.Options.make(compensated=compensated)
    h = L.nbody_hip_ensemble_create_opt(len(nsz), nsz.ctypes.data, pm.ctypes.data, vm.ctypes.data,
                                        e2.ctypes.data, 0.016, 0.95, ctypes.byref(opt))
    if not h:
        raise _native.NativeError(_native.last_error())
    force = np.zeros((total, 4), np.float32)
    try:
        _native.check(L.nbody_hip_ensemble_run(h, 1, _native.BATCHED, None), "ensemble step")
        _native.check(L.nbody_hip_ensemble_read(h, None, None, force.ctypes.data), "ensemble read")
    finally:
        L.nbody_hip_ensemble_destroy(h)
    dev = torch.device("cuda:0")
    pm_d = torch.from_numpy(pm).to(dev)
    f64 = torch.zeros((total, 4), dtype=torch.float64, device=dev)
    off = 0
    for n in nsz.tolist():
        _native.check_ms(L.nbody_hip_force_fp64(ctypes.c_void_p(f64.data_ptr() + 32 * off),
                                                ctypes.c_void_p(pm_d.data_ptr() + 16 * off), eps2, int(n)),
                         "fp64")
        off += int(n)
    torch.cuda.synchronize()
    ref = f64[:, :3].cpu().numpy()
    d = np.linalg.norm(force[:, :3].astype(np.float64) - ref, axis=1) / np.linalg.norm(ref, axis=1)
    del f64, pm_d
    return float(d.max()), float(np.sqrt(np.mean(d * d)))


_SETTLE = {}


def gpu_settle(ms):
    """About `ms` milliseconds of untimed force evaluations on a separate
    98304-body ball (made once), so that the measurement that follows starts
    at the GPU's sustained speed whatever ran (or idled) before it; the
    measured workload's own state is not touched. Returns the launches."""
    if ms <= 0:
        return 0
    import torch
    from nbody_amd import _native, ics
    L = _native.lib()
    if not _SETTLE:
        n = 98304
        pm_h, _ = ics.uniform_ball(n, seed=11)
        _SETTLE.update(n=n, pm=torch.from_numpy(pm_h).to("cuda:0"),
                       f=torch.zeros((n, 4), dtype=torch.float32, device="cuda:0"))
    n, pm, f = _SETTLE["n"], _SETTLE["pm"], _SETTLE["f"]
    launches = 0
    t_end = time.perf_counter() + ms / 1e3
    while time.perf_counter() < t_end:
        for _ in range(8):
            _native.check(L.nbody_hip_force(ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(pm.data_ptr()), 0.01, n,
                                            0, n, None, None), "settle")
        torch.cuda.synchronize()
        launches += 8
    return launches


def run_ensemble(count, steps, warmup, settle_ms=0.0):
    """`count` independent uniform balls, N log-uniform in [256, 16384]
    (seed 4), eps^2 = 0.01, dt = 0.016: batched steps replayed from a graph
    (the systems in two groups, each a persistent force launch and a
    slot-sum-and-integrate launch on its own stream), the same launches
    without a graph, and every system run on its own with its best
    single-system plan; then the batched graph and the per-system runs again
    with FP64 accumulation (compensated = 2). Each batched graph's force on
    its final state (and the FP32 one's on the initial state) is compared
    with the FP64 direct sum over every body."""
    import numpy as np
    from nbody_amd import _native
    L = _native.lib()
    sizes, systems = ensemble_systems(count)
    pm = np.ascontiguousarray(np.concatenate([s[0] for s in systems]))
    vm = np.ascontiguousarray(np.concatenate([s[1] for s in systems]))
    nsz = np.ascontiguousarray(sizes.astype(np.uint64))
    e2 = np.full(len(systems), 0.01, np.float32)
    inter = float((sizes.astype(np.float64) ** 2).sum())
    res, ends = {}, {}
    for name, mode, comp in (("batched_graph", _native.BATCHED_GRAPH, -1), ("batched", _native.BATCHED, -1),
                             ("per_system", _native.PER_SYSTEM, -1),
                             ("batched_graph_fp64", _native.BATCHED_GRAPH, 2),
                             ("per_system_fp64", _native.PER_SYSTEM, 2)):
        opt = _native.Options.make(compensated=comp)
        h = L.nbody_hip_ensemble_create_opt(len(systems), nsz.ctypes.data, pm.ctypes.data, vm.ctypes.data,
                                            e2.ctypes.data, 0.016, 0.95, ctypes.byref(opt))
        if not h:
            raise _native.NativeError(_native.last_error())
        try:
            ms = ctypes.c_float()
            _native.check(L.nbody_hip_ensemble_run(h, warmup, mode, ctypes.byref(ms)), name)
            gpu_settle(settle_ms)
            _native.check(L.nbody_hip_ensemble_run(h, steps, mode, ctypes.byref(ms)), name)
            res[name] = {"ms_per_step": round(ms.value / steps, 4),
                         "value": round(inter * steps / (ms.value * 1e-3) / 1e9, 2)}
            if name.startswith("batched_graph"):
                end_pm, end_vm = np.zeros_like(pm), np.zeros_like(vm)
                _native.check(L.nbody_hip_ensemble_read(h, end_pm.ctypes.data, end_vm.ctypes.data, None), "read")
                ends[name] = (end_pm, end_vm)
        finally:
            L.nbody_hip_ensemble_destroy(h)
    ic_max, ic_rms = batched_force_errors(sizes, pm, vm, 0.01)
    end_max, end_rms = batched_force_errors(sizes, *ends["batched_graph"], 0.01)
    end64_max, end64_rms = batched_force_errors(sizes, *ends["batched_graph_fp64"], 0.01, compensated=2)
    v = res["batched_graph"]["value"]
    v64 = res["batched_graph_fp64"]["value"]
    state = f"the batched force on every system after the {warmup} warm-up and {steps} timed steps"
    vs = "FP64 direct sum (GPU diagnostic) of every system, per-body |F-F64|/|F64| over all bodies"
    return {
        "value": v,
        "unit": "10^9 interactions/s",
        "ms_per_step": res["batched_graph"]["ms_per_step"],
        "steps": steps,
        "workload": f"{len(systems)} independent uniform balls, N log-uniform in [256, 16384] (seed 4; "
                    f"{int(sizes.sum())} bodies, sum N^2 = {inter:.4e}), eps^2=0.01, dt=0.016, damping=0.95",
        "modes": res,
        "settle_ms": settle_ms,
        "speedup_vs_per_system": round(v / res["per_system"]["value"], 2),
        "achieved_tflops": round(20 * v / 1e3, 2),
        "accuracy": {
            "max_rel": end_max, "rms_rel": end_rms, "tolerance": 1e-5, "pass": end_max <= 1e-5,
            "vs": vs, "state": state,
            "initial_state": {"max_rel": ic_max, "rms_rel": ic_rms, "pass": ic_max <= 1e-5},
        },
        "fp64_accumulation": {
            "value": v64,
            "ms_per_step": res["batched_graph_fp64"]["ms_per_step"],
            "speedup_vs_per_system": round(v64 / res["per_system_fp64"]["value"], 2),
            "what": "compensated = 2 in both modes: every term added in FP64, FP64 block partials, one rounding",
            "accuracy": {"max_rel": end64_max, "rms_rel": end64_rms, "tolerance": 1e-5, "pass": end64_max <= 1e-5,
                         "vs": vs, "state": state},
        },
    }


def served_trace(count, rate):
    from nbody_amd import serve
    return serve.poisson_trace(count, rate, 1024, 131072, 10, 200, seed=5)


def run_served(count, rate, max_active, compensated=-1):
    """A seeded Poisson stream (mean `rate` requests/s) of simulations with N
    log-uniform in [1024, 131072], steps uniform in [10, 200], uniform-ball or
    Plummer ICs, eps^2 = 0.01, under a 200 GB HBM budget, replayed in real
    time through the continuous-batching server. At 200/s the offered load
    (~3.7e13 interactions/s) exceeds one GPU several times over, so the
    latencies include the queue that builds up over the finite trace."""
    from nbody_amd import serve
    reqs = served_trace(count, rate)
    srv = serve.Server(budget_bytes=200 * 10 ** 9, max_active=max_active, compensated=compensated)
    try:
        res, wall, work = srv.run(reqs)
        peak = srv.peak_active()
    finally:
        srv.close()
    summ = serve.latency_summary(res, wall, work)
    offered = sum(float(r.N) ** 2 * r.steps for r in reqs) / max(reqs[len(reqs) - 1].arrival_s, 1e-9)
    acc = served_accuracy(reqs, compensated=compensated)
    return {
        "value": round(summ["interactions_per_s"] / 1e9, 2),
        "unit": "10^9 interactions/s",
        "p50_latency_s": round(summ["p50_latency_s"], 3),
        "p99_latency_s": round(summ["p99_latency_s"], 3),
        "p50_queue_wait_s": round(summ["p50_queue_wait_s"], 3),
        "p99_queue_wait_s": round(summ["p99_queue_wait_s"], 3),
        "completed": summ["completed"],
        "rejected": summ["rejected"],
        "wall_s": round(wall, 3),
        "peak_active": peak,
        "workload": f"{count} requests, Poisson {rate:g}/s, N log-uniform [1024, 131072], steps U[10, 200], ball "
                    f"or Plummer ICs, eps^2=0.01, dt=0.016, damping=0.95, 200 GB budget, {max_active} slots"
                    + (", FP64 accumulation" if compensated == 2 else ""),
        "offered_interactions_per_s": offered,
        "offered_over_served": round(offered / max(summ["interactions_per_s"], 1.0), 3),
        "achieved_tflops": round(20 * summ["interactions_per_s"] / 1e12, 2),
        "accuracy": acc,
    }, reqs


def served_accuracy(reqs, every=25, compensated=-1):
    """Force error on the final states of every `every`-th request of the
    trace. A request's final state does not depend on what else is resident,
    so these requests are replayed after the timed run, untimed, all
    arriving at once into a server that keeps their states; the batched
    force on each final state is then compared with the FP64 direct sum."""
    import numpy as np
    from nbody_amd import _native, serve
    picks = list(range(0, len(reqs), every))
    sub = (_native.Request * len(picks))()
    for j, k in enumerate(picks):
        r = reqs[k]
        sub[j] = _native.Request(r.id, r.N, r.steps, r.seed, r.ic, r.softeningSquared, 0.0)
    srv = serve.Server(budget_bytes=200 * 10 ** 9, max_active=len(picks), keep_states=True, compensated=compensated)
    try:
        res, _, _ = srv.run(sub)
        if any(x.status != 0 for x in res):
            raise RuntimeError("served accuracy replay: a request failed")
        states = [srv.final_state(int(sub[j].id), int(sub[j].N)) for j in range(len(picks))]
    finally:
        srv.close()
    eps2 = {float(sub[j].softeningSquared) for j in range(len(picks))}
    if len(eps2) != 1:
        raise RuntimeError("served accuracy replay: one softening expected")
    sizes = [int(sub[j].N) for j in range(len(picks))]
    mx, rms = batched_force_errors(sizes, np.concatenate([st[0] for st in states]),
                                   np.concatenate([st[1] for st in states]), eps2.pop(), compensated=compensated)
    return {"max_rel": mx, "rms_rel": rms, "tolerance": 1e-5, "pass": mx <= 1e-5,
            "vs": "FP64 direct sum (GPU diagnostic), per-body |F-F64|/|F64| over all bodies of the sampled requests",
            "sample": f"requests 0, {every}, {2 * every}, ... ({len(picks)} of {len(reqs)}; N {min(sizes)}.."
                      f"{max(sizes)}, {sum(sizes)} bodies): the batched force on each final state, the requests "
                      f"replayed untimed after the timed run"}


def run_plummer(steps):
    """N = 1048576 Plummer sphere (a = 1, total mass 1, velInvMass.w = 1),
    eps^2 = 1e-4, `steps` steps (force + integrate, graph replay), for: the
    default plan, the LDS kernel with TwoSum-compensated 128-body chunks, and
    the LDS kernel with one plain FP32 sum over all N per body. Each plan's
    force on the initial state is compared with the FP64 direct sum (GPU
    diagnostic) over all N bodies."""
    import torch
    from nbody_amd import _native, ics
    L = _native.lib()
    n = 1 << 20
    eps2 = 1e-4
    pm_h, vm_h = ics.plummer(n, seed=1)
    dev = torch.device("cuda:0")
    pm = torch.from_numpy(pm_h).to(dev)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    f64 = torch.zeros((n, 4), dtype=torch.float64, device=dev)
    _native.check_ms(L.nbody_hip_force_fp64(P(f64), P(pm), eps2, n), "fp64")
    ref = f64[:, :3]
    ref_norm = ref.norm(dim=1)
    plans = {
        "default": _native.Options.make(),
        "lds_compensated": _native.Options.make(variant=_native.LDS, compensated=1),
        "lds_plain_single_chain": _native.Options.make(variant=_native.LDS, compensated=0, i_per_lane=2, jslice=n),
    }
    out_modes = {}
    for name, opt in plans.items():
        force = torch.zeros((n, 4), dtype=torch.float32, device=dev)
        _native.check(L.nbody_hip_force(P(force), P(pm), eps2, n, 0, n, ctypes.byref(opt), None), name)
        torch.cuda.synchronize()
        d = (force[:, :3].double() - ref).norm(dim=1) / ref_norm
        h = L.nbody_hip_system_create(pm_h.ctypes.data, vm_h.ctypes.data, n, eps2, 0.016, 0.95, ctypes.byref(opt))
        if not h:
            raise _native.NativeError(_native.last_error())
        try:
            ms = ctypes.c_float()
            _native.check(L.nbody_hip_system_run(h, 1, 1, ctypes.byref(ms)), name)
            _native.check(L.nbody_hip_system_run(h, steps, 1, ctypes.byref(ms)), name)
            plan = _native.Options()
            L.nbody_hip_system_plan(h, ctypes.byref(plan))
        finally:
            L.nbody_hip_system_destroy(h)
        out_modes[name] = {
            "ms_per_step": round(ms.value / steps, 3),
            "value": round(float(n) * n * steps / (ms.value * 1e-3) / 1e9, 1),
            "max_rel": float(d.max()), "rms_rel": float(d.pow(2).mean().sqrt()),
            "plan": {"variant": _native.VARIANT_NAMES[plan.variant], "i_per_lane": plan.i_per_lane,
                     "jslice": plan.jslice, "compensated": int(plan.compensated)},
        }
    del f64, pm
    torch.cuda.empty_cache()
    best = out_modes["default"]
    return {
        "value": best["value"], "unit": "10^9 interactions/s", "steps": steps, "ms_per_step": best["ms_per_step"],
        "workload": f"single system, N={n} Plummer sphere (a=1, M=1, invMass=1), eps^2={eps2}, dt=0.016, "
                    f"damping=0.95; errors of the initial force vs the FP64 direct sum over all N",
        "modes": out_modes,
    }


def run_general_mass(n, launches, settle_ms=0.0):
    """The reference's own initial conditions (seedRandom(7) +
    randomUnitBodies(scale 0.9, velscale 8), src/nbody_util.cpp:102-144,
    masses U(0.5, 1)) at the headline N, softening 0.1^2: the default plan's
    force kernel timed over `launches` launches (events around the kernel),
    and its error against the FP64 direct sum. Unequal masses take the
    symmetric kernel's general pair weights (16 VALU + 1 rsq per pair against
    14 + 1 for one common mass)."""
    import numpy as np
    import torch
    from nbody_amd import _native, ics
    L = _native.lib()
    eps2 = float(np.float32(0.1) * np.float32(0.1))
    pm_h, _ = ics.random_unit_bodies(n, seed=7)
    dev = torch.device("cuda:0")
    pm = torch.from_numpy(pm_h).to(dev)
    force = torch.zeros((n, 4), dtype=torch.float32, device=dev)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    opt = _native.Options.make()
    for _ in range(3):
        _native.check(L.nbody_hip_force(P(force), P(pm), eps2, n, 0, n, ctypes.byref(opt), None), "force")
    torch.cuda.synchronize()
    settle = gpu_settle(settle_ms)
    _native.check(L.nbody_hip_kernel_timer(1), "timer")
    t0 = time.perf_counter()
    for _ in range(launches):
        _native.check(L.nbody_hip_force(P(force), P(pm), eps2, n, 0, n, ctypes.byref(opt), None), "force")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms, nl = ctypes.c_double(), ctypes.c_int()
    _native.check(L.nbody_hip_kernel_timer_read(ctypes.byref(kms), ctypes.byref(nl)), "timer read")
    L.nbody_hip_kernel_timer(0)
    kernel_ms = kms.value / max(nl.value, 1)
    f64 = torch.zeros((n, 4), dtype=torch.float64, device=dev)
    _native.check_ms(L.nbody_hip_force_fp64(P(f64), P(pm), eps2, n), "fp64")
    torch.cuda.synchronize()
    d = (force[:, :3].double() - f64[:, :3]).norm(dim=1) / f64[:, :3].norm(dim=1)
    plan = _native.Options()
    _native.check(L.nbody_hip_plan(n, n, ctypes.byref(opt), ctypes.byref(plan)), "plan")
    inter = float(n) * n
    tf = 20.0 * inter / (kernel_ms * 1e-3) / 1e12
    return {
        "value": round(inter * launches / wall / 1e9, 2), "unit": "10^9 interactions/s",
        "workload": f"reference ICs randomUnitBodies(seed 7), N={n}, masses U(0.5, 1), eps^2=0.1^2; force "
                    f"evaluations (kernel + slot sum) back to back",
        "kernel_ms": round(kernel_ms, 4),
        "settle_launches": settle,
        "roofline": {"achieved": round(tf, 2), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tf / PEAK_FP32_TFLOPS, 4), "kernel": KERNEL_NAMES[plan.variant]},
        "plan": {"variant": _native.VARIANT_NAMES[plan.variant], "i_per_lane": plan.i_per_lane,
                 "jslice": plan.jslice},
        "accuracy": {"max_rel": float(d.max()), "rms_rel": float(d.pow(2).mean().sqrt()),
                     "vs": "FP64 direct sum over all N"},
    }


def bench_share(args):
    """One GPU's share of a single system split by i-range over --shares GPUs
    (the multi-GPU decomposition of src/nbody_multiGPU.cu:87-124): bodies
    [base, base + N/G) against all N through nbody_hip_force's base/n, the
    one-sided plans timed over ipl and j-slice lengths, the fastest reported.
    Positions would be all-gathered between steps; that exchange is not part
    of this one-GPU measurement."""
    import torch
    from nbody_amd import _native, ics
    L = _native.lib()
    n, g = args.n, args.shares
    if g < 2 or n % g:
        raise SystemExit(f"--shares must be >= 2 and divide N (got {g}, N = {n})")
    count = n // g
    base = (g // 2) * count
    eps2 = 0.01
    pm_h, _ = ics.uniform_ball(n, seed=1)
    dev = torch.device("cuda:0")
    pm = torch.from_numpy(pm_h).to(dev)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    force = torch.zeros((count, 4), dtype=torch.float32, device=dev)
    launches = max(args.steps, 5)
    results = []
    shapes = []
    for variant in (_native.LDS, _native.SGPR, _native.DPP):
        for ipl in (1, 2, 4):
            auto = _native.Options()
            o = _native.Options.make(variant=variant, i_per_lane=ipl, compensated=1)
            _native.check(L.nbody_hip_plan(n, count, ctypes.byref(o), ctypes.byref(auto)), "plan")
            for mul in (1, 2, 4, 8):
                if mul == 1 or auto.jslice * mul <= n:
                    shapes.append((variant, ipl, min(auto.jslice * mul, n)))
    for variant, ipl, jslice in shapes:
        opt = _native.Options.make(variant=variant, i_per_lane=ipl, jslice=jslice, compensated=1)
        for _ in range(3):
            _native.check(L.nbody_hip_force(P(force), P(pm), eps2, n, base, count, ctypes.byref(opt), None), "share")
        torch.cuda.synchronize()
        _native.check(L.nbody_hip_kernel_timer(1), "timer")
        t0 = time.perf_counter()
        for _ in range(launches):
            _native.check(L.nbody_hip_force(P(force), P(pm), eps2, n, base, count, ctypes.byref(opt), None), "share")
        torch.cuda.synchronize()
        call_ms = (time.perf_counter() - t0) * 1e3 / launches
        kms, nl = ctypes.c_double(), ctypes.c_int()
        _native.check(L.nbody_hip_kernel_timer_read(ctypes.byref(kms), ctypes.byref(nl)), "timer read")
        L.nbody_hip_kernel_timer(0)
        plan = _native.Options()
        _native.check(L.nbody_hip_plan(n, count, ctypes.byref(opt), ctypes.byref(plan)), "plan")
        results.append((call_ms, kms.value / nl.value, plan))
    # the symmetric plan split by block columns instead: share k of g, every
    # share's kernel time (the slowest one sets an 8-GPU step), FP64 partial
    # forces summed over the shares against the FP64 direct sum
    f64share = torch.zeros((n, 4), dtype=torch.float64, device=dev)

    def pair_shares(sopt):
        """Every share's call and kernel time with symmetric options sopt."""
        calls, kernels = [], []
        for k in range(g):
            for _ in range(3):
                _native.check(L.nbody_hip_force_pair_share(P(f64share), P(pm), eps2, n, k, g, ctypes.byref(sopt),
                                                           None), "pair share")
            torch.cuda.synchronize()
            _native.check(L.nbody_hip_kernel_timer(1), "timer")
            t0 = time.perf_counter()
            for _ in range(launches):
                _native.check(L.nbody_hip_force_pair_share(P(f64share), P(pm), eps2, n, k, g, ctypes.byref(sopt),
                                                           None), "pair share")
            torch.cuda.synchronize()
            calls.append((time.perf_counter() - t0) * 1e3 / launches)
            kms, nl = ctypes.c_double(), ctypes.c_int()
            _native.check(L.nbody_hip_kernel_timer_read(ctypes.byref(kms), ctypes.byref(nl)), "timer read")
            L.nbody_hip_kernel_timer(0)
            kernels.append(kms.value / nl.value)
        return calls, kernels

    # the block shapes of the symmetric plan, the one whose slowest share is fastest kept
    sweep = {}
    for ipl, block in ((8, 512), (4, 256), (2, 128), (4, 512), (8, 1024)):
        sopt = _native.Options.make(variant=_native.SYMMETRIC, i_per_lane=ipl, jslice=block)
        sweep[(ipl, block)] = pair_shares(sopt)
    best_shape = min(sweep, key=lambda s: max(sweep[s][1]))
    share_ms, share_kernel_ms = sweep[best_shape]
    sopt = _native.Options.make(variant=_native.SYMMETRIC, i_per_lane=best_shape[0], jslice=best_shape[1])
    total = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    for k in range(g):
        _native.check(L.nbody_hip_force_pair_share(P(f64share), P(pm), eps2, n, k, g, ctypes.byref(sopt), None),
                      "pair share")
        total += f64share[:, :3]
    ms, kernel_ms, plan = min(results, key=lambda r: r[0])
    opt = _native.Options.make(variant=plan.variant, i_per_lane=plan.i_per_lane, jslice=plan.jslice,
                               compensated=plan.compensated)
    _native.check(L.nbody_hip_force(P(force), P(pm), eps2, n, base, count, ctypes.byref(opt), None), "share")
    f64 = torch.zeros((n, 4), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    _native.check_ms(L.nbody_hip_force_fp64(P(f64), P(pm), eps2, n), "fp64")
    ref = f64[base:base + count, :3]
    d = (force[:, :3].double() - ref).norm(dim=1) / ref.norm(dim=1)
    inter = float(count) * n
    tflops = 20.0 * inter / (kernel_ms * 1e-3) / 1e12
    ds = (total.float().double() - f64[:, :3]).norm(dim=1) / f64[:, :3].norm(dim=1)
    slowest = max(share_kernel_ms)
    pair_share = {
        "what": f"symmetric plan split over {g} GPUs by block columns (nbody_hip_force_pair_share): share k "
                f"evaluates about 1/{g} of the block pairs for both bodies and writes FP64 partial forces of all "
                f"N bodies, to be summed across GPUs by a reduce-scatter",
        "shape": {"i_per_lane": best_shape[0], "block": best_shape[1]},
        "slowest_kernel_ms_by_shape": {f"ipl{a}/B{b}": round(max(v[1]), 4) for (a, b), v in sweep.items()},
        "ms_per_call": [round(t, 4) for t in share_ms],
        "kernel_ms": [round(t, 4) for t in share_kernel_ms],
        "slowest_kernel_ms": round(slowest, 4),
        "credited_tflops_slowest": round(20.0 * float(n) * n / g / (slowest * 1e-3) / 1e12, 2),
        "frac_slowest": round(20.0 * float(n) * n / g / (slowest * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
        "accuracy_of_the_sum": {"max_rel": float(ds.max()), "rms_rel": float(ds.pow(2).mean().sqrt())},
    }
    return {
        "metric": METRIC, "value": round(inter / (ms * 1e-3) / 1e9, 2), "unit": "10^9 interactions/s",
        "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic (seeded uniform ball, unit masses)",
        "config": {"workload": f"one GPU's share of N={n} split over {g} GPUs: bodies [{base}, {base + count}) "
                               f"against all N (nbody_hip_force base/n), eps^2={eps2}; whole calls (force "
                               f"kernel + slice sum), best of the one-sided plans over {launches} calls each"},
        "ms_per_call": round(ms, 4),
        "kernel_ms": round(kernel_ms, 4),
        "roofline": {"bound": "valu", "achieved": round(tflops, 2), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tflops / PEAK_FP32_TFLOPS, 4), "kernel": KERNEL_NAMES[plan.variant]},
        "plan": {"variant": _native.VARIANT_NAMES[plan.variant], "i_per_lane": plan.i_per_lane,
                 "jslice": plan.jslice, "compensated": bool(plan.compensated)},
        "accuracy": {"max_rel": float(d.max()), "rms_rel": float(d.pow(2).mean().sqrt()),
                     "vs": "FP64 direct sum over the share's bodies"},
        "all_plans_ms_per_call": {f"{_native.VARIANT_NAMES[p.variant]}/ipl{p.i_per_lane}/js{p.jslice}": round(t, 4)
                                  for t, _, p in results},
        "pair_share": pair_share,
    }


def standalone(body, data, workload):
    """A workload's own line (--workload ensemble / served / plummer)."""
    out = {"metric": METRIC, "value": body["value"], "unit": body["unit"], "n_gpus": 1, "higher_is_better": True,
           "dtype": "f32", "data": data, "config": {"workload": body["workload"]}}
    out.update({k: v for k, v in body.items() if k not in ("value", "unit", "workload")})
    return out


def single_system(args, L, _native, ics, np, torch):
    """The headline line: one system of N bodies, K timed steps."""
    n, K, W = args.n, args.steps, args.warmup
    eps2, dt, damping = 0.01, 0.016, 0.95
    pm_h, vm_h = ics.uniform_ball(n, seed=1)
    dev = torch.device("cuda:0")
    pm = torch.from_numpy(pm_h).to(dev)
    vm = torch.from_numpy(vm_h).to(dev)
    force = torch.zeros((n, 4), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    variant = {"auto": _native.AUTO, "sgpr": _native.SGPR, "lds": _native.LDS, "dpp": _native.DPP,
               "aos": _native.AOS, "symmetric": _native.SYMMETRIC}[args.variant]
    opt = _native.Options.make(variant=variant, i_per_lane=args.ipl, jslice=args.jslice,
                               compensated=args.compensated, waves_per_simd=args.waves)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    pinned = args.ipl > 0 or args.jslice > 0 or args.waves > 0

    if not args.no_autotune and not pinned and variant != _native.AOS:
        _native.check_ms(L.nbody_hip_autotune(P(force), P(pm), eps2, n, variant), "autotune")
    plan = _native.Options()
    _native.check(L.nbody_hip_plan(n, n, ctypes.byref(opt), ctypes.byref(plan)), "plan")

    def step():
        _native.check(L.nbody_hip_force(P(force), P(pm), eps2, n, 0, n, ctypes.byref(opt), stream), "force")
        _native.check(L.nbody_hip_integrate(P(pm), P(vm), P(force), dt, damping, n, stream), "integrate")

    def accuracy():
        """This plan's force on the current positions against the FP64 direct sum."""
        f32 = torch.zeros_like(force)
        _native.check(L.nbody_hip_force(P(f32), P(pm), eps2, n, 0, n, ctypes.byref(opt), stream), "force")
        f64 = torch.zeros((n, 4), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        _native.check_ms(L.nbody_hip_force_fp64(P(f64), P(pm), eps2, n), "fp64")
        torch.cuda.synchronize()
        d = (f32[:, :3].double() - f64[:, :3]).norm(dim=1) / f64[:, :3].norm(dim=1)
        return float(d.max()), float(d.pow(2).mean().sqrt()), f32, f64

    # the configuration as specified (initial conditions), before any step
    ic_max, ic_rms, ic_force, ic_f64 = accuracy()
    ic_pm = pm.cpu().numpy()
    ic_force, ic_f64 = ic_force.cpu().numpy(), ic_f64.cpu().numpy()

    # settle: force evaluations on the initial positions into a spare buffer
    # (the state does not move), until about settle_ms have passed
    settle = 0
    if args.settle_ms > 0:
        spare = torch.zeros_like(force)
        t_end = time.perf_counter() + args.settle_ms / 1e3
        while time.perf_counter() < t_end:
            for _ in range(8):
                _native.check(L.nbody_hip_force(P(spare), P(pm), eps2, n, 0, n, ctypes.byref(opt), stream), "settle")
            torch.cuda.synchronize()
            settle += 8
        del spare
    for _ in range(W):
        step()
    torch.cuda.synchronize()
    _native.check(L.nbody_hip_kernel_timer(1), "timer")
    t0 = time.perf_counter()
    for _ in range(K):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kms = ctypes.c_double()
    launches = ctypes.c_int()
    _native.check(L.nbody_hip_kernel_timer_read(ctypes.byref(kms), ctypes.byref(launches)), "timer read")
    L.nbody_hip_kernel_timer(0)

    # accuracy on the final state (the last step's positions are gone; recompute there)
    max_rel, rms_rel, step_force, f64 = accuracy()

    inter = float(n) * n
    value = inter * K / elapsed / 1e9
    kernel_ms = kms.value / max(launches.value, 1)
    achieved = 20.0 * inter / (kernel_ms * 1e-3) / 1e12
    cus = ctypes.c_int()
    clk = ctypes.c_int()
    arch = ctypes.create_string_buffer(64)
    L.nbody_hip_device_info(ctypes.byref(cus), ctypes.byref(clk), arch, 64)
    workload = f"uniform_ball_n{n}"

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "10^9 interactions/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(elapsed * 1e3 / K, 4),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": round(value / REF_BEST_GPU, 2),
        "dtype": "f32",
        "data": "synthetic (seeded uniform ball, unit masses)",
        "settle_launches": settle,
        "config": {
            "workload": f"single system, N={n} uniform ball radius {ics.ball_radius(n):.4f}, unit masses, "
                        f"eps^2={eps2}, dt={dt}, damping={damping}; step = force + integrate, state in HBM",
            "n_bodies": n,
            "interactions_per_step": int(inter),
            "plan": {"variant": _native.VARIANT_NAMES[plan.variant], "i_per_lane": plan.i_per_lane,
                     "jslice": plan.jslice, "compensated": int(plan.compensated), "newton": bool(plan.newton),
                     "waves_per_simd": plan.waves_per_simd},
            "vs_baseline_ref": "GPU_Shared, RTX 2080 Ti, clang build: 501.5 x10^9 int/s (README.md:230-233)",
            "device": f"{arch.value.decode()} {cus.value} CUs {clk.value // 1000} MHz",
        },
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved, 2),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic_from_profiles(workload),
            "kernel_ms": round(kernel_ms, 4),
            "kernel": KERNEL_NAMES[plan.variant],
            "flops_per_launch": 20 * int(inter),
            "pair_evaluations_per_launch": pair_evaluations(plan, n),
        },
        "accuracy": {"max_rel": max_rel, "rms_rel": rms_rel, "tolerance": TOL_FP64,
                     "pass": bool(max_rel <= TOL_FP64),
                     "vs": "FP64 direct sum (GPU diagnostic), per-body |F-F64|/|F64|",
                     "state": f"after the {W} warm-up and {K} timed steps",
                     "initial_state": {"max_rel": ic_max, "rms_rel": ic_rms, "pass": bool(ic_max <= TOL_FP64)}},
    }
    # the host-pointer table entry (reference AOS contract): H2D of posMass, the
    # force kernel and D2H of the forces, the PCIe-inclusive rate of one evaluation
    host_f = np.zeros((n, 4), np.float32)
    host_pm = np.ascontiguousarray(pm.cpu().numpy())
    hms = []
    for _ in range(6):
        t0 = time.perf_counter()
        _native.check_ms(L.ComputeGravitation_HIP_Host(host_f.ctypes.data, host_pm.ctypes.data, eps2, n),
                         "ComputeGravitation_HIP_Host")
        hms.append((time.perf_counter() - t0) * 1e3)
    host_ms = sorted(hms[1:])[len(hms[1:]) // 2]
    out["host_entry"] = {"ms_per_call": round(host_ms, 4), "value": round(inter / (host_ms * 1e-3) / 1e9, 2),
                         "unit": "10^9 interactions/s",
                         "what": "ComputeGravitation_HIP_Host on host buffers: H2D of posMass, force kernel, D2H of "
                                 "x, y, z; median of 5"}
    crosscheck = None
    if not args.no_cpu_baseline:
        crosscheck = (
            crosscheck_cpu_aos(ic_pm, ic_force, ic_f64, np.float32(eps2)),
            crosscheck_cpu_aos(pm.cpu().numpy(), step_force.cpu().numpy(), f64.cpu().numpy(), np.float32(eps2)))
        out["accuracy"]["vs_cpu_aos"] = crosscheck[1]
        out["accuracy"]["initial_state"]["vs_cpu_aos"] = crosscheck[0]
    del force, pm, vm, f64, step_force
    torch.cuda.empty_cache()
    return out, pm_h, vm_h


def main():
    args = parse()
    # the CPU topology before anything binds this thread (see cpu_baseline)
    topology = physical_cores()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path. "
              "The single-system work would shard by i-range (nbody_hip_force base/n) with an all-gather "
              "of positions per step.", file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("OMP_PROC_BIND", "spread")

    import numpy as np
    import torch
    sys.path.insert(0, HERE)
    from nbody_amd import _native, ics

    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": src.numel() * 4, "bytes_written": dst.numel() * 4,
                          "launches": 5}))
        return

    if args.workload == "ensemble":
        body = run_ensemble(args.systems, args.ensemble_steps, 5, args.settle_ms)
        print(json.dumps(standalone(body, "synthetic (seeded uniform balls, unit masses)", "ensemble")))
        return
    if args.workload == "served":
        body, _ = run_served(args.requests, args.rate, args.max_active, args.compensated)
        print(json.dumps(standalone(body, "synthetic (seeded uniform-ball / Plummer requests)", "served")))
        return
    if args.workload == "plummer":
        body = run_plummer(args.plummer_steps)
        print(json.dumps(standalone(body, "synthetic (seeded Plummer sphere)", "plummer")))
        return
    if args.workload == "share":
        print(json.dumps(bench_share(args)))
        return

    L = _native.lib()
    out, pm_h, vm_h = single_system(args, L, _native, ics, np, torch)
    ens_sizes = reqs = None
    if not args.single_only:
        out["general_mass"] = run_general_mass(args.n, 20, args.settle_ms)
        out["ensemble"] = run_ensemble(args.systems, args.ensemble_steps, 5, args.settle_ms)
        out["served"], reqs = run_served(args.requests, args.rate, args.max_active)
        served64, _ = run_served(args.requests, args.rate, args.max_active, compensated=2)
        out["served"]["fp64_accumulation"] = {
            k: served64[k] for k in ("value", "p50_latency_s", "p99_latency_s", "p99_queue_wait_s", "completed",
                                     "wall_s", "offered_over_served", "accuracy")}
        out["served"]["fp64_accumulation"]["what"] = "the same trace with compensated = 2: FP64 block partials"
        out["plummer"] = run_plummer(args.plummer_steps)
        ens_sizes, _ = ensemble_systems(args.systems)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pm_h, vm_h, np.float32(0.01), 0.016, 0.95, args.cpu_budget_s, topology,
                                           ensemble_sizes=ens_sizes, served_reqs=reqs, plummer=ens_sizes is not None)
        if ens_sizes is not None:
            out["ensemble"]["vs_cpu"] = round(out["ensemble"]["value"] / out["cpu_baseline"]["ensemble"]["value"], 1)
            out["served"]["vs_cpu"] = round(out["served"]["value"] / out["cpu_baseline"]["served"]["value"], 1)
            out["plummer"]["vs_cpu"] = round(out["plummer"]["value"] / out["cpu_baseline"]["plummer"]["value"], 1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
