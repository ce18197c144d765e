"""Benchmark of the MI355X force path on the single-system workload.

    python bench.py [--steps K] [--warmup W] [--n N] [--variant NAME] [--no-cpu-baseline]

Workload: one system of N = 98304 bodies, positions uniform in a ball of
radius 1.54 (N/16384)^(1/3), unit masses, velocities uniform in a unit ball
(seeded), softeningSquared = 0.01, dt = 0.016, damping 0.95 (src/nbody.cu:152).
A step is one force evaluation over all N^2 pairs plus the integrator, with
the state resident in HBM. The timed region is K steps between device
synchronisations; the line printed is

    value = N^2 K / time / 1e9   (10^9 interactions/s, the reference's figure,
                                  src/nbody.cu:1026-1027)

plus the roofline of the force kernel (20 flops per interaction against the
FP32 vector peak; the kernel time comes from events the library records
around the force kernel itself, on its stream, over the timed region), the
accuracy of the last force against the FP64 direct sum (the library's GPU
diagnostic), and a CPU baseline: the reference's CPU_SOA_tiled
(src/nbody_CPU_SOA_tiled.cpp:51-115) compiled from its sources when that
build is present, else the oracle's restatement of it, timed on this host's
cores on a bounded number of the same steps.

Other workloads (--workload): ensemble (512 systems, batched persistent
launches per step), served (a Poisson request stream), plummer (N = 2^20), share (one
GPU's i-range share of the N = 98304 system split over --shares GPUs).
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
    ap.add_argument("--compensated", type=int, default=-1, choices=[-1, 0, 1])
    ap.add_argument("--waves", type=int, default=0, help="pin the symmetric plan's waves per SIMD (skips the autotuner)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy 5 times (counter calibration for --pmc runs)")
    ap.add_argument("--workload", default="single", choices=["single", "ensemble", "served", "plummer", "share"],
                    help="single: one N-body system (the headline line); ensemble: 512 independent systems; "
                         "served: a Poisson stream of simulation requests; plummer: N = 2^20 Plummer sphere, "
                         "plain vs compensated sums")
    ap.add_argument("--shares", type=int, default=8, help="share: GPUs the single system would be split over")
    ap.add_argument("--systems", type=int, default=512)
    ap.add_argument("--requests", type=int, default=400)
    ap.add_argument("--max-active", type=int, default=64)
    ap.add_argument("--rate", type=float, default=200.0, help="served: mean request arrivals per second")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget-s", type=float, default=15.0)
    return ap.parse_args()


def cpu_baseline(pm_host, vm_host, eps2, dt, damping, budget_s):
    """Time CPU_SOA_tiled + integrate on this host (rank 0, N = 1 only)."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import numpy as np
    from oracle_lib import oracle, ref_cpu
    n = pm_host.shape[0]
    ref = ref_cpu()
    kind, impl = ("reference", ref) if ref is not None else ("port", oracle())
    integ = oracle()
    pm, vm = pm_host.copy(), vm_host.copy()
    zero = np.zeros((n, 4), np.float32)
    times = []
    # the reference's CPU_SOA_tiled adds into a force zeroed by the caller (src/nbody.cu:328-331)
    f, ms = impl.cpu_soa_tiled(pm, eps2) if n % 4096 == 0 else impl.cpu_soa(pm, eps2)  # warm-up
    spent = 0.0
    steps = 0
    while steps < 20:
        t0 = time.perf_counter()
        f, ms = impl.cpu_soa_tiled(pm, eps2) if n % 4096 == 0 else impl.cpu_soa(pm, eps2)
        pm, vm = integ.integrate_aos(pm, vm, f if f is not None else zero, dt, damping)
        times.append(ms / 1e3)
        spent += time.perf_counter() - t0
        steps += 1
        if steps >= 3 and spent + spent / steps > budget_s:
            break
    med = float(np.median(times))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    algo = "CPU_SOA_tiled" if n % 4096 == 0 else "CPU_SOA"
    # the reference's other CPU rows on the same positions, three calls each
    others = {}
    for name, fn in (("CPU_SOA", impl.cpu_soa), ("AVX", impl.cpu_avx), ("CPU_AOS", impl.cpu_aos)):
        t = sorted(fn(pm, eps2)[1] for _ in range(3))
        others[name] = round(n * n / (t[1] / 1e3) / 1e9, 3)
    return {
        "value": round(n * n / med / 1e9, 3),
        "unit": "10^9 interactions/s",
        "cores": threads,
        "kind": kind,
        "sample": f"{steps} of the workload's steps ({algo} force, median of its own steady_clock ms, "
                  f"+ host integrate), N={n}, OMP_NUM_THREADS={threads}, OMP_PROC_BIND="
                  f"{os.environ.get('OMP_PROC_BIND', '')}",
        "ms_per_force": round(med * 1e3, 2),
        "cpu": cpu_model(),
        "other_cpu_rows": others,
    }


def crosscheck_cpu_aos(pm_host, force, f64, eps2):
    """The reference's crosscheck (src/nbody.cu:313-325,406-418) of the GPU
    force against CPU_AOS (its own build, or the oracle's restatement) on the
    same positions: its componentwise relError maximum, and the per-body
    normwise difference scaled by |F64|."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import numpy as np
    from oracle_lib import oracle, ref_cpu
    impl = ref_cpu() or oracle()
    aos, _ = impl.cpu_aos(np.ascontiguousarray(pm_host, np.float32), eps2)
    d = np.linalg.norm(force[:, :3].astype(np.float64) - aos[:, :3], axis=1) / np.linalg.norm(f64[:, :3], axis=1)
    return {"rel_error_max": float(oracle().max_rel_error(force, aos)), "normwise_max": float(d.max()),
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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def traffic_from_profiles(workload_key):
    """HBM bytes per force launch measured by rocprofv3 counters (FETCH_SIZE,
    WRITE_SIZE, calibrated on a copy kernel), committed under profiles/."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    try:
        data = json.load(open(path))
        return data[workload_key]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def bench_ensemble(args):
    """512 independent uniform balls, N log-uniform in [256, 16384] (seeded),
    eps^2 = 0.01, dt = 0.016, 50 steps: batched steps replayed from a graph
    (the systems in two groups, each a persistent force launch and a
    slot-sum-and-integrate launch on its own stream), against every system
    run on its own with its best single-system plan."""
    import numpy as np
    from nbody_amd import _native, ics
    L = _native.lib()
    steps = args.steps if args.steps != 20 else 50
    sizes = ics.log_uniform_sizes(args.systems, 256, 16384, seed=4)
    systems = [ics.uniform_ball(int(n), seed=100000 + k) for k, n in enumerate(sizes)]
    pm = np.ascontiguousarray(np.concatenate([s[0] for s in systems]))
    vm = np.ascontiguousarray(np.concatenate([s[1] for s in systems]))
    nsz = np.ascontiguousarray(sizes.astype(np.uint64))
    e2 = np.full(len(systems), 0.01, np.float32)
    inter = float((sizes.astype(np.float64) ** 2).sum())
    res = {}
    for name, mode in (("batched_graph", _native.BATCHED_GRAPH), ("batched", _native.BATCHED),
                       ("per_system", _native.PER_SYSTEM)):
        h = L.nbody_hip_ensemble_create(len(systems), nsz.ctypes.data, pm.ctypes.data, vm.ctypes.data,
                                        e2.ctypes.data, 0.016, 0.95)
        if not h:
            raise _native.NativeError(_native.last_error())
        try:
            ms = ctypes.c_float()
            _native.check(L.nbody_hip_ensemble_run(h, max(args.warmup // 10, 2), mode, ctypes.byref(ms)), name)
            _native.check(L.nbody_hip_ensemble_run(h, steps, mode, ctypes.byref(ms)), name)
            res[name] = {"ms": ms.value, "value": inter * steps / (ms.value * 1e-3) / 1e9}
        finally:
            L.nbody_hip_ensemble_destroy(h)
    v = res["batched_graph"]["value"]
    out = {
        "metric": METRIC,
        "value": round(v, 2),
        "unit": "10^9 interactions/s",
        "n_gpus": 1,
        "steps": steps,
        "ms_per_step": round(res["batched_graph"]["ms"] / steps, 4),
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic (seeded uniform balls, unit masses)",
        "config": {"workload": f"ensemble of {len(systems)} independent systems, N log-uniform in [256, 16384] "
                               f"(seed 4; total {int(sizes.sum())} bodies, sum N^2 = {inter:.3e}), eps^2=0.01, "
                               f"dt=0.016, damping=0.95, {steps} steps",
                   "systems": len(systems)},
        "modes": {k: {"ms_per_step": round(r["ms"] / steps, 4), "value": round(r["value"], 2)}
                  for k, r in res.items()},
        "speedup_vs_per_system": round(v / res["per_system"]["value"], 2),
        "achieved_tflops": round(20 * v / 1e3, 2),
    }
    print(json.dumps(out))


def bench_served(args):
    """A seeded Poisson stream (mean --rate requests/s, default 200) of
    simulations with N log-uniform in [1024, 131072], steps uniform in
    [10, 200], uniform-ball or Plummer ICs, eps^2 = 0.01, under a 200 GB HBM
    budget, replayed in real time through the continuous-batching server. At
    200/s the offered load (~3.7e13 interactions/s) exceeds one GPU several
    times over, so the latencies include the queue that builds up over the
    finite trace; --rate 20 offers about 85% of one GPU."""
    from nbody_amd import serve
    reqs = serve.poisson_trace(args.requests, args.rate, 1024, 131072, 10, 200, seed=5)
    srv = serve.Server(budget_bytes=200 * 10 ** 9, max_active=args.max_active)
    try:
        res, wall, work = srv.run(reqs)
        peak = srv.peak_active()
    finally:
        srv.close()
    summ = serve.latency_summary(res, wall, work)
    offered = sum(float(r.N) ** 2 * r.steps for r in reqs) / max(reqs[len(reqs) - 1].arrival_s, 1e-9)
    out = {
        "metric": METRIC,
        "value": round(summ["interactions_per_s"] / 1e9, 2),
        "unit": "10^9 interactions/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic (seeded uniform-ball / Plummer requests)",
        "config": {"workload": f"served stream: {args.requests} requests, Poisson {args.rate:g}/s, N log-uniform "
                               "[1024, 131072], steps U[10, 200], ball or Plummer ICs, eps^2=0.01, dt=0.016, "
                               f"damping=0.95, 200 GB budget, {args.max_active} slots",
                   "offered_interactions_per_s": offered,
                   "offered_over_served": round(offered / max(summ["interactions_per_s"], 1.0), 3)},
        "p50_latency_s": round(summ["p50_latency_s"], 3),
        "p99_latency_s": round(summ["p99_latency_s"], 3),
        "p50_queue_wait_s": round(summ["p50_queue_wait_s"], 3),
        "p99_queue_wait_s": round(summ["p99_queue_wait_s"], 3),
        "completed": summ["completed"],
        "rejected": summ["rejected"],
        "wall_s": round(wall, 3),
        "peak_active": peak,
        "achieved_tflops": round(20 * summ["interactions_per_s"] / 1e12, 2),
    }
    print(json.dumps(out))


def bench_plummer(args):
    """N = 1048576 Plummer sphere (a = 1, total mass 1, velInvMass.w = 1),
    eps^2 = 1e-4, 5 steps (force + integrate, graph replay), for: the default
    plan, the LDS kernel with TwoSum-compensated 128-body chunks, and the LDS
    kernel with one plain FP32 sum over all N per body. Each plan's force on
    the initial state is compared with the FP64 direct sum (GPU diagnostic)
    over all N bodies."""
    import numpy as np
    import torch
    from nbody_amd import _native, ics
    L = _native.lib()
    n = 1 << 20
    eps2 = 1e-4
    steps = args.steps if args.steps != 20 else 5
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
                     "jslice": plan.jslice, "compensated": bool(plan.compensated)},
        }
    best = out_modes["default"]
    print(json.dumps({
        "metric": METRIC, "value": best["value"], "unit": "10^9 interactions/s", "n_gpus": 1, "steps": steps,
        "ms_per_step": best["ms_per_step"], "higher_is_better": True, "dtype": "f32",
        "data": "synthetic (seeded Plummer sphere)",
        "config": {"workload": f"single system, N={n} Plummer sphere (a=1, M=1, invMass=1), eps^2={eps2}, "
                               f"dt=0.016, damping=0.95, {steps} steps; errors vs FP64 direct sum over all N"},
        "modes": out_modes,
    }))


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
    print(json.dumps({
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
    }))


def main():
    args = parse()
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
        return bench_ensemble(args)
    if args.workload == "served":
        return bench_served(args)
    if args.workload == "plummer":
        return bench_plummer(args)
    if args.workload == "share":
        return bench_share(args)

    L = _native.lib()
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
    ic_max, ic_rms, _, _ = accuracy()

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
        "config": {
            "workload": f"single system, N={n} uniform ball radius {ics.ball_radius(n):.4f}, unit masses, "
                        f"eps^2={eps2}, dt={dt}, damping={damping}; step = force + integrate, state in HBM",
            "n_bodies": n,
            "interactions_per_step": int(inter),
            "plan": {"variant": _native.VARIANT_NAMES[plan.variant], "i_per_lane": plan.i_per_lane,
                     "jslice": plan.jslice, "compensated": bool(plan.compensated), "newton": bool(plan.newton),
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
        "accuracy": {"max_rel": max_rel, "rms_rel": rms_rel,
                     "vs": "FP64 direct sum (GPU diagnostic), per-body |F-F64|/|F64|",
                     "state": f"after the {W} warm-up and {K} timed steps",
                     "initial_state": {"max_rel": ic_max, "rms_rel": ic_rms}},
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
                         "what": "ComputeGravitation_HIP_Host on host buffers: H2D of posMass, force kernel, D2H of x, y, z; median of 5"}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pm_h, vm_h, np.float32(eps2), dt, damping, args.cpu_budget_s)
        out["accuracy"]["vs_cpu_aos"] = crosscheck_cpu_aos(pm.cpu().numpy(), step_force.cpu().numpy(),
                                                           f64.cpu().numpy(), np.float32(eps2))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
