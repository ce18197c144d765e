"""Python driver of the request server (include/nbody_serve.h): build a
seeded request trace, replay it, summarise latencies."""
import ctypes
import math

import numpy as np

from . import _native


def poisson_trace(count, rate_hz, n_lo, n_hi, steps_lo, steps_hi, seed, eps2=0.01, plummer_fraction=0.5):
    """`count` requests with exponential inter-arrival times of mean 1/rate_hz,
    N log-uniform in [n_lo, n_hi], steps uniform in [steps_lo, steps_hi],
    initial conditions a uniform ball or (with probability plummer_fraction)
    a Plummer sphere, each with its own seed."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_hz, size=count))
    arrivals -= arrivals[0]
    sizes = np.exp(rng.uniform(math.log(n_lo), math.log(n_hi + 1), size=count)).astype(np.int64).clip(n_lo, n_hi)
    steps = rng.integers(steps_lo, steps_hi + 1, size=count)
    ics = (rng.uniform(size=count) < plummer_fraction).astype(np.int32)
    seeds = rng.integers(1, 2 ** 31 - 1, size=count)
    reqs = (_native.Request * count)()
    for k in range(count):
        reqs[k] = _native.Request(k, int(sizes[k]), int(steps[k]), int(seeds[k]), int(ics[k]), eps2,
                                  float(arrivals[k]))
    return reqs


def request_bytes(n, compensated=-1):
    """Device bytes a resident request of n bodies holds (what the budget
    counts); compensated=2 for a server with FP64 accumulation."""
    opt = _native.Options.make(compensated=compensated)
    return int(_native.lib().nbody_hip_request_bytes_opt(int(n), ctypes.byref(opt)))


class Server:
    """compensated=2 steps every request with FP64 accumulation
    (nbody_hip_server_create_opt); the default keeps FP32 block partials."""

    def __init__(self, budget_bytes=200 * 10 ** 9, max_active=64, dt=0.016, damping=0.95, keep_states=False,
                 compensated=-1):
        self.L = _native.lib()
        opt = _native.Options.make(compensated=compensated)
        self.h = self.L.nbody_hip_server_create_opt(int(budget_bytes), max_active, dt, damping,
                                                    1 if keep_states else 0, ctypes.byref(opt))
        if not self.h:
            raise _native.NativeError(_native.last_error())

    def run(self, reqs):
        count = len(reqs)
        res = (_native.Result * count)()
        wall = ctypes.c_double()
        work = ctypes.c_double()
        _native.check(self.L.nbody_hip_server_run_trace(self.h, reqs, count, res, ctypes.byref(wall),
                                                        ctypes.byref(work)), "server run")
        return list(res), wall.value, work.value

    def final_state(self, req_id, n):
        pm = np.zeros((n, 4), np.float32)
        vm = np.zeros((n, 4), np.float32)
        _native.check(self.L.nbody_hip_server_final_state(self.h, req_id, pm.ctypes.data, vm.ctypes.data), "state")
        return pm, vm

    def peak_active(self):
        return self.L.nbody_hip_server_peak_active(self.h)

    def close(self):
        if self.h:
            self.L.nbody_hip_server_destroy(self.h)
            self.h = None


def latency_summary(results, wall_s, interactions):
    done = [r for r in results if r.status == 0]
    lat = np.array([r.finish_s - r.arrival_s for r in done])
    wait = np.array([r.admit_s - r.arrival_s for r in done])
    return {
        "requests": len(results),
        "completed": len(done),
        "rejected": sum(1 for r in results if r.status == 1),
        "failed": sum(1 for r in results if r.status == 2),
        "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
        "p50_queue_wait_s": float(np.percentile(wait, 50)) if len(wait) else None,
        "p99_queue_wait_s": float(np.percentile(wait, 99)) if len(wait) else None,
        "wall_s": wall_s,
        "interactions": interactions,
        "interactions_per_s": interactions / wall_s if wall_s > 0 else None,
    }
