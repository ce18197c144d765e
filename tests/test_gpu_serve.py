"""The continuous-batching request server (include/nbody_serve.h) on the GPU.

Tolerance: a request's result is bit-identical alone and in any mix (the
batched force of a body depends only on its own system), and bit-identical
to the same system stepped by the ensemble engine's batched path; that path
is checked against the FP64 oracle in test_gpu_engine.py.
"""
import ctypes

import numpy as np
import pytest

from nbody_amd import _native, serve

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def small_trace(count=18, seed=3):
    reqs = serve.poisson_trace(count, 2000.0, 300, 6000, 2, 6, seed)
    return reqs


def one(req):
    r = (_native.Request * 1)()
    r[0] = req
    r[0].arrival_s = 0.0
    return r


def test_results_alone_equal_in_mix():
    reqs = small_trace()
    srv = serve.Server(max_active=8, keep_states=True)
    try:
        res, wall, work = srv.run(reqs)
        assert all(r.status == 0 for r in res)
        assert srv.peak_active() <= 8
        for r in res:
            assert r.arrival_s <= r.admit_s <= r.finish_s
        mixed = {k: srv.final_state(k, reqs[k].N) for k in range(len(reqs))}
        expect_work = sum(float(q.N) ** 2 * q.steps for q in reqs)
        assert work == pytest.approx(expect_work, rel=1e-12)
    finally:
        srv.close()
    # one slot: each request is uploaded while the one before runs its last
    # step and joins the next step, so the requests step one at a time
    single = serve.Server(max_active=1, keep_states=True)
    try:
        res, _, _ = single.run(reqs)
        assert all(r.status == 0 for r in res)
        assert single.peak_active() == 1
        # one at a time: requests finish in the order they were admitted
        done = sorted((r.admit_s, r.finish_s) for r in res)
        assert all(b[1] > a[1] for a, b in zip(done, done[1:]))
        for k in range(len(reqs)):
            pm, vm = single.final_state(k, reqs[k].N)
            assert np.array_equal(bits(pm), bits(mixed[k][0])), k
            assert np.array_equal(bits(vm), bits(mixed[k][1])), k
    finally:
        single.close()
    for k in (0, 5, len(reqs) - 1):
        alone = serve.Server(max_active=1, keep_states=True)
        try:
            res, _, _ = alone.run(one(reqs[k]))
            assert res[0].status == 0
            pm, vm = alone.final_state(k, reqs[k].N)
        finally:
            alone.close()
        assert np.array_equal(bits(pm), bits(mixed[k][0])), k
        assert np.array_equal(bits(vm), bits(mixed[k][1])), k


def test_large_request_alone_equals_in_mix():
    """A request above 2^18 bodies (batched as compensated one-sided rows)
    next to small block-pair requests: its final state is bit-identical to
    serving it alone, and so are the small ones'."""
    big = _native.Request(0, 2 ** 18 + 300, 2, 11, _native.IC_PLUMMER, 1e-4, 0.0)
    reqs = (_native.Request * 3)()
    reqs[0] = big
    reqs[1] = _native.Request(1, 700, 3, 12, _native.IC_UNIFORM_BALL, 0.01, 0.0)
    reqs[2] = _native.Request(2, 5000, 2, 13, _native.IC_UNIFORM_BALL, 0.01, 0.0)
    srv = serve.Server(max_active=4, keep_states=True)
    try:
        res, _, _ = srv.run(reqs)
        assert all(r.status == 0 for r in res)
        mixed = {k: srv.final_state(k, reqs[k].N) for k in range(3)}
    finally:
        srv.close()
    for k in range(3):
        alone = serve.Server(max_active=1, keep_states=True)
        try:
            res, _, _ = alone.run(one(reqs[k]))
            assert res[0].status == 0
            pm, vm = alone.final_state(k, reqs[k].N)
        finally:
            alone.close()
        assert np.array_equal(bits(pm), bits(mixed[k][0])), k
        assert np.array_equal(bits(vm), bits(mixed[k][1])), k


def test_served_equals_ensemble_batched():
    """A served uniform-ball request against the same bodies stepped by the
    ensemble engine's batched path."""
    from nbody_amd import ics
    req = _native.Request(7, 3000, 4, 99, _native.IC_UNIFORM_BALL, 0.01, 0.0)
    srv = serve.Server(keep_states=True)
    try:
        res, _, _ = srv.run(one(req))
        pm_s, vm_s = srv.final_state(7, 3000)
    finally:
        srv.close()
    pm, vm = ics.uniform_ball(3000, seed=99)
    L = _native.lib()
    n = np.array([3000], np.uint64)
    e2 = np.array([0.01], np.float32)
    h = L.nbody_hip_ensemble_create(1, n.ctypes.data, pm.ctypes.data, vm.ctypes.data, e2.ctypes.data, 0.016, 0.95)
    assert h, _native.last_error()
    try:
        _native.check(L.nbody_hip_ensemble_run(h, 4, _native.BATCHED, None), "run")
        out_p = np.zeros_like(pm)
        out_v = np.zeros_like(vm)
        _native.check(L.nbody_hip_ensemble_read(h, out_p.ctypes.data, out_v.ctypes.data, None), "read")
    finally:
        L.nbody_hip_ensemble_destroy(h)
    assert np.array_equal(bits(pm_s), bits(out_p)) and np.array_equal(bits(vm_s), bits(out_v))


def test_budget_rejects_and_bounds():
    reqs = (_native.Request * 3)()
    reqs[0] = _native.Request(0, 1000, 2, 1, 0, 0.01, 0.0)
    reqs[1] = _native.Request(1, 5000, 2, 2, 1, 0.01, 0.0)   # more than the budget: rejected
    reqs[2] = _native.Request(2, 2000, 3, 3, 0, 0.01, 0.001)
    budget = serve.request_bytes(2000)
    assert serve.request_bytes(1000) + budget > budget >= serve.request_bytes(1000)
    assert serve.request_bytes(5000) > budget
    srv = serve.Server(budget_bytes=budget, max_active=4)
    try:
        res, _, work = srv.run(reqs)
    finally:
        srv.close()
    assert [r.status for r in res] == [0, 1, 0]
    # only one of the two fits the budget at a time
    assert res[2].admit_s >= res[0].finish_s or res[0].admit_s >= res[2].finish_s
    assert work == pytest.approx(1000.0 ** 2 * 2 + 2000.0 ** 2 * 3)


def test_bad_request_is_an_error():
    reqs = (_native.Request * 1)()
    reqs[0] = _native.Request(0, 0, 2, 1, 0, 0.01, 0.0)
    srv = serve.Server()
    try:
        with pytest.raises(_native.NativeError, match="N = 0"):
            srv.run(reqs)
    finally:
        srv.close()
