"""Product host code on the CPU: the C-ABI library loads and exports every
symbol include/*.h declares (no GPU call is made), the initial-condition
generators, and the algorithm table."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import nbody_amd
from nbody_amd import _native, algorithms, ics


def header_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                             text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = header_functions()
    for must in ("ComputeGravitation_HIP_SGPR", "integrateGravitation_HIP", "nbody_hip_force",
                 "nbody_random_unit_bodies"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = _native.lib()  # raises when the library is missing: no fallback
    missing = [n for n in header_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(header_functions()) <= set(_native.SIGNATURES)


def test_abi_version():
    assert _native.lib().nbody_hip_abi_version() == _native.ABI_VERSION


def test_missing_library_is_an_error(monkeypatch, tmp_path):
    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(_native.NativeLibraryMissing):
        _native.lib()


@pytest.mark.parametrize("n", [1, 1024, 4096, 98304])
def test_random_unit_bodies_bitexact_vs_oracle(oracle, n):
    pm, vm = ics.random_unit_bodies(n, seed=7)
    po, vo = oracle.random_unit_bodies(n, 7)
    assert np.array_equal(pm.view(np.uint32), po.view(np.uint32))
    assert np.array_equal(vm.view(np.uint32), vo.view(np.uint32))


def test_random_unit_bodies_other_seed_differs(oracle):
    a, _ = ics.random_unit_bodies(1024, seed=8)
    b, _ = oracle.random_unit_bodies(1024, 8)
    c, _ = ics.random_unit_bodies(1024, seed=7)
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def test_uniform_ball_statistics():
    n = 16384
    pm, vm = ics.uniform_ball(n, seed=1)
    assert ics.ball_radius(16384) == pytest.approx(1.54)
    assert ics.ball_radius(98304) == pytest.approx(1.54 * 6 ** (1 / 3))
    r = np.linalg.norm(pm[:, :3].astype(np.float64), axis=1) / 1.54
    assert r.max() <= 1 + 1e-6
    # uniform density: P(r < x) = x^3, so E[r^3] = 1/2 and the median radius is 2^(-1/3)
    assert abs(np.mean(r ** 3) - 0.5) < 0.01
    assert abs(np.median(r) - 2 ** (-1 / 3)) < 0.01
    assert np.all(pm[:, 3] == 1.0) and np.all(vm[:, 3] == 1.0)
    v = np.linalg.norm(vm[:, :3].astype(np.float64), axis=1)
    assert v.max() <= 1 + 1e-6 and abs(np.mean(v ** 3) - 0.5) < 0.01
    assert abs(pm[:, :3].mean(0)).max() < 0.03


def test_plummer_statistics(oracle):
    n = 8192
    pm, vm = ics.plummer(n, seed=3)
    assert pm[:, 3].sum() == pytest.approx(1.0, rel=1e-5)
    assert np.all(vm[:, 3] == 1.0)
    r = np.linalg.norm(pm[:, :3].astype(np.float64), axis=1)
    # M(<a)/M = a^3/(a^2+a^2)^(3/2) = 2^(-3/2); M(<3a)/M = 27/10^(3/2)
    for x, frac in [(1.0, 2 ** -1.5), (3.0, 27 / 10 ** 1.5)]:
        assert abs(np.mean(r < x) - frac) < 0.02
    assert r.max() <= 10.0 * (1 + 1e-6)
    # virial equilibrium 2K + W = 0 (G = 1, softening small next to a = 1)
    ke, pe, _ = oracle.fp64_energy(pm, np.column_stack([vm[:, :3], np.full(n, 1.0, np.float32)]), 1e-6)
    assert abs(2 * ke / -pe - 1.0) < 0.08
    pm2, _ = ics.plummer(n, seed=3, unit_inv_mass=False)
    assert np.array_equal(pm2, pm)


def test_plummer_inv_mass_option():
    _, vm = ics.plummer(1000, seed=3, unit_inv_mass=False)
    assert np.allclose(vm[:, 3], 1000.0)


def test_log_uniform_sizes():
    s = ics.log_uniform_sizes(512, 256, 16384, seed=4)
    assert s.min() >= 256 and s.max() <= 16384 and len(s) == 512
    l = np.log(s)
    assert abs(np.median(l) - 0.5 * (np.log(256) + np.log(16384))) < 0.3


def test_algorithm_table():
    names = [a.name for a in algorithms.ALGORITHMS]
    assert len(set(names)) == len(names)
    for a in algorithms.ALGORITHMS:
        assert hasattr(_native.lib(), a.symbol)
    assert algorithms.find("hip_sgpr").name == "HIP_SGPR"
    assert algorithms.find("HIP_Sym").name == "HIP_Symmetric"
    with pytest.raises(KeyError):
        algorithms.find("GPU_Shared")
    assert nbody_amd.FLOPS_PER_INTERACTION == 20


def test_poisson_trace_statistics():
    from nbody_amd import serve
    reqs = serve.poisson_trace(4000, 200.0, 1024, 131072, 10, 200, seed=9)
    arr = np.array([r.arrival_s for r in reqs])
    gaps = np.diff(arr)
    assert arr[0] == 0.0 and np.all(gaps >= 0)
    assert abs(gaps.mean() - 1 / 200.0) < 0.1 / 200.0          # mean inter-arrival 5 ms
    assert abs(np.std(gaps) / gaps.mean() - 1.0) < 0.1         # exponential: cv = 1
    n = np.array([r.N for r in reqs])
    st = np.array([r.steps for r in reqs])
    assert n.min() >= 1024 and n.max() <= 131072 and st.min() >= 10 and st.max() <= 200
    assert abs(np.median(np.log(n)) - 0.5 * (np.log(1024) + np.log(131072))) < 0.2
    ics_kind = np.array([r.ic for r in reqs])
    assert 0.45 < ics_kind.mean() < 0.55
    again = serve.poisson_trace(4000, 200.0, 1024, 131072, 10, 200, seed=9)
    assert all(a.N == b.N and a.seed == b.seed and a.arrival_s == b.arrival_s for a, b in zip(reqs, again))


def test_request_bytes_closed_form():
    from nbody_amd import serve
    # 48 B per body, plus 3 float planes of ceil(N/512) slots x ceil(N/512)*512 bodies
    for n in (1, 512, 513, 1000, 16384, 2 ** 18):
        nb = -(-n // 512)
        assert serve.request_bytes(n) == 48 * n + 12 * nb * nb * 512, n
    # beyond 2^18 bodies a single compensated slot per body
    n = 2 ** 18 + 1
    assert serve.request_bytes(n) == 48 * n + 12 * n


def test_request_bytes_fp64_closed_form():
    from nbody_amd import serve
    # FP64 accumulation: 3 double planes of ceil(N/256) slots x ceil(N/256)*256 bodies
    for n in (1, 256, 257, 1000, 16384, 2 ** 17):
        nb = -(-n // 256)
        assert serve.request_bytes(n, compensated=2) == 48 * n + 24 * nb * nb * 256, n
    # beyond 2^17 bodies one FP64 slot per body
    n = 2 ** 17 + 1
    assert serve.request_bytes(n, compensated=2) == 48 * n + 24 * n
    # compensated 0 and 1 keep the FP32 partials
    for c in (0, 1):
        assert serve.request_bytes(5000, compensated=c) == serve.request_bytes(5000)


def test_size_and_range_limits_fail_before_any_device_call():
    """Argument checks run before the library touches a device: N above the
    2^28 cap, N = 0, and i-ranges outside [0, N) return the failure value
    with a message (the pointers are never dereferenced)."""
    L = _native.lib()
    fake = ctypes.c_void_p(16)
    assert L.ComputeGravitation_HIP_LDS(fake, fake, 0.01, 2 ** 28 + 1) == 0.0
    assert "outside" in _native.last_error()
    assert L.ComputeGravitation_HIP(fake, fake, 0.01, 0) == 0.0
    assert "N = 0" in _native.last_error()
    o = _native.Options.make()
    assert L.nbody_hip_force(fake, fake, 0.01, 100, 90, 11, ctypes.byref(o), None) != 0
    assert "i range" in _native.last_error()
    assert L.nbody_hip_force(fake, fake, 0.01, 100, 101, 0, ctypes.byref(o), None) != 0


C_CLIENT = r"""
#include "nbody_hip.h"
#include "nbody_ics.h"
#include "nbody_engine.h"
#include "nbody_serve.h"
#include <stdio.h>
#include <stdlib.h>

/* A C caller of the library: the calls below touch no GPU. */
int main(void)
{
    size_t n = 1000;
    float *pm = (float *)malloc(16 * n), *vm = (float *)malloc(16 * n);
    nbody_hip_options opt = {0};
    opt.compensated = 2;
    if (nbody_hip_abi_version() != NBODY_HIP_ABI_VERSION) return 2;
    if (nbody_uniform_ball(pm, vm, n, 1u, 1.0f, 1.0f, 1.0f) != 0) return 3;
    printf("%zu %zu %.9g\n", nbody_hip_request_bytes(n), nbody_hip_request_bytes_opt(n, &opt), pm[0]);
    free(pm);
    free(vm);
    return 0;
}
"""


def test_headers_are_plain_c_and_link_from_c(tmp_path):
    """include/*.h compile as C99 (pedantic) and C++11, and a C program built
    against them links libnbody_hip.so and runs calls that need no GPU."""
    import shutil
    import subprocess
    gcc, gxx = shutil.which("gcc"), shutil.which("g++")
    if not gcc or not gxx:
        pytest.skip("gcc/g++ not on PATH")
    inc = os.path.join(ROOT, "include")
    libdir = os.path.dirname(_native.LIB_PATH)
    src = tmp_path / "client.c"
    src.write_text(C_CLIENT)
    for cmd in ([gcc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I", inc, "-fsyntax-only", "-x", "c",
                 str(src)],
                [gxx, "-std=c++11", "-Wall", "-Werror", "-I", inc, "-fsyntax-only", "-x", "c++", str(src)]):
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    exe = tmp_path / "client"
    r = subprocess.run([gcc, "-std=c99", "-I", inc, str(src), "-o", str(exe), "-L", libdir, "-lnbody_hip",
                        "-Wl,-rpath," + libdir], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.returncode, r.stderr)
    fp32_bytes, fp64_bytes, x0 = r.stdout.split()
    from nbody_amd import serve
    assert int(fp32_bytes) == serve.request_bytes(1000) and int(fp64_bytes) == serve.request_bytes(1000, 2)
    pm, _ = ics.uniform_ball(1000, seed=1, radius=1.0)
    assert np.float32(x0) == pm[0, 0]
