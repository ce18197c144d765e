"""Build the native parts of nbody_amd in-tree.

    python -m nbody_amd.build          # or nbody_amd.build.build()

Outputs (git-ignored, shipped with the tree):
    nbody_amd/lib/libnbody_hip.so   HIP kernels + the C ABI of include/*.h
    nbody_amd/bin/nbody             the command-line driver (reference flags and output)
Objects go to nbody_amd/_build/. Host-only C++ (initial conditions) is built
with g++, strict IEEE. Every HIP translation unit is compiled for
gfx950 with packed FP32 disabled (-Xclang -target-feature -Xclang -packed-fp32-ops):
all FP32 arithmetic is plain v_fma/v_mul/v_add/v_sub.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
BINDIR = os.path.join(HERE, "bin")
INCLUDE = os.path.join(ROOT, "include")

ARCH = os.environ.get("NBODY_HIP_ARCH", "gfx950")

HIP_SOURCES = ["nbody_kernels.hip", "nbody_batch.hip", "nbody_abi.hip", "nbody_engine.hip"]
# host-only C++, strict IEEE (its arithmetic is written out explicitly)
CXX_SOURCES = ["nbody_ics.cpp"]
CXX_FLAGS = ["-O2", "-fno-fast-math", "-ffp-contract=off", "-std=c++17", "-fPIC", "-Wall", "-I", INCLUDE]

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-command-line-argument",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", CSRC,
    "-I", INCLUDE,
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found (need ROCm at /opt/rocm)")
    return path


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # the host pass does not know the packed-fp32 feature and says so; drop that noise
    out = "\n".join(l for l in r.stdout.splitlines() if "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False):
    """Compile every HIP source and link libnbody_hip.so. Always recompiles."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    cc = hipcc()
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        jobs.append((obj, [cc, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]))
    cxx = shutil.which("g++") or "g++"
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        jobs.append((obj, [cxx, *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]))
    with ThreadPoolExecutor(max_workers=min(4, len(jobs))) as ex:
        outs = list(ex.map(lambda j: _run(j[1]), jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    lib = os.path.join(LIBDIR, "libnbody_hip.so")
    tmp = lib + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *[j[0] for j in jobs]])
    os.replace(tmp, lib)
    # the command-line driver (host code only), linked against the library
    os.makedirs(BINDIR, exist_ok=True)
    exe = os.path.join(BINDIR, "nbody")
    _run([cc, "-O2", "-std=c++17", "-pthread", "-Wall", "-Wno-unused-command-line-argument", "-I", INCLUDE,
          os.path.join(CSRC, "nbody_main.cpp"), "-o", exe + ".tmp", "-L", LIBDIR, "-lnbody_hip",
          "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(exe + ".tmp", exe)
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
