"""nbody_amd — MI355X (gfx950) force evaluation for tycho/nbody's O(N^2)
softened-gravity path.

The compute lives in nbody_amd/lib/libnbody_hip.so (hand-written HIP kernels
behind the C ABI of include/nbody_hip.h); this package is its Python binding:

    _native      ctypes declarations of every C entry point
    algorithms   the algorithm table (src/nbody.cu:83-103 shape)
    engine       device-resident stepping: one system, or an ensemble
    ics          seeded initial conditions (reference-compatible, ball, Plummer)
    serve        request traces for the continuous-batching server
"""
from . import _native, algorithms, engine, ics, serve  # noqa: F401
from ._native import NativeError, NativeLibraryMissing, Options, lib  # noqa: F401

SOFTENING = 0.1          # src/nbody.cu:151 (softeningSquared = 0.1^2)
DAMPING = 0.95           # src/nbody.cu:152
TIMESTEP = 0.016         # src/nbody.cu:153
FLOPS_PER_INTERACTION = 20  # src/nbody.cu:1026-1027
