"""The algorithm table, mirroring the reference's s_algorithms
(src/nbody.cu:83-103, algorithm_def_t at src/nbody.h:71-78): each entry has a
name, a type deciding the calling contract (src/nbody.cu:333-383) and the
C entry point. Every entry here is ALGORITHM_AOS_GPU (device pointers) except
HIP_Host, which is ALGORITHM_AOS (host pointers).
"""
from dataclasses import dataclass

from . import _native

ALGORITHM_AOS = "AOS"          # host float4 pointers (src/nbody.cu:343-349)
ALGORITHM_AOS_GPU = "AOS_GPU"  # device float4 pointers (src/nbody.cu:351-371)


@dataclass(frozen=True)
class Algorithm:
    name: str
    type: str
    symbol: str

    def __call__(self, force, pos_mass, softening_squared, n):
        """Call the entry with raw pointers (int addresses); returns ms, 0.0 on failure."""
        fn = getattr(_native.lib(), self.symbol)
        return fn(force, pos_mass, softening_squared, n)


ALGORITHMS = [
    Algorithm("HIP_AOS", ALGORITHM_AOS_GPU, "ComputeGravitation_HIP_AOS"),
    Algorithm("HIP_LDS", ALGORITHM_AOS_GPU, "ComputeGravitation_HIP_LDS"),
    Algorithm("HIP_DPP", ALGORITHM_AOS_GPU, "ComputeGravitation_HIP_DPP"),
    Algorithm("HIP_SGPR", ALGORITHM_AOS_GPU, "ComputeGravitation_HIP_SGPR"),
    Algorithm("HIP_Symmetric", ALGORITHM_AOS_GPU, "ComputeGravitation_HIP_Symmetric"),
    Algorithm("HIP", ALGORITHM_AOS_GPU, "ComputeGravitation_HIP"),
    Algorithm("HIP_Host", ALGORITHM_AOS, "ComputeGravitation_HIP_Host"),
]


def find(name):
    """Exact (case-insensitive) match first, else the last prefix match, as the
    reference's --algorithm parsing (src/nbody.cu:861-901)."""
    hit = None
    for a in ALGORITHMS:
        if a.name.lower() == name.lower():
            return a
        if a.name.lower().startswith(name.lower()):
            hit = a
    if hit is None:
        raise KeyError(name)
    return hit
