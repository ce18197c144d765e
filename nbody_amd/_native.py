"""ctypes binding of libnbody_hip.so (include/nbody_hip.h, include/nbody_ics.h).

The library is loaded from nbody_amd/lib/ only. A missing or stale library is
an error: there is no CPU or PyTorch fallback for any entry point.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnbody_hip.so")

ABI_VERSION = 4

# bytes of the fixed device array of nbody_hip_force_noalloc (NBODY_HIP_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 64 << 20

# ensemble modes (include/nbody_engine.h)
BATCHED_GRAPH, BATCHED, PER_SYSTEM = 0, 1, 2

# variant ids (NBODY_HIP_* in include/nbody_hip.h)
AOS, LDS, DPP, SGPR, SYMMETRIC, AUTO = 0, 1, 2, 3, 4, -1
VARIANT_NAMES = {AOS: "AOS", LDS: "LDS", DPP: "DPP", SGPR: "SGPR", SYMMETRIC: "Symmetric", AUTO: "auto"}


class Options(ctypes.Structure):
    """nbody_hip_options."""

    _fields_ = [
        ("variant", ctypes.c_int),
        ("i_per_lane", ctypes.c_int),
        ("jslice", ctypes.c_int),
        ("compensated", ctypes.c_int),
        ("newton", ctypes.c_int),
        ("waves_per_simd", ctypes.c_int),
        ("slot_bytes", ctypes.c_size_t),
    ]

    @classmethod
    def make(cls, variant=AUTO, i_per_lane=0, jslice=0, compensated=-1, newton=0, waves_per_simd=0, slot_bytes=0):
        return cls(variant, i_per_lane, jslice, compensated, newton, waves_per_simd, slot_bytes)


class Request(ctypes.Structure):
    """nbody_hip_request (include/nbody_serve.h)."""

    _fields_ = [
        ("id", ctypes.c_uint64),
        ("N", ctypes.c_size_t),
        ("steps", ctypes.c_int),
        ("seed", ctypes.c_uint32),
        ("ic", ctypes.c_int),
        ("softeningSquared", ctypes.c_float),
        ("arrival_s", ctypes.c_double),
    ]


class Result(ctypes.Structure):
    """nbody_hip_result (include/nbody_serve.h)."""

    _fields_ = [
        ("id", ctypes.c_uint64),
        ("N", ctypes.c_size_t),
        ("steps", ctypes.c_int),
        ("status", ctypes.c_int),
        ("arrival_s", ctypes.c_double),
        ("admit_s", ctypes.c_double),
        ("finish_s", ctypes.c_double),
    ]


IC_UNIFORM_BALL, IC_PLUMMER = 0, 1

# C signatures: name -> (restype, argtypes)
_f, _d, _i, _sz, _p, _u32 = (ctypes.c_float, ctypes.c_double, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p,
                             ctypes.c_uint32)
_OPT = ctypes.POINTER(Options)
SIGNATURES = {
    "ComputeGravitation_HIP_AOS": (_f, [_p, _p, _f, _sz]),
    "ComputeGravitation_HIP_LDS": (_f, [_p, _p, _f, _sz]),
    "ComputeGravitation_HIP_DPP": (_f, [_p, _p, _f, _sz]),
    "ComputeGravitation_HIP_SGPR": (_f, [_p, _p, _f, _sz]),
    "ComputeGravitation_HIP_Symmetric": (_f, [_p, _p, _f, _sz]),
    "ComputeGravitation_HIP": (_f, [_p, _p, _f, _sz]),
    "ComputeGravitation_HIP_Host": (_f, [_p, _p, _f, _sz]),
    "nbody_hip_force": (_i, [_p, _p, _f, _sz, _sz, _sz, _OPT, _p]),
    "nbody_hip_force_noalloc": (_i, [_p, _p, _f, _sz, _sz, _sz, _OPT, _p]),
    "nbody_hip_force_pair_share": (_i, [_p, _p, _f, _sz, _i, _i, _OPT, _p]),
    "integrateGravitation_HIP": (_f, [_p, _p, _p, _f, _f, _sz]),
    "nbody_hip_integrate": (_i, [_p, _p, _p, _f, _f, _sz, _p]),
    "nbody_hip_force_fp64": (_f, [_p, _p, _d, _sz]),
    "nbody_hip_energy_fp64": (_i, [_p, _p, _p, _d, _sz]),
    "nbody_hip_plan": (_i, [_sz, _sz, _OPT, _OPT]),
    "nbody_hip_autotune": (_f, [_p, _p, _f, _sz, _i]),
    "nbody_hip_kernel_timer": (_i, [_i]),
    "nbody_hip_kernel_timer_read": (_i, [ctypes.POINTER(_d), ctypes.POINTER(_i)]),
    "nbody_hip_last_error": (ctypes.c_char_p, []),
    "nbody_hip_abi_version": (_i, []),
    "nbody_hip_device_info": (_i, [ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.c_char_p, _sz]),
    "nbody_hip_system_create": (_p, [_p, _p, _sz, _f, _f, _f, _OPT]),
    "nbody_hip_system_run": (_i, [_p, _i, _i, ctypes.POINTER(_f)]),
    "nbody_hip_system_read": (_i, [_p, _p, _p, _p]),
    "nbody_hip_system_plan": (_i, [_p, _OPT]),
    "nbody_hip_system_destroy": (None, [_p]),
    "nbody_hip_ensemble_create": (_p, [_sz, _p, _p, _p, _p, _f, _f]),
    "nbody_hip_ensemble_create_opt": (_p, [_sz, _p, _p, _p, _p, _f, _f, _OPT]),
    "nbody_hip_ensemble_run": (_i, [_p, _i, _i, ctypes.POINTER(_f)]),
    "nbody_hip_ensemble_read": (_i, [_p, _p, _p, _p]),
    "nbody_hip_ensemble_destroy": (None, [_p]),
    "nbody_hip_server_create": (_p, [_sz, _i, _f, _f, _i]),
    "nbody_hip_server_create_opt": (_p, [_sz, _i, _f, _f, _i, _OPT]),
    "nbody_hip_server_run_trace": (_i, [_p, ctypes.POINTER(Request), _sz, ctypes.POINTER(Result),
                                        ctypes.POINTER(_d), ctypes.POINTER(_d)]),
    "nbody_hip_server_final_state": (_i, [_p, ctypes.c_uint64, _p, _p]),
    "nbody_hip_server_peak_active": (_i, [_p]),
    "nbody_hip_request_bytes": (_sz, [_sz]),
    "nbody_hip_request_bytes_opt": (_sz, [_sz, _OPT]),
    "nbody_hip_server_destroy": (None, [_p]),
    "nbody_random_unit_bodies": (_i, [_p, _p, _sz, _u32, _f, _f]),
    "nbody_uniform_ball": (_i, [_p, _p, _sz, _u32, _f, _f, _f]),
    "nbody_plummer": (_i, [_p, _p, _sz, _u32, _f, _f, _i]),
}

_lib = None


class NativeLibraryMissing(RuntimeError):
    pass


def lib():
    """The loaded library with every C signature declared."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"{LIB_PATH} is not built; run `python -m nbody_amd.build` (needs hipcc). "
            "nbody_amd has no CPU fallback.")
    # PyTorch, when installed, is imported first: it loads its own HIP runtime,
    # which carries the same soname (libamdhip64.so.7) as the one this library
    # links, so the loader binds the library to it and the process holds one
    # HIP runtime. Loaded the other way round, torch would bring a second
    # runtime into the process, and its device initialisation can then fail.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    v = L.nbody_hip_abi_version()
    if v != ABI_VERSION:
        raise NativeLibraryMissing(f"{LIB_PATH} has ABI {v}, expected {ABI_VERSION}; rebuild it")
    _lib = L
    return L


def last_error():
    return lib().nbody_hip_last_error().decode(errors="replace")


class NativeError(RuntimeError):
    pass


def check(rc, what):
    """Raise NativeError for a nonzero return code of an int entry point."""
    if rc != 0:
        raise NativeError(f"{what} failed (code {rc}): {last_error()}")


def check_ms(ms, what):
    """Raise NativeError for a 0.0f return of a float (milliseconds) entry point."""
    if not ms > 0.0:
        raise NativeError(f"{what} failed: {last_error()}")
    return ms
