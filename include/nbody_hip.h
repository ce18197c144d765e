/*
 * nbody_hip.h — C ABI of libnbody_hip.so: MI355X (gfx950) force evaluation
 * for tycho/nbody's O(N^2) softened-gravity path.
 *
 * No HIP, CUDA or torch types appear here: plain pointers, sizes and floats.
 * Linkage is extern "C" (the reference's CPU algorithm headers use extern "C",
 * e.g. src/nbody_CPU_SOA.h:39-47; its GPU headers are C++ linkage, e.g.
 * src/nbody_GPU_AOS.h:37 — a table entry only needs the address).
 *
 * Every ComputeGravitation_HIP_* entry has the algorithm-table signature of
 * DECLARE_AOS (src/nbody_util.h:113-119, pfnComputeGravitation_AOS_t at
 * src/nbody.h:68):
 *
 *     float fn(float *force, float *posMass, float softeningSquared, size_t N)
 *
 * and the ALGORITHM_AOS_GPU contract (src/nbody.cu:351-371):
 *   - force and posMass are DEVICE pointers to N float4 records;
 *     posMass = (x, y, z, m); force receives (fx, fy, fz) and its w is left
 *     alone;
 *   - force_i = sum_j m_j d / (|d|^2 + softeningSquared)^(3/2), d = x_j - x_i
 *     (src/bodybodyInteraction.cuh:40-65: no m_i, no G; the self term is 0);
 *   - runs on the null stream and is synchronous; returns the kernels' elapsed
 *     milliseconds measured with events (src/nbody_GPU_Shared.cu:88-98);
 *   - returns 0.0f on failure, which the reference driver reads as "skip this
 *     algorithm" (src/nbody.cu:385-386); nbody_hip_last_error() says why;
 *   - any N >= 1 (the reference kernels need N % 256 or N % 32 == 0).
 * The entries keep their own scratch (slice partial sums, events) per device,
 * grown on demand, so they may allocate device memory on a first call or a
 * larger N. Not re-entrant across threads for the same device (neither is the
 * reference, src/nbody.cu:472-476).
 */
#ifndef NBODY_HIP_H
#define NBODY_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBODY_HIP_ABI_VERSION 4

/* Variant ids for nbody_hip_options.variant. */
enum {
    NBODY_HIP_AOS = 0,       /* one i per lane, whole j range (src/nbody_GPU_AOS.cu:42-94) */
    NBODY_HIP_LDS = 1,       /* j tiles in LDS (src/nbody_GPU_Shared.cu:42-103) */
    NBODY_HIP_DPP = 2,       /* j tile passed around the wave64 by DPP (src/nbody_GPU_Shuffle.cu:44-107) */
    NBODY_HIP_SGPR = 3,      /* scalar-memory broadcast (src/nbody_GPU_AOS_const.cu:45-121) */
    NBODY_HIP_SYMMETRIC = 4, /* each pair once, no global atomics (src/nbody_GPU_AOS_tiled.cu:44-220) */
    NBODY_HIP_AUTO = -1      /* the autotuned plan, else SYMMETRIC for whole systems of N >= 1024, else LDS */
};

/* Options of the general entry; zero / -1 fields mean "choose".
 * For NBODY_HIP_SYMMETRIC, i_per_lane is 2, 4 or 8 and jslice is the block
 * size B = 64 * i_per_lane * {1, 2, 4} (2 with 1 or 4); its block slots are
 * always summed with TwoSum, and it computes whole systems only (base 0, n N).
 * waves_per_simd caps its resident waves at about that many per SIMD: the
 * launch reserves extra LDS per workgroup, which bounds the workgroups per CU
 * (the dispatcher may still load the four SIMDs unevenly); how the block-pair
 * tasks divide into the resident waves changes how evenly the launch ends. */
typedef struct nbody_hip_options {
    int variant;        /* NBODY_HIP_* or NBODY_HIP_AUTO */
    int i_per_lane;     /* 1, 2, 4 (DPP also 8; symmetric 2, 4, 8), or 0 = auto */
    int jslice;         /* j bodies per slice (symmetric: block size), or 0 = auto */
    int compensated;    /* 1 TwoSum-compensated sums, 0 plain, -1 auto (on for N > 65536);
                           2 every term added in FP64 and the partial sums kept in FP64, rounded
                           once to the FP32 force (LDS and symmetric variants; symmetric with
                           i_per_lane 4 or 8 and blocks of 64 * i_per_lane, default 4) */
    int newton;         /* 1 adds a Newton step after v_rsq_f32 */
    int waves_per_simd; /* symmetric only: 1..8, or 0 = as many as registers and LDS allow */
    size_t slot_bytes;  /* symmetric only: bytes of block-slot partials one launch may keep, or 0 = 4 GiB.
                           A system whose slots (12 N^2 / B bytes) need more is evaluated in column
                           bands of blocks, one launch each, with 24 more bytes per body for the
                           running sums between them */
} nbody_hip_options;

/* ---- algorithm-table entries (AOS_GPU contract above) -------------------- */

/* Rung 1: replaces ComputeGravitation_GPU_AOS (src/nbody_GPU_AOS.cu:77-94). */
float ComputeGravitation_HIP_AOS(float *force, float *posMass, float softeningSquared, size_t N);
/* LDS-tiled: replaces ComputeGravitation_GPU_Shared (src/nbody_GPU_Shared.cu:83-103). */
float ComputeGravitation_HIP_LDS(float *force, float *posMass, float softeningSquared, size_t N);
/* DPP tile rotation: replaces ComputeGravitation_GPU_Shuffle (src/nbody_GPU_Shuffle.cu:91-107). */
float ComputeGravitation_HIP_DPP(float *force, float *posMass, float softeningSquared, size_t N);
/* Scalar-memory broadcast: replaces ComputeGravitation_GPU_AOS_const (src/nbody_GPU_AOS_const.cu:84-121). */
float ComputeGravitation_HIP_SGPR(float *force, float *posMass, float softeningSquared, size_t N);
/* Symmetric tiles: replaces ComputeGravitation_GPU_AOS_tiled (src/nbody_GPU_AOS_tiled.cu:196-220),
 * with the reaction scaled by m_i (the reference scales it by m_j). */
float ComputeGravitation_HIP_Symmetric(float *force, float *posMass, float softeningSquared, size_t N);
/* The library's choice for N: the plan nbody_hip_autotune found for this N
 * if it ran, else the symmetric kernel (whole systems, N >= 1024) or LDS. */
float ComputeGravitation_HIP(float *force, float *posMass, float softeningSquared, size_t N);

/* Host-pointer twin with the ALGORITHM_AOS contract (src/nbody.cu:343-349): force and posMass are
 * HOST pointers; the entry stages them through device memory itself. Returns milliseconds including
 * both transfers, like the host-pointer MultiGPU entry (src/nbody_multiGPU.cu:65-144). Slots into a
 * CUDA-less build of the reference driver, whose GPU types are gated off (src/nbody.cu:310-311). */
float ComputeGravitation_HIP_Host(float *force, float *posMass, float softeningSquared, size_t N);

/* ---- general entry ---------------------------------------------------------- */

/* Force on bodies [base, base+n) of an N-body system, written to force[0 .. 4n)
 * (local index, as src/nbody_multiGPU_shared.cuh:42-84), on `stream` (a
 * hipStream_t, NULL = null stream), asynchronously. opt may be NULL. Returns 0
 * or a nonzero error code (message in nbody_hip_last_error()).
 * Plans with j slices or symmetric blocks keep partial sums in one scratch
 * buffer per device, shared by all calls: a call on another stream than the
 * previous user of that buffer first waits (hipStreamWaitEvent) for the
 * previous user's kernels, so calls on different streams serialise on it.
 * The buffer grows (is reallocated) when a call needs more, so such a plan is
 * refused on a stream that is capturing a graph. */
int nbody_hip_force(float *force, const float *posMass, float softeningSquared, size_t N, size_t base,
                    size_t n, const nbody_hip_options *opt, void *stream);

/* One GPU's share of a whole system's block pairs, for running the symmetric
 * plan on G GPUs. It has no counterpart in the reference, whose multi-GPU path
 * splits the i range instead (src/nbody_multiGPU.cu:87-124; that split is
 * nbody_hip_force with base, n). Share k of G evaluates the block pairs whose
 * second block lies in the share's range of block columns, about 1/G of the
 * work, for both bodies of each pair, and writes every body's partial force
 * in FP64 to force64 (4 doubles per body; x, y, z written, w untouched). The
 * G shares' partial forces add up to the force: across G GPUs, one
 * reduce-scatter (sum, FP64) of them gives each GPU its own bodies' forces.
 * FP64 keeps each share's running TwoSum value whole: a share's partial force
 * on a body can be much larger than the net force. opt as for nbody_hip_force,
 * with variant NBODY_HIP_AUTO or NBODY_HIP_SYMMETRIC. Asynchronous on stream,
 * ordered with the library's other uses of its shared scratch as
 * nbody_hip_force. Returns 0 or an error code. */
int nbody_hip_force_pair_share(double *force64, const float *posMass, float softeningSquared, size_t N, int share,
                               int shares, const nbody_hip_options *opt, void *stream);

/* Bytes of the fixed device array that holds nbody_hip_force_noalloc's
 * partial sums: one per GPU the library runs on, held until the process ends. */
#define NBODY_HIP_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* No counterpart in the reference. For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc. The same computation as nbody_hip_force, with the same
 * arguments and kernels, but it allocates nothing: its partial sums live in
 * a fixed array of NBODY_HIP_NOALLOC_SCRATCH_BYTES per GPU. Its plan is
 * nbody_hip_force's when that plan's partial sums fit there. Otherwise the
 * symmetric plan keeps its slots in narrower column bands (slot_bytes) and,
 * unless opt fixes the shape (i_per_lane, jslice or waves_per_simd), takes
 * larger blocks; with variant NBODY_HIP_AUTO a one-sided plan follows, with
 * fewer, longer j slices, down to one slice, which keeps no partial sums. A
 * fixed shape whose partial sums cannot fit is refused with an error and
 * nothing is launched.
 * Every call on a GPU shares that GPU's array, which keeps the last call's
 * contents; the kernels write every partial sum they read, so nothing needs
 * to start at zero. Calls that overlap on one GPU (on different streams, or
 * graph replays that do) race on the array and give wrong results: order
 * them. The first call on a device looks up the array's address
 * (hipGetSymbolAddress) and creates the library's per-device events; make it
 * before any graph capture. Returns 0 or an error code, as nbody_hip_force. */
int nbody_hip_force_noalloc(float *force, const float *posMass, float softeningSquared, size_t N, size_t base,
                            size_t n, const nbody_hip_options *opt, void *stream);

/* integrateGravitation (src/nbody.cu:158-236) on device: for each body,
 * v += invMass * (F dt); v *= damping; x += v dt, with invMass = velInvMass.w.
 * posMass/velInvMass/force are device float4 arrays. Synchronous on the null
 * stream; returns elapsed ms, 0.0f on failure. */
float integrateGravitation_HIP(float *posMass, float *velInvMass, const float *force, float dt, float damping,
                               size_t N);
/* Asynchronous twin on `stream`; returns 0 or an error code. */
int nbody_hip_integrate(float *posMass, float *velInvMass, const float *force, float dt, float damping, size_t N,
                        void *stream);

/* ---- FP64 diagnostics ---------------------------------------------------- */

/* FP64 direct sum of the same force (device pointers; force64 holds 4 doubles
 * per body, w untouched), softening taken as the double given. Synchronous;
 * returns ms or 0.0f on failure. */
float nbody_hip_force_fp64(double *force64, const float *posMass, double softeningSquared, size_t N);
/* FP64 kinetic energy, softened potential energy and momentum of a device
 * state: out[0] kinetic, out[1] potential, out[2..4] momentum (host array of
 * 5). Synchronous; returns 0 or an error code. */
int nbody_hip_energy_fp64(double *out, const float *posMass, const float *velInvMass, double softeningSquared,
                          size_t N);

/* ---- plans and tuning ---------------------------------------------------- */

/* The plan the library would use: fills variant, i_per_lane, jslice,
 * compensated, newton for an n-body i range of an N-body system. */
int nbody_hip_plan(size_t N, size_t n, const nbody_hip_options *opt, nbody_hip_options *plan);
/* Time candidate plans of `variant` (NBODY_HIP_AUTO: the LDS, SGPR and
 * symmetric families) on the given device data, force kernel plus slice
 * reduction, and keep the fastest for this N and variant; later calls with
 * the same variant and default shape use it. Writes force. Returns the best
 * ms, 0.0f on failure. */
float nbody_hip_autotune(float *force, const float *posMass, float softeningSquared, size_t N, int variant);

/* ---- measurement --------------------------------------------------------- */

/* Kernel timer: while enabled, every force evaluation records a pair of
 * events on its own stream around the force kernel alone (not the slice
 * reduction). enable = 1 clears the record and starts; 0 stops. At most
 * 16384 launches are recorded between reads. */
int nbody_hip_kernel_timer(int enable);
/* Sum of the recorded force-kernel times and the number of launches;
 * waits for the recorded events. */
int nbody_hip_kernel_timer_read(double *total_ms, int *launches);

/* ---- misc ---------------------------------------------------------------- */

const char *nbody_hip_last_error(void);
int nbody_hip_abi_version(void);
/* Peak-relevant device facts: compute units and max engine clock (kHz). */
int nbody_hip_device_info(int *compute_units, int *clock_khz, char *arch, size_t arch_len);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_HIP_H */
