// Host-visible description of the force kernels and their launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <utility>
#include <vector>

namespace nbody {

// How the wave-uniform j-body reaches the 64 lanes of a wave.
enum Variant : int {
    VAR_AOS = 0,        // one i per lane over the whole j range (the ladder's first rung)
    VAR_LDS = 1,        // 256-body j tiles staged in LDS, broadcast ds_read_b128
    VAR_DPP = 2,        // 64-body j tile in VGPRs, rotated one lane per step (DPP wave_ror:1)
    VAR_SGPR = 3,       // scalar loads: j lives in SGPRs, used as VALU operands
    VAR_SYMMETRIC = 4,  // each pair once, reaction kept in LDS (no global atomics)
    VAR_COUNT = 5
};

// One launch's decomposition: i bodies per lane, j range split into slices
// whose partial sums are reduced in slice order (deterministic).
// For VAR_SYMMETRIC, jslice is the block size B = 64 * ipl * waves (waves =
// 1, 2 or 4 per workgroup) and nslices the number of blocks.
struct ForcePlan {
    int variant = VAR_SGPR;
    int ipl = 1;        // i bodies per lane: 1, 2, 4 (symmetric: 2, 4 or 8)
    int jslice = 0;     // j bodies per slice (symmetric: bodies per block)
    int nslices = 1;    // ceil(n / jslice)
    bool comp = false;  // TwoSum-compensated accumulation
    bool acc64 = false; // every term added in FP64, partial sums kept in FP64 (LDS and symmetric)
    bool newton = false;
    int occ = 0;        // symmetric: waves per SIMD cap (0 = register-limited occupancy)
    size_t band_cap = 0;  // symmetric: bytes of slot partials one band may keep (0 = kSymBandCap)
    int col0 = 0, col1 = 0;  // symmetric: the column blocks [col0, col1) of the block triangle (0, 0 = all)
};

// Bytes of block-slot partials one launch of the symmetric plan keeps by
// default; a system whose slots need more runs in several column bands.
constexpr size_t kSymBandCap = size_t(4) << 30;

// The column bands [q0, q1) of a symmetric plan with nb blocks of B bodies
// whose partials (elem bytes each) fit in cap bytes each (a band of one
// column always runs).
std::vector<std::pair<int, int>> symmetric_bands(int nb, int B, size_t cap, int elem = 4, int c0 = 0, int c1 = -1);

// The column blocks [c0, c1) of share k of G of a symmetric plan with nb
// blocks: consecutive column ranges of about equal work (column Q holds Q
// block pairs and one diagonal block).
void symmetric_share_columns(int nb, int k, int G, int* c0, int* c1);

struct ForceParams {
    const float4* pm = nullptr;  // posMass (x, y, z, m) of the whole system: j source and i source
    int n = 0;                   // bodies in the system
    int i0 = 0;                  // first i body
    int ni = 0;                  // number of i bodies; outputs are local: force[4 * (i - i0)]
    float eps2 = 0.f;
    int jslice = 0;
    int nslices = 1;
    float4* hi = nullptr;        // slice partials [nslices][ni]; nullptr -> write force directly
    float4* lo = nullptr;        // compensation partials [nslices][ni] (comp only)
    double* hi64 = nullptr;      // FP64 partials (acc64): slices [nslices][3][ni], or the symmetric slot planes
    double* force64 = nullptr;   // symmetric column share: FP64 partial forces, 4 doubles per body (w untouched),
                                 // written instead of force
    float* force = nullptr;      // 4 floats per i body; x, y, z written, w left alone
};

// Whether (i per lane, block size) is a compiled symmetric shape.
bool symmetric_shape_ok(int ipl, int block);

// Waves per SIMD the plan's symmetric kernel reaches when its registers and
// static LDS alone limit it (0 for other variants or on error).
int symmetric_natural_occupancy(const ForcePlan& plan);

// Bytes of slice-partial scratch a plan needs for ni bodies (0 when direct).
size_t plan_scratch_bytes(const ForcePlan& plan, int ni);

// Queue the force kernel and, for sliced plans, the slice reduction on
// `stream`. When before/after are given they are recorded on the stream
// around the force kernel alone.
hipError_t launch_force(const ForcePlan& plan, const ForceParams& p, hipStream_t stream,
                        hipEvent_t before = nullptr, hipEvent_t after = nullptr);

// integrateGravitation (src/nbody.cu:158-236) on float4 records.
hipError_t launch_integrate(float4* posMass, float4* velInvMass, const float4* force, float dt,
                            float damping, int n, hipStream_t stream);

// FP64 direct sum (diagnostic): out holds 4 doubles per i body (w untouched).
// partial must hold nslices * ni * 3 doubles.
hipError_t launch_force_fp64(double* out, const float4* pm, int n, int i0, int ni, double eps2,
                             int jslice, double* partial, hipStream_t stream);

// FP64 energy/momentum diagnostic: acc[0] kinetic, acc[1] potential,
// acc[2..4] momentum; partial holds 5 * blocks doubles (see fp64_energy_blocks).
int fp64_energy_blocks(int n);
hipError_t launch_energy_fp64(double* acc, const float4* pm, const float4* vim, int n, double eps2,
                              double* partial, hipStream_t stream);

}  // namespace nbody

namespace nbody {

// ---- batches of independent systems (ensembles, served requests) ----------
//
// One persistent launch computes the forces of many systems: one-wave
// workgroups take ForceTasks from a device work queue (an atomic counter).
// A system of n <= kBatchPairMaxN bodies is cut into blocks of kBatchBlock
// and runs as the symmetric plan does (nbody_sym.h): one PAIR task per block
// pair P < Q and one DIAG task per diagonal block, each writing its own slot
// of the system's partial planes part[3][nb][npad]. A
// larger system (its slot planes would grow as n^2/kBatchBlock) runs as
// ROW_COMP tasks: 64 bodies against all n with TwoSum-folded chunks, into a
// single slot. A second launch adds each body's slots in slot order (TwoSum)
// and integrates it. A body's result therefore depends only on its own
// system, never on what else is in the batch or on scheduling.
//
// A batch with FP64 accumulation (compensated = 2) runs the FP64 symmetric
// plan's pieces instead (nbody_acc64.h): blocks of kBatchBlock64 (4 bodies
// per lane, one wave), PAIR64 tasks, four DIAG64 pieces per diagonal block
// (64 bodies one-sided against the block), FP64 slot planes summed in slot
// order and rounded once; a system above kBatchPairMaxN64 runs as ROW64
// tasks (64 bodies against all n, one-sided, FP64 sums) into one slot. A
// system of 1024 .. kBatchPairMaxN64 bodies then gets exactly the forces of
// the single-system FP64 symmetric plan (4 bodies per lane, blocks of 256).

constexpr int kBatchBlock = 512;                     // 8 bodies per lane, one wave
constexpr size_t kBatchPairMaxN = size_t(1) << 18;   // largest system run as block pairs
constexpr int kBatchBlock64 = 256;                   // FP64 accumulation: 4 bodies per lane, one wave
constexpr size_t kBatchPairMaxN64 = size_t(1) << 17; // FP64 accumulation: largest system run as block pairs

enum BatchTaskKind : int {
    TASK_PAIR = 0,
    TASK_DIAG = 1,
    TASK_ROW_COMP = 2,
    TASK_PAIR64 = 3,
    TASK_DIAG64 = 4,
    TASK_ROW64 = 5,
};

struct ForceTask {
    const float4* pm;  // the system's bodies
    float* part;       // the system's slot planes [3][nb][npad] (FP32 tasks)
    double* part64;    // the same in FP64 (FP64 tasks)
    int n;             // bodies in the system
    int nb;            // slots per body
    int npad;          // length of one slot row
    float eps2;
    int kind;          // BatchTaskKind
    int a;             // PAIR, DIAG: block P; ROW_COMP, ROW64: first i body (64 of them)
    int b;             // PAIR: block Q; DIAG64: piece (0..3); ROW_COMP, ROW64: first j body
    int len;           // DIAG, DIAG64: bodies in the block; ROW_COMP, ROW64: number of j bodies
    int slot;          // slot written (DIAG, DIAG64: P; ROW_COMP, ROW64: 0)
    int reserved;
};

// Reduction + integration of up to 256 consecutive bodies of one system.
struct BodyTask {
    float4* pm;
    float4* vm;
    float4* force;
    const float* part;     // FP32 slot planes, or
    const double* part64;  // FP64 slot planes
    int i0;
    int ni;
    int nb;
    int npad;
};

// Slots per body and slot row length of a batched system of n bodies.
void batch_layout(size_t n, bool acc64, int* nb, int* npad);
// Bytes of slot planes a batched system of n bodies needs (3 * nb * npad
// floats, or doubles with FP64 accumulation).
size_t batch_part_bytes(size_t n, bool acc64);
// Append one system's force tasks / body tasks; `part` is the system's slot
// planes (floats, or doubles when acc64).
void append_force_tasks(std::vector<ForceTask>& out, const float4* pm, void* part, size_t n, float eps2, bool acc64);
void append_body_tasks(std::vector<BodyTask>& out, float4* pm, float4* vm, float4* force, const void* part,
                       size_t n, bool acc64);
// Most expensive tasks first (stable), so the end of the launch holds the
// cheapest ones: compensated rows by cost, then the block tasks by their
// tiles (a block pair whose second block is a system's ragged last block
// skips its padding tiles; a diagonal block is about half a pair), the
// latter by a counting sort, linear in the task count.
void order_force_tasks(std::vector<ForceTask>& tasks);

// One system's force tasks split by kind: compensated rows of a large system
// (each outweighs a block task), block pairs, diagonal blocks. A server that
// rebuilds its task list whenever a request joins or leaves keeps these per
// request and concatenates them in the same order as order_force_tasks, in
// time linear in the task count.
struct ForceTaskBands {
    std::vector<ForceTask> big, pairs, diag;
};
void build_force_task_bands(ForceTaskBands& out, const float4* pm, void* part, size_t n, float eps2, bool acc64);
void concat_force_task_bands(std::vector<ForceTask>& out, const std::vector<const ForceTaskBands*>& systems);
// Resident one-wave force workgroups per CU (the persistent grid is this
// times the CU count) of the FP32 or the FP64-accumulation force kernel.
int batch_force_blocks_per_cu(bool acc64);

// Persistent force launch: `blocks` workgroups each start with the task of
// their own index, then take tasks from `counter` (which must read 0 at
// launch) until all ntasks are done. The tasks are all FP32 kinds, or all
// FP64 kinds with acc64.
hipError_t launch_force_batch(const ForceTask* tasks, int ntasks, unsigned* counter, int blocks, hipStream_t stream,
                              bool acc64);

// Sum every body's slots, write force (x, y, z; w untouched), integrate
// (dt, damping as integrateGravitation), and reset `counter` to 0 for the
// next force launch.
hipError_t launch_reduce_integrate_batch(const BodyTask* tasks, int ntasks, float dt, float damping,
                                         unsigned* counter, hipStream_t stream, bool acc64);

}  // namespace nbody
