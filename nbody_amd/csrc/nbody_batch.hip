// Batched force and integrate kernels for many independent systems: a
// persistent grid of one-wave workgroups pulls ForceTasks (block pairs,
// diagonal blocks, compensated rows) from a device work queue, then one
// launch adds every body's slots and integrates it. Task lists are built on
// the host, most expensive first.
#include "nbody_device.h"
#include "nbody_kernels.h"
#include "nbody_sym.h"

#include <algorithm>

namespace nbody {

namespace {

constexpr int kWave = 64;
constexpr int kChunk = 128;                  // j bodies between TwoSum folds of a ROW_COMP task
constexpr int kPairA = kBatchBlock / kWave;  // bodies per lane of a PAIR task
constexpr int kBodyTaskBodies = 256;         // bodies per BodyTask (one thread each, per component)

static_assert(kBatchBlock % kWave == 0 && kPairA == 8, "PAIR tasks are compiled for 8 bodies per lane");

__global__ __launch_bounds__(kWave) void k_force_batch(const ForceTask* __restrict__ tasks, int ntasks,
                                                       unsigned* __restrict__ counter)
{
    __shared__ float jacc[3][kBatchBlock];
    // Workgroup b's first task is task b; later tickets come from the
    // counter, offset by the grid. All workgroups start together, and their
    // first tickets from one atomic counter would queue up behind each other
    // (the launch measured 3 % slower that way).
    for (bool first = true;; first = false) {
        int t = (int)blockIdx.x;
        if (!first && threadIdx.x == 0) t = (int)atomicAdd(counter, 1u) + (int)gridDim.x;
        // one wave per workgroup: lane 0's ticket is the wave's task, kept scalar
        t = __builtin_amdgcn_readfirstlane(t);
        if (t >= ntasks) return;
        const ForceTask task = tasks[t];
        if (task.kind == TASK_PAIR) {
            sym_pair<kPairA, 1, false, true>(task.pm, task.n, task.eps2, SlotMap::whole(task.part, task.nb, task.npad),
                                             task.a, task.b, jacc);
            continue;
        }
        if (task.kind == TASK_DIAG) {
            sym_diag_block(task.pm, task.n, task.npad, task.eps2, task.part, task.a);
            continue;
        }
        // TASK_ROW_COMP: 64 bodies against task.len bodies, TwoSum-folded chunks
        const int i = task.a + (int)threadIdx.x;
        const float4 q = task.pm[min(i, task.n - 1)];
        Lanes<1> me, hi, lo;
        me.x[0] = q.x;
        me.y[0] = q.y;
        me.z[0] = q.z;
        zero<1>(hi);
        zero<1>(lo);
        const float4* src = task.pm + task.b;
        for (int c = 0; c < task.len; c += kChunk) {
            Lanes<1> part;
            zero<1>(part);
            sweep_const<1, false>(src + c, min(kChunk, task.len - c), task.eps2, me, part);
            fold<1>(hi, lo, part);
        }
        hi.x[0] += lo.x[0];
        hi.y[0] += lo.y[0];
        hi.z[0] += lo.z[0];
        if (i < task.n) {
            const size_t plane = (size_t)task.nb * task.npad;
            const size_t o = (size_t)task.slot * task.npad + i;
            task.part[o] = hi.x[0];
            task.part[plane + o] = hi.y[0];
            task.part[2 * plane + o] = hi.z[0];
        }
    }
}

// 3 x kBodyTaskBodies threads per task: the threads of component c add that
// component of the task's bodies over their slots (eight slots in flight,
// each slot row read as one contiguous 1 KiB run), then the first
// kBodyTaskBodies threads integrate. The slot sum is k_reduce_band's (slot
// order, TwoSum, rounded once) and the update is k_integrate's, both in
// nbody_kernels.hip.
__global__ __launch_bounds__(3 * kBodyTaskBodies) void k_reduce_integrate_batch(const BodyTask* __restrict__ tasks,
                                                                                float dt, float damping,
                                                                                unsigned* __restrict__ counter)
{
    constexpr int G = 8;
    __shared__ float sum[3][kBodyTaskBodies];
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
    const BodyTask task = tasks[blockIdx.x];
    const int lane = threadIdx.x % kBodyTaskBodies;
    const int c = threadIdx.x / kBodyTaskBodies;  // wave-uniform component
    if (lane < task.ni) {
        const float* src = task.part + (size_t)c * task.nb * task.npad + task.i0 + lane;
        float s = 0.f, e = 0.f;
        int k = 0;
        for (; k + G <= task.nb; k += G) {
            float v[G];
#pragma unroll
            for (int g = 0; g < G; ++g) v[g] = src[(size_t)(k + g) * task.npad];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                float err;
                s = two_sum(s, v[g], err);
                e += err;
            }
        }
        for (; k < task.nb; ++k) {
            float err;
            s = two_sum(s, src[(size_t)k * task.npad], err);
            e += err;
        }
        sum[c][lane] = s + e;
    }
    __syncthreads();
    if (c != 0 || lane >= task.ni) return;
    const int i = task.i0 + lane;
    float4 g = task.force[i];
    g.x = sum[0][lane];
    g.y = sum[1][lane];
    g.z = sum[2][lane];
    task.force[i] = g;
    float4 p = task.pm[i];
    float4 v = task.vm[i];
    v.x = __builtin_fmaf(v.w, g.x * dt, v.x);
    v.y = __builtin_fmaf(v.w, g.y * dt, v.y);
    v.z = __builtin_fmaf(v.w, g.z * dt, v.z);
    v.x *= damping;
    v.y *= damping;
    v.z *= damping;
    p.x = __builtin_fmaf(v.x, dt, p.x);
    p.y = __builtin_fmaf(v.y, dt, p.y);
    p.z = __builtin_fmaf(v.z, dt, p.z);
    task.pm[i] = p;
    task.vm[i] = v;
}

// Work of a ROW_COMP task in VALU instructions per lane (12 + rsq per
// one-sided interaction), for ordering those tasks.
double row_cost(const ForceTask& t) { return (double)t.len * 13.0; }

// Cost of a block task in halves of a 64 x 64 tile block: a pair meets the
// tiles of block Q that hold bodies (one-wave pairs skip the padding tiles)
// with all 8 tiles of block P, 16 per Q tile, 128 for a full pair; a
// diagonal block of T tiles takes each pair once, T^2, 64 when full.
constexpr int kMaxTileKey = 16 * (kBatchBlock / kWave);

int tile_key(const ForceTask& t)
{
    if (t.kind == TASK_PAIR) {
        const int qlen = std::min(kBatchBlock, t.n - t.b * kBatchBlock);
        return 16 * ((qlen + kWave - 1) / kWave);
    }
    const int tiles = (t.len + kWave - 1) / kWave;
    return tiles * tiles;
}

// Append block tasks to `out` most expensive first: a stable counting sort
// on tile_key, linear in the task count. `each(f)` calls f on every task.
template <class Each>
void append_by_tile_key(std::vector<ForceTask>& out, const Each& each)
{
    size_t start[kMaxTileKey + 1] = {};
    each([&](const ForceTask& t) { ++start[tile_key(t)]; });
    size_t pos = out.size();
    for (int k = kMaxTileKey; k >= 0; --k) {
        const size_t c = start[k];
        start[k] = pos;
        pos += c;
    }
    out.resize(pos);
    each([&](const ForceTask& t) { out[start[tile_key(t)]++] = t; });
}

}  // namespace

void batch_layout(size_t n, int* nb, int* npad)
{
    if (n <= kBatchPairMaxN) {
        *nb = (int)((n + kBatchBlock - 1) / kBatchBlock);
        *npad = *nb * kBatchBlock;
    } else {
        *nb = 1;
        *npad = (int)n;
    }
}

size_t batch_part_floats(size_t n)
{
    int nb, npad;
    batch_layout(n, &nb, &npad);
    return (size_t)3 * nb * npad;
}

void append_force_tasks(std::vector<ForceTask>& out, const float4* pm, float* part, size_t n, float eps2)
{
    int nb, npad;
    batch_layout(n, &nb, &npad);
    ForceTask t{};
    t.pm = pm;
    t.part = part;
    t.n = (int)n;
    t.nb = nb;
    t.npad = npad;
    t.eps2 = eps2;
    if (n > kBatchPairMaxN) {
        t.kind = TASK_ROW_COMP;
        t.b = 0;
        t.len = (int)n;
        t.slot = 0;
        for (size_t i = 0; i < n; i += kWave) {
            t.a = (int)i;
            out.push_back(t);
        }
        return;
    }
    t.kind = TASK_PAIR;
    for (int q = 1; q < nb; ++q)
        for (int p = 0; p < q; ++p) {
            t.a = p;
            t.b = q;
            out.push_back(t);
        }
    t.kind = TASK_DIAG;
    t.b = 0;
    for (int p = 0; p < nb; ++p) {
        t.a = p;
        t.len = (int)std::min<size_t>(kBatchBlock, n - (size_t)p * kBatchBlock);
        t.slot = p;
        out.push_back(t);
    }
}

void append_body_tasks(std::vector<BodyTask>& out, float4* pm, float4* vm, float4* force, const float* part,
                       size_t n)
{
    int nb, npad;
    batch_layout(n, &nb, &npad);
    for (size_t i = 0; i < n; i += kBodyTaskBodies) {
        BodyTask t{};
        t.pm = pm;
        t.vm = vm;
        t.force = force;
        t.part = part;
        t.i0 = (int)i;
        t.ni = (int)std::min<size_t>(kBodyTaskBodies, n - i);
        t.nb = nb;
        t.npad = npad;
        out.push_back(t);
    }
}

void order_force_tasks(std::vector<ForceTask>& tasks)
{
    // A ROW_COMP task (64 bodies against more than 2^18) outweighs any block
    // task: those go first, by cost; the block pairs and diagonal blocks
    // follow by tile_key, so the cheapest tasks close the launch.
    std::vector<ForceTask> big, blocks;
    for (const ForceTask& t : tasks) (t.kind == TASK_ROW_COMP ? big : blocks).push_back(t);
    std::stable_sort(big.begin(), big.end(),
                     [](const ForceTask& a, const ForceTask& b) { return row_cost(a) > row_cost(b); });
    tasks.assign(big.begin(), big.end());
    append_by_tile_key(tasks, [&](const auto& f) {
        for (const ForceTask& t : blocks) f(t);
    });
}

void build_force_task_bands(ForceTaskBands& out, const float4* pm, float* part, size_t n, float eps2)
{
    std::vector<ForceTask> all;
    append_force_tasks(all, pm, part, n, eps2);
    out.big.clear();
    out.pairs.clear();
    out.diag.clear();
    for (const ForceTask& t : all) {
        if (t.kind == TASK_PAIR) out.pairs.push_back(t);
        else if (t.kind == TASK_ROW_COMP) out.big.push_back(t);
        else out.diag.push_back(t);
    }
    // within one system a ragged last block (fewer bodies) goes last
    std::stable_sort(out.diag.begin(), out.diag.end(),
                     [](const ForceTask& a, const ForceTask& b) { return a.len > b.len; });
}

void concat_force_task_bands(std::vector<ForceTask>& out, const std::vector<const ForceTaskBands*>& systems)
{
    out.clear();
    size_t total = 0;
    for (const ForceTaskBands* b : systems) total += b->big.size() + b->pairs.size() + b->diag.size();
    out.reserve(total);
    // the large systems' rows, largest system first
    std::vector<const ForceTaskBands*> by_size;
    for (const ForceTaskBands* b : systems)
        if (!b->big.empty()) by_size.push_back(b);
    std::stable_sort(by_size.begin(), by_size.end(),
                     [](const ForceTaskBands* a, const ForceTaskBands* b) { return a->big[0].len > b->big[0].len; });
    for (const ForceTaskBands* b : by_size) out.insert(out.end(), b->big.begin(), b->big.end());
    // then every system's block tasks, most expensive first
    append_by_tile_key(out, [&](const auto& f) {
        for (const ForceTaskBands* b : systems)
            for (const ForceTask& t : b->pairs) f(t);
        for (const ForceTaskBands* b : systems)
            for (const ForceTask& t : b->diag) f(t);
    });
}

int batch_force_blocks_per_cu()
{
    static int cached = 0;
    if (cached == 0) {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_force_batch, kWave, 0) != hipSuccess ||
            blocks < 1)
            blocks = 16;
        cached = blocks;
    }
    return cached;
}

hipError_t launch_force_batch(const ForceTask* tasks, int ntasks, unsigned* counter, int blocks, hipStream_t stream)
{
    if (ntasks <= 0) return hipSuccess;
    blocks = blocks < ntasks ? blocks : ntasks;
    k_force_batch<<<blocks, kWave, 0, stream>>>(tasks, ntasks, counter);
    return hipGetLastError();
}

hipError_t launch_reduce_integrate_batch(const BodyTask* tasks, int ntasks, float dt, float damping,
                                         unsigned* counter, hipStream_t stream)
{
    if (ntasks <= 0) return hipSuccess;
    k_reduce_integrate_batch<<<ntasks, 3 * kBodyTaskBodies, 0, stream>>>(tasks, dt, damping, counter);
    return hipGetLastError();
}

}  // namespace nbody
