// Batched force and integrate kernels for many independent systems: a
// persistent grid of one-wave workgroups pulls ForceTasks (block pairs,
// diagonal blocks, compensated rows) from a device work queue, then one
// launch adds every body's slots and integrates it. Task lists are built on
// the host, most expensive first. With FP64 accumulation the same scheme
// runs the FP64 symmetric plan's pieces (k_force_batch64).
#include "nbody_acc64.h"
#include "nbody_device.h"
#include "nbody_kernels.h"
#include "nbody_sym.h"

#include <algorithm>
#include <type_traits>

namespace nbody {

namespace {

constexpr int kWave = 64;
constexpr int kChunk = 128;                  // j bodies between TwoSum folds of a ROW_COMP task
constexpr int kPairA = kBatchBlock / kWave;  // bodies per lane of a PAIR task
constexpr int kBodyTaskBodies = 256;         // bodies per BodyTask (one thread each, per component)
constexpr int kPairA64 = kBatchBlock64 / kWave;  // bodies per lane of a PAIR64 task

static_assert(kBatchBlock % kWave == 0 && kPairA == 8, "PAIR tasks are compiled for 8 bodies per lane");
static_assert(kBatchBlock64 % kWave == 0 && kPairA64 == 4, "PAIR64 tasks are compiled for 4 bodies per lane");

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

// The FP64-accumulation tasks: PAIR64 block pairs and DIAG64 diagonal
// pieces are k_force_sym64's (nbody_kernels.hip, 4 bodies per lane, blocks
// of 256), ROW64 is 64 bodies against all n in j order into slot 0. No LDS:
// every tile's reactions go from the lanes to their slots.
__global__ __launch_bounds__(kWave) void k_force_batch64(const ForceTask* __restrict__ tasks, int ntasks,
                                                         unsigned* __restrict__ counter)
{
    for (bool first = true;; first = false) {
        int t = (int)blockIdx.x;
        if (!first && threadIdx.x == 0) t = (int)atomicAdd(counter, 1u) + (int)gridDim.x;
        t = __builtin_amdgcn_readfirstlane(t);
        if (t >= ntasks) return;
        const ForceTask task = tasks[t];
        const SlotMap64 slots = SlotMap64::whole(task.part64, task.nb, task.npad);
        if (task.kind == TASK_PAIR64) {
            sym_pair64<kPairA64, false>(task.pm, task.n, task.eps2, slots, task.a, task.b);
            continue;
        }
        if (task.kind == TASK_DIAG64) {
            sym_diag_piece64<kBatchBlock64, false>(task.pm, task.n, task.eps2, slots, task.a, task.b);
            continue;
        }
        // TASK_ROW64
        const int i = task.a + (int)threadIdx.x;
        double ax, ay, az;
        row64<false>(task.pm, task.n, task.eps2, i, task.b, task.len, ax, ay, az);
        if (i < task.n) {
            const size_t plane = (size_t)task.nb * task.npad;
            const size_t o = (size_t)task.slot * task.npad + i;
            task.part64[o] = ax;
            task.part64[plane + o] = ay;
            task.part64[2 * plane + o] = az;
        }
    }
}

// 3 x kBodyTaskBodies threads per task: the threads of component c add that
// component of the task's bodies over their slots (eight slots in flight,
// each slot row read as one contiguous run), then the first kBodyTaskBodies
// threads integrate. The slot sum is k_reduce_band's (slot order, TwoSum,
// rounded once) for FP32 slots and k_reduce_band64's (slot order, FP64,
// rounded once) for FP64 slots; the update is k_integrate's, all in
// nbody_kernels.hip.
template <class T>
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
    if constexpr (std::is_same<T, double>::value) {
        if (lane < task.ni) {
            const double* src = task.part64 + (size_t)c * task.nb * task.npad + task.i0 + lane;
            double s = 0.0;
            int k = 0;
            for (; k + G <= task.nb; k += G) {
                double v[G];
#pragma unroll
                for (int g = 0; g < G; ++g) v[g] = src[(size_t)(k + g) * task.npad];
#pragma unroll
                for (int g = 0; g < G; ++g) s += v[g];
            }
            for (; k < task.nb; ++k) s += src[(size_t)k * task.npad];
            sum[c][lane] = (float)s;
        }
    } else if (lane < task.ni) {
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
// one-sided interaction), for ordering those tasks; a ROW64 interaction
// takes about 20 (FP64 products and sums at half rate, conversions).
double row_cost(const ForceTask& t) { return (double)t.len * (t.kind == TASK_ROW64 ? 20.0 : 13.0); }

bool is_row(const ForceTask& t) { return t.kind == TASK_ROW_COMP || t.kind == TASK_ROW64; }

// Cost of a block task in halves of a 64 x 64 tile block: a pair meets the
// tiles of block Q that hold bodies (one-wave pairs skip the padding tiles)
// with all 8 tiles of block P, 16 per Q tile, 128 for a full pair; a
// diagonal block of T tiles takes each pair once, T^2, 64 when full. A
// PAIR64 task meets block P's 4 tiles, 8 per Q tile; a DIAG64 piece (one
// tile against the block, one-sided) 2 per tile of its block.
constexpr int kMaxTileKey = 16 * (kBatchBlock / kWave);

int tile_key(const ForceTask& t)
{
    const int ceil_tiles = (t.len + kWave - 1) / kWave;
    switch (t.kind) {
    case TASK_PAIR: return 16 * ((std::min(kBatchBlock, t.n - t.b * kBatchBlock) + kWave - 1) / kWave);
    case TASK_PAIR64: return 8 * ((std::min(kBatchBlock64, t.n - t.b * kBatchBlock64) + kWave - 1) / kWave);
    case TASK_DIAG64: return 2 * ceil_tiles;
    default: return ceil_tiles * ceil_tiles;
    }
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

void batch_layout(size_t n, bool acc64, int* nb, int* npad)
{
    const int B = acc64 ? kBatchBlock64 : kBatchBlock;
    if (n <= (acc64 ? kBatchPairMaxN64 : kBatchPairMaxN)) {
        *nb = (int)((n + B - 1) / B);
        *npad = *nb * B;
    } else {
        *nb = 1;
        *npad = (int)n;
    }
}

size_t batch_part_bytes(size_t n, bool acc64)
{
    int nb, npad;
    batch_layout(n, acc64, &nb, &npad);
    return (size_t)3 * nb * npad * (acc64 ? sizeof(double) : sizeof(float));
}

void append_force_tasks(std::vector<ForceTask>& out, const float4* pm, void* part, size_t n, float eps2, bool acc64)
{
    int nb, npad;
    batch_layout(n, acc64, &nb, &npad);
    const int B = acc64 ? kBatchBlock64 : kBatchBlock;
    ForceTask t{};
    t.pm = pm;
    if (acc64) t.part64 = static_cast<double*>(part);
    else t.part = static_cast<float*>(part);
    t.n = (int)n;
    t.nb = nb;
    t.npad = npad;
    t.eps2 = eps2;
    if (n > (acc64 ? kBatchPairMaxN64 : kBatchPairMaxN)) {
        t.kind = acc64 ? TASK_ROW64 : TASK_ROW_COMP;
        t.b = 0;
        t.len = (int)n;
        t.slot = 0;
        for (size_t i = 0; i < n; i += kWave) {
            t.a = (int)i;
            out.push_back(t);
        }
        return;
    }
    t.kind = acc64 ? TASK_PAIR64 : TASK_PAIR;
    for (int q = 1; q < nb; ++q)
        for (int p = 0; p < q; ++p) {
            t.a = p;
            t.b = q;
            out.push_back(t);
        }
    t.kind = acc64 ? TASK_DIAG64 : TASK_DIAG;
    t.b = 0;
    for (int p = 0; p < nb; ++p) {
        t.a = p;
        t.len = (int)std::min<size_t>(B, n - (size_t)p * B);
        t.slot = p;
        if (!acc64) {
            out.push_back(t);
            continue;
        }
        // the pieces of 64 bodies that hold any (a ragged last block has fewer)
        for (int k = 0; k * kWave < t.len; ++k) {
            t.b = k;
            out.push_back(t);
        }
        t.b = 0;
    }
}

void append_body_tasks(std::vector<BodyTask>& out, float4* pm, float4* vm, float4* force, const void* part,
                       size_t n, bool acc64)
{
    int nb, npad;
    batch_layout(n, acc64, &nb, &npad);
    for (size_t i = 0; i < n; i += kBodyTaskBodies) {
        BodyTask t{};
        t.pm = pm;
        t.vm = vm;
        t.force = force;
        if (acc64) t.part64 = static_cast<const double*>(part);
        else t.part = static_cast<const float*>(part);
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
    for (const ForceTask& t : tasks) (is_row(t) ? big : blocks).push_back(t);
    std::stable_sort(big.begin(), big.end(),
                     [](const ForceTask& a, const ForceTask& b) { return row_cost(a) > row_cost(b); });
    tasks.assign(big.begin(), big.end());
    append_by_tile_key(tasks, [&](const auto& f) {
        for (const ForceTask& t : blocks) f(t);
    });
}

void build_force_task_bands(ForceTaskBands& out, const float4* pm, void* part, size_t n, float eps2, bool acc64)
{
    std::vector<ForceTask> all;
    append_force_tasks(all, pm, part, n, eps2, acc64);
    out.big.clear();
    out.pairs.clear();
    out.diag.clear();
    for (const ForceTask& t : all) {
        if (t.kind == TASK_PAIR || t.kind == TASK_PAIR64) out.pairs.push_back(t);
        else if (is_row(t)) out.big.push_back(t);
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

int batch_force_blocks_per_cu(bool acc64)
{
    static int cached[2] = {0, 0};
    int& c = cached[acc64 ? 1 : 0];
    if (c == 0) {
        int blocks = 0;
        const hipError_t e =
            acc64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_force_batch64, kWave, 0)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_force_batch, kWave, 0);
        if (e != hipSuccess || blocks < 1) blocks = 16;
        c = blocks;
    }
    return c;
}

hipError_t launch_force_batch(const ForceTask* tasks, int ntasks, unsigned* counter, int blocks, hipStream_t stream,
                              bool acc64)
{
    if (ntasks <= 0) return hipSuccess;
    blocks = blocks < ntasks ? blocks : ntasks;
    if (acc64) k_force_batch64<<<blocks, kWave, 0, stream>>>(tasks, ntasks, counter);
    else k_force_batch<<<blocks, kWave, 0, stream>>>(tasks, ntasks, counter);
    return hipGetLastError();
}

hipError_t launch_reduce_integrate_batch(const BodyTask* tasks, int ntasks, float dt, float damping,
                                         unsigned* counter, hipStream_t stream, bool acc64)
{
    if (ntasks <= 0) return hipSuccess;
    if (acc64)
        k_reduce_integrate_batch<double><<<ntasks, 3 * kBodyTaskBodies, 0, stream>>>(tasks, dt, damping, counter);
    else
        k_reduce_integrate_batch<float><<<ntasks, 3 * kBodyTaskBodies, 0, stream>>>(tasks, dt, damping, counter);
    return hipGetLastError();
}

}  // namespace nbody
