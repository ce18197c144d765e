// Batched force and integrate kernels for many independent systems in one
// launch: a persistent grid pulls (system, i-block) tasks from a device work
// queue (an atomic counter), largest systems first.
#include "nbody_device.h"
#include "nbody_kernels.h"

namespace nbody {

namespace {

constexpr int kBlock = 256;
constexpr int kChunk = 128;

template <int I>
__global__ __launch_bounds__(kBlock) void k_force_batch(const float4* __restrict__ pm, float* __restrict__ force,
                                                        const BatchTask* __restrict__ tasks, int ntasks,
                                                        unsigned* __restrict__ counter)
{
    __shared__ int s_task;
    for (;;) {
        if (threadIdx.x == 0) s_task = (int)atomicAdd(counter, 1u);
        __syncthreads();
        // the task index is the same for the whole workgroup: keep it scalar
        const int t = __builtin_amdgcn_readfirstlane(s_task);
        __syncthreads();
        if (t >= ntasks) return;
        const BatchTask task = tasks[t];
        Lanes<I> me, hi, lo;
#pragma unroll
        for (int u = 0; u < I; ++u) {
            const int m = min((int)threadIdx.x + u * kBlock, task.ni - 1);
            const float4 q = pm[task.i0 + m];
            me.x[u] = q.x;
            me.y[u] = q.y;
            me.z[u] = q.z;
        }
        zero<I>(hi);
        zero<I>(lo);
        const float4* __restrict__ src = pm + task.j0;
        for (int c = 0; c < task.nj; c += kChunk) {
            Lanes<I> part;
            zero<I>(part);
            sweep_sgpr<I, false>(src + c, min(kChunk, task.nj - c), task.eps2, me, part);
            fold<I>(hi, lo, part);
        }
#pragma unroll
        for (int u = 0; u < I; ++u) {
            const int m = threadIdx.x + u * kBlock;
            if (m < task.ni) {
                float* f = force + 4 * (size_t)(task.i0 + m);
                f[0] = hi.x[u] + lo.x[u];
                f[1] = hi.y[u] + lo.y[u];
                f[2] = hi.z[u] + lo.z[u];
            }
        }
    }
}

// One workgroup per task; same arithmetic as k_integrate (nbody_kernels.hip).
__global__ __launch_bounds__(kBlock) void k_integrate_batch(float4* __restrict__ pm, float4* __restrict__ vim,
                                                            const float4* __restrict__ f,
                                                            const BatchTask* __restrict__ tasks, float dt,
                                                            float damping, unsigned* __restrict__ counter)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
    const BatchTask task = tasks[blockIdx.x];
    for (int m = threadIdx.x; m < task.ni; m += kBlock) {
        const int i = task.i0 + m;
        float4 p = pm[i];
        float4 v = vim[i];
        const float4 g = f[i];
        v.x = __builtin_fmaf(v.w, g.x * dt, v.x);
        v.y = __builtin_fmaf(v.w, g.y * dt, v.y);
        v.z = __builtin_fmaf(v.w, g.z * dt, v.z);
        v.x *= damping;
        v.y *= damping;
        v.z *= damping;
        p.x = __builtin_fmaf(v.x, dt, p.x);
        p.y = __builtin_fmaf(v.y, dt, p.y);
        p.z = __builtin_fmaf(v.z, dt, p.z);
        pm[i] = p;
        vim[i] = v;
    }
}

}  // namespace

hipError_t launch_force_batch(const float4* pm, float* force, const BatchTask* tasks, int ntasks, int ipl,
                              unsigned* counter, int blocks, hipStream_t stream)
{
    if (ntasks <= 0) return hipSuccess;
    blocks = blocks < ntasks ? blocks : ntasks;
    switch (ipl) {
    case 1: k_force_batch<1><<<blocks, kBlock, 0, stream>>>(pm, force, tasks, ntasks, counter); break;
    case 2: k_force_batch<2><<<blocks, kBlock, 0, stream>>>(pm, force, tasks, ntasks, counter); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_integrate_batch(float4* pm, float4* vim, const float4* force, const BatchTask* tasks, int ntasks,
                                  float dt, float damping, unsigned* counter, hipStream_t stream)
{
    if (ntasks <= 0) return hipSuccess;
    k_integrate_batch<<<ntasks, kBlock, 0, stream>>>(pm, vim, force, tasks, dt, damping, counter);
    return hipGetLastError();
}

}  // namespace nbody
