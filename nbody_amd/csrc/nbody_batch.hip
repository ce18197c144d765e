// Batched force and integrate kernels for many independent systems in one
// launch: a persistent grid of one-wave workgroups pulls (system, 64-body
// block) tasks from a device work queue (an atomic counter); the caller
// orders the tasks largest system first.
#include "nbody_device.h"
#include "nbody_kernels.h"

namespace nbody {

namespace {

constexpr int kWave = 64;
constexpr int kChunk = 128;

__global__ __launch_bounds__(kWave) void k_force_batch(const BatchTask* __restrict__ tasks, int ntasks,
                                                       unsigned* __restrict__ counter)
{
    for (;;) {
        int t = 0;
        if (threadIdx.x == 0) t = (int)atomicAdd(counter, 1u);
        // one wave per workgroup: lane 0's ticket is the wave's task, kept scalar
        t = __builtin_amdgcn_readfirstlane(t);
        if (t >= ntasks) return;
        const BatchTask task = tasks[t];
        const float4* __restrict__ pm = task.pm;
        Lanes<1> me, hi, lo;
        const float4 q = pm[task.i0 + min((int)threadIdx.x, task.ni - 1)];
        me.x[0] = q.x;
        me.y[0] = q.y;
        me.z[0] = q.z;
        zero<1>(hi);
        zero<1>(lo);
        for (int c = 0; c < task.n; c += kChunk) {
            Lanes<1> part;
            zero<1>(part);
            sweep_const<1, false>(pm + c, min(kChunk, task.n - c), task.eps2, me, part);
            fold<1>(hi, lo, part);
        }
        if ((int)threadIdx.x < task.ni) {
            float* f = reinterpret_cast<float*>(task.force + task.i0 + threadIdx.x);
            f[0] = hi.x[0] + lo.x[0];
            f[1] = hi.y[0] + lo.y[0];
            f[2] = hi.z[0] + lo.z[0];
        }
    }
}

// One wave per task; same arithmetic as k_integrate (nbody_kernels.hip).
__global__ __launch_bounds__(kWave) void k_integrate_batch(const BatchTask* __restrict__ tasks, float dt,
                                                           float damping, unsigned* __restrict__ counter)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
    const BatchTask task = tasks[blockIdx.x];
    if ((int)threadIdx.x >= task.ni) return;
    const int i = task.i0 + threadIdx.x;
    float4 p = task.pm[i];
    float4 v = task.vm[i];
    const float4 g = task.force[i];
    v.x = __builtin_fmaf(v.w, g.x * dt, v.x);
    v.y = __builtin_fmaf(v.w, g.y * dt, v.y);
    v.z = __builtin_fmaf(v.w, g.z * dt, v.z);
    v.x *= damping;
    v.y *= damping;
    v.z *= damping;
    p.x = __builtin_fmaf(v.x, dt, p.x);
    p.y = __builtin_fmaf(v.y, dt, p.y);
    p.z = __builtin_fmaf(v.z, dt, p.z);
    // the batch's own integrate writes pm: it runs after the force launch has finished
    const_cast<float4*>(task.pm)[i] = p;
    task.vm[i] = v;
}

}  // namespace

hipError_t launch_force_batch(const BatchTask* tasks, int ntasks, unsigned* counter, int blocks, hipStream_t stream)
{
    if (ntasks <= 0) return hipSuccess;
    blocks = blocks < ntasks ? blocks : ntasks;
    k_force_batch<<<blocks, kWave, 0, stream>>>(tasks, ntasks, counter);
    return hipGetLastError();
}

hipError_t launch_integrate_batch(const BatchTask* tasks, int ntasks, float dt, float damping, unsigned* counter,
                                  hipStream_t stream)
{
    if (ntasks <= 0) return hipSuccess;
    k_integrate_batch<<<ntasks, kWave, 0, stream>>>(tasks, dt, damping, counter);
    return hipGetLastError();
}

}  // namespace nbody
