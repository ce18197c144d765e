// Device-resident stepping (include/nbody_engine.h): one system with a
// captured per-step graph, and ensembles of independent systems run either
// batched (a persistent work-queue force launch over all of them) or one
// system after another.
#include "nbody_engine.h"
#include "nbody_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

using namespace nbody;
using namespace nbody::detail;

namespace {

// Workgroups of the persistent batch kernel per CU: its ~50 VGPRs allow 8
// waves per SIMD, i.e. 8 four-wave workgroups per CU.
constexpr int kBatchBlocksPerCU = 8;
// i bodies per lane in a batch task (256 per task).
constexpr int kBatchIpl = 1;

template <class T>
hipError_t dev_alloc(T** p, size_t count)
{
    return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
}

struct Timing {
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t init()
    {
        hipError_t e = hipEventCreate(&a);
        return e == hipSuccess ? hipEventCreate(&b) : e;
    }
    void destroy()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        a = b = nullptr;
    }
};

// Start/stop of a synchronous timed region on a stream.
int timed_wait(hipStream_t s, Timing& t, float* ms)
{
    HIP_TRY(hipEventRecord(t.b, s));
    HIP_TRY(hipEventSynchronize(t.b));
    float v = 0.f;
    HIP_TRY(hipEventElapsedTime(&v, t.a, t.b));
    if (ms) *ms = v;
    return OK;
}

}  // namespace

struct nbody_hip_system {
    size_t n = 0;
    float eps2 = 0.f, dt = 0.f, damping = 1.f;
    hipStream_t stream = nullptr;
    float4* pm = nullptr;
    float4* vm = nullptr;
    float4* force = nullptr;
    Buffer scratch;
    ForcePlan plan;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    Timing timing;

    void release()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (pm) (void)hipFree(pm);
        if (vm) (void)hipFree(vm);
        if (force) (void)hipFree(force);
        if (scratch.ptr) (void)hipFree(scratch.ptr);
        if (stream) (void)hipStreamDestroy(stream);
        timing.destroy();
    }

    // One step queued on the stream: force (+ reduction) then integrate.
    int queue_step(DeviceState& d)
    {
        if (int r = run_force(d, plan, reinterpret_cast<float*>(force), reinterpret_cast<const float*>(pm), eps2, n,
                              0, n, stream, &scratch, false))
            return r;
        HIP_TRY(launch_integrate(pm, vm, force, dt, damping, (int)n, stream));
        return OK;
    }
};

struct nbody_hip_ensemble {
    size_t count = 0, total = 0;
    std::vector<size_t> n, off;
    std::vector<float> eps2;
    float dt = 0.f, damping = 1.f;
    hipStream_t stream = nullptr;
    float4* pm = nullptr;
    float4* vm = nullptr;
    float4* force = nullptr;
    BatchTask* tasks = nullptr;
    unsigned* counter = nullptr;
    int ntasks = 0;
    int blocks = 0;
    std::vector<ForcePlan> plans;  // per-system mode
    Buffer scratch;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    Timing timing;

    void release()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : {(void*)pm, (void*)vm, (void*)force, (void*)tasks, (void*)counter, scratch.ptr})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
        timing.destroy();
    }

    int queue_batched()
    {
        HIP_TRY(launch_force_batch(pm, reinterpret_cast<float*>(force), tasks, ntasks, kBatchIpl, counter, blocks,
                                   stream));
        HIP_TRY(launch_integrate_batch(pm, vm, force, tasks, ntasks, dt, damping, counter, stream));
        return OK;
    }

    int queue_per_system(DeviceState& d)
    {
        for (size_t s = 0; s < count; ++s) {
            float4* p = pm + off[s];
            if (int r = run_force(d, plans[s], reinterpret_cast<float*>(force + off[s]),
                                  reinterpret_cast<const float*>(p), eps2[s], n[s], 0, n[s], stream, &scratch, false))
                return r;
            HIP_TRY(launch_integrate(p, vm + off[s], force + off[s], dt, damping, (int)n[s], stream));
        }
        return OK;
    }
};

extern "C" {

nbody_hip_system* nbody_hip_system_create(const float* posMass, const float* velInvMass, size_t N,
                                          float softeningSquared, float dt, float damping,
                                          const nbody_hip_options* opt)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!velInvMass || check_args(posMass, posMass, N, 0, N)) {
        if (!velInvMass) set_error("null buffer");
        return nullptr;
    }
    DeviceState* d;
    if (device_state(&d)) return nullptr;
    auto* s = new nbody_hip_system();
    s->n = N;
    s->eps2 = softeningSquared;
    s->dt = dt;
    s->damping = damping;
    hipError_t e = hipSuccess;
    if (make_plan(*d, N, N, opt, &s->plan)) goto fail_plan;
    e = hipStreamCreate(&s->stream);
    if (e == hipSuccess) e = dev_alloc(&s->pm, N);
    if (e == hipSuccess) e = dev_alloc(&s->vm, N);
    if (e == hipSuccess) e = dev_alloc(&s->force, N);
    if (e == hipSuccess) e = s->scratch.reserve(plan_scratch_bytes(s->plan, (int)N));
    if (e == hipSuccess) e = hipMemcpy(s->pm, posMass, N * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->vm, velInvMass, N * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->force, 0, N * sizeof(float4));
    if (e == hipSuccess) e = s->timing.init();
    if (e != hipSuccess) {
        set_error("system create failed: %s", hipGetErrorString(e));
        goto fail_plan;
    }
    return s;
fail_plan:
    s->release();
    delete s;
    return nullptr;
}

int nbody_hip_system_run(nbody_hip_system* s, int steps, int use_graph, float* ms)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!s || steps < 0) {
        set_error("bad system or step count");
        return E_ARG;
    }
    DeviceState* d;
    if (int r = device_state(&d)) return r;
    if (use_graph && !s->exec) {
        // capture one step; the scratch was reserved at creation, so nothing allocates
        HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        const int r = s->queue_step(*d);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s->stream, &g);
        if (r) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        HIP_TRY(e);
        s->graph = g;
        HIP_TRY(hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipEventRecord(s->timing.a, s->stream));
    for (int k = 0; k < steps; ++k) {
        if (use_graph) {
            HIP_TRY(hipGraphLaunch(s->exec, s->stream));
        } else if (int r = s->queue_step(*d)) {
            return r;
        }
    }
    return timed_wait(s->stream, s->timing, ms);
}

int nbody_hip_system_read(const nbody_hip_system* s, float* posMass, float* velInvMass, float* force)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!s) {
        set_error("null system");
        return E_ARG;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    const size_t bytes = s->n * sizeof(float4);
    if (posMass) HIP_TRY(hipMemcpy(posMass, s->pm, bytes, hipMemcpyDeviceToHost));
    if (velInvMass) HIP_TRY(hipMemcpy(velInvMass, s->vm, bytes, hipMemcpyDeviceToHost));
    if (force) HIP_TRY(hipMemcpy(force, s->force, bytes, hipMemcpyDeviceToHost));
    return OK;
}

int nbody_hip_system_plan(const nbody_hip_system* s, nbody_hip_options* plan)
{
    if (!s || !plan) {
        set_error("null argument");
        return E_ARG;
    }
    plan->variant = s->plan.variant;
    plan->i_per_lane = s->plan.ipl;
    plan->jslice = s->plan.jslice;
    plan->compensated = s->plan.comp ? 1 : 0;
    plan->newton = s->plan.newton ? 1 : 0;
    return OK;
}

void nbody_hip_system_destroy(nbody_hip_system* s)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!s) return;
    (void)hipStreamSynchronize(s->stream);
    s->release();
    delete s;
}

nbody_hip_ensemble* nbody_hip_ensemble_create(size_t count, const size_t* N, const float* posMass,
                                              const float* velInvMass, const float* softeningSquared, float dt,
                                              float damping)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (count == 0 || !N || !posMass || !velInvMass || !softeningSquared) {
        set_error("empty ensemble or null argument");
        return nullptr;
    }
    DeviceState* d;
    if (device_state(&d)) return nullptr;
    auto* e = new nbody_hip_ensemble();
    e->count = count;
    e->dt = dt;
    e->damping = damping;
    e->n.assign(N, N + count);
    e->eps2.assign(softeningSquared, softeningSquared + count);
    e->off.resize(count);
    size_t total = 0;
    for (size_t s = 0; s < count; ++s) {
        if (N[s] < 1 || N[s] > (size_t(1) << 24)) {
            set_error("system %zu has N = %zu (1 .. 2^24 per system)", s, N[s]);
            delete e;
            return nullptr;
        }
        e->off[s] = total;
        total += N[s];
    }
    e->total = total;
    if (total > (size_t(1) << 28)) {
        set_error("ensemble of %zu bodies exceeds 2^28", total);
        delete e;
        return nullptr;
    }
    // (system, 256-body i block) tasks, largest systems first so the tail of
    // the persistent launch is made of the cheapest tasks
    std::vector<BatchTask> tasks;
    const int per = batch_task_bodies(kBatchIpl);
    for (size_t s = 0; s < count; ++s)
        for (size_t b = 0; b < N[s]; b += per) {
            BatchTask t{};
            t.i0 = (int)(e->off[s] + b);
            t.ni = (int)std::min<size_t>(per, N[s] - b);
            t.j0 = (int)e->off[s];
            t.nj = (int)N[s];
            t.eps2 = softeningSquared[s];
            tasks.push_back(t);
        }
    std::stable_sort(tasks.begin(), tasks.end(), [](const BatchTask& a, const BatchTask& b) { return a.nj > b.nj; });
    e->ntasks = (int)tasks.size();
    e->blocks = d->cus * kBatchBlocksPerCU;
    size_t scratch = 0;
    e->plans.resize(count);
    for (size_t s = 0; s < count; ++s) {
        if (make_plan(*d, N[s], N[s], nullptr, &e->plans[s])) {
            delete e;
            return nullptr;
        }
        scratch = std::max(scratch, plan_scratch_bytes(e->plans[s], (int)N[s]));
    }
    hipError_t err = hipStreamCreate(&e->stream);
    if (err == hipSuccess) err = dev_alloc(&e->pm, total);
    if (err == hipSuccess) err = dev_alloc(&e->vm, total);
    if (err == hipSuccess) err = dev_alloc(&e->force, total);
    if (err == hipSuccess) err = dev_alloc(&e->tasks, tasks.size());
    if (err == hipSuccess) err = dev_alloc(&e->counter, 1);
    if (err == hipSuccess) err = e->scratch.reserve(scratch);
    if (err == hipSuccess) err = hipMemcpy(e->pm, posMass, total * sizeof(float4), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(e->vm, velInvMass, total * sizeof(float4), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemset(e->force, 0, total * sizeof(float4));
    if (err == hipSuccess)
        err = hipMemcpy(e->tasks, tasks.data(), tasks.size() * sizeof(BatchTask), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemset(e->counter, 0, sizeof(unsigned));
    if (err == hipSuccess) err = e->timing.init();
    if (err != hipSuccess) {
        set_error("ensemble create failed: %s", hipGetErrorString(err));
        e->release();
        delete e;
        return nullptr;
    }
    return e;
}

int nbody_hip_ensemble_run(nbody_hip_ensemble* e, int steps, int mode, float* ms)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!e || steps < 0 || mode < NBODY_HIP_BATCHED_GRAPH || mode > NBODY_HIP_PER_SYSTEM) {
        set_error("bad ensemble, step count or mode");
        return E_ARG;
    }
    DeviceState* d;
    if (int r = device_state(&d)) return r;
    if (mode == NBODY_HIP_BATCHED_GRAPH && !e->exec) {
        HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        const int r = e->queue_batched();
        hipGraph_t g = nullptr;
        const hipError_t err = hipStreamEndCapture(e->stream, &g);
        if (r) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        HIP_TRY(err);
        e->graph = g;
        HIP_TRY(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    }
    HIP_TRY(hipEventRecord(e->timing.a, e->stream));
    for (int k = 0; k < steps; ++k) {
        int r = OK;
        if (mode == NBODY_HIP_BATCHED_GRAPH) HIP_TRY(hipGraphLaunch(e->exec, e->stream));
        else if (mode == NBODY_HIP_BATCHED) r = e->queue_batched();
        else r = e->queue_per_system(*d);
        if (r) return r;
    }
    return timed_wait(e->stream, e->timing, ms);
}

int nbody_hip_ensemble_read(const nbody_hip_ensemble* e, float* posMass, float* velInvMass, float* force)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!e) {
        set_error("null ensemble");
        return E_ARG;
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    const size_t bytes = e->total * sizeof(float4);
    if (posMass) HIP_TRY(hipMemcpy(posMass, e->pm, bytes, hipMemcpyDeviceToHost));
    if (velInvMass) HIP_TRY(hipMemcpy(velInvMass, e->vm, bytes, hipMemcpyDeviceToHost));
    if (force) HIP_TRY(hipMemcpy(force, e->force, bytes, hipMemcpyDeviceToHost));
    return OK;
}

void nbody_hip_ensemble_destroy(nbody_hip_ensemble* e)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!e) return;
    (void)hipStreamSynchronize(e->stream);
    e->release();
    delete e;
}

}  // extern "C"
