// Device-resident stepping (include/nbody_engine.h): one system with a
// captured per-step graph, and ensembles of independent systems run either
// batched (a persistent work-queue force launch over all of them) or one
// system after another; and the request server (include/nbody_serve.h), which
// runs the same batched launch over a changing set of resident requests.
#include "nbody_engine.h"
#include "nbody_ics.h"
#include "nbody_internal.h"
#include "nbody_serve.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <numeric>
#include <thread>
#include <vector>

using namespace nbody;
using namespace nbody::detail;

namespace {

template <class T>
hipError_t dev_alloc(T** p, size_t count)
{
    return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
}

// Copy a task list to a device buffer that grows (to twice the need) when
// short, on `stream`; the host vector must stay unchanged until it is done.
template <class T>
int upload(T** dev, size_t* cap, const std::vector<T>& host, hipStream_t stream)
{
    if (host.size() > *cap) {
        if (*dev) (void)hipFree(*dev);
        *dev = nullptr;
        *cap = 0;
        if (hipError_t e = dev_alloc(dev, 2 * host.size()); e != hipSuccess) {
            set_error("task buffer: %s", hipGetErrorString(e));
            return E_HIP;
        }
        *cap = 2 * host.size();
    }
    if (host.empty()) return OK;
    if (hipError_t e = hipMemcpyAsync(*dev, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, stream);
        e != hipSuccess) {
        set_error("task upload: %s", hipGetErrorString(e));
        return E_HIP;
    }
    return OK;
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
    char* part = nullptr;          // every system's slot planes, back to back
    bool acc64 = false;            // FP64 accumulation (compensated = 2)
    // The systems in kGroups groups of about equal work, each with its own
    // task lists and work-queue counter; a batched step forks, runs each
    // group's force and slot-sum launches on its own stream, and joins, so
    // one group's launches fill another's drain (two groups measured best:
    // four were no faster than one, eight slower).
    static constexpr int kGroups = 2;
    int ngroups = 1;
    ForceTask* ftasks[kGroups] = {};
    BodyTask* btasks[kGroups] = {};
    int nft[kGroups] = {}, nbt[kGroups] = {};
    unsigned* counter = nullptr;   // kGroups counters, kCounterStride words apart
    static constexpr int kCounterStride = 64;
    hipStream_t side[kGroups] = {};    // group g > 0 runs on side[g]
    hipEvent_t fork = nullptr, join[kGroups] = {};
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
        for (void* p : {(void*)pm, (void*)vm, (void*)force, (void*)part, (void*)counter, scratch.ptr})
            if (p) (void)hipFree(p);
        for (int g = 0; g < kGroups; ++g) {
            if (ftasks[g]) (void)hipFree(ftasks[g]);
            if (btasks[g]) (void)hipFree(btasks[g]);
            if (side[g]) (void)hipStreamDestroy(side[g]);
            if (join[g]) (void)hipEventDestroy(join[g]);
        }
        if (stream) (void)hipStreamDestroy(stream);
        if (fork) (void)hipEventDestroy(fork);
        timing.destroy();
    }

    int queue_group(int g, hipStream_t s)
    {
        unsigned* c = counter + g * kCounterStride;
        HIP_TRY(launch_force_batch(ftasks[g], nft[g], c, blocks, s, acc64));
        HIP_TRY(launch_reduce_integrate_batch(btasks[g], nbt[g], dt, damping, c, s, acc64));
        return OK;
    }

    int queue_batched()
    {
        if (ngroups == 1) return queue_group(0, stream);
        HIP_TRY(hipEventRecord(fork, stream));
        for (int g = 1; g < ngroups; ++g) HIP_TRY(hipStreamWaitEvent(side[g], fork, 0));
        if (int r = queue_group(0, stream)) return r;
        for (int g = 1; g < ngroups; ++g) {
            if (int r = queue_group(g, side[g])) return r;
            HIP_TRY(hipEventRecord(join[g], side[g]));
        }
        for (int g = 1; g < ngroups; ++g) HIP_TRY(hipStreamWaitEvent(stream, join[g], 0));
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
    plan->compensated = s->plan.acc64 ? 2 : s->plan.comp ? 1 : 0;
    plan->newton = s->plan.newton ? 1 : 0;
    plan->waves_per_simd = s->plan.occ;
    plan->slot_bytes = s->plan.band_cap;
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
    return nbody_hip_ensemble_create_opt(count, N, posMass, velInvMass, softeningSquared, dt, damping, nullptr);
}

nbody_hip_ensemble* nbody_hip_ensemble_create_opt(size_t count, const size_t* N, const float* posMass,
                                                  const float* velInvMass, const float* softeningSquared, float dt,
                                                  float damping, const nbody_hip_options* opt)
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
    e->acc64 = opt && opt->compensated == 2;
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
    std::vector<size_t> part_off(count);  // bytes
    size_t part_total = 0;
    for (size_t s = 0; s < count; ++s) {
        part_off[s] = part_total;
        part_total += batch_part_bytes(N[s], e->acc64);
    }
    e->blocks = d->cus * batch_force_blocks_per_cu(e->acc64);
    size_t scratch = 0;
    e->plans.resize(count);
    for (size_t s = 0; s < count; ++s) {
        if (make_plan(*d, N[s], N[s], opt, &e->plans[s])) {
            delete e;
            return nullptr;
        }
        scratch = std::max(scratch, plan_scratch_bytes(e->plans[s], (int)N[s]));
    }
    hipError_t err = hipStreamCreate(&e->stream);
    if (err == hipSuccess) err = dev_alloc(&e->pm, total);
    if (err == hipSuccess) err = dev_alloc(&e->vm, total);
    if (err == hipSuccess) err = dev_alloc(&e->force, total);
    if (err == hipSuccess) err = dev_alloc(&e->part, part_total);
    if (err == hipSuccess) err = dev_alloc(&e->counter, (size_t)nbody_hip_ensemble::kGroups * nbody_hip_ensemble::kCounterStride);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
    for (int g = 1; g < nbody_hip_ensemble::kGroups && err == hipSuccess; ++g) {
        err = hipStreamCreate(&e->side[g]);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&e->join[g], hipEventDisableTiming);
    }
    if (err == hipSuccess) err = e->scratch.reserve(scratch);
    if (err == hipSuccess) err = hipMemcpy(e->pm, posMass, total * sizeof(float4), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(e->vm, velInvMass, total * sizeof(float4), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemset(e->force, 0, total * sizeof(float4));
    if (err == hipSuccess) {
        // two groups of about equal work (sum of N^2), the largest systems
        // dealt first to the lighter group
        std::vector<size_t> by_work(count);
        std::iota(by_work.begin(), by_work.end(), 0);
        std::stable_sort(by_work.begin(), by_work.end(), [&](size_t a, size_t b) { return N[a] > N[b]; });
        std::vector<int> group(count, 0);
        double load[nbody_hip_ensemble::kGroups] = {};
        for (size_t s : by_work) {
            int g = 0;
            for (int k = 1; k < nbody_hip_ensemble::kGroups; ++k)
                if (load[k] < load[g]) g = k;
            group[s] = g;
            load[g] += (double)N[s] * (double)N[s];
        }
        e->ngroups = (int)std::min<size_t>(count, nbody_hip_ensemble::kGroups);
        for (int g = 0; g < e->ngroups && err == hipSuccess; ++g) {
            std::vector<ForceTask> ft;
            std::vector<BodyTask> bt;
            for (size_t s = 0; s < count; ++s) {
                if (group[s] != g) continue;
                append_force_tasks(ft, e->pm + e->off[s], e->part + part_off[s], N[s], softeningSquared[s],
                                   e->acc64);
                append_body_tasks(bt, e->pm + e->off[s], e->vm + e->off[s], e->force + e->off[s],
                                  e->part + part_off[s], N[s], e->acc64);
            }
            order_force_tasks(ft);
            e->nft[g] = (int)ft.size();
            e->nbt[g] = (int)bt.size();
            err = dev_alloc(&e->ftasks[g], ft.size());
            if (err == hipSuccess) err = dev_alloc(&e->btasks[g], bt.size());
            if (err == hipSuccess)
                err = hipMemcpy(e->ftasks[g], ft.data(), ft.size() * sizeof(ForceTask), hipMemcpyHostToDevice);
            if (err == hipSuccess)
                err = hipMemcpy(e->btasks[g], bt.data(), bt.size() * sizeof(BodyTask), hipMemcpyHostToDevice);
        }
    }
    if (err == hipSuccess)
        err = hipMemset(e->counter, 0,
                        (size_t)nbody_hip_ensemble::kGroups * nbody_hip_ensemble::kCounterStride * sizeof(unsigned));
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
    for (int g = 1; g < nbody_hip_ensemble::kGroups; ++g)
        if (e->side[g]) (void)hipStreamSynchronize(e->side[g]);
    e->release();
    delete e;
}

}  // extern "C"

// ---- request server ----------------------------------------------------------

namespace {

struct ServeReq {
    nbody_hip_request r{};
    double admit = -1.0, finish = -1.0;
    int done = 0;
    int status = 0;
    float4* buf = nullptr;  // [posMass | velInvMass | force], 3 N records, then the slot planes
    bool reserved = false;  // its bytes count against the budget
    int ics = 0;            // host initial conditions: 0 not built, 1 built, 2 failed
    int group = 0;          // which of the server's two launch groups steps it
    bool acc64 = false;     // FP64 accumulation (the server's mode)
    std::vector<float> hpm, hvm;  // initial conditions built on the host, until uploaded
    ForceTaskBands tasks;         // its force tasks, built once it is on the device
    std::vector<BodyTask> body_tasks;
    size_t bytes() const { return 3 * r.N * sizeof(float4) + batch_part_bytes(r.N, acc64); }
    void* part() const { return buf + 3 * r.N; }
};

// Seeded initial conditions of a request (nbody_ics.h).
int request_ics(const nbody_hip_request& r, std::vector<float>& pm, std::vector<float>& vm)
{
    pm.assign(4 * r.N, 0.f);
    vm.assign(4 * r.N, 0.f);
    if (r.ic == NBODY_IC_PLUMMER) return nbody_plummer(pm.data(), vm.data(), r.N, r.seed, 1.0f, 1.0f, 1);
    const float radius = (float)(1.54 * std::pow((double)r.N / 16384.0, 1.0 / 3.0));
    return nbody_uniform_ball(pm.data(), vm.data(), r.N, r.seed, radius, 1.0f, 1.0f);
}

}  // namespace

struct nbody_hip_server {
    size_t budget = 0;
    int max_active = 1;
    float dt = 0.f, damping = 1.f;
    bool keep = false;
    bool acc64 = false;            // FP64 accumulation (compensated = 2)
    hipStream_t stream = nullptr;  // the steps (group 0)
    hipStream_t side = nullptr;    // group 1's launches
    hipStream_t up = nullptr;      // uploads during a step (non-blocking: it does not wait for the step)
    hipEvent_t fork = nullptr, join = nullptr;
    static constexpr int kCounterStride = 64;
    unsigned* counter = nullptr;   // one per group, kCounterStride words apart
    // The resident requests in two groups of about equal work, each with its
    // own task lists and counter, launched on two streams so one group's
    // launches fill the other's drain. Two task lists per group: a step
    // reads one while the next step's is uploaded into the other.
    ForceTask* ftasks[2][2] = {};
    BodyTask* btasks[2][2] = {};
    size_t ft_cap[2][2] = {}, bt_cap[2][2] = {};
    int blocks = 0;
    int peak_active = 0;
    std::map<uint64_t, std::pair<std::vector<float>, std::vector<float>>> states;

    void release()
    {
        for (int k = 0; k < 2; ++k)
            for (int g = 0; g < 2; ++g) {
                if (ftasks[k][g]) (void)hipFree(ftasks[k][g]);
                if (btasks[k][g]) (void)hipFree(btasks[k][g]);
            }
        if (counter) (void)hipFree(counter);
        if (stream) (void)hipStreamDestroy(stream);
        if (side) (void)hipStreamDestroy(side);
        if (up) (void)hipStreamDestroy(up);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
    }
};

extern "C" {

nbody_hip_server* nbody_hip_server_create(size_t budget_bytes, int max_active, float dt, float damping,
                                          int keep_states)
{
    return nbody_hip_server_create_opt(budget_bytes, max_active, dt, damping, keep_states, nullptr);
}

nbody_hip_server* nbody_hip_server_create_opt(size_t budget_bytes, int max_active, float dt, float damping,
                                              int keep_states, const nbody_hip_options* opt)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (budget_bytes == 0 || max_active < 1) {
        set_error("server needs a nonzero budget and at least one slot");
        return nullptr;
    }
    DeviceState* d;
    if (device_state(&d)) return nullptr;
    auto* s = new nbody_hip_server();
    s->budget = budget_bytes;
    s->max_active = max_active;
    s->dt = dt;
    s->damping = damping;
    s->keep = keep_states != 0;
    s->acc64 = opt && opt->compensated == 2;
    s->blocks = d->cus * batch_force_blocks_per_cu(s->acc64);
    hipError_t e = hipStreamCreate(&s->stream);
    if (e == hipSuccess) e = hipStreamCreate(&s->side);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->join, hipEventDisableTiming);
    if (e == hipSuccess) e = dev_alloc(&s->counter, 2 * nbody_hip_server::kCounterStride);
    if (e == hipSuccess) e = hipMemset(s->counter, 0, 2 * nbody_hip_server::kCounterStride * sizeof(unsigned));
    if (e != hipSuccess) {
        set_error("server create failed: %s", hipGetErrorString(e));
        s->release();
        delete s;
        return nullptr;
    }
    return s;
}

int nbody_hip_server_run_trace(nbody_hip_server* s, const nbody_hip_request* reqs, size_t count,
                               nbody_hip_result* results, double* wall_s, double* interactions)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!s || (count && (!reqs || !results))) {
        set_error("null server or buffers");
        return E_ARG;
    }
    for (size_t k = 0; k < count; ++k)
        if (reqs[k].N < 1 || reqs[k].N > (size_t(1) << 24) || reqs[k].steps < 0 || !(reqs[k].arrival_s >= 0.0)) {
            set_error("request %zu: N = %zu, steps = %d, arrival %g", k, reqs[k].N, reqs[k].steps, reqs[k].arrival_s);
            return E_ARG;
        }
    std::vector<ServeReq> rq(count);
    std::vector<size_t> order(count);
    for (size_t k = 0; k < count; ++k) {
        rq[k].r = reqs[k];
        rq[k].acc64 = s->acc64;
        order[k] = k;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return rq[a].r.arrival_s < rq[b].r.arrival_s; });
    s->states.clear();
    s->peak_active = 0;
    const auto t0 = std::chrono::steady_clock::now();
    auto now = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    std::deque<ServeReq*> waiting;
    std::vector<ServeReq*> staged;   // admitted, initial conditions on the host, not yet on the device
    std::vector<ServeReq*> joining;  // on the device (uploads queued on s->up), in the next step's list
    std::vector<ServeReq*> active;   // in the current task list
    std::vector<ForceTask> host_ft[2];
    std::vector<BodyTask> host_bt[2];
    int cur = 0;                     // the task lists that describe `active`
    int nft[2][2] = {}, nbt[2][2] = {};  // [list][group]
    double load[2] = {0.0, 0.0};     // resident work (sum of N^2) per group
    size_t next = 0, finished = 0, used = 0, freed = 0;
    double work = 0.0;
    int rc = OK;
    auto retire = [&](ServeReq* q, double t) {
        q->finish = t;
        if (q->buf) {
            load[q->group] -= (double)q->r.N * (double)q->r.N;
            (void)hipFree(q->buf);
            q->buf = nullptr;
        }
        if (q->reserved) {
            used -= q->bytes();
            q->reserved = false;
        }
        ++finished;
    };
    auto arrivals = [&] {
        const double t = now();
        while (next < count && rq[order[next]].r.arrival_s <= t) waiting.push_back(&rq[order[next++]]);
    };
    // A request's initial conditions, built on the host once.
    auto build_ics = [&](ServeReq* q) {
        if (q->ics == 0) q->ics = request_ics(q->r, q->hpm, q->hvm) ? 2 : 1;
        return q->ics == 1;
    };
    // Admit in arrival order while a slot is free and the budget holds: the
    // slot and the device bytes are reserved and the initial conditions are
    // built on the host unless they already were. A request's slot is free
    // once its last step is launched (`freed` counts those of the running
    // step); its bytes count until it retires after that step.
    auto admit = [&] {
        while (!waiting.empty() &&
               active.size() - freed + joining.size() + staged.size() < (size_t)s->max_active) {
            ServeReq* q = waiting.front();
            if (q->bytes() > s->budget) {
                q->status = 1;
                waiting.pop_front();
                retire(q, now());
                continue;
            }
            if (used + q->bytes() > s->budget) break;
            waiting.pop_front();
            if (!build_ics(q)) {
                q->status = 2;
                retire(q, now());
                continue;
            }
            q->admit = now();
            if (q->r.steps == 0) {
                retire(q, q->admit);
                continue;
            }
            used += q->bytes();
            q->reserved = true;
            staged.push_back(q);
        }
    };
    // Staged requests onto the device: allocate, queue the uploads of their
    // initial conditions on s->up, build their task bands. They join the
    // next task list.
    auto stage = [&] {
        for (ServeReq* q : staged) {
            hipError_t e = hipMalloc(reinterpret_cast<void**>(&q->buf), q->bytes());
            if (e != hipSuccess) q->buf = nullptr;
            if (e == hipSuccess)
                e = hipMemcpyAsync(q->buf, q->hpm.data(), q->r.N * sizeof(float4), hipMemcpyHostToDevice, s->up);
            if (e == hipSuccess)
                e = hipMemcpyAsync(q->buf + q->r.N, q->hvm.data(), q->r.N * sizeof(float4), hipMemcpyHostToDevice,
                                   s->up);
            joining.push_back(q);  // from here its buffer is released with the others on any error
            if (e != hipSuccess) {
                set_error("admitting request %llu failed: %s", (unsigned long long)q->r.id, hipGetErrorString(e));
                return E_HIP;
            }
            build_force_task_bands(q->tasks, q->buf, q->part(), q->r.N, q->r.softeningSquared, q->acc64);
            q->group = load[1] < load[0] ? 1 : 0;
            load[q->group] += (double)q->r.N * (double)q->r.N;
            q->body_tasks.clear();
            append_body_tasks(q->body_tasks, q->buf, q->buf + q->r.N, q->buf + 2 * q->r.N, q->part(), q->r.N,
                              q->acc64);
        }
        staged.clear();
        return OK;
    };
    // The task lists of `set`, per group, queued on s->up into list k.
    auto build_list = [&](const std::vector<ServeReq*>& set, int k) -> int {
        for (int g = 0; g < 2; ++g) {
            std::vector<const ForceTaskBands*> bands;
            bands.reserve(set.size());
            host_bt[g].clear();
            for (ServeReq* q : set) {
                if (q->group != g) continue;
                bands.push_back(&q->tasks);
                host_bt[g].insert(host_bt[g].end(), q->body_tasks.begin(), q->body_tasks.end());
            }
            concat_force_task_bands(host_ft[g], bands);
            nft[k][g] = (int)host_ft[g].size();
            nbt[k][g] = (int)host_bt[g].size();
            if (int r = upload(&s->ftasks[k][g], &s->ft_cap[k][g], host_ft[g], s->up)) return r;
            if (int r = upload(&s->btasks[k][g], &s->bt_cap[k][g], host_bt[g], s->up)) return r;
        }
        return OK;
    };
    // One step of list k: each group's force and slot-sum launches, group 1
    // on s->side between a fork and a join when both groups hold requests.
    auto launch_step = [&](int k) -> hipError_t {
        auto group = [&](int g, hipStream_t st) {
            unsigned* c = s->counter + g * nbody_hip_server::kCounterStride;
            hipError_t e = launch_force_batch(s->ftasks[k][g], nft[k][g], c, s->blocks, st, s->acc64);
            if (e == hipSuccess)
                e = launch_reduce_integrate_batch(s->btasks[k][g], nbt[k][g], s->dt, s->damping, c, st, s->acc64);
            return e;
        };
        if (nft[k][0] == 0 || nft[k][1] == 0) {
            hipError_t e = group(0, s->stream);
            return e == hipSuccess ? group(1, s->stream) : e;
        }
        hipError_t e = hipEventRecord(s->fork, s->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(s->side, s->fork, 0);
        if (e == hipSuccess) e = group(0, s->stream);
        if (e == hipSuccess) e = group(1, s->side);
        if (e == hipSuccess) e = hipEventRecord(s->join, s->side);
        if (e == hipSuccess) e = hipStreamWaitEvent(s->stream, s->join, 0);
        return e;
    };
    // The uploads are done: the host copies of the initial conditions can go.
    auto uploaded = [&]() -> int {
        HIP_TRY(hipStreamSynchronize(s->up));
        for (ServeReq* q : joining) {
            std::vector<float>().swap(q->hpm);
            std::vector<float>().swap(q->hvm);
        }
        return OK;
    };
    while (finished < count && rc == OK) {
        arrivals();
        if (active.empty()) {
            // nothing on the device: admit and upload before a step (with
            // requests resident, admission happens while their step runs)
            admit();
            if (staged.empty()) {
                if (next < count) {
                    const double wait = rq[order[next]].r.arrival_s - now();
                    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
                }
                continue;
            }
            if ((rc = stage()) || (rc = build_list(joining, cur)) || (rc = uploaded())) break;
            active.swap(joining);
            joining.clear();
        }
        s->peak_active = std::max(s->peak_active, (int)active.size());
        // One step of every active request from list `cur`. While it runs
        // the host takes arrivals and admissions, builds the initial
        // conditions of the waiting requests next in line for the slots this
        // step frees (slots free only after the step, and a 131072-body
        // Plummer sphere takes ~45 ms on the host), uploads the admitted
        // requests, and builds and uploads the next step's task list into
        // the other list. Only retiring the finished requests waits for the
        // step.
        hipError_t e = launch_step(cur);
        if (e != hipSuccess) {
            set_error("server step: %s", hipGetErrorString(e));
            rc = E_HIP;
            break;
        }
        std::vector<ServeReq*> next_active;
        next_active.reserve(active.size() + s->max_active);
        for (ServeReq* q : active)
            if (q->done + 1 < q->r.steps) next_active.push_back(q);
        freed = active.size() - next_active.size();
        arrivals();
        admit();
        for (size_t k = 0; k < waiting.size() && k < freed; ++k) (void)build_ics(waiting[k]);
        bool flip = false;
        if ((rc = stage())) break;
        if (freed || !joining.empty()) {
            next_active.insert(next_active.end(), joining.begin(), joining.end());
            if (!next_active.empty() && (rc = build_list(next_active, 1 - cur))) break;
            flip = true;
        }
        e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) {
            set_error("server step: %s", hipGetErrorString(e));
            rc = E_HIP;
            break;
        }
        if ((rc = uploaded())) break;
        freed = 0;
        const double t = now();
        for (ServeReq* q : active) {
            work += (double)q->r.N * (double)q->r.N;
            if (++q->done < q->r.steps) continue;
            if (s->keep) {
                auto& st = s->states[q->r.id];
                st.first.resize(4 * q->r.N);
                st.second.resize(4 * q->r.N);
                e = hipMemcpy(st.first.data(), q->buf, q->r.N * sizeof(float4), hipMemcpyDeviceToHost);
                if (e == hipSuccess)
                    e = hipMemcpy(st.second.data(), q->buf + q->r.N, q->r.N * sizeof(float4), hipMemcpyDeviceToHost);
                if (e != hipSuccess) {
                    set_error("reading request %llu: %s", (unsigned long long)q->r.id, hipGetErrorString(e));
                    rc = E_HIP;
                }
            }
            std::vector<ForceTask>().swap(q->tasks.big);
            std::vector<ForceTask>().swap(q->tasks.pairs);
            std::vector<ForceTask>().swap(q->tasks.diag);
            std::vector<BodyTask>().swap(q->body_tasks);
            retire(q, t);
        }
        active.swap(next_active);
        joining.clear();
        if (flip) cur = 1 - cur;
    }
    // Release what is still on the device once nothing runs on it. After a
    // failure a request can sit in more than one list, so each buffer is
    // freed once and forgotten.
    (void)hipStreamSynchronize(s->stream);
    (void)hipStreamSynchronize(s->side);
    (void)hipStreamSynchronize(s->up);
    auto drop = [](ServeReq* q) {
        if (q->buf) (void)hipFree(q->buf);
        q->buf = nullptr;
    };
    for (ServeReq* q : staged) drop(q);
    for (ServeReq* q : joining) drop(q);
    for (ServeReq* q : active) drop(q);
    for (size_t k = 0; k < count; ++k) {
        const ServeReq& q = rq[k];
        results[k].id = q.r.id;
        results[k].N = q.r.N;
        results[k].steps = q.r.steps;
        results[k].status = q.finish < 0 ? 2 : q.status;
        results[k].arrival_s = q.r.arrival_s;
        results[k].admit_s = q.admit;
        results[k].finish_s = q.finish;
    }
    if (wall_s) *wall_s = now();
    if (interactions) *interactions = work;
    return rc;
}

int nbody_hip_server_final_state(const nbody_hip_server* s, uint64_t id, float* posMass, float* velInvMass)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!s) {
        set_error("null server");
        return E_ARG;
    }
    auto it = s->states.find(id);
    if (it == s->states.end()) {
        set_error("no kept state for request %llu", (unsigned long long)id);
        return E_ARG;
    }
    if (posMass) std::copy(it->second.first.begin(), it->second.first.end(), posMass);
    if (velInvMass) std::copy(it->second.second.begin(), it->second.second.end(), velInvMass);
    return OK;
}

int nbody_hip_server_peak_active(const nbody_hip_server* s) { return s ? s->peak_active : 0; }

size_t nbody_hip_request_bytes(size_t N) { return nbody_hip_request_bytes_opt(N, nullptr); }

size_t nbody_hip_request_bytes_opt(size_t N, const nbody_hip_options* opt)
{
    ServeReq q;
    q.r.N = N;
    q.acc64 = opt && opt->compensated == 2;
    return q.bytes();
}

void nbody_hip_server_destroy(nbody_hip_server* s)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!s) return;
    (void)hipStreamSynchronize(s->stream);
    (void)hipStreamSynchronize(s->side);
    s->release();
    delete s;
}

}  // extern "C"
