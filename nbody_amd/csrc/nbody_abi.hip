// extern "C" entry points of libnbody_hip.so (declared in include/nbody_hip.h).
//
// Error convention of the reference (src/chError.h:86-91 + src/nbody.cu:385-386):
// a failing table entry returns 0.0f and the reference driver skips it. Here the reason is
// also kept per thread for nbody_hip_last_error().
#include "nbody_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

using namespace nbody;

namespace nbody {
namespace detail {

thread_local std::string g_last_error;
std::mutex g_mutex;
std::vector<DeviceState> g_devices;


void set_error(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

int device_state(DeviceState** out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if ((int)g_devices.size() <= dev) g_devices.resize(dev + 1);
    DeviceState& d = g_devices[dev];
    if (!d.init) {
        hipDeviceProp_t prop;
        HIP_TRY(hipGetDeviceProperties(&prop, dev));
        d.cus = prop.multiProcessorCount;
        d.clock_khz = prop.clockRate;
        snprintf(d.arch, sizeof d.arch, "%s", prop.gcnArchName);
        HIP_TRY(hipEventCreate(&d.ev0));
        HIP_TRY(hipEventCreate(&d.ev1));
        HIP_TRY(hipEventCreateWithFlags(&d.partial_done, hipEventDisableTiming));
        d.init = true;
    }
    *out = &d;
    return OK;
}

int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Default symmetric plans aim for at most this many blocks (slot planes of
// 12 N * blocks bytes; past kSymBandCap they are kept one column band at a
// time).
constexpr int kSymSlotsTarget = 512;
// Below N = 512 * kSmallSymBlocks the default symmetric plan takes smaller
// blocks (256, then 128 bodies below N = 256 * kSmallSymBlocks), keeping
// about this many: at N = 8192, 16384 and 32768 the autotuner's choices
// (blocks of 128, 256, 512) all have 64.
constexpr int kSmallSymBlocks = 64;

// Default decomposition: enough blocks for ~48 waves per SIMD (the dispatcher
// balances them across the 256 CUs), j slices of at least 1024 bodies in
// multiples of 256.
ForcePlan default_plan(int variant, size_t N, size_t n, int cus, const nbody_hip_options* opt)
{
    ForcePlan p;
    p.variant = variant;
    p.newton = opt && opt->newton > 0;
    const int comp_opt = opt ? opt->compensated : -1;
    p.comp = comp_opt < 0 ? (N > 65536) : comp_opt > 0;
    p.acc64 = comp_opt == 2;
    if (variant == VAR_SYMMETRIC && p.acc64) {
        // one wave per block pair, 4 (or 8) bodies per lane, blocks of 64 per lane
        p.ipl = (opt && opt->i_per_lane > 0) ? opt->i_per_lane : 4;
        p.jslice = (opt && opt->jslice > 0) ? opt->jslice : 64 * p.ipl;
        p.nslices = (int)((N + p.jslice - 1) / p.jslice);
        p.band_cap = opt ? opt->slot_bytes : 0;
        return p;
    }
    if (variant == VAR_SYMMETRIC) {
        // blocks of B = 64 * ipl * waves (default 8 bodies per lane and the
        // fewest waves that keep at most kSymSlotsTarget blocks: B = 512 from
        // N = 32K up to 256K, 2048 at N = 1M; smaller systems below); the block
        // slots are always summed with TwoSum
        p.ipl = (opt && opt->i_per_lane > 0) ? opt->i_per_lane : 8;
        const bool shape_given = opt && (opt->i_per_lane > 0 || opt->jslice > 0);
        if (!shape_given && N < (size_t)kSmallSymBlocks * 512) {
            // small systems: keep about kSmallSymBlocks blocks (the shapes the
            // autotuner picks there), so the pair tasks fill the chip
            p.ipl = N < (size_t)kSmallSymBlocks * 256 ? 2 : 4;
            p.jslice = 64 * p.ipl;
        } else if (opt && opt->jslice > 0) {
            p.jslice = opt->jslice;
        } else if (p.ipl == 2) {
            p.jslice = 512;
        } else {
            p.jslice = 64 * p.ipl;
            while (p.jslice < 64 * p.ipl * 4 && (N + p.jslice - 1) / p.jslice > (size_t)kSymSlotsTarget) p.jslice *= 2;
        }
        p.nslices = (int)((N + p.jslice - 1) / p.jslice);
        p.comp = true;
        p.occ = opt ? opt->waves_per_simd : 0;
        p.band_cap = opt ? opt->slot_bytes : 0;
        return p;
    }
    if (variant == VAR_AOS) {
        p.ipl = 1;
        p.jslice = (int)N;
        p.nslices = 1;
        p.comp = false;
        return p;
    }
    p.ipl = (opt && opt->i_per_lane > 0) ? opt->i_per_lane : 1;
    if (variant == VAR_DPP && !(opt && opt->i_per_lane > 0)) {
        // the rotation's 4 moves per step are shared by the lane's i bodies:
        // as many as leave enough waves (measured best: 8 at N = 98304 and
        // 196608, 4 at 32768 and 65536, 2 at 16384, 1 at 8192)
        const double work = (double)n * (double)N;
        p.ipl = work >= 6.0e9 ? 8 : work >= 1.0e9 ? 4 : work >= 2.0e8 ? 2 : 1;
    }
    if (opt && opt->jslice > 0) {
        p.jslice = opt->jslice;
    } else {
        const long blocks_i = (long)((n + 256 * p.ipl - 1) / (256 * p.ipl));
        const long target = (long)std::max(cus, 1) * 48;
        long ns = std::max(1L, (target + blocks_i - 1) / blocks_i);
        const long max_ns = std::max(1L, (long)(N / 1024));
        ns = std::min(ns, max_ns);
        p.jslice = round_up((int)((N + ns - 1) / ns), 256);
    }
    p.jslice = std::max(1, std::min(p.jslice, (int)N));
    p.nslices = (int)((N + p.jslice - 1) / p.jslice);
    return p;
}

// The untuned choice for NBODY_HIP_AUTO: the symmetric kernel for whole
// systems of at least two of its blocks (half the pairs, 14-16 VALU + 1 rsq
// per pair); otherwise the LDS broadcast, whose VALU operands all sit in VGPRs
// (gfx950 issues those faster than ops reading an SGPR).
// Largest block count of a symmetric plan (the pair index must fit an int).
constexpr int kMaxSymBlocks = 16384;

int auto_variant(size_t N, size_t n)
{
    if (n != N || N < 1024) return VAR_LDS;
    const ForcePlan p = default_plan(VAR_SYMMETRIC, N, N, 1, nullptr);
    return p.nslices <= kMaxSymBlocks ? VAR_SYMMETRIC : VAR_LDS;
}

std::tuple<int, size_t, size_t, int, int> tune_key(const nbody_hip_options* opt, size_t N, size_t n)
{
    const int v = opt ? opt->variant : NBODY_HIP_AUTO;
    const int comp = opt ? opt->compensated : -1;
    const int newton = opt && opt->newton > 0;
    return std::make_tuple(v, N, n, comp, newton);
}

int validate_plan(const ForcePlan& p, size_t N)
{
    if (p.acc64) {
        if (p.variant != VAR_LDS && p.variant != VAR_SYMMETRIC) {
            set_error("FP64 accumulation (compensated = 2) is built for the LDS and symmetric variants");
            return E_ARG;
        }
        if (p.variant == VAR_SYMMETRIC && ((p.ipl != 4 && p.ipl != 8) || p.jslice != 64 * p.ipl)) {
            set_error("the FP64-accumulation symmetric plan needs i_per_lane 4 or 8 with blocks of 64*i_per_lane "
                      "(got %d, %d)", p.ipl, p.jslice);
            return E_ARG;
        }
    }
    if (p.variant == VAR_SYMMETRIC) {
        if (!symmetric_shape_ok(p.ipl, p.jslice)) {
            set_error("symmetric plan needs i_per_lane 2 with blocks of 128 or 512 bodies, or 4 or 8 with "
                      "blocks of 64*i_per_lane*{1,2,4} (got %d, %d)", p.ipl, p.jslice);
            return E_ARG;
        }
        if (p.occ < 0 || p.occ > 8) {
            set_error("waves_per_simd must be 0..8 (got %d)", p.occ);
            return E_ARG;
        }
        if (p.nslices > kMaxSymBlocks) {
            set_error("N = %zu needs %d symmetric blocks of %d (at most %d)", N, p.nslices, p.jslice, kMaxSymBlocks);
            return E_UNSUPPORTED;
        }
        return OK;
    }
    if (p.ipl != 1 && p.ipl != 2 && p.ipl != 4 && !(p.ipl == 8 && p.variant == VAR_DPP)) {
        set_error("i_per_lane must be 1, 2 or 4 (8 also for the DPP variant; got %d)", p.ipl);
        return E_ARG;
    }
    return OK;
}

int make_plan(DeviceState& d, size_t N, size_t n, const nbody_hip_options* opt, ForcePlan* plan)
{
    const int requested = opt ? opt->variant : NBODY_HIP_AUTO;
    const int variant = requested == NBODY_HIP_AUTO ? auto_variant(N, n) : requested;
    if (variant < 0 || variant >= VAR_COUNT) {
        set_error("unknown variant %d", requested);
        return E_ARG;
    }
    ForcePlan p = default_plan(variant, N, n, d.cus, opt);
    const bool explicit_shape = opt && (opt->i_per_lane > 0 || opt->jslice > 0 || opt->waves_per_simd > 0);
    if (!explicit_shape) {
        auto it = d.tuned.find(tune_key(opt, N, n));
        if (it != d.tuned.end()) p = it->second;
        if (p.variant == VAR_SYMMETRIC && opt) p.band_cap = opt->slot_bytes;
    }
    if (int r = validate_plan(p, N)) return r;
    *plan = p;
    return OK;
}

int check_args(const void* force, const void* pm, size_t N, size_t base, size_t n)
{
    if (!force || !pm) {
        set_error("null buffer");
        return E_ARG;
    }
    if (N < 1 || N > kMaxBodies) {
        set_error("N = %zu outside [1, %zu]", N, kMaxBodies);
        return E_ARG;
    }
    if (base > N || n > N - base) {
        set_error("i range [%zu, %zu) outside [0, %zu)", base, base + n, N);
        return E_ARG;
    }
    return OK;
}

// Queue one force evaluation on `stream` (caller holds g_mutex).
int run_force(DeviceState& d, const ForcePlan& plan, float* force, const float* pm, float eps2, size_t N, size_t base,
              size_t n, hipStream_t stream, Buffer* own, bool timed, double* force64)
{
    Buffer& buf = own ? *own : d.partial;
    if (plan.variant == VAR_SYMMETRIC && (base != 0 || n != N)) {
        set_error("the symmetric variant computes whole systems only (base = 0, n = N)");
        return E_UNSUPPORTED;
    }
    ForceParams p;
    p.pm = reinterpret_cast<const float4*>(pm);
    p.n = (int)N;
    p.i0 = (int)base;
    p.ni = (int)n;
    p.eps2 = eps2;
    p.jslice = plan.jslice;
    p.nslices = plan.nslices;
    p.force = force;
    p.force64 = force64;
    const size_t bytes = plan_scratch_bytes(plan, (int)n);
    const bool shared = bytes && !own;
    if (shared) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIP_TRY(hipStreamIsCapturing(stream, &cs));
        if (cs != hipStreamCaptureStatusNone) {
            set_error("this plan needs the library's shared scratch, which cannot be captured into a graph "
                      "(a later call may reallocate it)");
            return E_UNSUPPORTED;
        }
    }
    if (bytes) {
        HIP_TRY(buf.reserve(bytes));
        if (plan.acc64) {
            p.hi64 = static_cast<double*>(buf.ptr);
        } else {
            p.hi = static_cast<float4*>(buf.ptr);
            p.lo = plan.comp ? p.hi + (size_t)plan.nslices * n : nullptr;
        }
    }
    if (shared && d.partial_pending && d.partial_stream != stream)
        HIP_TRY(hipStreamWaitEvent(stream, d.partial_done, 0));
    hipEvent_t before = nullptr, after = nullptr;
    if (timed && d.timing && d.timer_used < kMaxTimedLaunches) {
        if (d.timer_used == d.timer_events.size()) {
            hipEvent_t a, b;
            HIP_TRY(hipEventCreate(&a));
            HIP_TRY(hipEventCreate(&b));
            d.timer_events.emplace_back(a, b);
        }
        before = d.timer_events[d.timer_used].first;
        after = d.timer_events[d.timer_used].second;
        d.timer_used++;
    }
    HIP_TRY(launch_force(plan, p, stream, before, after));
    if (shared) {
        HIP_TRY(hipEventRecord(d.partial_done, stream));
        d.partial_stream = stream;
        d.partial_pending = true;
    }
    return OK;
}

}  // namespace detail
}  // namespace nbody

using namespace nbody::detail;

// Partial sums of nbody_hip_force_noalloc: one fixed array per GPU, which the
// device holds from the library's load to the end of the process. Host code
// never passes this name to a kernel (on the host it is only a stand-in for
// the symbol); it uses the address hipGetSymbolAddress returns for the
// current device, kept in DeviceState::noalloc_scratch.
static __device__ __align__(256) float4 scratch[NBODY_HIP_NOALLOC_SCRATCH_BYTES / sizeof(float4)];

namespace {

// This device's address of `scratch`, looked up once per device with that
// device current.
int noalloc_scratch(DeviceState& d, void** out)
{
    if (!d.noalloc_scratch) {
        void* ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        d.noalloc_scratch = ptr;
    }
    *out = d.noalloc_scratch;
    return OK;
}

// A plan whose partial sums fit in `cap` bytes. The plan nbody_hip_force
// would use when it fits. Otherwise a symmetric plan keeps its slots one
// narrower column band at a time, then (unless the caller fixed the shape)
// takes larger blocks, and for NBODY_HIP_AUTO the LDS plan after that; a
// one-sided plan with no fixed shape gets fewer, longer j slices, down to a
// single slice, which keeps no partial sums at all.
int fit_plan(const DeviceState& d, size_t N, size_t n, const nbody_hip_options* opt, size_t cap, ForcePlan& p)
{
    const size_t need = plan_scratch_bytes(p, (int)n);
    if (need <= cap) return OK;
    const bool pinned = opt && (opt->i_per_lane > 0 || opt->jslice > 0 || opt->waves_per_simd > 0);
    const int requested = opt ? opt->variant : NBODY_HIP_AUTO;
    if (p.variant == VAR_SYMMETRIC) {
        // narrower column bands (the running sums, 24 bytes per body, come
        // out of the array too), then larger blocks unless the shape is fixed
        ForcePlan q = p;
        for (int block = p.jslice; block <= 64 * p.ipl * 4; block *= 2) {
            if (!symmetric_shape_ok(q.ipl, block) || (pinned && block != p.jslice)) continue;
            q.jslice = block;
            q.nslices = (int)((N + block - 1) / block);
            q.band_cap = 0;
            if (plan_scratch_bytes(q, (int)n) <= cap) {
                p = q;
                return OK;
            }
            if (cap > 24 * n) {
                q.band_cap = cap - 24 * n;
                if (plan_scratch_bytes(q, (int)n) <= cap) {
                    p = q;
                    return OK;
                }
            }
        }
        if (pinned || requested != NBODY_HIP_AUTO) {
            set_error("no symmetric block size keeps the partial sums of N = %zu within the %zu bytes of the "
                      "fixed scratch array", N, cap);
            return E_ARG;
        }
        nbody_hip_options o{VAR_LDS, 0, 0, opt ? opt->compensated : -1, opt ? opt->newton : 0, 0};
        p = default_plan(VAR_LDS, N, n, d.cus, &o);
        if (plan_scratch_bytes(p, (int)n) <= cap) return OK;
    } else if (pinned) {
        set_error("this plan keeps %zu bytes of partial sums, more than the %zu of the fixed scratch array "
                  "(NBODY_HIP_NOALLOC_SCRATCH_BYTES)", need, cap);
        return E_ARG;
    }
    const size_t per_slice = n * sizeof(float4) * (p.comp ? 2 : 1);
    const size_t max_slices = cap / per_slice;
    if (max_slices >= 2) {
        p.jslice = round_up((int)((N + max_slices - 1) / max_slices), 256);
        p.jslice = std::max(1, std::min(p.jslice, (int)N));
        p.nslices = (int)((N + p.jslice - 1) / p.jslice);
    }
    if (plan_scratch_bytes(p, (int)n) > cap) {
        p.jslice = (int)N;
        p.nslices = 1;
    }
    return OK;
}

// Synchronous, event-timed evaluation on the null stream; returns ms or 0.
float timed_force(const nbody_hip_options* opt, float* force, const float* pm, float eps2, size_t N)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (check_args(force, pm, N, 0, N)) return 0.0f;
    DeviceState* d;
    if (device_state(&d)) return 0.0f;
    ForcePlan plan;
    if (make_plan(*d, N, N, opt, &plan)) return 0.0f;
    // scratch is reserved outside the timed region
    const size_t bytes = plan_scratch_bytes(plan, (int)N);
    if (bytes && d->partial.reserve(bytes) != hipSuccess) {
        set_error("cannot allocate %zu bytes of slice scratch", bytes);
        return 0.0f;
    }
    if (hipError_t e = hipEventRecord(d->ev0, nullptr); e != hipSuccess) {
        set_error("timing failed: %s", hipGetErrorString(e));
        return 0.0f;
    }
    if (run_force(*d, plan, force, pm, eps2, N, 0, N, nullptr)) return 0.0f;
    float ms = 0.0f;
    hipError_t e = hipEventRecord(d->ev1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(d->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, d->ev0, d->ev1);
    if (e != hipSuccess) {
        set_error("timing failed: %s", hipGetErrorString(e));
        return 0.0f;
    }
    // a kernel that ran in no time at all would read as "skipped" to the caller
    return std::max(ms, 1e-6f);
}

float variant_entry(int variant, float* force, float* pm, float eps2, size_t N)
{
    nbody_hip_options o{variant, 0, 0, -1, 0};
    return timed_force(&o, force, pm, eps2, N);
}

}  // namespace

extern "C" {

float ComputeGravitation_HIP_AOS(float* force, float* posMass, float softeningSquared, size_t N)
{
    return variant_entry(NBODY_HIP_AOS, force, posMass, softeningSquared, N);
}

float ComputeGravitation_HIP_LDS(float* force, float* posMass, float softeningSquared, size_t N)
{
    return variant_entry(NBODY_HIP_LDS, force, posMass, softeningSquared, N);
}

float ComputeGravitation_HIP_DPP(float* force, float* posMass, float softeningSquared, size_t N)
{
    return variant_entry(NBODY_HIP_DPP, force, posMass, softeningSquared, N);
}

float ComputeGravitation_HIP_SGPR(float* force, float* posMass, float softeningSquared, size_t N)
{
    return variant_entry(NBODY_HIP_SGPR, force, posMass, softeningSquared, N);
}

float ComputeGravitation_HIP_Symmetric(float* force, float* posMass, float softeningSquared, size_t N)
{
    return variant_entry(NBODY_HIP_SYMMETRIC, force, posMass, softeningSquared, N);
}

float ComputeGravitation_HIP(float* force, float* posMass, float softeningSquared, size_t N)
{
    return variant_entry(NBODY_HIP_AUTO, force, posMass, softeningSquared, N);
}

float ComputeGravitation_HIP_Host(float* force, float* posMass, float softeningSquared, size_t N)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (check_args(force, posMass, N, 0, N)) return 0.0f;
    DeviceState* d;
    if (device_state(&d)) return 0.0f;
    ForcePlan plan;
    if (make_plan(*d, N, N, nullptr, &plan)) return 0.0f;
    const size_t bytes = N * 4 * sizeof(float);
    const size_t sbytes = plan_scratch_bytes(plan, (int)N);
    if (d->stage_pm.reserve(bytes) != hipSuccess || d->stage_force.reserve(bytes) != hipSuccess ||
        d->host_stage.reserve(2 * bytes) != hipSuccess || (sbytes && d->partial.reserve(sbytes) != hipSuccess)) {
        set_error("cannot allocate staging buffers for N = %zu", N);
        return 0.0f;
    }
    // The caller's buffers are pageable: posMass goes through page-locked
    // staging so that the transfers run at full link speed; only x, y, z of
    // the forces come back into the caller's buffer (w stays untouched). The
    // time returned is the whole call's, transfers included, like the
    // reference's host-pointer entries.
    const auto t0 = std::chrono::steady_clock::now();
    float* dpm = static_cast<float*>(d->stage_pm.ptr);
    float* dforce = static_cast<float*>(d->stage_force.ptr);
    float* hin = static_cast<float*>(d->host_stage.ptr);
    float* hout = hin + 4 * N;
    std::memcpy(hin, posMass, bytes);
    hipError_t e = hipMemcpyAsync(dpm, hin, bytes, hipMemcpyHostToDevice, nullptr);
    if (e != hipSuccess) {
        set_error("host-to-device copy failed: %s", hipGetErrorString(e));
        return 0.0f;
    }
    if (run_force(*d, plan, dforce, dpm, softeningSquared, N, 0, N, nullptr)) return 0.0f;
    e = hipMemcpyAsync(hout, dforce, bytes, hipMemcpyDeviceToHost, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
        set_error("device-to-host copy failed: %s", hipGetErrorString(e));
        return 0.0f;
    }
    for (size_t i = 0; i < N; ++i) {
        force[4 * i + 0] = hout[4 * i + 0];
        force[4 * i + 1] = hout[4 * i + 1];
        force[4 * i + 2] = hout[4 * i + 2];
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return std::max((float)ms, 1e-6f);
}

int nbody_hip_force(float* force, const float* posMass, float softeningSquared, size_t N, size_t base, size_t n,
                    const nbody_hip_options* opt, void* stream)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (int r = check_args(force, posMass, N, base, n)) return r;
    if (n == 0) return OK;
    DeviceState* d;
    if (int r = device_state(&d)) return r;
    ForcePlan plan;
    if (int r = make_plan(*d, N, n, opt, &plan)) return r;
    return run_force(*d, plan, force, posMass, softeningSquared, N, base, n, static_cast<hipStream_t>(stream));
}

int nbody_hip_force_pair_share(double* force64, const float* posMass, float softeningSquared, size_t N, int share,
                               int shares, const nbody_hip_options* opt, void* stream)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (int r = check_args(force64, posMass, N, 0, N)) return r;
    if (shares < 1 || share < 0 || share >= shares) {
        set_error("share %d of %d: need 0 <= share < shares", share, shares);
        return E_ARG;
    }
    const int requested = opt ? opt->variant : NBODY_HIP_AUTO;
    if (requested != NBODY_HIP_AUTO && requested != NBODY_HIP_SYMMETRIC) {
        set_error("a share of the block pairs needs the symmetric plan (variant %d given)", requested);
        return E_ARG;
    }
    DeviceState* d;
    if (int r = device_state(&d)) return r;
    nbody_hip_options o{};
    if (opt) o = *opt;
    else o.compensated = -1;
    o.variant = NBODY_HIP_SYMMETRIC;
    ForcePlan plan;
    if (int r = make_plan(*d, N, N, &o, &plan)) return r;
    symmetric_share_columns(plan.nslices, share, shares, &plan.col0, &plan.col1);
    return run_force(*d, plan, nullptr, posMass, softeningSquared, N, 0, N, static_cast<hipStream_t>(stream), nullptr,
                     true, force64);
}

int nbody_hip_force_noalloc(float* force, const float* posMass, float softeningSquared, size_t N, size_t base,
                            size_t n, const nbody_hip_options* opt, void* stream)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (int r = check_args(force, posMass, N, base, n)) return r;
    if (n == 0) return OK;
    DeviceState* d;
    if (int r = device_state(&d)) return r;
    void* ptr = nullptr;
    if (int r = noalloc_scratch(*d, &ptr)) return r;
    ForcePlan plan;
    if (int r = make_plan(*d, N, n, opt, &plan)) return r;
    if (int r = fit_plan(*d, N, n, opt, NBODY_HIP_NOALLOC_SCRATCH_BYTES, plan)) return r;
    if (int r = validate_plan(plan, N)) return r;
    // the array as a buffer that is already as large as it can grow: run_force
    // never allocates for it (the plan's partial sums fit, checked above)
    Buffer fixed;
    fixed.ptr = ptr;
    fixed.cap = NBODY_HIP_NOALLOC_SCRATCH_BYTES;
    return run_force(*d, plan, force, posMass, softeningSquared, N, base, n, static_cast<hipStream_t>(stream), &fixed,
                     false);
}

int nbody_hip_integrate(float* posMass, float* velInvMass, const float* force, float dt, float damping, size_t N,
                        void* stream)
{
    if (!posMass || !velInvMass || !force) {
        set_error("null buffer");
        return E_ARG;
    }
    if (N > kMaxBodies) {
        set_error("N = %zu too large", N);
        return E_ARG;
    }
    HIP_TRY(launch_integrate(reinterpret_cast<float4*>(posMass), reinterpret_cast<float4*>(velInvMass),
                             reinterpret_cast<const float4*>(force), dt, damping, (int)N,
                             static_cast<hipStream_t>(stream)));
    return OK;
}

float integrateGravitation_HIP(float* posMass, float* velInvMass, const float* force, float dt, float damping,
                               size_t N)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* d;
    if (device_state(&d)) return 0.0f;
    if (hipError_t e = hipEventRecord(d->ev0, nullptr); e != hipSuccess) {
        set_error("integrate failed: %s", hipGetErrorString(e));
        return 0.0f;
    }
    if (nbody_hip_integrate(posMass, velInvMass, force, dt, damping, N, nullptr)) return 0.0f;
    float ms = 0.0f;
    hipError_t e = hipEventRecord(d->ev1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(d->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, d->ev0, d->ev1);
    if (e != hipSuccess) {
        set_error("integrate failed: %s", hipGetErrorString(e));
        return 0.0f;
    }
    return std::max(ms, 1e-6f);
}

float nbody_hip_force_fp64(double* force64, const float* posMass, double softeningSquared, size_t N)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (check_args(force64, posMass, N, 0, N)) return 0.0f;
    DeviceState* d;
    if (device_state(&d)) return 0.0f;
    // at most 64 slices: the FP64 partials take 24 bytes per body and slice
    const int jslice = (int)std::max<size_t>(2048, (N + 63) / 64);
    const size_t nslices = (N + jslice - 1) / jslice;
    if (d->partial64.reserve(nslices * N * 3 * sizeof(double)) != hipSuccess) {
        set_error("cannot allocate FP64 scratch");
        return 0.0f;
    }
    float ms = 0.0f;
    hipError_t e = hipEventRecord(d->ev0, nullptr);
    if (e == hipSuccess)
        e = launch_force_fp64(force64, reinterpret_cast<const float4*>(posMass), (int)N, 0, (int)N, softeningSquared,
                              jslice, static_cast<double*>(d->partial64.ptr), nullptr);
    if (e == hipSuccess) e = hipEventRecord(d->ev1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(d->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, d->ev0, d->ev1);
    if (e != hipSuccess) {
        set_error("FP64 force failed: %s", hipGetErrorString(e));
        return 0.0f;
    }
    return std::max(ms, 1e-6f);
}

int nbody_hip_energy_fp64(double* out, const float* posMass, const float* velInvMass, double softeningSquared,
                          size_t N)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!out || !velInvMass) {
        set_error("null buffer");
        return E_ARG;
    }
    if (int r = check_args(posMass, posMass, N, 0, N)) return r;
    DeviceState* d;
    if (int r = device_state(&d)) return r;
    const int blocks = fp64_energy_blocks((int)N);
    HIP_TRY(d->partial64.reserve((size_t)blocks * 5 * sizeof(double)));
    HIP_TRY(d->small.reserve(8 * sizeof(double)));
    double* acc = static_cast<double*>(d->small.ptr);
    HIP_TRY(launch_energy_fp64(acc, reinterpret_cast<const float4*>(posMass),
                               reinterpret_cast<const float4*>(velInvMass), (int)N, softeningSquared,
                               static_cast<double*>(d->partial64.ptr), nullptr));
    HIP_TRY(hipMemcpy(out, acc, 5 * sizeof(double), hipMemcpyDeviceToHost));
    return OK;
}

int nbody_hip_plan(size_t N, size_t n, const nbody_hip_options* opt, nbody_hip_options* out)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!out) {
        set_error("null plan");
        return E_ARG;
    }
    if (N < 1 || N > kMaxBodies || n > N) {
        set_error("bad sizes N = %zu, n = %zu", N, n);
        return E_ARG;
    }
    DeviceState* d;
    if (int r = device_state(&d)) return r;
    ForcePlan p;
    if (int r = make_plan(*d, N, n, opt, &p)) return r;
    out->variant = p.variant;
    out->i_per_lane = p.ipl;
    out->jslice = p.jslice;
    out->compensated = p.acc64 ? 2 : p.comp ? 1 : 0;
    out->newton = p.newton ? 1 : 0;
    out->waves_per_simd = p.occ;
    out->slot_bytes = p.band_cap;
    return OK;
}

float nbody_hip_autotune(float* force, const float* posMass, float softeningSquared, size_t N, int variant)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (check_args(force, posMass, N, 0, N)) return 0.0f;
    DeviceState* d;
    if (device_state(&d)) return 0.0f;
    nbody_hip_options base{variant, 0, 0, -1, 0};
    std::vector<int> families;
    if (variant == NBODY_HIP_AUTO) families = {VAR_LDS, VAR_SGPR, VAR_SYMMETRIC};
    else if (variant >= 0 && variant < VAR_COUNT) families = {variant};
    else {
        set_error("unknown variant %d", variant);
        return 0.0f;
    }
    std::vector<ForcePlan> cands;
    for (int fam : families) {
        nbody_hip_options o = base;
        o.variant = fam;
        if (fam == VAR_AOS) {
            cands.push_back(default_plan(fam, N, N, d->cus, &o));
        } else if (fam == VAR_SYMMETRIC) {
            const int shapes[][2] = {{2, 128}, {2, 512}, {4, 256}, {4, 512}, {4, 1024}, {8, 512}, {8, 1024}, {8, 2048}};
            for (const auto& sh : shapes) {
                o.i_per_lane = sh[0];
                o.jslice = sh[1];
                ForcePlan c = default_plan(fam, N, N, d->cus, &o);
                if (c.nslices > kMaxSymBlocks) continue;
                cands.push_back(c);
                // the same shape with one and two fewer resident waves per
                // SIMD, and with two: a launch of a few tasks per SIMD ends
                // unevenly when they all start at once, while at two resident
                // waves (about the issue rate of six) the later tasks start
                // as slots free
                const int natural = symmetric_natural_occupancy(c);
                int last = natural;
                for (int waves : {natural - 1, natural - 2, 2}) {
                    if (waves < 2 || waves >= last) continue;
                    last = waves;
                    c.occ = waves;
                    cands.push_back(c);
                }
            }
        } else {
            for (int ipl : {1, 2, 4, 8}) {
                if (ipl == 8 && fam != VAR_DPP) continue;
                o.i_per_lane = ipl;
                o.jslice = 0;
                ForcePlan q = default_plan(fam, N, N, d->cus, &o);
                for (int mul : {1, 2, 4}) {
                    for (int div : {1, 2, 4}) {
                        if (mul > 1 && div > 1) continue;
                        ForcePlan c = q;
                        c.jslice = std::min((int)N, round_up(std::max(256, q.jslice * mul / div), 256));
                        c.nslices = (int)((N + c.jslice - 1) / c.jslice);
                        cands.push_back(c);
                    }
                }
            }
        }
    }
    if (cands.empty()) {
        set_error("no plan of variant %d fits N = %zu", variant, N);
        return 0.0f;
    }
    float best = 0.0f;
    ForcePlan best_plan = cands.front();
    for (const ForcePlan& c : cands) {
        const size_t bytes = plan_scratch_bytes(c, (int)N);
        if (bytes && d->partial.reserve(bytes) != hipSuccess) continue;
        // one untimed launch, one to size the batches (about 2 ms of launches
        // each, so short kernels are not timed at the resolution of a single
        // launch), then the best per-launch average of three batches
        float t = 0.0f;
        int batch = 1;
        for (int rep = 0; rep < 5; ++rep) {
            if (hipEventRecord(d->ev0, nullptr) != hipSuccess) {
                set_error("autotune timing failed");
                return 0.0f;
            }
            for (int k = 0; k < batch; ++k)
                if (run_force(*d, c, force, posMass, softeningSquared, N, 0, N, nullptr)) return 0.0f;
            float ms = 0.0f;
            if (hipEventRecord(d->ev1, nullptr) != hipSuccess || hipEventSynchronize(d->ev1) != hipSuccess ||
                hipEventElapsedTime(&ms, d->ev0, d->ev1) != hipSuccess) {
                set_error("autotune timing failed");
                return 0.0f;
            }
            ms /= (float)batch;
            if (rep == 1) batch = std::max(1, std::min(64, (int)(2.0f / std::max(ms, 1e-3f))));
            if (rep >= 2 && (t == 0.0f || ms < t)) t = ms;
        }
        if (t > 0.0f && (best == 0.0f || t < best)) {
            best = t;
            best_plan = c;
        }
    }
    if (best == 0.0f) {
        set_error("autotune: no candidate plan could allocate its scratch for N = %zu", N);
        return 0.0f;
    }
    d->tuned[tune_key(&base, N, N)] = best_plan;
    return best;
}

int nbody_hip_kernel_timer(int enable)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* d;
    if (int r = device_state(&d)) return r;
    d->timing = enable != 0;
    if (enable) d->timer_used = 0;
    return OK;
}

int nbody_hip_kernel_timer_read(double* total_ms, int* launches)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* d;
    if (int r = device_state(&d)) return r;
    double sum = 0.0;
    for (size_t k = 0; k < d->timer_used; ++k) {
        float ms = 0.0f;
        HIP_TRY(hipEventSynchronize(d->timer_events[k].second));
        HIP_TRY(hipEventElapsedTime(&ms, d->timer_events[k].first, d->timer_events[k].second));
        sum += ms;
    }
    if (total_ms) *total_ms = sum;
    if (launches) *launches = (int)d->timer_used;
    return OK;
}

const char* nbody_hip_last_error(void) { return g_last_error.c_str(); }

int nbody_hip_abi_version(void) { return NBODY_HIP_ABI_VERSION; }

int nbody_hip_device_info(int* compute_units, int* clock_khz, char* arch, size_t arch_len)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState* d;
    if (int r = device_state(&d)) return r;
    if (compute_units) *compute_units = d->cus;
    if (clock_khz) *clock_khz = d->clock_khz;
    if (arch && arch_len) snprintf(arch, arch_len, "%s", d->arch);
    return OK;
}

}  // extern "C"
