// Host-side internals shared by the C-ABI translation units (not installed).
#pragma once

#include "nbody_hip.h"
#include "nbody_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

namespace nbody {
namespace detail {

enum : int { OK = 0, E_ARG = 1, E_HIP = 2, E_UNSUPPORTED = 3 };

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                      __LINE__);                                                       \
            return E_HIP;                                                              \
        }                                                                              \
    } while (0)

// Largest N accepted: float4 offsets and 4*N fit comfortably in int.
constexpr size_t kMaxBodies = size_t(1) << 28;

// Device buffer that only grows. Growing waits for the device first, so a
// buffer still read by queued kernels is never freed under them.
struct Buffer {
    void* ptr = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (ptr) {
            hipError_t e = hipDeviceSynchronize();
            if (e != hipSuccess) return e;
            (void)hipFree(ptr);
            ptr = nullptr;
            cap = 0;
        }
        size_t want = std::max(bytes, cap + cap / 2);
        hipError_t e = hipMalloc(&ptr, want);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        cap = want;
        return hipSuccess;
    }
};

// Grow-only page-locked host buffer (hipHostMalloc), the staging area of
// the host-pointer entry: DMA runs at full link speed from and to it.
struct PinnedBuffer {
    void* ptr = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (ptr) {
            hipError_t e = hipDeviceSynchronize();
            if (e != hipSuccess) return e;
            (void)hipHostFree(ptr);
            ptr = nullptr;
            cap = 0;
        }
        hipError_t e = hipHostMalloc(&ptr, bytes, hipHostMallocDefault);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        cap = bytes;
        return hipSuccess;
    }
};

struct DeviceState {
    bool init = false;
    int cus = 0;
    int clock_khz = 0;
    char arch[64] = {0};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Buffer partial;   // FP32 slice partials, shared by the calls that bring no scratch of their own
    // The last queued use of `partial`: its stream and an event recorded
    // after its kernels. A call on another stream waits for that event
    // before its kernels touch the buffer.
    hipEvent_t partial_done = nullptr;
    hipStream_t partial_stream = nullptr;
    bool partial_pending = false;
    Buffer partial64; // FP64 slice partials / energy partials
    Buffer stage_pm, stage_force, small;
    // this device's copy of the no-allocation entry's fixed scratch array
    // (its address from hipGetSymbolAddress, looked up on the first call)
    void* noalloc_scratch = nullptr;
    PinnedBuffer host_stage;  // posMass in, force out, for the host-pointer entry
    // tuned plans: (requested variant, N, n, requested compensation, newton) -> plan
    std::map<std::tuple<int, size_t, size_t, int, int>, ForcePlan> tuned;
    // kernel timer: event pairs recorded around each force kernel while enabled
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> timer_events;
    size_t timer_used = 0;
};

// Largest number of force launches the kernel timer records between reads.
constexpr size_t kMaxTimedLaunches = 1 << 14;



// Last error message of this thread (nbody_hip_last_error) and the lock every
// entry point holds while it touches device state.
extern thread_local std::string g_last_error;
extern std::mutex g_mutex;

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int device_state(DeviceState** out);
int round_up(int v, int m);
ForcePlan default_plan(int variant, size_t N, size_t n, int cus, const nbody_hip_options* opt);
int make_plan(DeviceState& d, size_t N, size_t n, const nbody_hip_options* opt, ForcePlan* plan);
int check_args(const void* force, const void* pm, size_t N, size_t base, size_t n);
// Queue one force evaluation on `stream`. Slice/slot scratch comes from
// `own` (the device's shared buffer when null; a caller that captures the
// launch into a graph passes its own, already large enough). Uses of the
// shared buffer are ordered across streams (each waits for the previous
// one's kernels), and refused on a stream that is capturing a graph, since a
// later, larger call may reallocate the buffer under the graph. The kernel
// timer records only when `timed` is set. With force64 (a symmetric plan over
// a share of the columns) the partial forces go there in FP64 instead.
int run_force(DeviceState& d, const ForcePlan& plan, float* force, const float* pm, float eps2, size_t N, size_t base,
              size_t n, hipStream_t stream, Buffer* own = nullptr, bool timed = true, double* force64 = nullptr);

}  // namespace detail
}  // namespace nbody
