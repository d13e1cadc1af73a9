// runtime.hpp -- host-side runtime shared by the C ABI translation units:
// the thread-local error message, device selection, the per-device
// release/setup streams and stream-ordered lifetimes of device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/intp_amd.h"

// A device context (intp_ctx_create).
struct intp_ctx_s {
    int device;
};

namespace intp_amd {

inline thread_local std::string g_error;

inline int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

inline int hip_fail(hipError_t e, const char* what) {
    const int code = e == hipErrorOutOfMemory ? INTP_E_NOMEM : INTP_E_HIP;
    return fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                    \
    do {                                       \
        hipError_t e_ = (expr);                \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

// Makes `device` current for the scope, restoring the caller's device.
class DeviceScope {
   public:
    explicit DeviceScope(int device) {
        ok_ = hipGetDevice(&prev_) == hipSuccess;
        if (ok_ && prev_ != device) ok_ = hipSetDevice(device) == hipSuccess;
    }
    ~DeviceScope() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev_) (void)hipSetDevice(prev_);
    }
    bool ok() const { return ok_; }

   private:
    int prev_ = 0;
    bool ok_ = false;
};

// ---- device memory lifetimes -------------------------------------------
//
// Every device buffer comes from the stream-ordered allocator (the
// device's default memory pool, kept cached up to kPoolKeepBytes) and is
// released with hipFreeAsync on a per-device release stream that first
// waits for every recorded use of the buffer.  Nothing on the build or
// evaluation path calls hipMalloc/hipFree, which synchronize the device.

constexpr int kMaxDevices = 64;
constexpr uint64_t kPoolKeepBytes = uint64_t{16} << 30;

struct DeviceRuntime {
    bool ready = false;
    hipStream_t release = nullptr;  // frees, ordered after recorded uses
    hipStream_t setup = nullptr;    // uploads of calls that take no stream
};

// Per-device streams and pool settings, created on first use and kept
// until the process ends.  Both streams are non-blocking with default
// priority.
inline DeviceRuntime* device_runtime(int device) {
    static std::mutex mu;
    static DeviceRuntime rt[kMaxDevices];
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    DeviceRuntime& r = rt[device];
    if (!r.ready) {
        DeviceScope scope(device);
        if (!scope.ok()) return nullptr;
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
            uint64_t keep = 0;
            if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess &&
                keep < kPoolKeepBytes) {
                keep = kPoolKeepBytes;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            }
            // A block freed on one stream is handed to an allocation on
            // another stream only once the free has completed, never by
            // dependencies the pool inserts between the two streams.
            int no = 0;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolReuseAllowInternalDependencies, &no);
        }
        if (!r.release && hipStreamCreateWithFlags(&r.release, hipStreamNonBlocking) != hipSuccess) return nullptr;
        if (!r.setup && hipStreamCreateWithFlags(&r.setup, hipStreamNonBlocking) != hipSuccess) return nullptr;
        r.ready = true;
    }
    return &r;
}

// The streams an object's device memory was used on, each with an event
// recorded after its latest use there.
class UseTracker {
   public:
    UseTracker() = default;
    UseTracker(const UseTracker&) = delete;
    UseTracker& operator=(const UseTracker&) = delete;
    ~UseTracker() {
        for (const Use& u : uses_) (void)hipEventDestroy(u.event);
    }

    hipError_t record(hipStream_t s) {
        std::lock_guard<std::mutex> lock(mu_);
        for (const Use& u : uses_)
            if (u.stream == s) return hipEventRecord(u.event, s);
        if (uses_.size() >= kMaxStreams) {
            // many distinct streams: settle the old ones instead of growing
            for (const Use& u : uses_) {
                (void)hipEventSynchronize(u.event);
                (void)hipEventDestroy(u.event);
            }
            uses_.clear();
        }
        hipEvent_t ev = nullptr;
        hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return e;
        uses_.push_back({s, ev});
        return hipEventRecord(ev, s);
    }

    // Makes `s` wait for every recorded use.
    void order_before(hipStream_t s) {
        std::lock_guard<std::mutex> lock(mu_);
        for (const Use& u : uses_) (void)hipStreamWaitEvent(s, u.event, 0);
    }

    hipError_t wait_host() {
        std::lock_guard<std::mutex> lock(mu_);
        for (const Use& u : uses_) {
            const hipError_t e = hipEventSynchronize(u.event);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }

   private:
    struct Use {
        hipStream_t stream;
        hipEvent_t event;
    };
    static constexpr size_t kMaxStreams = 16;
    std::mutex mu_;
    std::vector<Use> uses_;
};

// Frees `p` (from hipMallocAsync) once every use recorded in `uses` is
// done, without blocking the host.
inline void release_after(int device, void* p, UseTracker& uses) {
    if (!p) return;
    DeviceScope scope(device);
    DeviceRuntime* rt = device_runtime(device);
    if (rt) {
        uses.order_before(rt->release);
        if (hipFreeAsync(p, rt->release) == hipSuccess) return;
    }
    (void)uses.wait_host();
    (void)hipFree(p);
}


}  // namespace intp_amd
