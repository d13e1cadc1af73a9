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
        // An error an earlier call on this thread left in the runtime's
        // last-error slot (the caller's, or a release whose failure the
        // library ignores on purpose) would otherwise be read back by the
        // first launch check of this call and reported as its own.
        (void)hipGetLastError();
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
// Every device buffer comes from the library's own stream-ordered memory
// pool on its device (device_alloc; kept cached up to kPoolKeepBytes) and
// is released with hipFreeAsync, either on the stream that used it or on
// a per-device release stream that first waits for every recorded use of
// the buffer.  Nothing on the build or evaluation path calls
// hipMalloc/hipFree, which synchronize the device.
//
// Reuse policy of the pool: a freed block goes to a later allocation on
// the stream that freed it (stream order makes that safe), or to another
// stream once the free has completed on the device.  Reuse through event
// dependencies is off: with it on, two refined builds on different
// streams of the serving loop computed into one scratch block; the
// second was launched after its stream had waited on an event of the
// first one's stream that was recorded before the first build's free.
// Reuse through dependencies the pool inserts itself is off as well.

constexpr int kMaxDevices = 64;
constexpr uint64_t kPoolKeepBytes = uint64_t{16} << 30;

struct DeviceRuntime {
    bool ready = false;
    hipStream_t release = nullptr;  // frees, ordered after recorded uses
    hipStream_t setup = nullptr;    // uploads of calls that take no stream
    hipMemPool_t pool = nullptr;    // every device buffer of the library
};

inline hipError_t create_pool(int device, hipMemPool_t* out) {
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = device;
    hipMemPool_t pool = nullptr;
    hipError_t e = hipMemPoolCreate(&pool, &props);
    if (e != hipSuccess) return e;
    uint64_t keep = kPoolKeepBytes;
    int no = 0, yes = 1;
    e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    if (e == hipSuccess) e = hipMemPoolSetAttribute(pool, hipMemPoolReuseFollowEventDependencies, &no);
    if (e == hipSuccess) e = hipMemPoolSetAttribute(pool, hipMemPoolReuseAllowInternalDependencies, &no);
    if (e == hipSuccess) e = hipMemPoolSetAttribute(pool, hipMemPoolReuseAllowOpportunistic, &yes);
    if (e != hipSuccess) {
        (void)hipMemPoolDestroy(pool);
        return e;
    }
    *out = pool;
    return hipSuccess;
}

// Per-device streams and memory pool, created on first use and kept until
// the process ends.  Both streams are non-blocking with default priority.
inline DeviceRuntime* device_runtime(int device) {
    static std::mutex mu;
    static DeviceRuntime rt[kMaxDevices];
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    DeviceRuntime& r = rt[device];
    if (!r.ready) {
        DeviceScope scope(device);
        if (!scope.ok()) return nullptr;
        if (!r.pool && create_pool(device, &r.pool) != hipSuccess) return nullptr;
        if (!r.release && hipStreamCreateWithFlags(&r.release, hipStreamNonBlocking) != hipSuccess) return nullptr;
        if (!r.setup && hipStreamCreateWithFlags(&r.setup, hipStreamNonBlocking) != hipSuccess) return nullptr;
        r.ready = true;
    }
    return &r;
}

// Stream-ordered allocation of `bytes` on the current device from the
// library's pool, usable on `s` (and, after an event, on other streams).
inline hipError_t device_alloc(void** p, size_t bytes, hipStream_t s) {
    int device = -1;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    DeviceRuntime* rt = device_runtime(device);
    if (!rt) return hipErrorInvalidDevice;
    return hipMallocFromPoolAsync(p, bytes, rt->pool, s);
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

// Frees `p` (from device_alloc) once every use recorded in `uses` is
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
