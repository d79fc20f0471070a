// Device selection, stream-ordered parameter upload and the error channel.
#include <cstring>
#include <mutex>

#include "common.hpp"

namespace tigre {

namespace {
thread_local std::string g_last_error;

// Per-device pinned staging for parameter uploads.  Two slots, each guarded
// by an event recorded after the copy that reads it, so the host never
// rewrites a staging buffer that an enqueued copy has not consumed yet.
struct DeviceState {
    std::mutex mu;
    bool ready = false;
    void* staging[2] = {nullptr, nullptr};
    size_t staging_bytes[2] = {0, 0};
    hipEvent_t done[2] = {nullptr, nullptr};
    int next = 0;
};

constexpr int kMaxDevices = 64;
DeviceState g_dev[kMaxDevices];

int current_state(DeviceState** st) {
    int dev = 0;
    TIGRE_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device ordinal out of range");
        return TIGRE_ERR_NO_CAPABLE_DEVICES;
    }
    DeviceState& s = g_dev[dev];
    if (!s.ready) {
        for (int i = 0; i < 2; ++i) TIGRE_HIP_TRY(hipEventCreateWithFlags(&s.done[i], hipEventDisableTiming));
        // Keep stream-ordered allocations cached in the pool between calls.
        hipMemPool_t pool;
        TIGRE_HIP_TRY(hipDeviceGetDefaultMemPool(&pool, dev));
        uint64_t threshold = UINT64_MAX;
        TIGRE_HIP_TRY(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold));
        s.ready = true;
    }
    *st = &s;
    return TIGRE_OK;
}

struct TimedLaunch {
    hipEvent_t start, stop;
};
std::mutex g_timer_mu;
bool g_timing = false;
std::vector<TimedLaunch> g_timed[TIGRE_KERNEL_COUNT];
}  // namespace

KernelTimerScope::KernelTimerScope(int kernel, hipStream_t stream) : kernel_(kernel), stream_(stream) {
    {
        std::lock_guard<std::mutex> lock(g_timer_mu);
        if (!g_timing || kernel < 0 || kernel >= TIGRE_KERNEL_COUNT) return;
    }
    if (hipEventCreate(&start_) != hipSuccess) {
        start_ = nullptr;
        return;
    }
    if (hipEventRecord(start_, stream_) != hipSuccess) {
        (void)hipEventDestroy(start_);
        start_ = nullptr;
    }
}

KernelTimerScope::~KernelTimerScope() {
    if (start_ == nullptr) return;
    hipEvent_t stop = nullptr;
    if (hipEventCreate(&stop) != hipSuccess || hipEventRecord(stop, stream_) != hipSuccess) {
        if (stop) (void)hipEventDestroy(stop);
        (void)hipEventDestroy(start_);
        return;
    }
    std::lock_guard<std::mutex> lock(g_timer_mu);
    g_timed[kernel_].push_back({start_, stop});
}

int kernel_timing_enable(int on) {
    std::lock_guard<std::mutex> lock(g_timer_mu);
    g_timing = on != 0;
    return TIGRE_OK;
}

int kernel_time_collect(int kernel, double* total_ms, int* launches) {
    TIGRE_REQUIRE(kernel >= 0 && kernel < TIGRE_KERNEL_COUNT, "unknown kernel id");
    TIGRE_REQUIRE(total_ms != nullptr && launches != nullptr, "null output pointer");
    std::vector<TimedLaunch> recorded;
    {
        std::lock_guard<std::mutex> lock(g_timer_mu);
        recorded.swap(g_timed[kernel]);
    }
    double sum = 0.0;
    int status = TIGRE_OK;
    for (const TimedLaunch& t : recorded) {
        float ms = 0.0f;
        if (status == TIGRE_OK) {
            hipError_t e = hipEventSynchronize(t.stop);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, t.start, t.stop);
            if (e != hipSuccess) {
                set_error(std::string("kernel timing: ") + hipGetErrorString(e));
                status = TIGRE_ERR_HIP;
            }
        }
        sum += ms;
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    *total_ms = sum;
    *launches = (int)recorded.size();
    return status;
}

void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

int select_device(const int* gpuids, int ngpus) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        set_error("no HIP capable device available");
        return TIGRE_ERR_NO_CAPABLE_DEVICES;
    }
    int dev = 0;
    if (gpuids != nullptr && ngpus > 0) dev = gpuids[0];
    if (dev < 0 || dev >= count) {
        set_error("requested device id " + std::to_string(dev) + " not present");
        return TIGRE_ERR_NO_CAPABLE_DEVICES;
    }
    TIGRE_HIP_TRY(hipSetDevice(dev));
    return TIGRE_OK;
}

int upload_async(void** d_ptr, const void* host, size_t bytes, hipStream_t stream) {
    DeviceState* st = nullptr;
    TIGRE_CHECK(current_state(&st));
    std::lock_guard<std::mutex> lock(st->mu);
    const int slot = st->next;
    st->next ^= 1;
    TIGRE_HIP_TRY(hipEventSynchronize(st->done[slot]));
    if (st->staging_bytes[slot] < bytes) {
        if (st->staging[slot]) TIGRE_HIP_TRY(hipHostFree(st->staging[slot]));
        st->staging[slot] = nullptr;
        size_t cap = bytes < 65536 ? 65536 : bytes * 2;
        TIGRE_HIP_TRY(hipHostMalloc(&st->staging[slot], cap, hipHostMallocDefault));
        st->staging_bytes[slot] = cap;
    }
    std::memcpy(st->staging[slot], host, bytes);
    TIGRE_HIP_TRY(hipMallocAsync(d_ptr, bytes, stream));
    TIGRE_HIP_TRY(hipMemcpyAsync(*d_ptr, st->staging[slot], bytes, hipMemcpyHostToDevice, stream));
    TIGRE_HIP_TRY(hipEventRecord(st->done[slot], stream));
    return TIGRE_OK;
}

}  // namespace tigre
