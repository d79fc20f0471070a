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
}  // namespace

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
