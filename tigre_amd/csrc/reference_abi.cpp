// The reference's C++ entry points (include/tigre_reference_abi.hpp), each a
// forwarder to the C ABI.  The detector layout is fixed per build of this
// file, as the reference fixes it per build (IS_FOR_PYTIGRE,
// TIGRE_common.hpp:10-23): TIGRE_REFERENCE_LAYOUT 0 (Python, u fastest) goes
// into libtigre_amd.so, 1 (MATLAB, v fastest) into libtigre_amd_matlab.so.
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <hip/hip_runtime.h>

#include "../../include/tigre_reference_abi.hpp"

#ifndef TIGRE_REFERENCE_LAYOUT
#define TIGRE_REFERENCE_LAYOUT TIGRE_LAYOUT_PYTHON
#endif

// Geometry and TigreGeometry are the same record.
#define TIGRE_SAME_OFFSET(f) static_assert(offsetof(Geometry, f) == offsetof(TigreGeometry, f), #f)
TIGRE_SAME_OFFSET(nVoxelX);
TIGRE_SAME_OFFSET(sVoxelX);
TIGRE_SAME_OFFSET(dVoxelX);
TIGRE_SAME_OFFSET(offOrigX);
TIGRE_SAME_OFFSET(DSO);
TIGRE_SAME_OFFSET(nDetecU);
TIGRE_SAME_OFFSET(sDetecU);
TIGRE_SAME_OFFSET(dDetecU);
TIGRE_SAME_OFFSET(offDetecU);
TIGRE_SAME_OFFSET(DSD);
TIGRE_SAME_OFFSET(dRoll);
TIGRE_SAME_OFFSET(dYaw);
TIGRE_SAME_OFFSET(unitX);
TIGRE_SAME_OFFSET(alpha);
TIGRE_SAME_OFFSET(COR);
TIGRE_SAME_OFFSET(maxLength);
TIGRE_SAME_OFFSET(accuracy);
#undef TIGRE_SAME_OFFSET
static_assert(sizeof(Geometry) == sizeof(TigreGeometry), "Geometry size");
static_assert(offsetof(GpuIds, m_iCount) == sizeof(int*), "GpuIds layout");

namespace {

const TigreGeometry* as_c(const Geometry& g) { return reinterpret_cast<const TigreGeometry*>(&g); }

void report(const char* where, int status) {
    if (status != TIGRE_OK) std::fprintf(stderr, "%s: %s (status %d)\n", where, tigre_last_error(), status);
}

}  // namespace

// ---- GPU list (GpuIds.hpp:3-14) -------------------------------------------
GpuIds::GpuIds() : m_piDeviceIds(nullptr), m_iCount(0) {}

GpuIds::~GpuIds() {
    std::free(m_piDeviceIds);
    m_piDeviceIds = nullptr;
    m_iCount = 0;
}

void GpuIds::SetIds(int iCount, int* piDeviceIds) {
    if (iCount <= 0 || piDeviceIds == nullptr) return;
    int* ids = static_cast<int*>(std::malloc((size_t)iCount * sizeof(int)));
    if (ids == nullptr) return;
    std::memcpy(ids, piDeviceIds, (size_t)iCount * sizeof(int));
    std::free(m_piDeviceIds);
    m_piDeviceIds = ids;
    m_iCount = iCount;
}

int GpuIds::GetLength() const { return m_iCount; }

// Devices 0 .. iTotalDeviceCount - 1.
void GpuIds::SetAllGpus(int iTotalDeviceCount) {
    if (iTotalDeviceCount <= 0) return;
    int* ids = static_cast<int*>(std::malloc((size_t)iTotalDeviceCount * sizeof(int)));
    if (ids == nullptr) return;
    for (int i = 0; i < iTotalDeviceCount; ++i) ids[i] = i;
    SetIds(iTotalDeviceCount, ids);
    std::free(ids);
}

int& GpuIds::operator[](int iIndex) { return m_piDeviceIds[iIndex]; }
int GpuIds::operator[](int iIndex) const { return m_piDeviceIds[iIndex]; }

bool GpuIds::AreEqualDevices() const {
    char first[256] = {0};
    for (int i = 0; i < m_iCount; ++i) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, m_piDeviceIds[i]) != hipSuccess) return false;
        if (i == 0)
            std::strncpy(first, prop.name, sizeof(first) - 1);
        else if (std::strncmp(first, prop.name, sizeof(first)) != 0)
            return false;
    }
    return true;
}

// ---- forward projectors ---------------------------------------------------
int siddon_ray_projection(float* img, Geometry geo, float** result, float const* const angles, int nangle,
                          const GpuIds& gpuids) {
    return tigre_ax(img, as_c(geo), result, angles, nangle, TIGRE_PROJ_SIDDON, TIGRE_REFERENCE_LAYOUT,
                    gpuids.m_piDeviceIds, gpuids.m_iCount);
}

int interpolation_projection(float* img, Geometry geo, float** result, float const* const alphas, int nalpha,
                             const GpuIds& gpuids) {
    return tigre_ax(img, as_c(geo), result, alphas, nalpha, TIGRE_PROJ_INTERPOLATED, TIGRE_REFERENCE_LAYOUT,
                    gpuids.m_piDeviceIds, gpuids.m_iCount);
}

int siddon_ray_projection_parallel(float* img, Geometry geo, float** result, float const* const alphas, int nalpha,
                                   const GpuIds& gpuids) {
    return tigre_ax_parallel(img, as_c(geo), result, alphas, nalpha, TIGRE_PROJ_SIDDON, TIGRE_REFERENCE_LAYOUT,
                             gpuids.m_piDeviceIds, gpuids.m_iCount);
}

int interpolation_projection_parallel(float* img, Geometry geo, float** result, float const* const alphas,
                                      int nalpha, const GpuIds& gpuids) {
    return tigre_ax_parallel(img, as_c(geo), result, alphas, nalpha, TIGRE_PROJ_INTERPOLATED,
                             TIGRE_REFERENCE_LAYOUT, gpuids.m_piDeviceIds, gpuids.m_iCount);
}

// ---- backprojectors -------------------------------------------------------
int voxel_backprojection(float* projections, Geometry geo, float* result, float const* const alphas, int nalpha,
                         const GpuIds& gpuids) {
    return tigre_atb(projections, as_c(geo), result, alphas, nalpha, TIGRE_BACKPROJ_FDK, TIGRE_REFERENCE_LAYOUT,
                     gpuids.m_piDeviceIds, gpuids.m_iCount);
}

int voxel_backprojection2(float* projections, Geometry geo, float* result, float const* const alphas, int nalpha,
                          const GpuIds& gpuids) {
    return tigre_atb(projections, as_c(geo), result, alphas, nalpha, TIGRE_BACKPROJ_MATCHED, TIGRE_REFERENCE_LAYOUT,
                     gpuids.m_piDeviceIds, gpuids.m_iCount);
}

int voxel_backprojection_parallel(float* projections, Geometry geo, float* result, float const* const alphas,
                                  int nalpha, const GpuIds& gpuids) {
    return tigre_atb_parallel(projections, as_c(geo), result, alphas, nalpha, TIGRE_REFERENCE_LAYOUT,
                              gpuids.m_piDeviceIds, gpuids.m_iCount);
}

// ---- TV steps -------------------------------------------------------------
void pocs_tv(float* img, float* dst, float alpha, const long* image_size, int maxIter, const GpuIds& gpuids) {
    report("pocs_tv", tigre_pocs_tv(img, dst, alpha, image_size, maxIter, gpuids.m_piDeviceIds, gpuids.m_iCount));
}

void aw_pocs_tv(float* img, float* dst, float alpha, const long* image_size, int maxIter, const float delta,
                const GpuIds& gpuids) {
    report("aw_pocs_tv", tigre_aw_pocs_tv(img, dst, alpha, image_size, maxIter, delta, gpuids.m_piDeviceIds,
                                          gpuids.m_iCount));
}

void piccs_tv(const float* img, const float* prior, float* dst, float alpha, float ratio, const long* image_size,
              int maxIter, const GpuIds& gpuids) {
    report("piccs_tv", tigre_piccs_tv(img, prior, dst, alpha, ratio, image_size, maxIter, gpuids.m_piDeviceIds,
                                      gpuids.m_iCount));
}

void tvdenoising(float* src, float* dst, float lambda, const float* spacing, const long* image_size, int maxIter,
                 const GpuIds& gpuids) {
    report("tvdenoising", tigre_tvdenoising(src, dst, lambda, spacing, image_size, maxIter, gpuids.m_piDeviceIds,
                                            gpuids.m_iCount));
}

// ---- device queries (gpuUtils.hpp:5-15) -----------------------------------
int GetGpuCount() { return tigre_get_gpu_count(); }

// Ids of the devices named kacGPUName (every device for an empty name), at
// most iIdCountMax of them; pcMessage (65535 bytes, or null) receives a log.
int GetGpuIdArray(const char* kacGPUName, int* piDeviceIds, int iIdCountMax, char* pcMessage) {
    const size_t cap = 65535;
    size_t pos = 0;
    auto say = [&](const char* fmt, auto... args) {
        if (pcMessage == nullptr || pos >= cap) return;
        const int n = std::snprintf(pcMessage + pos, cap - pos, fmt, args...);
        if (n > 0) pos += (size_t)n < cap - pos ? (size_t)n : cap - pos - 1;
    };
    if (pcMessage) std::memset(pcMessage, 0, cap);
    if (piDeviceIds == nullptr || iIdCountMax <= 0) return 0;
    int count = GetGpuCount();
    say("Found GPUs: %d\n", count);
    if (count <= 0) return 0;
    if (count > iIdCountMax) count = iIdCountMax;
    say("Max GPUs: %d\n", count);
    const bool any = kacGPUName == nullptr || kacGPUName[0] == '\0';
    int found = 0;
    for (int id = 0; id < count; ++id) {
        char name[256] = {0};
        if (tigre_get_gpu_name(id, name, (int)sizeof(name)) != TIGRE_OK) continue;
        say("propDevice.name = %s\n", name);
        if (any || std::strcmp(name, kacGPUName) == 0) piDeviceIds[found++] = id;
    }
    for (int i = 0; i < found; ++i) say("%d, ", piDeviceIds[i]);
    return found;
}

// pcName: 128 bytes, zero-filled, then the device name (at most 127 characters).
void GetGpuName(int iDeviceId, char* pcName) {
    if (pcName == nullptr) return;
    std::memset(pcName, 0, 128);
    (void)tigre_get_gpu_name(iDeviceId, pcName, 128);
}
