/*
 * tigre_amd -- the reference's own C++ entry points, for gateways that were
 * written against them (the Cython extensions under
 * Python/tigre/utilities/cuda_interface/_*.pyx and the MEX files under
 * MATLAB/Source/Ax_mex.cpp and its siblings).  Those gateways include the reference headers and call
 * C++ functions (not extern "C") that take the geometry by value and the GPU
 * list by reference; libtigre_amd.so exports the same names with the same
 * signatures, so a gateway links against it unchanged.  Every function
 * forwards to the C ABI in tigre_amd.h.
 *
 * The two structs below are declared with the reference's names, members and
 * member order (types_TIGRE.hpp:48-85, GpuIds.hpp:3-14): their names enter the
 * mangled symbol names, and their layout is the binary contract.  Only one
 * device is used per call: the first id of the list (device 0 when the list is
 * empty).
 *
 * Errors: the int functions return 0 on success and the status codes of
 * tigre_amd.h otherwise.  The void functions (pocs_tv, aw_pocs_tv, piccs_tv,
 * tvdenoising) print the failure on stderr and return; the reference prints
 * and exits the process instead (TIGRE_common.cpp:13-16).  tigre_last_error()
 * holds the message in both cases.
 */
#ifndef TIGRE_REFERENCE_ABI_HPP
#define TIGRE_REFERENCE_ABI_HPP

#include "tigre_amd.h"

/* Same members, order and layout as TigreGeometry (checked by static_assert in
 * the implementation). */
struct Geometry {
    int nVoxelX, nVoxelY, nVoxelZ;
    float sVoxelX, sVoxelY, sVoxelZ;
    float dVoxelX, dVoxelY, dVoxelZ;
    float *offOrigX, *offOrigY, *offOrigZ;
    float* DSO;
    int nDetecU, nDetecV;
    float sDetecU, sDetecV;
    float dDetecU, dDetecV;
    float *offDetecU, *offDetecV;
    float* DSD;
    float* dRoll;
    float* dPitch;
    float* dYaw;
    float unitX;
    float unitY;
    float unitZ;
    float alpha;
    float theta;
    float psi;
    float* COR;
    float maxLength;
    float accuracy;
};

/* Device list.  m_piDeviceIds is malloc'd memory owned by the object once set
 * through SetIds / SetAllGpus (released by the destructor with free); the
 * Cython gateways build the struct themselves with malloc and never call the
 * members. */
struct GpuIds {
    int* m_piDeviceIds;
    int m_iCount;
    ~GpuIds();
    GpuIds();
    void SetIds(int iCount, int* piDeviceIds);
    int GetLength() const;
    void SetAllGpus(int iTotalDeviceCount);
    int& operator[](int iIndex);
    int operator[](int iIndex) const;
    bool AreEqualDevices() const;
};

/* Siddon_projection.hpp:54 */
int siddon_ray_projection(float* img, Geometry geo, float** result, float const* const angles, int nangle,
                          const GpuIds& gpuids);
/* ray_interpolated_projection.hpp:55 */
int interpolation_projection(float* img, Geometry geo, float** result, float const* const alphas, int nalpha,
                             const GpuIds& gpuids);
/* Siddon_projection_parallel.hpp:56 */
int siddon_ray_projection_parallel(float* img, Geometry geo, float** result, float const* const alphas, int nalpha,
                                   const GpuIds& gpuids);
/* ray_interpolated_projection_parallel.hpp:57 */
int interpolation_projection_parallel(float* img, Geometry geo, float** result, float const* const alphas,
                                      int nalpha, const GpuIds& gpuids);
/* voxel_backprojection.hpp:52 (FDK weights) */
int voxel_backprojection(float* projections, Geometry geo, float* result, float const* const alphas, int nalpha,
                         const GpuIds& gpuids);
/* voxel_backprojection2.hpp:57 (matched weights) */
int voxel_backprojection2(float* projections, Geometry geo, float* result, float const* const alphas, int nalpha,
                          const GpuIds& gpuids);
/* voxel_backprojection_parallel.hpp:53 */
int voxel_backprojection_parallel(float* projections, Geometry geo, float* result, float const* const alphas,
                                  int nalpha, const GpuIds& gpuids);
/* GD_TV.hpp:58 */
void pocs_tv(float* img, float* dst, float alpha, const long* image_size, int maxIter, const GpuIds& gpuids);
/* GD_AwTV.hpp:58 */
void aw_pocs_tv(float* img, float* dst, float alpha, const long* image_size, int maxIter, const float delta,
                const GpuIds& gpuids);
/* PICCS.hpp:58 */
void piccs_tv(const float* img, const float* prior, float* dst, float alpha, float ratio, const long* image_size,
              int maxIter, const GpuIds& gpuids);
/* tv_proximal.hpp:52 */
void tvdenoising(float* src, float* dst, float lambda, const float* spacing, const long* image_size, int maxIter,
                 const GpuIds& gpuids);
/* gpuUtils.hpp:5-15 */
int GetGpuCount();
int GetGpuIdArray(const char* kacGPUName, int* piDeviceIds, int iIdCountMax, char* pcMessage);
void GetGpuName(int iDeviceId, char* pcName);

#endif /* TIGRE_REFERENCE_ABI_HPP */
