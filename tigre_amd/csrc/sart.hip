// Elementwise step of the SART family on device buffers.
//
// r = W * (p - Ax): the weighted data residual of one subset
// (iterative_recon_alg.py:326-345, art_data_minimizing), written over the
// projection buffer.  The two rounding steps are kept apart (no fused
// multiply-add), so the result equals the unfused tensor expression bit for bit.
#include "common.hpp"

namespace tigre {

namespace {

// Elements (as float4) first, first + stride, ... and the scalar tail.
__device__ __forceinline__ void sart_residual_range(const float* __restrict__ proj, const float* __restrict__ w,
                                                    float* __restrict__ ax, long long n, long long first,
                                                    long long stride) {
    const long long n4 = n >> 2;
    const float4* p4 = reinterpret_cast<const float4*>(proj);
    const float4* w4 = reinterpret_cast<const float4*>(w);
    float4* a4 = reinterpret_cast<float4*>(ax);
    for (long long i = first; i < n4; i += stride) {
        const float4 p = p4[i], q = w4[i];
        float4 a = a4[i];
        a.x = __fmul_rn(q.x, __fsub_rn(p.x, a.x));
        a.y = __fmul_rn(q.y, __fsub_rn(p.y, a.y));
        a.z = __fmul_rn(q.z, __fsub_rn(p.z, a.z));
        a.w = __fmul_rn(q.w, __fsub_rn(p.w, a.w));
        a4[i] = a;
    }
    for (long long i = (n4 << 2) + first; i < n; i += stride) ax[i] = __fmul_rn(w[i], __fsub_rn(proj[i], ax[i]));
}

__global__ __launch_bounds__(256) void sart_residual_kernel(const float* __restrict__ proj,
                                                            const float* __restrict__ w, float* __restrict__ ax,
                                                            long long n) {
    sart_residual_range(proj, w, ax, n, (long long)blockIdx.x * blockDim.x + threadIdx.x,
                        (long long)gridDim.x * blockDim.x);
}

__global__ __launch_bounds__(256) void sart_residual_multi_kernel(const ResidualJob* __restrict__ jobs, int njobs) {
    const int b = blockIdx.x;
    const int j = find_job(jobs, njobs, b);
    const ResidualJob& J = jobs[j];
    sart_residual_range(J.proj, J.w, J.ax, J.n, (long long)(b - J.begin) * 256 + threadIdx.x,
                        (long long)J.blocks * 256);
}

}  // namespace

int residual_job_blocks(long long n) {
    long long blocks = (n / 4 + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 8192) blocks = 8192;
    return (int)blocks;
}

int launch_residual_multi(const ResidualJob* d_jobs, int njobs, long long blocks, hipStream_t stream) {
    TIGRE_REQUIRE(njobs > 0 && blocks > 0 && blocks < (1LL << 31), "bad multi-job launch");
    hipLaunchKernelGGL(sart_residual_multi_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, d_jobs, njobs);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

int launch_sart_residual(const float* proj, const float* w, float* ax, long long n, hipStream_t stream) {
    TIGRE_REQUIRE(((reinterpret_cast<uintptr_t>(proj) | reinterpret_cast<uintptr_t>(w) |
                    reinterpret_cast<uintptr_t>(ax)) & 15) == 0,
                  "residual buffers must be 16-byte aligned");
    if (n <= 0) return TIGRE_OK;
    const long long blocks = residual_job_blocks(n);
    hipLaunchKernelGGL(sart_residual_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, proj, w, ax, n);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

}  // namespace tigre
