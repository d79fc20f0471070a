// Elementwise step of the SART family on device buffers.
//
// r = W * (p - Ax): the weighted data residual of one subset
// (iterative_recon_alg.py:326-345, art_data_minimizing), written over the
// projection buffer.  The two rounding steps are kept apart (no fused
// multiply-add), so the result equals the unfused tensor expression bit for bit.
#include "common.hpp"

namespace tigre {

namespace {

__global__ __launch_bounds__(256) void sart_residual_kernel(const float* __restrict__ proj,
                                                            const float* __restrict__ w, float* __restrict__ ax,
                                                            long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long n4 = n >> 2;
    const float4* p4 = reinterpret_cast<const float4*>(proj);
    const float4* w4 = reinterpret_cast<const float4*>(w);
    float4* a4 = reinterpret_cast<float4*>(ax);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        const float4 p = p4[i], q = w4[i];
        float4 a = a4[i];
        a.x = __fmul_rn(q.x, __fsub_rn(p.x, a.x));
        a.y = __fmul_rn(q.y, __fsub_rn(p.y, a.y));
        a.z = __fmul_rn(q.z, __fsub_rn(p.z, a.z));
        a.w = __fmul_rn(q.w, __fsub_rn(p.w, a.w));
        a4[i] = a;
    }
    for (long long i = (n4 << 2) + (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        ax[i] = __fmul_rn(w[i], __fsub_rn(proj[i], ax[i]));
}

}  // namespace

int launch_sart_residual(const float* proj, const float* w, float* ax, long long n, hipStream_t stream) {
    TIGRE_REQUIRE(((reinterpret_cast<uintptr_t>(proj) | reinterpret_cast<uintptr_t>(w) |
                    reinterpret_cast<uintptr_t>(ax)) & 15) == 0,
                  "residual buffers must be 16-byte aligned");
    if (n <= 0) return TIGRE_OK;
    long long blocks = (n / 4 + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(sart_residual_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, proj, w, ax, n);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

}  // namespace tigre
