// Rounding of float results to the storage type of the caller's buffers.
#include "common.hpp"

namespace tigre {

namespace {

// float -> bfloat16, round to nearest even; NaN stays a (quiet) NaN.
__device__ __forceinline__ unsigned short to_bf16(float x) {
    const unsigned u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (unsigned short)((u >> 16) | 0x0040u);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__global__ __launch_bounds__(256) void store_converted_kernel(const float* __restrict__ src, void* __restrict__ dst,
                                                              long long n, int dtype) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float x = src[i];
        if (dtype == TIGRE_DTYPE_BF16)
            static_cast<unsigned short*>(dst)[i] = to_bf16(x);
        else
            static_cast<_Float16*>(dst)[i] = (_Float16)x;  // v_cvt_f16_f32: round to nearest even
    }
}

}  // namespace

int launch_store_converted(const float* src, void* dst, long long n, int dtype, hipStream_t stream) {
    TIGRE_REQUIRE(dtype == TIGRE_DTYPE_BF16 || dtype == TIGRE_DTYPE_F16, "conversion target must be bf16 or f16");
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) return TIGRE_OK;
    KernelTimerScope timer(TIGRE_KERNEL_RELAYOUT, stream);
    hipLaunchKernelGGL(store_converted_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, src, dst, n, dtype);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

}  // namespace tigre
