// Issue-rate probe for counter calibration: a kernel bound by VALU issue
// alone, with a known instruction count.  Every lane runs `iters` rounds of
// eight independent FMA chains (eight v_fma_f32 per round, no memory access in
// the loop), so with several wavefronts per SIMD nothing but the VALU issue
// rate limits it.  Its duration prices the operator kernels' VALU instruction
// counts (SQ_INSTS_VALU) against what the hardware issues per second.
#include "common.hpp"

namespace tigre {

namespace {

// kMode 0: all 64 lanes active; 1: lanes 0..15 only (one 16-lane group of the
// wavefront, the rest masked off); 2: four independent integer chains of add,
// compare and select (12 instructions per round); 3: four FMA chains interleaved
// with a ring of four integer adds (8 instructions per round: 4 v_fma_f32,
// 3 v_add_u32, 1 v_xad_u32).
template <int kMode>
__global__ __launch_bounds__(256) void valu_probe_kernel(float* __restrict__ out, int iters, float b, float c) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    float a0 = (float)threadIdx.x, a1 = a0 + 1.0f, a2 = a0 + 2.0f, a3 = a0 + 3.0f;
    float a4 = a0 + 4.0f, a5 = a0 + 5.0f, a6 = a0 + 6.0f, a7 = a0 + 7.0f;
    if (kMode != 1 || (threadIdx.x & 63) < 16) {
        if (kMode == 2) {
            unsigned u0 = threadIdx.x, u1 = u0 + 1, u2 = u0 + 2, u3 = u0 + 3;
            const unsigned k = __float_as_uint(b), h = __float_as_uint(c);
            for (int i = 0; i < iters; ++i) {
                u0 += k;
                u1 += k;
                u2 += k;
                u3 += k;
                u0 = u0 > h ? u0 : k;
                u1 = u1 > h ? u1 : k;
                u2 = u2 > h ? u2 : k;
                u3 = u3 > h ? u3 : k;
            }
            a0 = __uint_as_float(u0 ^ u1 ^ u2 ^ u3);
        } else if (kMode == 3) {
            unsigned u0 = threadIdx.x, u1 = u0 + 1, u2 = u0 + 2, u3 = u0 + 3;
            const unsigned k = __float_as_uint(b);
            for (int i = 0; i < iters; ++i) {
                a0 = fmaf(a0, b, c);
                u0 += u1;
                a1 = fmaf(a1, b, c);
                u1 += u2;
                a2 = fmaf(a2, b, c);
                u2 += u3;
                a3 = fmaf(a3, b, c);
                u3 += u0 ^ k;
            }
            a4 = __uint_as_float(u0 ^ u1 ^ u2 ^ u3);
        } else {
            for (int i = 0; i < iters; ++i) {
                a0 = fmaf(a0, b, c);
                a1 = fmaf(a1, b, c);
                a2 = fmaf(a2, b, c);
                a3 = fmaf(a3, b, c);
                a4 = fmaf(a4, b, c);
                a5 = fmaf(a5, b, c);
                a6 = fmaf(a6, b, c);
                a7 = fmaf(a7, b, c);
            }
        }
    }
    out[id] = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
}

}  // namespace

}  // namespace tigre

extern "C" int tigre_valu_probe(float* d_out, int blocks, int iters, int mode, void* stream) {
    using namespace tigre;
    set_error("");
    TIGRE_REQUIRE(d_out != nullptr && blocks > 0 && iters > 0, "bad probe arguments");
    TIGRE_REQUIRE(mode >= 0 && mode <= 3, "unknown probe mode");
    KernelTimerScope timer(TIGRE_KERNEL_PROBE, (hipStream_t)stream);
    const hipStream_t s = (hipStream_t)stream;
    if (mode == 0)
        hipLaunchKernelGGL(valu_probe_kernel<0>, dim3((unsigned)blocks), dim3(256), 0, s, d_out, iters, 0.999f, 1e-3f);
    else if (mode == 1)
        hipLaunchKernelGGL(valu_probe_kernel<1>, dim3((unsigned)blocks), dim3(256), 0, s, d_out, iters, 0.999f, 1e-3f);
    else if (mode == 3)
        hipLaunchKernelGGL(valu_probe_kernel<3>, dim3((unsigned)blocks), dim3(256), 0, s, d_out, iters, 0.999f, 1e-3f);
    else
        hipLaunchKernelGGL(valu_probe_kernel<2>, dim3((unsigned)blocks), dim3(256), 0, s, d_out, iters, 0.999f, 1e-3f);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}
