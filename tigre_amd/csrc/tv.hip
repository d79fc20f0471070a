// TV steps for gfx950: POCS-TV steepest descent (minimizeTV) and the ROF
// primal-dual denoiser.
//
// POCS-TV, GD_TV.cu:489-597 (and its adaptive-weighted and PICCS variants,
// GD_AwTV.cu, PICCS.cu).  One iteration is two passes over the volume:
//   pass 1  g = grad TV(f) at every voxel, sum of g^2 in FP64 per workgroup;
//   pass 2  every workgroup sums the partials in a fixed order (the same
//           result in every workgroup and every run), then writes
//           f' = f - alpha * g / sqrt(sum + 1e-7), recomputing g from f.
// The gradient is never stored: 12 bytes of HBM traffic per voxel and
// iteration instead of the reference's read-gradient-copy-reduce-scale chain,
// and no host round trip for the norm.  The grids are fixed-size grid-stride
// launches so the partial array stays small.
#include <cmath>

#include "common.hpp"

namespace tigre {

namespace {
constexpr int kTvBlocks = 1024;
constexpr int kTvThreads = 256;

// Voxel loaders: the image, or the image minus a prior (PICCS).
struct Plain {
    const float* __restrict__ f;
    __device__ float operator()(long i) const { return f[i]; }
};
struct MinusPrior {
    const float* __restrict__ f;
    const float* __restrict__ p;
    __device__ float operator()(long i) const { return f[i] - p[i]; }
};

// Backward differences at (z, y, x) (GD_TV.cu:95-115): y/x terms dropped at the
// low edge.  kPadZ: the image sits between zero slices (the single-device
// pocs_tv / aw_pocs_tv layout, GD_TV.cu:367-368,456-465), so z always has a
// (possibly zero) neighbour; otherwise (PICCS.cu:109-128) the z term is
// dropped at the low edge like the others.
template <bool kPadZ, class L>
__device__ inline void bdiff(const L& f, long nx, long ny, long nz, long z, long y, long x, float d[3]) {
    const long i = (z * ny + y) * nx + x;
    const bool in = z >= 0 && z < nz;
    const float c = in ? f(i) : 0.0f;
    if (kPadZ)
        d[0] = c - ((z >= 1 && z - 1 < nz) ? f(i - nx * ny) : 0.0f);
    else
        d[0] = (z >= 1) ? c - f(i - nx * ny) : 0.0f;
    d[1] = (y >= 1 && in) ? c - f(i - nx) : 0.0f;
    d[2] = (x >= 1 && in) ? c - f(i - 1) : 0.0f;
}

// (Weighted) TV sub-gradient at one voxel, GD_TV.cu:117-143 / GD_AwTV.cu:117-170
// / PICCS.cu:132-157.  kWeighted: each difference component d carries the
// weight exp(-(d / delta)^2) (AwTV; the reference uses the fast __expf, here
// the accurate expf).
template <bool kPadZ, bool kWeighted, class L>
__device__ inline float tv_grad(const L& f, long nx, long ny, long nz, long z, long y, long x, float eps,
                                float inv_delta) {
    float df[3], di[3] = {0.f, 0.f, 0.f}, dj[3] = {0.f, 0.f, 0.f}, dk[3] = {0.f, 0.f, 0.f};
    bdiff<kPadZ>(f, nx, ny, nz, z, y, x, df);
    if (x + 1 < nx) bdiff<kPadZ>(f, nx, ny, nz, z, y, x + 1, di);
    if (y + 1 < ny) bdiff<kPadZ>(f, nx, ny, nz, z, y + 1, x, dj);
    if (kPadZ || z + 1 < nz) bdiff<kPadZ>(f, nx, ny, nz, z + 1, y, x, dk);
    auto w = [&](float d) { return kWeighted ? expf(-(d * inv_delta) * (d * inv_delta)) : 1.0f; };
    const float w0 = w(df[0]), w1 = w(df[1]), w2 = w(df[2]);
    const float wi0 = w(di[0]), wi1 = w(di[1]), wi2 = w(di[2]);
    const float wj0 = w(dj[0]), wj1 = w(dj[1]), wj2 = w(dj[2]);
    const float wk0 = w(dk[0]), wk1 = w(dk[1]), wk2 = w(dk[2]);
    float g = (w0 * df[0] + w1 * df[1] + w2 * df[2]) /
              (sqrtf(w0 * df[0] * df[0] + w1 * df[1] * df[1] + w2 * df[2] * df[2]) + eps);
    g -= wi2 * di[2] / (sqrtf(wi0 * di[0] * di[0] + wi1 * di[1] * di[1] + wi2 * di[2] * di[2]) + eps);
    g -= wj1 * dj[1] / (sqrtf(wj0 * dj[0] * dj[0] + wj1 * dj[1] * dj[1] + wj2 * dj[2] * dj[2]) + eps);
    g -= wk0 * dk[0] / (sqrtf(wk0 * dk[0] * dk[0] + wk1 * dk[1] * dk[1] + wk2 * dk[2] * dk[2]) + eps);
    return g;
}

// Descent direction of one iteration at one voxel.  kKind 0: TV (GD_TV),
// 1: adaptive-weighted TV (GD_AwTV), 2: PICCS, (1 - ratio) grad TV(f) + ratio
// grad TV(f - prior) (PICCS.cu:322-347).
struct TvParams {
    const float* prior;
    float inv_delta, ratio;
};

template <int kKind>
__device__ inline float tv_direction(const float* __restrict__ f, const TvParams& tp, long nx, long ny, long nz,
                                     long z, long y, long x) {
    if (kKind == 0) return tv_grad<true, false>(Plain{f}, nx, ny, nz, z, y, x, 1e-8f, 0.0f);
    if (kKind == 1) return tv_grad<true, true>(Plain{f}, nx, ny, nz, z, y, x, 1e-8f, tp.inv_delta);
    const float a = tv_grad<false, false>(Plain{f}, nx, ny, nz, z, y, x, 1e-6f, 0.0f);
    const float b = tv_grad<false, false>(MinusPrior{f, tp.prior}, nx, ny, nz, z, y, x, 1e-6f, 0.0f);
    return (1.0f - tp.ratio) * a + tp.ratio * b;
}

// Workgroup sum of one double per lane (wavefront shuffles, then LDS).
__device__ inline double block_sum(double v) {
    __shared__ double part[kTvThreads / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kTvThreads / 64; ++i) s += part[i];
    __syncthreads();
    return s;  // valid in thread 0
}
}  // namespace

template <int kKind>
__global__ __launch_bounds__(kTvThreads) void tv_norm_kernel(const float* __restrict__ f, TvParams tp, long nx, long ny,
                                                             long nz, double* __restrict__ partial) {
    const long n = nx * ny * nz;
    double s = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const long x = q % nx, y = (q / nx) % ny, z = q / (nx * ny);
        const float g = tv_direction<kKind>(f, tp, nx, ny, nz, z, y, x);
        s += (double)g * (double)g;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// f' = f - alpha g / sqrt(sum g^2 + norm_eps); norm_eps = 1e-7 for TV and AwTV
// (GD_TV.cu:561), 0 for PICCS (PICCS.cu:370).
template <int kKind>
__global__ __launch_bounds__(kTvThreads) void tv_step_kernel(const float* __restrict__ f, float* __restrict__ out,
                                                             TvParams tp, long nx, long ny, long nz,
                                                             const double* __restrict__ partial, int npartial,
                                                             float alpha, double norm_eps) {
    __shared__ float s_scale;
    double s = 0.0;
    for (int i = threadIdx.x; i < npartial; i += blockDim.x) s += partial[i];
    s = block_sum(s);
    if (threadIdx.x == 0) s_scale = (float)((double)alpha / sqrt(s + norm_eps));
    __syncthreads();
    const float scale = s_scale;
    const long n = nx * ny * nz;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const long x = q % nx, y = (q / nx) % ny, z = q / (nx * ny);
        out[q] = f[q] - tv_direction<kKind>(f, tp, nx, ny, nz, z, y, x) * scale;
    }
}

namespace {
template <int kKind>
int run_tv_descent(const float* img, float* dst, float alpha, const TvParams& tp, double norm_eps, long nx, long ny,
                   long nz, int maxiter, hipStream_t stream) {
    const size_t n = (size_t)nx * ny * nz;
    if (maxiter <= 0) {
        if (dst != img) TIGRE_HIP_TRY(hipMemcpyAsync(dst, img, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
        return TIGRE_OK;
    }
    float* buf[2] = {nullptr, nullptr};
    double* partial = nullptr;
    TIGRE_HIP_TRY(hipMallocAsync((void**)&buf[0], n * sizeof(float), stream));
    TIGRE_HIP_TRY(hipMallocAsync((void**)&buf[1], n * sizeof(float), stream));
    TIGRE_HIP_TRY(hipMallocAsync((void**)&partial, kTvBlocks * sizeof(double), stream));
    const float* src = img;
    int cur = 0;
    for (int it = 0; it < maxiter; ++it) {
        float* o = (it == maxiter - 1 && dst != img) ? dst : buf[cur];
        KernelTimerScope timer(TIGRE_KERNEL_POCS_TV, stream);
        hipLaunchKernelGGL(tv_norm_kernel<kKind>, dim3(kTvBlocks), dim3(kTvThreads), 0, stream, src, tp, nx, ny, nz,
                           partial);
        hipLaunchKernelGGL(tv_step_kernel<kKind>, dim3(kTvBlocks), dim3(kTvThreads), 0, stream, src, o, tp, nx, ny,
                           nz, partial, kTvBlocks, alpha, norm_eps);
        TIGRE_HIP_TRY(hipGetLastError());
        src = o;
        cur ^= 1;
    }
    if (src != dst) TIGRE_HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
    TIGRE_HIP_TRY(hipFreeAsync(buf[0], stream));
    TIGRE_HIP_TRY(hipFreeAsync(buf[1], stream));
    TIGRE_HIP_TRY(hipFreeAsync(partial, stream));
    return TIGRE_OK;
}
}  // namespace

int run_pocs_tv(const float* img, float* dst, float alpha, long nx, long ny, long nz, int maxiter,
                hipStream_t stream) {
    return run_tv_descent<0>(img, dst, alpha, TvParams{nullptr, 0.0f, 0.0f}, 1e-7, nx, ny, nz, maxiter, stream);
}

int run_aw_pocs_tv(const float* img, float* dst, float alpha, float delta, long nx, long ny, long nz, int maxiter,
                   hipStream_t stream) {
    TIGRE_REQUIRE(delta != 0.0f, "delta must be nonzero");
    return run_tv_descent<1>(img, dst, alpha, TvParams{nullptr, 1.0f / delta, 0.0f}, 1e-7, nx, ny, nz, maxiter,
                             stream);
}

int run_piccs_tv(const float* img, const float* prior, float* dst, float alpha, float ratio, long nx, long ny, long nz,
                 int maxiter, hipStream_t stream) {
    return run_tv_descent<2>(img, dst, alpha, TvParams{prior, 0.0f, ratio}, 0.0, nx, ny, nz, maxiter, stream);
}

// ---------------------------------------------------------------------------
// ROF primal-dual denoise, tv_proximal.cu:81-181 (kernels) and :436-502 (loop).
// Domain: nz + 2 slices; slice 0 mirrors image slice 1, slices 1..nz hold the
// image, slice nz + 1 is zero (the single-device layout of cpy_from_host,
// tv_proximal.cu:672-693).  All slices evolve; slices 1..nz are returned.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void tvd_init_kernel(const float* __restrict__ src, long nx, long ny, long nz,
                                                       float* __restrict__ f, float* __restrict__ u,
                                                       float* __restrict__ px, float* __restrict__ py,
                                                       float* __restrict__ pz) {
    const long sl = nx * ny, D = nz + 2, n = sl * D;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const long z = q / sl, r = q % sl;
        float v = 0.0f;
        if (z == 0) v = (nz > 1) ? src[sl + r] : 0.0f;
        else if (z <= nz) v = src[(z - 1) * sl + r];
        f[q] = v;
        u[q] = v;
        px[q] = 0.0f;
        py[q] = 0.0f;
        pz[q] = 0.0f;
    }
}

__global__ __launch_bounds__(256) void tvd_update_u_kernel(const float* __restrict__ f, const float* __restrict__ pz,
                                                           const float* __restrict__ py, const float* __restrict__ px,
                                                           float* __restrict__ u, float tau, float inv_lambda, long D,
                                                           long ny, long nx, float dz, float dy, float dx) {
    const long sl = nx * ny, n = sl * D;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const long x = q % nx, y = (q / nx) % ny, z = q / sl;
        float dv = 0.0f;
        dv += (z >= 1) ? (pz[q] - pz[q - sl]) / dz : pz[q];
        dv += (y >= 1) ? (py[q] - py[q - nx]) / dy : py[q];
        dv += (x >= 1) ? (px[q] - px[q - 1]) / dx : px[q];
        u[q] = u[q] * (1.0f - tau) + tau * (f[q] + inv_lambda * dv);
    }
}

__global__ __launch_bounds__(256) void tvd_update_p_kernel(const float* __restrict__ u, float* __restrict__ pz,
                                                           float* __restrict__ py, float* __restrict__ px, float tau,
                                                           long D, long ny, long nx, float dz, float dy, float dx) {
    const long sl = nx * ny, n = sl * D;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const long x = q % nx, y = (q / nx) % ny, z = q / sl;
        const float c = u[q];
        const float gz = (z + 1 < D) ? (u[q + sl] - c) / dz : 0.0f;
        const float gy = (y + 1 < ny) ? (u[q + nx] - c) / dy : 0.0f;
        const float gx = (x + 1 < nx) ? (u[q + 1] - c) / dx : 0.0f;
        const float qz = pz[q] + tau * gz, qy = py[q] + tau * gy, qx = px[q] + tau * gx;
        const float nrm = fmaxf(1.0f, sqrtf(qz * qz + qy * qy + qx * qx));
        pz[q] = qz / nrm;
        py[q] = qy / nrm;
        px[q] = qx / nrm;
    }
}

int run_tvdenoise(const float* src, float* dst, float lambda, const float* spacing, long nx, long ny, long nz,
                  int maxiter, hipStream_t stream) {
    const long D = nz + 2, sl = nx * ny;
    const size_t n = (size_t)D * sl;
    float* buf = nullptr;
    TIGRE_HIP_TRY(hipMallocAsync((void**)&buf, 5 * n * sizeof(float), stream));
    float *f = buf, *u = buf + n, *px = buf + 2 * n, *py = buf + 3 * n, *pz = buf + 4 * n;
    long blocks = ((long)n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(tvd_init_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, src, nx, ny, nz, f, u, px, py,
                       pz);
    const float dx = spacing[0], dy = spacing[1], dz = spacing[2];
    for (int it = 0; it < maxiter; ++it) {
        // step sizes of tv_proximal.cu:478-479, evaluated in float as there
        const float tau2 = 0.3f + 0.02f * (float)it;
        const float tau1 = (1.f / tau2) * ((1.f / 6.f) - (5.f / (15.f + (float)it)));
        KernelTimerScope timer(TIGRE_KERNEL_TVDENOISE, stream);
        hipLaunchKernelGGL(tvd_update_u_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, f, pz, py, px, u, tau1,
                           1.0f / lambda, D, ny, nx, dz, dy, dx);
        hipLaunchKernelGGL(tvd_update_p_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, u, pz, py, px, tau2, D,
                           ny, nx, dz, dy, dx);
    }
    TIGRE_HIP_TRY(hipGetLastError());
    TIGRE_HIP_TRY(hipMemcpyAsync(dst, u + sl, (size_t)nz * sl * sizeof(float), hipMemcpyDeviceToDevice, stream));
    TIGRE_HIP_TRY(hipFreeAsync(buf, stream));
    return TIGRE_OK;
}

}  // namespace tigre
