// TV steps for gfx950: POCS-TV steepest descent (minimizeTV) and the ROF
// primal-dual denoiser.
//
// POCS-TV, GD_TV.cu:489-597 (and its adaptive-weighted and PICCS variants,
// GD_AwTV.cu, PICCS.cu).  The TV sub-gradient of GD_TV.cu:117-143 at voxel v
// is the divergence of the normalised backward-difference field
//     n_c(v) = w_c d_c(v) / (sqrt(sum_k w_k d_k(v)^2) + eps),
//     g(v)   = n_x(v) + n_y(v) + n_z(v) - n_x(v + x) - n_y(v + y) - n_z(v + z),
// (w = 1 for TV, exp(-(d / delta)^2) per component for AwTV, with the fast
// exponential as the reference's __expf), so each n is
// computed once per voxel instead of four times.  One iteration is two passes:
//   pass 1  g at every voxel, sum of g^2 in FP64 per workgroup;
//   scale   one workgroup sums the partials in a fixed order (the same result
//           every run) into alpha / sqrt(sum + norm_eps);
//   pass 2  f' = f - scale * g, recomputing g from f.
// Each pass is a 64 x 8 column tile that marches along z through a chunk of
// planes: the f planes (with a one-voxel halo) and two planes of n live in
// LDS, so f is read from HBM about once per pass: 12 bytes of traffic per
// voxel and iteration, no host round trip for the norm.
#include <cmath>

#include "common.hpp"

namespace tigre {

namespace {
constexpr int kTX = 64, kTY = 8, kZC = 16;  // tile columns, rows, planes per chunk
constexpr int kTvThreads = 256;

struct TvParams {
    const float* prior;  // PICCS prior image
    float inv_delta;     // AwTV 1 / delta
    float ratio;         // PICCS prior ratio
};

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

// kKind 0: TV (GD_TV; the image sits between zero slices, so z always has a
// neighbour, GD_TV.cu:367-368,456-465), 1: adaptive-weighted TV (GD_AwTV,
// same padding), 2: PICCS, (1 - ratio) g(f) + ratio g(f - prior) with the z
// term dropped at the edges like x and y (PICCS.cu:109-157,322-347).  At the
// low x / y edge the difference is 0; beyond the high edge n is 0
// (GD_TV.cu:104-114).  kStep false: pass 1 (partials); true: pass 2.
template <int kKind, bool kStep>
__global__ __launch_bounds__(kTvThreads) void tv_tile_kernel(const float* __restrict__ f, float* __restrict__ out,
                                                             TvParams tp, int nx, int ny, int nz,
                                                             double* __restrict__ partial,
                                                             const float* __restrict__ scale_in) {
    constexpr int NF = kKind == 2 ? 2 : 1;
    constexpr bool kPadZ = kKind != 2;
    constexpr float eps = kKind == 2 ? 1e-6f : 1e-8f;
    constexpr int FW = kTX + 2, FH = kTY + 2, NW = kTX + 1, NH = kTY + 1;
    __shared__ float F[NF][2][FH * FW];      // f planes p - 1, p (ring), halo of one voxel
    __shared__ float Nn[NF][2][3][NH * NW];  // n planes p - 1, p (ring)
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * kTX, y0 = blockIdx.y * kTY, z0 = blockIdx.z * kZC;
    const int z1 = min(z0 + kZC, nz);
    const long long plane = (long long)nx * ny;
    const float scale = kStep ? scale_in[0] : 0.0f;

    // This thread's entries of the f tile (with halo) and of the n tile, and
    // their edge tests, are fixed across the planes of the march: set up once.
    constexpr int kPerF = (FH * FW + kTvThreads - 1) / kTvThreads;
    constexpr int kPerN = (NH * NW + kTvThreads - 1) / kTvThreads;
    int fofs[kPerF];  // in-plane offset of the entry, or -1 outside the tile / volume
#pragma unroll
    for (int j = 0; j < kPerF; ++j) {
        const int i = tid + j * kTvThreads;
        const int ly = i / FW, lx = i - ly * FW;
        const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
        fofs[j] = (i < FH * FW && gx >= 0 && gx < nx && gy >= 0 && gy < ny) ? gy * nx + gx : -1;
    }
    // n entries: bit 0 the voxel's column exists, bit 1 a y neighbour below
    // (gy >= 1), bit 2 an x neighbour (gx >= 1); -1 outside the tile
    int nflag[kPerN];
#pragma unroll
    for (int j = 0; j < kPerN; ++j) {
        const int i = tid + j * kTvThreads;
        const int ly = i / NW, lx = i - ly * NW;
        const int gx = x0 + lx, gy = y0 + ly;
        nflag[j] = i < NH * NW ? ((gx < nx && gy < ny) ? 1 : 0) | (gy >= 1 ? 2 : 0) | (gx >= 1 ? 4 : 0) : -1;
    }

    // A plane (with halo) is fetched into registers one step before it is
    // stored to LDS, so its loads are in flight during the previous plane's
    // arithmetic.
    float pre[NF][kPerF];
    auto fetch_plane = [&](int p) {
        const bool pin = p >= 0 && p < nz;
        const long long base = (long long)p * plane;
#pragma unroll
        for (int j = 0; j < kPerF; ++j) {
            const bool ok = pin && fofs[j] >= 0;
            pre[0][j] = ok ? f[base + fofs[j]] : 0.0f;
            if (NF == 2) pre[NF - 1][j] = ok ? tp.prior[base + fofs[j]] : 0.0f;
        }
    };
    auto store_plane = [&](int slot) {
#pragma unroll
        for (int j = 0; j < kPerF; ++j) {
            const int i = tid + j * kTvThreads;
            if (i >= FH * FW) continue;
            F[0][slot][i] = pre[0][j];
            if (NF == 2) F[NF - 1][slot][i] = pre[0][j] - pre[NF - 1][j];
        }
    };
    auto compute_n = [&](int p, int fs, int ps, int ns) {
        const bool pex = p >= 0 && (kPadZ ? p <= nz : p < nz);
        const bool in = p < nz, has_below = p >= 1;
#pragma unroll
        for (int j = 0; j < kPerN; ++j) {
            if (nflag[j] < 0) continue;
            const int i = tid + j * kTvThreads;
            const int ly = i / NW, lx = i - ly * NW;
            const int ci = (ly + 1) * FW + lx + 1;  // the voxel in the f tile
#pragma unroll
            for (int k = 0; k < NF; ++k) {
                float n0 = 0.0f, n1 = 0.0f, n2 = 0.0f;
                if (pex && (nflag[j] & 1)) {
                    const float c = in ? F[k][fs][ci] : 0.0f;
                    const float below = has_below ? F[k][ps][ci] : 0.0f;
                    const float d0 = kPadZ ? c - below : (has_below ? c - below : 0.0f);
                    const float d1 = ((nflag[j] & 2) && in) ? c - F[k][fs][ci - FW] : 0.0f;
                    const float d2 = ((nflag[j] & 4) && in) ? c - F[k][fs][ci - 1] : 0.0f;
                    float w0 = 1.0f, w1 = 1.0f, w2 = 1.0f;
                    if (kKind == 1) {
                        const float a0 = d0 * tp.inv_delta, a1 = d1 * tp.inv_delta, a2 = d2 * tp.inv_delta;
                        w0 = __expf(-a0 * a0);
                        w1 = __expf(-a1 * a1);
                        w2 = __expf(-a2 * a2);
                    }
                    // hardware square root and reciprocal (1 ulp each)
                    const float inv = __builtin_amdgcn_rcpf(
                        __builtin_amdgcn_sqrtf(w0 * d0 * d0 + w1 * d1 * d1 + w2 * d2 * d2) + eps);
                    n0 = w0 * d0 * inv;
                    n1 = w1 * d1 * inv;
                    n2 = w2 * d2 * inv;
                }
                Nn[k][ns][0][i] = n0;
                Nn[k][ns][1][i] = n1;
                Nn[k][ns][2][i] = n2;
            }
        }
    };

    // g: two rows of the tile per thread, column tid % 64
    const int tx = tid & 63;
    constexpr int kRows = kTY / (kTvThreads / 64);
    int gofs[kRows];  // in-plane offset, or -1 outside the volume
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
        const int ty = (tid >> 6) + r * (kTvThreads / 64);
        const int gx = x0 + tx, gy = y0 + ty;
        gofs[r] = (gx < nx && gy < ny) ? gy * nx + gx : -1;
    }

    double acc = 0.0;
    fetch_plane(z0 - 1);
    store_plane(0);
    fetch_plane(z0);
    store_plane(1);
    fetch_plane(z0 + 1);
    __syncthreads();
    compute_n(z0, 1, 0, 0);
    __syncthreads();
    for (int p = z0 + 1; p <= z1; ++p) {
        const int fs = (p - z0 + 1) & 1, ns = (p - z0) & 1;  // slots of plane p
        store_plane(fs);
        if (p < z1) fetch_plane(p + 1);
        __syncthreads();
        compute_n(p, fs, fs ^ 1, ns);
        __syncthreads();
        // g at plane p - 1
        const int no = ns ^ 1;
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
            if (gofs[r] < 0) continue;
            const int ty = (tid >> 6) + r * (kTvThreads / 64);
            const int ni = ty * NW + tx;
            float g[NF];
#pragma unroll
            for (int k = 0; k < NF; ++k)
                g[k] = Nn[k][no][0][ni] + Nn[k][no][1][ni] + Nn[k][no][2][ni] - Nn[k][no][2][ni + 1] -
                       Nn[k][no][1][ni + NW] - Nn[k][ns][0][ni];
            const float gv = kKind == 2 ? (1.0f - tp.ratio) * g[0] + tp.ratio * g[NF - 1] : g[0];
            if (kStep) {
                out[(long long)(p - 1) * plane + gofs[r]] = F[0][fs ^ 1][(ty + 1) * FW + tx + 1] - gv * scale;
            } else {
                acc += (double)gv * (double)gv;
            }
        }
        __syncthreads();  // the next plane overwrites slot fs ^ 1 and n slot ns ^ 1
    }
    if (!kStep) {
        acc = block_sum(acc);
        if (tid == 0)
            partial[((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = acc;
    }
}

// scale = alpha / sqrt(sum of partials + norm_eps), summed in a fixed order;
// norm_eps = 1e-7 for TV and AwTV (GD_TV.cu:561), 0 for PICCS (PICCS.cu:370).
__global__ __launch_bounds__(kTvThreads) void tv_scale_kernel(const double* __restrict__ partial, int npartial,
                                                              float alpha, double norm_eps, float* __restrict__ scale) {
    double s = 0.0;
    for (int i = threadIdx.x; i < npartial; i += kTvThreads) s += partial[i];
    s = block_sum(s);
    if (threadIdx.x == 0) scale[0] = (float)((double)alpha / sqrt(s + norm_eps));
}

namespace {
template <int kKind>
int run_tv_descent(const float* img, float* dst, float alpha, const TvParams& tp, double norm_eps, long nx, long ny,
                   long nz, int maxiter, hipStream_t stream) {
    TIGRE_REQUIRE(nx > 0 && ny > 0 && nz > 0, "empty image");
    TIGRE_REQUIRE(nx * ny * nz < (1LL << 40) && nx * ny < (1LL << 31) && ny < 65535L * kTY && nx < 65535L * kTX,
                  "image too large");
    const size_t n = (size_t)nx * ny * nz;
    if (maxiter <= 0) {
        if (dst != img) TIGRE_HIP_TRY(hipMemcpyAsync(dst, img, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
        return TIGRE_OK;
    }
    const dim3 grid((unsigned)((nx + kTX - 1) / kTX), (unsigned)((ny + kTY - 1) / kTY),
                    (unsigned)((nz + kZC - 1) / kZC));
    const int nblocks = (int)(grid.x * grid.y * grid.z);
    float* buf[2] = {nullptr, nullptr};
    double* partial = nullptr;
    float* scale = nullptr;
    TIGRE_HIP_TRY(hipMallocAsync((void**)&buf[0], n * sizeof(float), stream));
    TIGRE_HIP_TRY(hipMallocAsync((void**)&buf[1], n * sizeof(float), stream));
    TIGRE_HIP_TRY(hipMallocAsync((void**)&partial, (size_t)nblocks * sizeof(double), stream));
    TIGRE_HIP_TRY(hipMallocAsync((void**)&scale, 256, stream));
    const float* src = img;
    int cur = 0;
    for (int it = 0; it < maxiter; ++it) {
        float* o = (it == maxiter - 1 && dst != img) ? dst : buf[cur];
        KernelTimerScope timer(TIGRE_KERNEL_POCS_TV, stream);
        hipLaunchKernelGGL((tv_tile_kernel<kKind, false>), grid, dim3(kTvThreads), 0, stream, src, nullptr, tp,
                           (int)nx, (int)ny, (int)nz, partial, nullptr);
        hipLaunchKernelGGL(tv_scale_kernel, dim3(1), dim3(kTvThreads), 0, stream, partial, nblocks, alpha, norm_eps,
                           scale);
        hipLaunchKernelGGL((tv_tile_kernel<kKind, true>), grid, dim3(kTvThreads), 0, stream, src, o, tp, (int)nx,
                           (int)ny, (int)nz, nullptr, scale);
        TIGRE_HIP_TRY(hipGetLastError());
        src = o;
        cur ^= 1;
    }
    if (src != dst) TIGRE_HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
    TIGRE_HIP_TRY(hipFreeAsync(buf[0], stream));
    TIGRE_HIP_TRY(hipFreeAsync(buf[1], stream));
    TIGRE_HIP_TRY(hipFreeAsync(partial, stream));
    TIGRE_HIP_TRY(hipFreeAsync(scale, stream));
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
