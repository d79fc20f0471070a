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
//   pass 1  g at every voxel, sum of g^2 per workgroup (FP32 per lane over
//           its 32 voxels at most, FP64 across lanes);
//   scale   one workgroup sums the partials in a fixed order (the same result
//           every run) into alpha / sqrt(sum + norm_eps);
//   pass 2  f' = f - scale * g, recomputing g from f.
// Each pass is a 128 x 8 column tile that marches along z through a chunk of
// planes, every lane two columns with their z neighbours in registers and
// their x / y neighbours through LDS, so f is read from HBM about once per pass:
// 12 bytes of traffic per voxel and iteration, no host round trip for the
// norm.
#include <cmath>

#include "common.hpp"
#include "ray.hpp"  // xcd_logical_block

namespace tigre {

namespace {
constexpr int kTY = 8, kZC = 16;            // tile rows (one wavefront each), planes per chunk
constexpr int kTX = 128;                    // tile columns: two adjacent ones per lane
constexpr int kAhead = 4;                   // planes of f in flight per lane
constexpr int kTvThreads = 64 * kTY;
// LDS planes.  f: rows -1 .. kTY, columns -1 .. kTX at index lx + 2 (an even
// row length, so every lane's column pair is 8-byte aligned); n: rows
// 0 .. kTY, columns 0 .. kTX (padded to an even length).
constexpr int kFW = kTX + 4, kFH = kTY + 2;
constexpr int kNW = kTX + 2, kNH = kTY + 1;
constexpr int kFHalo = 2 * (kTX + 2) + 2 * kTY;  // halo entries of an f plane
__device__ __forceinline__ int f_index(int lx, int ly) { return (ly + 1) * kFW + lx + 2; }
__device__ __forceinline__ int n_index(int lx, int ly) { return ly * kNW + lx; }

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

// Tile-relative position of f-plane halo entry e: row -1 (x0 - 1 .. x0 + kTX),
// row kTY, then the columns x0 - 1 and x0 + kTX of rows 0 .. kTY - 1.
__device__ __forceinline__ void f_halo_pos(int e, int& lx, int& ly) {
    if (e < kTX + 2) {
        lx = e - 1;
        ly = -1;
    } else if (e < 2 * (kTX + 2)) {
        lx = e - (kTX + 2) - 1;
        ly = kTY;
    } else {
        const int k = e - 2 * (kTX + 2);
        lx = k < kTY ? -1 : kTX;
        ly = k < kTY ? k : k - kTY;
    }
}

// The normalised difference field at one voxel from its value c, the value
// below it (z - 1) and its neighbours at x - 1 and y - 1 (GD_TV.cu:104-140:
// a missing neighbour's difference is 0).  No masks on the differences: the
// f planes hold, at x = -1 and y = -1, copies of the edge voxels (so those
// differences are exactly 0), and 0 in the zero slices (plane nz, and plane
// -1 below the first); `ok` says whether n exists at all (the column lies in
// the volume and, for PICCS, the plane too).
template <int kKind>
__device__ __forceinline__ void n_field(float c, float below, float cy, float cx, bool ok, bool has_below,
                                        const TvParams& tp, float& n0, float& n1, float& n2) {
    constexpr bool kPadZ = kKind != 2;
    constexpr float eps = kKind == 2 ? 1e-6f : 1e-8f;
    // branch-free: where n does not exist the reciprocal is replaced by 0, so
    // n is +-0 (every operand is finite: f planes hold loaded values or 0)
    const float d0 = kPadZ ? c - below : (has_below ? c - below : 0.0f);
    const float d1 = c - cy;
    const float d2 = c - cx;
    float w0 = 1.0f, w1 = 1.0f, w2 = 1.0f;
    if (kKind == 1) {
        const float a0 = d0 * tp.inv_delta, a1 = d1 * tp.inv_delta, a2 = d2 * tp.inv_delta;
        w0 = __expf(-a0 * a0);
        w1 = __expf(-a1 * a1);
        w2 = __expf(-a2 * a2);
    }
    // hardware square root and reciprocal (1 ulp each)
    const float inv = __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(w0 * d0 * d0 + w1 * d1 * d1 + w2 * d2 * d2) + eps);
    const float s = ok ? inv : 0.0f;
    n0 = w0 * d0 * s;
    n1 = w1 * d1 * s;
    n2 = w2 * d2 * s;
}
}  // namespace

// kKind 0: TV (GD_TV; the image sits between zero slices, so z always has a
// neighbour, GD_TV.cu:367-368,456-465), 1: adaptive-weighted TV (GD_AwTV,
// same padding), 2: PICCS, (1 - ratio) g(f) + ratio g(f - prior) with the z
// term dropped at the edges like x and y (PICCS.cu:109-157,322-347).  At the
// low x / y edge the difference is 0; beyond the high edge n is 0
// (GD_TV.cu:104-114).  kStep false: pass 1 (partials); true: pass 2.
//
// A workgroup is a 128 x 8 column tile, one wavefront per row, marching along
// z through a chunk of planes; every lane owns two adjacent columns.  Its f
// and n values of the previous plane stay in registers (the z neighbours), as
// do the x neighbours between its two columns; the f plane goes through LDS
// for the other x - 1 / y - 1 neighbours and the n_x / n_y planes for the
// x + 1 / y + 1 ones, in 8-byte pairs where both columns need one.  f and n
// planes are double-buffered so one barrier per plane orders everything:
// plane p's f is written before it, read after it; its n is written after it
// and read after the next plane's barrier.  The first lanes also own the halo
// entries of the f plane (prefetched in the same register ring as the own
// columns); the last three wavefronts compute the n-plane halo row y0 + 8 and
// column x0 + 128, one entry per lane (its f of the previous plane kept in a
// register).
template <int kKind, bool kStep>
__global__ __launch_bounds__(kTvThreads) void tv_tile_kernel(const float* __restrict__ f, float* __restrict__ out,
                                                             TvParams tp, int nx, int ny, int nz,
                                                             double* __restrict__ partial,
                                                             const float* __restrict__ scale_in) {
    constexpr int NF = kKind == 2 ? 2 : 1;
    constexpr bool kPadZ = kKind != 2;
    // 16-byte aligned: every lane reads and writes 8-byte pairs (even indices)
    __shared__ __align__(16) float F[2][NF][kFH * kFW];      // f planes (ring of two), halo of one voxel
    __shared__ __align__(16) float Nn[2][2][kNH * kNW];  // (combined) n_y (0) and n_x (1) planes (ring of two)
    const int tid = threadIdx.x, tx = tid & 63, ty = tid >> 6;
    // tiles in x-fastest order, consecutive ones on one XCD: the halos a tile
    // reads from its neighbours are then mostly in that XCD's L2
    const int ntx = (nx + kTX - 1) / kTX, nty = (ny + kTY - 1) / kTY;
    const int L = xcd_logical_block((int)blockIdx.x, (int)gridDim.x);
    const int x0 = (L % ntx) * kTX, y0 = ((L / ntx) % nty) * kTY, z0 = (L / (ntx * nty)) * kZC;
    const int z1 = min(z0 + kZC, nz);
    const long long plane = (long long)nx * ny;
    const float scale = kStep ? scale_in[0] : 0.0f;

    // own columns x0 + 2 tx + c, c = 0, 1
    const int gy = y0 + ty;
    int own_ofs[2];
    bool col[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int gx = x0 + 2 * tx + c;
        col[c] = gx < nx && gy < ny;
        own_ofs[c] = col[c] ? gy * nx + gx : -1;
    }
    const int own_f = f_index(2 * tx, ty), own_n = n_index(2 * tx, ty);  // even: 8-byte aligned pairs
    // halo entry of the f plane (lanes tid < kFHalo); the other lanes store
    // their unused halo value to their own first column just before the own
    // pair, so the store needs no branch (a store under a branch would pull
    // its load into the branch).  Positions outside the volume read the
    // nearest voxel: at x = -1 / y = -1 that makes the edge difference exactly
    // 0 (GD_TV.cu:104-114); beyond the high edges the value only enters n of
    // columns outside the volume, which is 0.
    int hf_idx = own_f, hf_ofs = 0;
    if (tid < kFHalo) {
        int lx, ly;
        f_halo_pos(tid, lx, ly);
        hf_idx = f_index(lx, ly);
        const int hx = min(max(x0 + lx, 0), nx - 1), hy = min(max(y0 + ly, 0), ny - 1);
        hf_ofs = hy * nx + hx;
    }
    // halo of the n plane, one entry per lane: row kTY, column 2 tx (last
    // wavefront) and 2 tx + 1 (the one before), and column kTX, row tx (first
    // kTY lanes of the third-to-last), so no wavefront does more than one extra
    // n evaluation per plane
    int hn_lx = -1, hn_ly = -1;
    if (ty == kTY - 1) {
        hn_lx = 2 * tx;
        hn_ly = kTY;
    } else if (ty == kTY - 2) {
        hn_lx = 2 * tx + 1;
        hn_ly = kTY;
    } else if (ty == kTY - 3 && tx < kTY) {
        hn_lx = kTX;
        hn_ly = tx;
    }
    const bool hn_on = hn_lx >= 0;
    int hn_ofs = -1, hn_f = 0, hn_n = 0;
    bool hn_col = false;
    if (hn_on) {
        const int hx = x0 + hn_lx, hy = y0 + hn_ly;
        hn_col = hx < nx && hy < ny;
        hn_ofs = hn_col ? hy * nx + hx : -1;
        hn_f = f_index(hn_lx, hn_ly);
        hn_n = n_index(hn_lx, hn_ly);
    }

    auto load = [&](int k, long long idx) -> float {
        return k == 0 ? f[idx] : f[idx] - tp.prior[idx];
    };
    // The image (and prior) of plane p at the lane's two columns and its
    // f-halo entry, fetched kAhead planes before use into a ring of registers.
    // The loads are unconditional, from clamped in-bounds addresses, and raw:
    // the plane mask and the PICCS difference are applied when the slot is
    // used, so nothing has to wait for a load before then.  Loads and stores
    // go through a buffer descriptor per plane (a uniform base) with a 32-bit
    // byte offset per lane (the plane is below 2^30 voxels), so no 64-bit
    // address arithmetic runs per access.
    const float* src[2] = {f, tp.prior};
    const unsigned ob0 = (unsigned)max(own_ofs[0], 0) * 4u, ob1 = (unsigned)max(own_ofs[1], 0) * 4u,
                   obh = (unsigned)hf_ofs * 4u;
    const int plane_bytes = (int)(plane * 4);
    auto rsrc = [&](const float* pl, int bytes) {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(pl), (short)0, bytes, 0x00020000);
    };
    auto at = [](__amdgpu_buffer_rsrc_t r, unsigned ob) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, ob, 0, 0));
    };
    float ring[kAhead][2][3];  // [slot][image, prior][column 0, column 1, halo]
    // The zero slice p = nz (the loop never fetches further) is read through
    // a descriptor of no records: every load of it returns 0, so no select
    // runs on the loaded values.
    auto fetch = [&](int p, float (*v)[3]) {
        const long long base = (long long)min(max(p, 0), nz - 1) * plane;
        const int bytes = plane_bytes;
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const __amdgpu_buffer_rsrc_t r = rsrc(src[k] + base, bytes);
            v[k][0] = at(r, ob0);
            v[k][1] = at(r, ob1);
            v[k][2] = at(r, obh);
        }
    };
    // field k of entry e of a fetched slot: f, or f - prior for PICCS
    // (columns outside the volume keep the clamped value: their n is 0)
    auto field = [&](float (*v)[3], int k, int e, bool pin) -> float {
        return pin ? (k == 0 ? v[0][e] : v[0][e] - v[1][e]) : 0.0f;
    };
    // 8-byte pair stores need col[1] and an 8-aligned address: the lane's
    // offset parity is fixed, the plane base's is uniform
    const bool pair_at0 = col[1] && (ob0 & 4u) == 0, pair_at4 = col[1] && (ob0 & 4u) != 0;

    // f of the previous plane: own columns, n-halo entries
    float fprev[NF][2], hprev[NF];
    {
        const int p = z0 - 1;
        const bool pin = p >= 0;
        const long long base = (long long)p * plane;
#pragma unroll
        for (int k = 0; k < NF; ++k) {
#pragma unroll
            for (int c = 0; c < 2; ++c) fprev[k][c] = (pin && own_ofs[c] >= 0) ? load(k, base + own_ofs[c]) : 0.0f;
            hprev[k] = (pin && hn_on && hn_ofs >= 0) ? load(k, base + hn_ofs) : 0.0f;
        }
    }
#pragma unroll
    for (int j = 0; j < kAhead; ++j) fetch(min(z0 + j, z1), ring[j]);
    float nprev[2][3];  // (combined) n of the previous plane at the own columns (z, y, x)
    // sum of g^2 over the lane's (at most 2 x 16) voxels in FP32, then FP64
    // across lanes and workgroups
    float acc = 0.0f;
    // fully unrolled (planes z0 .. z0 + kZC, at most): ring slot jj % kAhead is
    // a fixed register set and the compiler sees every load it must wait for
#pragma unroll
    for (int jj = 0; jj <= kZC; ++jj) {
        const int j = jj % kAhead;
        const int p = z0 + jj;
        if (p > z1) break;
        const int b = p & 1;
        const float fpp[2] = {fprev[0][0], fprev[0][1]};  // f at plane p - 1 (the image, also for PICCS)
        const bool pin = p < nz;
        float fc[NF][2];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            fc[k][0] = field(ring[j], k, 0, pin);
            fc[k][1] = field(ring[j], k, 1, pin);
            F[b][k][hf_idx] = field(ring[j], k, 2, pin);
            *reinterpret_cast<float2*>(&F[b][k][own_f]) = make_float2(fc[k][0], fc[k][1]);
        }
        // unconditional (a branch around the loads would make the compiler
        // wait for all of them); past the chunk it re-reads plane z1 from cache
        fetch(min(p + kAhead, z1), ring[j]);
        __syncthreads();
        // n exists in every plane up to the zero slice nz (TV, AwTV) or in the
        // volume's planes (PICCS); the loop never runs past plane nz
        const bool pex = kPadZ || p < nz;
        const bool has_below = p >= 1;
        const bool ok0 = pex && col[0], ok1 = pex && col[1];
        // n of the own columns and of the halo entries, combined over the
        // fields (PICCS: (1 - ratio) n(f) + ratio n(f - prior); g is linear in
        // n, so one n field goes through LDS; the weight is 1 for one field)
        float ncur[2][3], hn3[3];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const float wk = NF == 1 ? 1.0f : (k == 0 ? 1.0f - tp.ratio : tp.ratio);
            auto mix = [&](float* acc3, const float* t) {
#pragma unroll
                for (int q = 0; q < 3; ++q) acc3[q] = k == 0 ? wk * t[q] : fmaf(wk, t[q], acc3[q]);
            };
            const float2 up = *reinterpret_cast<const float2*>(&F[b][k][own_f - kFW]);  // y - 1 of both columns
            const float left = F[b][k][own_f - 1];                                      // x - 1 of column 0
            float t0[3], t1[3];
            n_field<kKind>(fc[k][0], fprev[k][0], up.x, left, ok0, has_below, tp, t0[0], t0[1], t0[2]);
            n_field<kKind>(fc[k][1], fprev[k][1], up.y, fc[k][0], ok1, has_below, tp, t1[0], t1[1], t1[2]);
            mix(ncur[0], t0);
            mix(ncur[1], t1);
            fprev[k][0] = fc[k][0];
            fprev[k][1] = fc[k][1];
            if (hn_on) {
                const float hc = F[b][k][hn_f];
                float m[3];
                n_field<kKind>(hc, hprev[k], F[b][k][hn_f - kFW], F[b][k][hn_f - 1], pex && hn_col, has_below, tp,
                               m[0], m[1], m[2]);
                mix(hn3, m);
                hprev[k] = hc;
            }
        }
        *reinterpret_cast<float2*>(&Nn[b][0][own_n]) = make_float2(ncur[0][1], ncur[1][1]);
        *reinterpret_cast<float2*>(&Nn[b][1][own_n]) = make_float2(ncur[0][2], ncur[1][2]);
        if (hn_on) {
            Nn[b][0][hn_n] = hn3[1];
            Nn[b][1][hn_n] = hn3[2];
        }
        // g at plane p - 1 from the n planes of p - 1 (buffer b ^ 1, complete
        // since this plane's barrier) and n_z of plane p
        if (p > z0) {
            float res[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                // n_x at x + 1: the lane's own second column, or the next lane's first
                const float nxr = c == 0 ? nprev[1][2] : Nn[b ^ 1][1][own_n + 2];
                const float nyu = Nn[b ^ 1][0][own_n + kNW + c];
                const float gv = nprev[c][0] + nprev[c][1] + nprev[c][2] - nxr - nyu - ncur[c][0];
                res[c] = fpp[c] - gv * scale;
                if (!kStep) acc = fmaf(col[c] ? gv : 0.0f, gv, acc);
            }
            if (kStep) {
                // the pair as one 8-byte store where both columns exist and the
                // address allows it (col[1] implies col[0])
                float* o = out + (long long)(p - 1) * plane;
                const __amdgpu_buffer_rsrc_t r = rsrc(o, plane_bytes);
                if ((reinterpret_cast<unsigned long long>(o) & 4ull) == 0 ? pair_at0 : pair_at4) {
                    __builtin_amdgcn_raw_buffer_store_b64(
                        __builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, make_float2(res[0], res[1])),
                        r, ob0, 0, 0);
                } else {
                    if (col[0]) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(res[0]), r, ob0, 0, 0);
                    if (col[1]) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(res[1]), r, ob1, 0, 0);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int q = 0; q < 3; ++q) nprev[c][q] = ncur[c][q];
    }
    if (!kStep) {
        const double sum = block_sum((double)acc);
        if (tid == 0)
            partial[L] = sum;
    }
}

// scale = alpha / sqrt(sum of partials + norm_eps), summed in a fixed order;
// norm_eps = 1e-7 for TV and AwTV (GD_TV.cu:561), 0 for PICCS (PICCS.cu:370).
__global__ __launch_bounds__(kTvThreads) void tv_scale_kernel(const double* __restrict__ partial, int npartial,
                                                              float alpha, double norm_eps, float* __restrict__ scale) {
    double s = 0.0;
    // unrolled so the loads are issued together; the additions keep their order
#pragma unroll 8
    for (int i = threadIdx.x; i < npartial; i += kTvThreads) s += partial[i];
    s = block_sum(s);
    if (threadIdx.x == 0) scale[0] = (float)((double)alpha / sqrt(s + norm_eps));
}

namespace {
template <int kKind>
int run_tv_descent(const float* img, float* dst, float alpha, const TvParams& tp, double norm_eps, long nx, long ny,
                   long nz, int maxiter, hipStream_t stream) {
    TIGRE_REQUIRE(nx > 0 && ny > 0 && nz > 0, "empty image");
    // 32-bit byte offsets within a plane
    TIGRE_REQUIRE(nx * ny * nz < (1LL << 40) && nx * ny < (1LL << 30), "image too large");
    const size_t n = (size_t)nx * ny * nz;
    if (maxiter <= 0) {
        if (dst != img) TIGRE_HIP_TRY(hipMemcpyAsync(dst, img, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
        return TIGRE_OK;
    }
    // one tile per workgroup in a 1-D grid (the kernel orders the tiles per XCD)
    const long long tiles = ((nx + kTX - 1) / kTX) * ((ny + kTY - 1) / kTY) * ((nz + kZC - 1) / kZC);
    TIGRE_REQUIRE(tiles < (1LL << 31), "image too large");
    const int nblocks = (int)tiles;
    const dim3 grid((unsigned)nblocks);
    StreamScratch scratch(stream);
    float* buf[2] = {nullptr, nullptr};
    double* partial = nullptr;
    float* scale = nullptr;
    TIGRE_CHECK(scratch.alloc(&buf[0], n * sizeof(float)));
    TIGRE_CHECK(scratch.alloc(&buf[1], n * sizeof(float)));
    TIGRE_CHECK(scratch.alloc(&partial, (size_t)nblocks * sizeof(double)));
    TIGRE_CHECK(scratch.alloc(&scale, 256));
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
    StreamScratch scratch(stream);
    float* buf = nullptr;
    TIGRE_CHECK(scratch.alloc(&buf, 5 * n * sizeof(float)));
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
    return TIGRE_OK;
}

}  // namespace tigre
