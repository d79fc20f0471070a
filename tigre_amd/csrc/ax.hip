// Cone-beam forward projectors for gfx950.
//
// Ray-driven, one ray per lane.  A 256-thread workgroup owns a 16 x 16
// detector tile of one view and each 64-lane wavefront an 8 x 8 sub-tile, so
// the rays of a wavefront are nearly parallel and walk the same voxels
// (coherent gathers, shared cache lines).  Per-ray set-up (frame, entry and
// exit, first crossings) runs in FP64 -- on CDNA4 FP64 FMA has the same vector
// rate as non-packed FP32 -- and the inner loop runs in FP32.
//
// The volume is read from re-laid-out, zero-bordered copies (one voxel below,
// two above on every axis), which replace the reference's border-addressed
// texture: taps that fall outside the volume read zeros without bounds checks.
//  - Siddon: 4^3 bricks (256 bytes), so the nearly parallel rays of a
//    wavefront share bricks whatever their direction.
//  - interpolated: 5^3 "apron" bricks, so the 8 trilinear taps share one base.
#include <cmath>

#include "common.hpp"
#include "ray.hpp"

namespace tigre {

namespace {

// Lane -> (view, detector pixel).  A wavefront owns an 8 x 8 pixel tile of
// one view; the four wavefronts of a workgroup take the same tile in four
// consecutive views, whose rays cross nearly the same voxels.
constexpr int kTileEdge = 8;
constexpr int kViewsPerBlock = 4;

struct RayCoords {
    int u, v, view;
};

__device__ inline RayCoords ray_of_lane(int nU, int nV, int nviews) {
    const int tilesU = (nU + kTileEdge - 1) / kTileEdge;
    const int tilesV = (nV + kTileEdge - 1) / kTileEdge;
    const int L = xcd_logical_block(blockIdx.x, gridDim.x);
    const int tile = L % (tilesU * tilesV), vgroup = L / (tilesU * tilesV);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    RayCoords c;
    c.u = (tile % tilesU) * kTileEdge + (lane & 7);
    c.v = (tile / tilesU) * kTileEdge + (lane >> 3);
    c.view = vgroup * kViewsPerBlock + wave;
    (void)nV;
    (void)nviews;
    return c;
}

// Offset of padded coordinate p along one axis of the 4^3-bricked volume
// (p >= 0; stride < 2^24, so the 24-bit multiplier suffices).
__device__ inline unsigned brick_off(int p, unsigned stride, unsigned unit) {
    return __umul24((unsigned)p >> 2, stride) + ((unsigned)p & 3u) * unit;
}

// Volume reads.  Below 4 GiB the volume is read through a raw buffer
// descriptor whose range is the volume: an element index outside it reads 0
// instead of touching memory.  Larger volumes use plain global loads.
template <bool kBuffer>
struct VolumeReader {
    __amdgpu_buffer_rsrc_t rsrc;
    const float* base;
    __device__ VolumeReader(const float* p, unsigned long long elements) : base(p) {
        rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0,
                                                 kBuffer ? (int)(unsigned)(elements * 4ull) : 0, 0x00020000);
    }
    __device__ float operator()(unsigned index) const {
        if (kBuffer) return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, index * 4u, 0, 0));
        return base[index];
    }
};

}  // namespace


// Siddon / Jacobs exact radiological path, Siddon_projection.cu:110-268.
// vol: 4^3-bricked copy (see brick_volume_kernel); pad[0] of each view record
// holds its projection index in `out`.
template <bool kBuffer>
__global__ __launch_bounds__(256) void ax_siddon_kernel(const float* __restrict__ vol, unsigned long long elements,
                                                        int nx, int ny, int nz, int nbx, int nby,
                                                        const FwdView* __restrict__ views, int nviews, int nU, int nV,
                                                        long long su, long long sv, long long sview,
                                                        float* __restrict__ out, double dvx, double dvy, double dvz) {
    const RayCoords c = ray_of_lane(nU, nV, nviews);
    if (c.u >= nU || c.v >= nV || c.view >= nviews) return;
    const FwdView& w = views[c.view];
    const double pv = (double)(nV - 1 - c.v);  // detector row flip, :141
    double sp[3];
    ray_source(w, c.u, pv, sp);
    double rx = ray_dir(w, c.u, pv, 0);
    double ry = ray_dir(w, c.u, pv, 1);
    double rz = ray_dir(w, c.u, pv, 2);
    if (w.pad[1] != 0.0) parallel_siddon_height(sp, rx, ry, rz, nx, ny);
    // components below 1e-3 voxel are treated as exactly zero (:155-159)
    if (fabs(rx) < 1e-3) rx = 0.0;
    if (fabs(ry) < 1e-3) ry = 0.0;
    if (fabs(rz) < 1e-3) rz = 0.0;
    float* dst = out + (long long)w.pad[0] * sview + (long long)c.u * su + (long long)c.v * sv;

    double am = -INFINITY, aM = INFINITY;
    bool hit = slab(sp[0], rx, nx, am, aM);
    hit = hit && slab(sp[1], ry, ny, am, aM);
    hit = hit && slab(sp[2], rz, nz, am, aM);
    if (!hit || !(am < aM)) {  // a ray that misses the volume projects to 0 (:179-180)
        *dst = 0.0f;
        return;
    }
    const Axis ax = axis_setup(sp[0], rx, am);
    const Axis ay = axis_setup(sp[1], ry, am);
    const Axis az = axis_setup(sp[2], rz, am);
    const float T = (float)(aM - am);

    // padded coordinates (index + 1) and their brick offsets
    const unsigned SY = 64u * (unsigned)nbx, SZ = SY * (unsigned)nby;
    const VolumeReader<kBuffer> vload(vol, elements);
    int px = ax.idx + 1, py = ay.idx + 1, pz = az.idx + 1;
    unsigned ox = brick_off(px, 64u, 1u), oy = brick_off(py, SY, 4u), oz = brick_off(pz, SZ, 16u);

    float tx = ax.t0, ty = ay.t0, tz = az.t0;
    float kx = 0.0f, ky = 0.0f, kz = 0.0f;
    float tc = 0.0f, sum = 0.0f;
    // Every step consumes one plane crossing; the crossings inside the volume
    // number at most nx + ny + nz + 3, so the guard always ends the walk.
    // Software pipeline: step k fetches the voxel of segment k + 1 and
    // accumulates segment k - 1, so every load has two steps to arrive; the
    // loop is unrolled three times so the three voxel registers rotate.  The
    // last segment, cut at the exit T, is added when the walk ends.
#define TIGRE_SIDDON_STEP(PREV, CUR, NEXT)                                 \
    {                                                                      \
        const float tm = fminf(tx, fminf(ty, tz));                         \
        if (tm >= T) {                                                     \
            sum = fmaf(lp, PREV, sum);                                     \
            sum = fmaf(T - tc, CUR, sum);                                  \
            break;                                                         \
        }                                                                  \
        if (tx == tm) {                                                    \
            kx += 1.0f;                                                    \
            tx = fmaf(kx, ax.dt, ax.t0);                                   \
            px += ax.step;                                                 \
            ox = brick_off(px, 64u, 1u);                                   \
        } else if (ty == tm) {                                             \
            ky += 1.0f;                                                    \
            ty = fmaf(ky, ay.dt, ay.t0);                                   \
            py += ay.step;                                                 \
            oy = brick_off(py, SY, 4u);                                    \
        } else {                                                           \
            kz += 1.0f;                                                    \
            tz = fmaf(kz, az.dt, az.t0);                                   \
            pz += az.step;                                                 \
            oz = brick_off(pz, SZ, 16u);                                   \
        }                                                                  \
        NEXT = vload(ox + oy + oz);                                        \
        sum = fmaf(lp, PREV, sum);                                         \
        lp = tm - tc;                                                      \
        tc = tm;                                                           \
    }
    float r0 = 0.0f, r1 = vload(ox + oy + oz), r2 = 0.0f, lp = 0.0f;
    for (int guard = (nx + ny + nz + 9) / 3; guard > 0; --guard) {
        TIGRE_SIDDON_STEP(r0, r1, r2)
        TIGRE_SIDDON_STEP(r1, r2, r0)
        TIGRE_SIDDON_STEP(r2, r0, r1)
    }
#undef TIGRE_SIDDON_STEP
    // scale by the ray length in mm (:241, :267)
    const double len = sqrt(rx * rx * dvx * dvx + ry * ry * dvy * dvy + rz * rz * dvz * dvz);
    *dst = (float)((double)sum * len);
}

// Uniform-step trilinear sampling, ray_interpolated_projection.cu:117-206.
// vol: apron-bricked volume: cell (lower corner) c in padded coordinates lies
// in brick c >> 2, and each brick stores the 5^3 samples 4b .. 4b + 4 of its
// axis ranges, so the 8 taps of a trilinear sample are one brick base plus the
// constant offsets {0, 1, 5, 6, 25, 26, 30, 31}.
template <bool kNearest, bool kBuffer>
__global__ __launch_bounds__(256) void ax_interp_kernel(const float* __restrict__ vol, int nx, int ny, int nz,
                                                        int nbx, int nby, int nbz, const FwdView* __restrict__ views,
                                                        int nviews, int nU, int nV, long long su, long long sv,
                                                        long long sview, float* __restrict__ out, double acc,
                                                        double dvx, double dvy, double dvz) {
    const RayCoords c = ray_of_lane(nU, nV, nviews);
    if (c.u >= nU || c.v >= nV || c.view >= nviews) return;
    const FwdView& w = views[c.view];
    const double pv = (double)(nV - 1 - c.v);
    double sp[3];
    ray_source(w, c.u, pv, sp);
    const double dx = ray_dir(w, c.u, pv, 0);
    const double dy = ray_dir(w, c.u, pv, 1);
    const double dz = ray_dir(w, c.u, pv, 2);
    float* dst = out + (long long)w.pad[0] * sview + (long long)c.u * su + (long long)c.v * sv;

    // samples per ray and step (:164-169)
    const double n = sample_count(w, sqrt(dx * dx + dy * dy + dz * dz), acc);
    if (!(n >= 1.0 && n < 1e9)) {  // degenerate ray (pixel at the source)
        *dst = 0.0f;
        return;
    }
    const double stx = dx / n, sty = dy / n, stz = dz / n;
    // sample index range: [first, last] of the reference (:182-193), clipped
    // to the samples inside the zero-bordered volume (-1, n) on every axis;
    // the clipped-away samples are exactly zero.
    const double last = (w.cap < n) ? ceil(w.cap) : n;
    double lo = -INFINITY, hi = INFINITY;
    const double sxyz[3] = {sp[0], sp[1], sp[2]};
    const double st[3] = {stx, sty, stz};
    const int nn[3] = {nx, ny, nz};
    bool empty = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (st[a] == 0.0) {
            if (!(sxyz[a] > -1.0 && sxyz[a] < (double)nn[a])) empty = true;
        } else {
            const double a0 = (-1.0 - sxyz[a]) / st[a], a1 = ((double)nn[a] - sxyz[a]) / st[a];
            lo = fmax(lo, fmin(a0, a1));
            hi = fmin(hi, fmax(a0, a1));
        }
    }
    const double i0 = fmax(w.first, floor(lo));
    const double i1 = fmin(last, ceil(hi));
    const double dl = sqrt(stx * dvx * stx * dvx + sty * dvy * sty * dvy + stz * dvz * stz * dvz);
    if (empty || !(i0 <= i1)) {
        *dst = 0.0f;
        return;
    }
    // Positions relative to the first sample keep the FP32 loop accurate;
    // +1 moves them to padded coordinates (index -1 -> 0).
    const float qx = (float)(sp[0] + i0 * stx + 1.0), qy = (float)(sp[1] + i0 * sty + 1.0),
                qz = (float)(sp[2] + i0 * stz + 1.0);
    const float fx = (float)stx, fy = (float)sty, fz = (float)stz;
    const int count = (int)(i1 - i0) + 1;

    const unsigned BY = 125u * (unsigned)nbx, BZ = BY * (unsigned)nby;
    const VolumeReader<kBuffer> vload(vol, (unsigned long long)BZ * (unsigned)nbz);
    // padded coordinates stay in [0, n + 1] so the cell and its +1 taps exist
    const float hx = (float)(nx + 1), hy = (float)(ny + 1), hz = (float)(nz + 1);
    float sum = 0.0f;
#pragma unroll 2
    for (int m = 0; m < count; ++m) {
        const float fm = (float)m;
        const float x = __builtin_amdgcn_fmed3f(fmaf(fm, fx, qx), 0.0f, hx);
        const float y = __builtin_amdgcn_fmed3f(fmaf(fm, fy, qy), 0.0f, hy);
        const float z = __builtin_amdgcn_fmed3f(fmaf(fm, fz, qz), 0.0f, hz);
        if (kNearest) {
            const int ix = (int)floorf(x + 0.5f), iy = (int)floorf(y + 0.5f), iz = (int)floorf(z + 0.5f);
            const unsigned o = __umul24((unsigned)ix >> 2, 125u) + __umul24((unsigned)iy >> 2, BY) +
                               __umul24((unsigned)iz >> 2, BZ) + ((unsigned)ix & 3u) +
                               __umul24((unsigned)iy & 3u, 5u) + __umul24((unsigned)iz & 3u, 25u);
            sum += vload(o);
        } else {
            const float flx = floorf(x), fly = floorf(y), flz = floorf(z);
            const float wx = x - flx, wy = y - fly, wz = z - flz;
            const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
            const unsigned o = __umul24((unsigned)ix >> 2, 125u) + __umul24((unsigned)iy >> 2, BY) +
                               __umul24((unsigned)iz >> 2, BZ) + ((unsigned)ix & 3u) +
                               __umul24((unsigned)iy & 3u, 5u) + __umul24((unsigned)iz & 3u, 25u);
            const float c000 = vload(o), c100 = vload(o + 1), c010 = vload(o + 5),
                        c110 = vload(o + 6);
            const float c001 = vload(o + 25), c101 = vload(o + 26), c011 = vload(o + 30),
                        c111 = vload(o + 31);
            const float c00 = fmaf(wx, c100 - c000, c000), c10 = fmaf(wx, c110 - c010, c010);
            const float c01 = fmaf(wx, c101 - c001, c001), c11 = fmaf(wx, c111 - c011, c011);
            const float c0 = fmaf(wy, c10 - c00, c00), c1 = fmaf(wy, c11 - c01, c01);
            sum += fmaf(wz, c1 - c0, c0);
        }
    }
    *dst = (float)((double)sum * dl);
}

// 4^3 bricks over padded coordinates [0, 4 nb): voxel index = padded - 1,
// zero outside the volume.  Output element q: brick q >> 6, local q & 63.
__global__ __launch_bounds__(256) void brick_volume_kernel(const void* __restrict__ vol, int dtype, int nx, int ny,
                                                           int nz, int nbx, int nby, int nbz,
                                                           float* __restrict__ bricked) {
    const long long total = (long long)nbx * nby * nbz * 64;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const long long b = q >> 6;
        const int l = (int)(q & 63);
        const long long bx = b % nbx, by = (b / nbx) % nby, bz = b / ((long long)nbx * nby);
        const long long i = bx * 4 + (l & 3) - 1, j = by * 4 + ((l >> 2) & 3) - 1, k = bz * 4 + (l >> 4) - 1;
        float v = 0.0f;
        if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) v = load_storage(vol, (k * ny + j) * nx + i, dtype);
        bricked[q] = v;
    }
}

// Apron bricks: brick b stores padded coordinates 4b .. 4b + 4 on each axis
// (5^3 values, x fastest), so neighbouring bricks share one face.
__global__ __launch_bounds__(256) void apron_volume_kernel(const void* __restrict__ vol, int dtype, int nx, int ny,
                                                           int nz, int nbx, int nby, int nbz,
                                                           float* __restrict__ bricked) {
    const long long total = (long long)nbx * nby * nbz * 125;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const long long b = q / 125;
        const int l = (int)(q % 125);
        const long long bx = b % nbx, by = (b / nbx) % nby, bz = b / ((long long)nbx * nby);
        const long long i = bx * 4 + (l % 5) - 1, j = by * 4 + ((l / 5) % 5) - 1, k = bz * 4 + (l / 25) - 1;
        float v = 0.0f;
        if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) v = load_storage(vol, (k * ny + j) * nx + i, dtype);
        bricked[q] = v;
    }
}

namespace {
long long grid_stride_blocks(long long total) {
    long long blocks = (total + 255) / 256;
    return blocks > 65536 ? 65536 : blocks;
}

void out_strides(const HostGeometry& g, int layout, long long* su, long long* sv, long long* sview) {
    const long long nU = g.nDetec[0], nV = g.nDetec[1];
    *sview = nU * nV;
    if (layout == TIGRE_LAYOUT_MATLAB) {
        *su = nV;
        *sv = 1;
    } else {
        *su = 1;
        *sv = nU;
    }
}

int ray_blocks(const HostGeometry& g, int nviews) {
    const long long tiles = (long long)((g.nDetec[0] + kTileEdge - 1) / kTileEdge) *
                            ((g.nDetec[1] + kTileEdge - 1) / kTileEdge);
    const long long groups = (nviews + kViewsPerBlock - 1) / kViewsPerBlock;
    return (int)(tiles * groups);
}
}  // namespace

BrickGrid interp_bricks(const HostGeometry& g) {
    // cells (lower taps) 0 .. n + 1
    return BrickGrid{(g.nVoxel[0] + 1) / 4 + 1, (g.nVoxel[1] + 1) / 4 + 1, (g.nVoxel[2] + 1) / 4 + 1, 125};
}

int launch_brick_volume(const void* vol, int dtype, const HostGeometry& g, const BrickGrid& b, float* bricked,
                        hipStream_t stream) {
    const long long total = b.elements();
    KernelTimerScope timer(TIGRE_KERNEL_RELAYOUT, stream);
    if (b.per_brick == 64)
        hipLaunchKernelGGL(brick_volume_kernel, dim3((unsigned)grid_stride_blocks(total)), dim3(256), 0, stream, vol,
                           dtype, g.nVoxel[0], g.nVoxel[1], g.nVoxel[2], b.nbx, b.nby, b.nbz, bricked);
    else
        hipLaunchKernelGGL(apron_volume_kernel, dim3((unsigned)grid_stride_blocks(total)), dim3(256), 0, stream, vol,
                           dtype, g.nVoxel[0], g.nVoxel[1], g.nVoxel[2], b.nbx, b.nby, b.nbz, bricked);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

BrickGrid siddon_bricks(const HostGeometry& g) {
    // padded coordinates 0 .. n + 2
    return BrickGrid{(g.nVoxel[0] + 2) / 4 + 1, (g.nVoxel[1] + 2) / 4 + 1, (g.nVoxel[2] + 2) / 4 + 1, 64};
}

int launch_ax_siddon(const float* bricked, const HostGeometry& g, const BrickGrid& b, const FwdView* d_views,
                     int nviews, int layout, float* proj, hipStream_t stream) {
    long long su, sv, sview;
    out_strides(g, layout, &su, &sv, &sview);
    // 24-bit multiplier operands and 32-bit offsets
    TIGRE_REQUIRE(b.elements() < (1LL << 32) && (long long)b.nbx * b.nby * 64 < (1LL << 24),
                  "volume too large for 32-bit brick offsets");
    const bool buf = b.elements() * 4 < (1LL << 32);
    auto kern = buf ? ax_siddon_kernel<true> : ax_siddon_kernel<false>;
    KernelTimerScope timer(TIGRE_KERNEL_AX_SIDDON, stream);
    hipLaunchKernelGGL(kern, dim3(ray_blocks(g, nviews)), dim3(256), 0, stream, bricked,
                       (unsigned long long)b.elements(), g.nVoxel[0], g.nVoxel[1], g.nVoxel[2], b.nbx, b.nby, d_views,
                       nviews, g.nDetec[0],
                       g.nDetec[1], su, sv, sview, proj, g.dVoxel[0], g.dVoxel[1], g.dVoxel[2]);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

int launch_ax_interp(const float* bricked, const HostGeometry& g, const BrickGrid& b, const FwdView* d_views,
                     int nviews, int layout, float* proj, hipStream_t stream) {
    long long su, sv, sview;
    out_strides(g, layout, &su, &sv, &sview);
    TIGRE_REQUIRE(b.elements() < (1LL << 32) && (long long)b.nbx * b.nby * 125 < (1LL << 24),
                  "volume too large for 32-bit brick offsets");
    // point filtering when accuracy > 1 (ray_interpolated_projection.cu:589-595)
    const bool buf = b.elements() * 4 < (1LL << 32);
    auto kern = (g.accuracy > 1.0) ? (buf ? ax_interp_kernel<true, true> : ax_interp_kernel<true, false>)
                                   : (buf ? ax_interp_kernel<false, true> : ax_interp_kernel<false, false>);
    KernelTimerScope timer(TIGRE_KERNEL_AX_INTERPOLATED, stream);
    hipLaunchKernelGGL(kern, dim3(ray_blocks(g, nviews)), dim3(256), 0, stream, bricked, g.nVoxel[0], g.nVoxel[1],
                       g.nVoxel[2], b.nbx, b.nby, b.nbz, d_views, nviews, g.nDetec[0], g.nDetec[1], su, sv, sview,
                       proj, g.accuracy, g.dVoxel[0], g.dVoxel[1], g.dVoxel[2]);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

}  // namespace tigre
