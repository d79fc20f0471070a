// Cone-beam forward projectors for gfx950.
//
// Ray-driven, one ray per lane.  A 256-thread workgroup owns a 16 x 16
// detector tile of one view and each 64-lane wavefront an 8 x 8 sub-tile, so
// the rays of a wavefront are nearly parallel and walk the same voxels
// (coherent gathers, shared cache lines).  Per-ray set-up (frame, entry and
// exit, first crossings) runs in FP64 -- on CDNA4 FP64 FMA has the same vector
// rate as non-packed FP32 -- and the inner loop runs in FP32.
//
// The volume is read from a zero-padded copy (one voxel below, two above on
// every axis), which replaces the reference's border-addressed texture: taps
// that fall outside the volume read zeros without bounds checks.
#include <cmath>

#include "common.hpp"

namespace tigre {

namespace {

constexpr int kTile = 16;

struct RayCoords {
    int u, v;
};

// Lane -> detector pixel: 2 x 2 wavefronts of 8 x 8 rays per 16 x 16 tile.
__device__ inline RayCoords tile_pixel(int tile, int nU) {
    const int tilesU = (nU + kTile - 1) / kTile;
    const int tu = tile % tilesU, tv = tile / tilesU;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    RayCoords c;
    c.u = tu * kTile + (wave & 1) * 8 + (lane & 7);
    c.v = tv * kTile + (wave >> 1) * 8 + (lane >> 3);
    return c;
}

// Slab [0, n] limits of the parametric ray s + a r (IEEE semantics of
// Siddon_projection.cu:168-176: a ray parallel to the slab and strictly inside
// is unconstrained, otherwise it misses).
__device__ inline bool slab(double s, double r, int n, double& am, double& aM) {
    if (r == 0.0) return s > 0.0 && s < (double)n;
    const double a0 = (0.0 - s) / r, a1 = ((double)n - s) / r;
    am = fmax(am, fmin(a0, a1));
    aM = fmin(aM, fmax(a0, a1));
    return true;
}

struct Axis {
    float t0, dt;    // first crossing (relative to the entry) and crossing spacing
    int idx;         // voxel index before the first crossing
    int step;        // +1 / -1 (0 if the ray is parallel to the axis)
};

// First plane crossed after the entry a = am and the voxel index before it.
// The index is derived from that plane, so index and crossing list can never
// disagree (the reference derives it from a midpoint instead).
__device__ inline Axis axis_setup(double s, double r, double am) {
    Axis a;
    if (r == 0.0) {
        a.t0 = INFINITY;
        a.dt = INFINITY;
        a.idx = (int)floor(s);
        a.step = 0;
        return a;
    }
    const double p = s + am * r;
    double plane;
    if (r > 0.0) {
        plane = floor(p) + 1.0;
        a.idx = (int)plane - 1;
        a.step = 1;
    } else {
        plane = ceil(p) - 1.0;
        a.idx = (int)plane;
        a.step = -1;
    }
    a.t0 = (float)((plane - s) / r - am);
    a.dt = (float)(1.0 / fabs(r));
    return a;
}

}  // namespace

// Siddon / Jacobs exact radiological path, Siddon_projection.cu:110-268.
__global__ __launch_bounds__(256) void ax_siddon_kernel(const float* __restrict__ vol, int nx, int ny, int nz,
                                                        const FwdView* __restrict__ views, int nU, int nV,
                                                        long long su, long long sv, long long sview,
                                                        float* __restrict__ out, double dvx, double dvy,
                                                        double dvz) {
    const RayCoords c = tile_pixel(blockIdx.x, nU);
    if (c.u >= nU || c.v >= nV) return;
    const int view = blockIdx.y;
    const FwdView& w = views[view];
    const double pv = (double)(nV - 1 - c.v);  // detector row flip, :141
    double rx = w.org[0] + c.u * w.du[0] + pv * w.dv[0] - w.src[0];
    double ry = w.org[1] + c.u * w.du[1] + pv * w.dv[1] - w.src[1];
    double rz = w.org[2] + c.u * w.du[2] + pv * w.dv[2] - w.src[2];
    // components below 1e-3 voxel are treated as exactly zero (:155-159)
    if (fabs(rx) < 1e-3) rx = 0.0;
    if (fabs(ry) < 1e-3) ry = 0.0;
    if (fabs(rz) < 1e-3) rz = 0.0;
    float* dst = out + (long long)view * sview + (long long)c.u * su + (long long)c.v * sv;

    double am = -INFINITY, aM = INFINITY;
    bool hit = slab(w.src[0], rx, nx, am, aM);
    hit = hit && slab(w.src[1], ry, ny, am, aM);
    hit = hit && slab(w.src[2], rz, nz, am, aM);
    if (!hit || !(am < aM)) {  // a ray that misses the volume projects to 0 (:179-180)
        *dst = 0.0f;
        return;
    }
    const Axis ax = axis_setup(w.src[0], rx, am);
    const Axis ay = axis_setup(w.src[1], ry, am);
    const Axis az = axis_setup(w.src[2], rz, am);
    const float T = (float)(aM - am);

    const long long PX = padded_dim(nx), PXY = PX * padded_dim(ny);
    long long addr = (long long)(az.idx + 1) * PXY + (long long)(ay.idx + 1) * PX + (ax.idx + 1);
    const long long sx = ax.step, sy = ay.step * PX, sz = az.step * PXY;

    float tx = ax.t0, ty = ay.t0, tz = az.t0;
    float kx = 0.0f, ky = 0.0f, kz = 0.0f;
    float tc = 0.0f, sum = 0.0f;
    // Every iteration consumes one plane crossing; the crossings inside the
    // volume number at most nx + ny + nz + 3, so the guard always ends the walk.
    for (int guard = nx + ny + nz + 4; guard > 0; --guard) {
        const float tm = fminf(tx, fminf(ty, tz));
        const float tn = fminf(tm, T);
        sum = fmaf(tn - tc, vol[addr], sum);
        if (tm >= T) break;
        tc = tn;
        if (tx == tm) {
            kx += 1.0f;
            tx = fmaf(kx, ax.dt, ax.t0);
            addr += sx;
        } else if (ty == tm) {
            ky += 1.0f;
            ty = fmaf(ky, ay.dt, ay.t0);
            addr += sy;
        } else {
            kz += 1.0f;
            tz = fmaf(kz, az.dt, az.t0);
            addr += sz;
        }
    }
    // scale by the ray length in mm (:241, :267)
    const double len = sqrt(rx * rx * dvx * dvx + ry * ry * dvy * dvy + rz * rz * dvz * dvz);
    *dst = (float)((double)sum * len);
}

// Uniform-step trilinear sampling, ray_interpolated_projection.cu:117-206.
template <bool kNearest>
__global__ __launch_bounds__(256) void ax_interp_kernel(const float* __restrict__ vol, int nx, int ny, int nz,
                                                        const FwdView* __restrict__ views, int nU, int nV,
                                                        long long su, long long sv, long long sview,
                                                        float* __restrict__ out, double acc, double dvx,
                                                        double dvy, double dvz) {
    const RayCoords c = tile_pixel(blockIdx.x, nU);
    if (c.u >= nU || c.v >= nV) return;
    const int view = blockIdx.y;
    const FwdView& w = views[view];
    const double pv = (double)(nV - 1 - c.v);
    const double dx = w.org[0] + c.u * w.du[0] + pv * w.dv[0] - w.src[0];
    const double dy = w.org[1] + c.u * w.du[1] + pv * w.dv[1] - w.src[1];
    const double dz = w.org[2] + c.u * w.du[2] + pv * w.dv[2] - w.src[2];
    float* dst = out + (long long)view * sview + (long long)c.u * su + (long long)c.v * sv;

    // samples per ray and step (:164-169)
    const double n = ceil(sqrt(dx * dx + dy * dy + dz * dz) / acc);
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
    const double sxyz[3] = {w.src[0], w.src[1], w.src[2]};
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
    // Positions relative to the first sample keep the FP32 loop accurate.
    const float qx = (float)(w.src[0] + i0 * stx), qy = (float)(w.src[1] + i0 * sty),
                qz = (float)(w.src[2] + i0 * stz);
    const float fx = (float)stx, fy = (float)sty, fz = (float)stz;
    const int count = (int)(i1 - i0) + 1;

    const long long PX = padded_dim(nx), PXY = PX * padded_dim(ny);
    const float hx = (float)nx, hy = (float)ny, hz = (float)nz;
    float sum = 0.0f;
    for (int m = 0; m < count; ++m) {
        const float fm = (float)m;
        float x = fmaf(fm, fx, qx), y = fmaf(fm, fy, qy), z = fmaf(fm, fz, qz);
        x = clampf(x, hx);
        y = clampf(y, hy);
        z = clampf(z, hz);
        if (kNearest) {
            const int ix = (int)floorf(x + 0.5f), iy = (int)floorf(y + 0.5f), iz = (int)floorf(z + 0.5f);
            sum += vol[(long long)(iz + 1) * PXY + (long long)(iy + 1) * PX + (ix + 1)];
        } else {
            const float flx = floorf(x), fly = floorf(y), flz = floorf(z);
            const float wx = x - flx, wy = y - fly, wz = z - flz;
            const float* p = vol + ((long long)((int)flz + 1) * PXY + (long long)((int)fly + 1) * PX + ((int)flx + 1));
            const float c000 = p[0], c100 = p[1], c010 = p[PX], c110 = p[PX + 1];
            const float c001 = p[PXY], c101 = p[PXY + 1], c011 = p[PXY + PX], c111 = p[PXY + PX + 1];
            const float c00 = fmaf(wx, c100 - c000, c000), c10 = fmaf(wx, c110 - c010, c010);
            const float c01 = fmaf(wx, c101 - c001, c001), c11 = fmaf(wx, c111 - c011, c011);
            const float c0 = fmaf(wy, c10 - c00, c00), c1 = fmaf(wy, c11 - c01, c01);
            sum += fmaf(wz, c1 - c0, c0);
        }
    }
    *dst = (float)((double)sum * dl);
}

// Zero-padded copy of a volume (padding 1 below, 2 above on every axis).
__global__ __launch_bounds__(256) void pad_volume_kernel(const float* __restrict__ vol, int nx, int ny, int nz,
                                                         float* __restrict__ padded) {
    const long long PX = padded_dim(nx), PY = padded_dim(ny), PZ = padded_dim(nz);
    const long long total = PX * PY * PZ;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const long long i = q % PX - 1, j = (q / PX) % PY - 1, k = q / (PX * PY) - 1;
        float v = 0.0f;
        if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) v = vol[(k * ny + j) * nx + i];
        padded[q] = v;
    }
}

int launch_pad_volume(const float* vol, int nx, int ny, int nz, float* padded, hipStream_t stream) {
    const long long total = padded_dim(nx) * padded_dim(ny) * padded_dim(nz);
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(pad_volume_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, vol, nx, ny, nz, padded);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

namespace {
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
}  // namespace

int launch_ax_siddon(const float* padded, const HostGeometry& g, const FwdView* d_views, int layout, float* proj,
                     hipStream_t stream) {
    long long su, sv, sview;
    out_strides(g, layout, &su, &sv, &sview);
    const int nU = g.nDetec[0], nV = g.nDetec[1];
    const int tiles = ((nU + kTile - 1) / kTile) * ((nV + kTile - 1) / kTile);
    TIGRE_REQUIRE(g.nangles <= 65535, "at most 65535 projections per call");
    hipLaunchKernelGGL(ax_siddon_kernel, dim3(tiles, g.nangles), dim3(256), 0, stream, padded, g.nVoxel[0],
                       g.nVoxel[1], g.nVoxel[2], d_views, nU, nV, su, sv, sview, proj, g.dVoxel[0], g.dVoxel[1],
                       g.dVoxel[2]);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

int launch_ax_interp(const float* padded, const HostGeometry& g, const FwdView* d_views, int layout, float* proj,
                     hipStream_t stream) {
    long long su, sv, sview;
    out_strides(g, layout, &su, &sv, &sview);
    const int nU = g.nDetec[0], nV = g.nDetec[1];
    const int tiles = ((nU + kTile - 1) / kTile) * ((nV + kTile - 1) / kTile);
    TIGRE_REQUIRE(g.nangles <= 65535, "at most 65535 projections per call");
    // point filtering when accuracy > 1 (ray_interpolated_projection.cu:589-595)
    if (g.accuracy > 1.0)
        hipLaunchKernelGGL(ax_interp_kernel<true>, dim3(tiles, g.nangles), dim3(256), 0, stream, padded, g.nVoxel[0],
                           g.nVoxel[1], g.nVoxel[2], d_views, nU, nV, su, sv, sview, proj, g.accuracy, g.dVoxel[0],
                           g.dVoxel[1], g.dVoxel[2]);
    else
        hipLaunchKernelGGL(ax_interp_kernel<false>, dim3(tiles, g.nangles), dim3(256), 0, stream, padded,
                           g.nVoxel[0], g.nVoxel[1], g.nVoxel[2], d_views, nU, nV, su, sv, sview, proj, g.accuracy,
                           g.dVoxel[0], g.dVoxel[1], g.dVoxel[2]);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

}  // namespace tigre
