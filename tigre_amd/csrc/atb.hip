// Cone-beam voxel-driven backprojectors for gfx950 ('FDK' and 'matched').
//
// A 256-thread workgroup owns a 16 x 16 column block of the volume and kZ
// consecutive slices; each lane keeps its kZ voxels in registers across all
// views, so the volume is read once and written once per call.  The four
// 64-lane wavefronts each cover an 8 x 8 column patch: the patch projects
// to a compact detector footprint, so the bilinear gathers of a wavefront
// hit a few cache lines.  Per-view parameters are uniform across the
// workgroup and are fetched with scalar loads.
//
// Projections are read from a "quad" copy: with the four taps a = p(u,v),
// b = p(u+1,v), c = p(u,v+1), d = p(u+1,v+1) (zero outside the detector),
// element (u, v) holds (a, b - a, c - a, (d - c) - (b - a)) for tap origins
// u in [-1, nU], v in [-1, nV], stored at index (u + 1, v + 1).  One 16-byte
// load per voxel update replaces four scattered taps, the zero border
// replaces the border-addressed linear texture of the reference without
// bounds checks, and the bilinear sample is three fused multiply-adds
// (bilerp below): the differences are formed once per detector pixel instead
// of once per voxel update.  Detector coordinates are carried in
// that storage frame (pixel index + 1, the offset folded into the per-view
// records), clamped with one med3 to [0, n + 1]: the tap origin is then the
// truncating conversion and the weight the hardware fraction.
//
// kAxial: when no view tilts the rotation axis (dZ has no x/y component, the
// case of every planar orbit without detector pitch/yaw) the column projects
// to one detector column, so the u interpolation weights, the 1/x division
// and the FDK weight are computed once per column and only v steps with z.
#include <cmath>

#include "common.hpp"

namespace tigre {

namespace {
// 1/x with the hardware reciprocal (1 ulp), instead of a correctly rounded division.
__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Waves per SIMD the gather kernel is compiled for: the register budget that
// keeps all kZ gathers of a column in flight (the compiler otherwise trades
// them for occupancy and serialises the loads).  The matched weight with
// tilted views and 16 slices needs more than the 168 registers of three
// waves (it spilled), so it gets two.
constexpr int atb_waves(int kW, bool kAxial, int kZ) {
    return kZ >= 16 ? (kW == 1 ? (kAxial ? 3 : 2) : 4) : kZ == 8 ? (kW == 1 ? 5 : 6) : 8;
}

// Matched weight l2^(3/2) / s2 from the DSD-normalised squares l2 (source to
// detector point, about 1) and s2 (source to voxel, about (DSO/DSD)^2): one
// reciprocal square root, l2^2 / sqrt(l2 s2^2), instead of a square root and a
// reciprocal.  The normalised operands keep every product near 1, whatever
// the length unit of the geometry.
__device__ __forceinline__ float matched_weight(float l2, float s2) {
    return (l2 * l2) * __builtin_amdgcn_rsqf(l2 * s2 * s2);
}

// Minimum / maximum of two wave-uniform values on the scalar unit.
__device__ __forceinline__ unsigned smin(unsigned a, unsigned b) {
    unsigned r;
    asm("s_min_u32 %0, %1, %2" : "=s"(r) : "s"(a), "s"(b));
    return r;
}
__device__ __forceinline__ unsigned smax(unsigned a, unsigned b) {
    unsigned r;
    asm("s_max_u32 %0, %1, %2" : "=s"(r) : "s"(a), "s"(b));
    return r;
}

// Bilinear sample from a quad (a, b - a, c - a, (d - c) - (b - a)) at weights
// (wu, wv): s = a0 + wv (a1 - a0) with a0 = a + wu (b - a) and
// a1 - a0 = (c - a) + wu ((d - c) - (b - a)).
__device__ __forceinline__ float bilerp(const float4& q, float wu, float wv) {
    const float a0 = fmaf(wu, q.y, q.x);
    return fmaf(wv, fmaf(wu, q.w, q.z), a0);
}
}

// Writes a lane's kZ accumulated voxels: the plain result acc * scale or,
// with ep.x set, the SART update x + inv_v * (acc * scale) (clamped at 0 when
// ep.noneg) with the rounding steps of that tensor expression.
template <int kZ>
__device__ __forceinline__ void store_column(const float (&acc)[kZ], float* vol, int ix, int iy, int z0, int nx,
                                             int ny, int nz, float scale, const AtbUpdate& ep) {
    const long long plane = (long long)nx * ny;
    const long long at = (long long)z0 * plane + (long long)iy * nx + ix;
    float* out = vol + at;
    if (ep.x == nullptr) {
#pragma unroll
        for (int k = 0; k < kZ; ++k)
            if (z0 + k < nz) out[(long long)k * plane] = acc[k] * scale;
    } else {
        // three roundings, no fused multiply-add (contraction is off in this block)
#pragma clang fp contract(off)
        const float inv = ep.inv_v[(long long)iy * nx + ix];
        const float* xin = ep.x + at;
#pragma unroll
        for (int k = 0; k < kZ; ++k) {
            if (z0 + k < nz) {
                const float b = acc[k] * scale;
                const float t = inv * b;
                float v = xin[(long long)k * plane] + t;
                if (ep.noneg) v = fmaxf(v, 0.0f);
                out[(long long)k * plane] = v;
            }
        }
    }
}

// kZ: slices per lane (4, 8 or 16; fewer gives more workgroups on small
// volumes).  With ep.x set, the result is not stored but applied as the SART
// update x + inv_v[iy, ix] * result, clamped at 0 when ep.noneg, written to
// vol (which may alias ep.x: each voxel is read and written by one lane).
// kW: 0 FDK weight (voxel_backprojection.cu:241-249), 1 matched weight
// (voxel_backprojection2.cu:201-279), 2 parallel beam: the voxel projects
// orthogonally, u = y + nU/2 - 0.5, v = z + nV/2 - 0.5, weight 1
// (voxel_backprojection_parallel.cu:221-251).
template <int kW, bool kAxial, int kZ>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(atb_waves(kW, kAxial, kZ), atb_waves(kW, kAxial, kZ)))) void atb_kernel(const float* __restrict__ proj, int nU, int nV, int nviews,
                                                  const BwdView* __restrict__ views, float* vol, int nx,
                                                  int ny, int nz, float scale, AtbUpdate ep) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ix = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
    const int iy = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
    const int z0 = blockIdx.z * kZ;
    if (ix >= nx || iy >= ny) return;

    const long long QU = quad_det(nU);
    const long long QP = QU * quad_det(nV);
    const float4* __restrict__ quads = reinterpret_cast<const float4*>(proj);
    const unsigned rowb = (unsigned)QU * 16u;  // bytes per quad row (< 2^24, checked by the launcher)
    const float hu = (float)(nU + 1), hv = (float)(nV + 1);  // storage-frame clamp bounds
    const float fx = (float)ix, fy = (float)iy;

    float acc[kZ];
#pragma unroll
    for (int k = 0; k < kZ; ++k) acc[k] = 0.0f;

    // one voxel update from a storage-frame detector position (u, v) already
    // clamped to [0, n + 1]: bilinear sample of the quad copy.  Offsets are
    // 32-bit bytes from the view's (uniform) base, formed with one 24-bit
    // multiply-add.
    auto sample = [&](const char* pjb, unsigned colb, float wu, float pv) {
        const float wv = __builtin_amdgcn_fractf(pv);
        const float4 q = *reinterpret_cast<const float4*>(pjb + (__umul24((unsigned)pv, rowb) + colb));
        return bilerp(q, wu, wv);
    };

    for (int p = 0; p < nviews; ++p) {
        const BwdView& w = views[p];
        const char* pjb = reinterpret_cast<const char*>(quads + (long long)p * QP);
        // source -> voxel (ix, iy, z0) vector, frame of computeDeltasCube
        const float bx = fmaf(fy, w.dY[0], fmaf(fx, w.dX[0], w.org[0]));
        const float by = fmaf(fy, w.dY[1], fmaf(fx, w.dX[1], w.org[1]));
        const float bz = fmaf(fy, w.dY[2], fmaf(fx, w.dX[2], w.org[2]));
        constexpr bool kMatched = kW == 1, kParallel = kW == 2;
        // FDK distance weight, constant along the column (voxel_backprojection.cu:241-249)
        float wfdk = 1.0f;
        if (kW == 0) {
            const float ww = fmaf(fy, w.wfdk[2], fmaf(fx, w.wfdk[1], w.wfdk[0]));
            wfdk = rcp(ww * ww);
        }
        // matched weight: squared in-plane source-voxel distance / DSD^2, constant along z
        float Axy = 0.0f;
        if (kMatched) {
            const float ex = w.mS[0] - fmaf(fx, w.mdv[0], w.mrv0[0]);
            const float ey = w.mS[1] - fmaf(fy, w.mdv[1], w.mrv0[1]);
            Axy = ex * ex + ey * ey;
        }
        if (kAxial) {
            const float t = kParallel ? 1.0f : w.tnum * rcp(bx);
            const float pu_raw = fmaf(by, t, w.uoff);
            const float pu = __builtin_amdgcn_fmed3f(pu_raw, 0.0f, hu);
            const float wu = __builtin_amdgcn_fractf(pu);
            const unsigned colb = (unsigned)pu * 16u;
            const float dpv = w.dZ[2] * t;
            const float pv0 = fmaf(fmaf((float)z0, w.dZ[2], bz), t, w.voff);
            float Day2 = 0.0f;
            if (kMatched) {
                const float Day = fmaf(pu_raw, w.mdet[0], w.mdet[2]);
                Day2 = fmaf(Day, Day, 1.0f);
            }
            // every gather of the column first, then the arithmetic (keeps the
            // kZ loads in flight together)
            float4 q[kZ];
            float wv[kZ];
#pragma unroll
            for (int k = 0; k < kZ; ++k) {
                const float pv = __builtin_amdgcn_fmed3f(fmaf((float)k, dpv, pv0), 0.0f, hv);
                wv[k] = __builtin_amdgcn_fractf(pv);
                q[k] = *reinterpret_cast<const float4*>(pjb + (__umul24((unsigned)pv, rowb) + colb));
            }
#pragma unroll
            for (int k = 0; k < kZ; ++k) {
                const float pv_raw = fmaf((float)k, dpv, pv0);
                const float s = bilerp(q[k], wu, wv[k]);
                if (kMatched) {
                    const float Dz = fmaf(pv_raw, w.mdet[1], w.mdet[3]);
                    const float rvz = fmaf((float)(z0 + k), w.mdv[2], w.mrv0[2]);
                    acc[k] = fmaf(s, matched_weight(fmaf(Dz, Dz, Day2), fmaf(rvz, rvz, Axy)), acc[k]);
                } else {
                    acc[k] = fmaf(s, wfdk, acc[k]);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < kZ; ++k) {
                const float fz = (float)(z0 + k);
                const float vx = fmaf(fz, w.dZ[0], bx), vy = fmaf(fz, w.dZ[1], by), vz = fmaf(fz, w.dZ[2], bz);
                const float t = kParallel ? 1.0f : w.tnum * rcp(vx);
                const float pu_raw = fmaf(vy, t, w.uoff), pv_raw = fmaf(vz, t, w.voff);
                const float pu = __builtin_amdgcn_fmed3f(pu_raw, 0.0f, hu);
                const float s = sample(pjb, (unsigned)pu * 16u, __builtin_amdgcn_fractf(pu),
                                       __builtin_amdgcn_fmed3f(pv_raw, 0.0f, hv));
                if (kMatched) {
                    const float Day = fmaf(pu_raw, w.mdet[0], w.mdet[2]);
                    const float Dz = fmaf(pv_raw, w.mdet[1], w.mdet[3]);
                    const float rvz = fmaf(fz, w.mdv[2], w.mrv0[2]);
                    acc[k] = fmaf(s, matched_weight(fmaf(Dz, Dz, fmaf(Day, Day, 1.0f)), fmaf(rvz, rvz, Axy)),
                                  acc[k]);
                } else {
                    acc[k] = fmaf(s, wfdk, acc[k]);
                }
            }
        }
    }
    store_column<kZ>(acc, vol, ix, iy, z0, nx, ny, nz, scale, ep);
}

// LDS-staged variant for axial views with the FDK or parallel-beam weight.
// The gather above moves 16 bytes through the texture-address path per voxel
// update, and that path (about 64 bytes per clock per CU) bounds it.  Here
// every wavefront copies the detector window its 8 x 8 x kZ voxel patch
// projects to (at most kWinRows x kWinCols quads, about 14 x 21 at unit
// magnification) into its own LDS region with direct global-to-LDS loads, and
// gathers from LDS (128 bytes per clock).  The window of view p + 1 is
// computed and its copy issued before the updates of view p, into the other
// half of a double buffer, so the copy overlaps the arithmetic; each
// wavefront only touches its own region, so no barrier is needed, only a wait
// for its own loads.  The window is the projection of the patch's corners:
// u and v are linear-fractional in the voxel position (the source is never
// inside the patch), so over the patch their extremes lie at corners, and
// one quad of margin on each side covers the rounding of the other lanes.
// A view whose window does not fit is gathered from global memory as in
// atb_kernel.  Both paths read the same quads with the same arithmetic, so
// the results do not depend on which one a view takes.
namespace {
constexpr int kWinCols = 16;                    // quads per staged row
constexpr int kWinLoads = 6;                    // direct loads per window, 4 rows each
constexpr int kWinRows = 4 * kWinLoads;
constexpr int kWinQuads = kWinCols * kWinRows;  // 384 quads, 6 KiB
}  // namespace

// The kZ samples of one column and view, from a staged window (LDS) or from
// the quad copy in global memory: every read is issued before any arithmetic
// needs it.  The LDS pointer is restrict-qualified: its reads then carry
// alias-scope information, and the compiler does not make them wait for the
// direct LDS loads of the next view (into the other half of the double
// buffer) that are still in flight; the kernel waits for a view's own loads
// explicitly before its updates.
template <int kZ>
__device__ __forceinline__ void samples_lds(const char* __restrict__ lds, int lb, float pv0, float dpv, float wu,
                                            float hv, float (&smp)[kZ]) {
    float4 q[kZ];
    float wv[kZ];
#pragma unroll
    for (int k = 0; k < kZ; ++k) {
        const float pv = __builtin_amdgcn_fmed3f(fmaf((float)k, dpv, pv0), 0.0f, hv);
        wv[k] = __builtin_amdgcn_fractf(pv);
        q[k] = *reinterpret_cast<const float4*>(lds + ((int)((unsigned)pv << 8) + lb));
    }
#pragma unroll
    for (int k = 0; k < kZ; ++k) smp[k] = bilerp(q[k], wu, wv[k]);
}

template <int kZ>
__device__ __forceinline__ void samples_global(const char* pjb, unsigned rowb, unsigned colb, float pv0, float dpv,
                                               float wu, float hv, float (&smp)[kZ]) {
    float4 q[kZ];
    float wv[kZ];
#pragma unroll
    for (int k = 0; k < kZ; ++k) {
        const float pv = __builtin_amdgcn_fmed3f(fmaf((float)k, dpv, pv0), 0.0f, hv);
        wv[k] = __builtin_amdgcn_fractf(pv);
        q[k] = *reinterpret_cast<const float4*>(pjb + (__umul24((unsigned)pv, rowb) + colb));
    }
#pragma unroll
    for (int k = 0; k < kZ; ++k) smp[k] = bilerp(q[k], wu, wv[k]);
}

// The workgroup of block coordinates (bx, by, bz): columns 16 bx .. + 15,
// rows 16 by .. + 15, slices kZ bz .. + kZ - 1.
template <int kW, int kZ>
__device__ __forceinline__ void atb_lds_block(const float* __restrict__ proj, int nU, int nV, int nviews,
                                              const BwdView* __restrict__ views, float* vol, int nx, int ny, int nz,
                                              float scale, const AtbUpdate& ep, int bx, int by, int bz) {
    static_assert(kW == 0 || kW == 2, "FDK or parallel-beam weight");
    __shared__ float4 win[4][2][kWinQuads];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: LDS bases stay scalar
    const int ix = bx * 16 + (wave & 1) * 8 + (lane & 7);
    const int iy = by * 16 + (wave >> 1) * 8 + (lane >> 3);
    const int z0 = bz * kZ;
    // no early exit: every lane takes part in the copies, and the window is
    // read from the corner lanes 0, 7, 56, 63; lanes outside the volume only
    // skip the store

    const long long QU = quad_det(nU);
    const long long QP = QU * quad_det(nV);
    const char* qbase = reinterpret_cast<const char*>(proj);
    const unsigned rowb = (unsigned)QU * 16u;
    const float hu = (float)(nU + 1), hv = (float)(nV + 1);
    const float fx = (float)ix, fy = (float)iy;
    // this lane's quad in each 4 x 16 block of a copy: row lane / 16, column lane % 16
    const unsigned laneoff = (unsigned)(lane >> 4) * rowb + (unsigned)(lane & 15) * 16u;
    const char* lds = reinterpret_cast<const char*>(&win[0][0][0]);

    // per view and column: u weight, quad column (bytes), v at slice z0 and
    // its step, distance weight; the window and the column's LDS offset
    struct Col {
        float wu, pv0, dpv, wt;
        unsigned colb;
        int lb;      // LDS byte offset of (row 0, this column) in the view's window; may be negative
        bool staged;
    };
    auto column = [&](int p, Col& c) {
        const BwdView& w = views[p];
        const float bx = fmaf(fy, w.dY[0], fmaf(fx, w.dX[0], w.org[0]));
        const float by = fmaf(fy, w.dY[1], fmaf(fx, w.dX[1], w.org[1]));
        const float bz = fmaf(fy, w.dY[2], fmaf(fx, w.dX[2], w.org[2]));
        c.wt = 1.0f;
        if (kW == 0) {  // voxel_backprojection.cu:241-249
            const float ww = fmaf(fy, w.wfdk[2], fmaf(fx, w.wfdk[1], w.wfdk[0]));
            c.wt = rcp(ww * ww);
        }
        const float t = kW == 2 ? 1.0f : w.tnum * rcp(bx);
        const float pu = __builtin_amdgcn_fmed3f(fmaf(by, t, w.uoff), 0.0f, hu);
        c.wu = __builtin_amdgcn_fractf(pu);
        c.colb = (unsigned)pu * 16u;
        c.dpv = w.dZ[2] * t;
        c.pv0 = fmaf(fmaf((float)z0, w.dZ[2], bz), t, w.voff);
        // rows of the first and last slice (v is monotonic along the column)
        const unsigned ra = (unsigned)__builtin_amdgcn_fmed3f(c.pv0, 0.0f, hv);
        const unsigned rb = (unsigned)__builtin_amdgcn_fmed3f(fmaf((float)(kZ - 1), c.dpv, c.pv0), 0.0f, hv);
        const unsigned rpack = min(ra, rb) | (max(ra, rb) << 16);
        // corner values, reduced on the scalar unit
        unsigned rlo = 0xffffu, rhi = 0u, clo = 0xffffffffu, chi = 0u;
        const int corner[4] = {0, 7, 56, 63};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const unsigned r = (unsigned)__builtin_amdgcn_readlane((int)rpack, corner[j]);
            const unsigned cb = (unsigned)__builtin_amdgcn_readlane((int)c.colb, corner[j]) >> 4;
            rlo = smin(rlo, r & 0xffffu);
            rhi = smax(rhi, r >> 16);
            clo = smin(clo, cb);
            chi = smax(chi, cb);
        }
        const int r0 = max((int)rlo - 1, 0), r1 = min((int)rhi + 1, nV + 1);
        const int c0 = max((int)clo - 1, 0), c1 = min((int)chi + 1, nU + 1);
        c.staged = r1 - r0 < kWinRows && c1 - c0 < kWinCols;
        const int buf = p & 1;
        c.lb = (int)c.colb - c0 * 16 - r0 * (kWinCols * 16) + (wave * 2 + buf) * (kWinQuads * 16);
        if (c.staged) {
            // rows r0 .. r0 + 4 nl - 1, columns c0 .. c0 + 15: the quads past
            // the window belong to the next row or view or to the slack after
            // the copy, and are never read.  The address is a uniform base
            // plus the lane's fixed offset.
            const char* blk = qbase + ((long long)p * QP + (long long)r0 * QU + c0) * 16;
            float4* dst = &win[wave][buf][0];
            const int nl = (r1 - r0 + 4) >> 2;
#pragma unroll
            for (int j = 0; j < kWinLoads; ++j) {
                if (j < nl) {
                    const char* bj = blk + (long long)(4 * j) * rowb;
                    __builtin_amdgcn_global_load_lds(bj + laneoff, dst + 64 * j, 16, 0, 0);
                }
            }
        }
    };

    float acc[kZ];
#pragma unroll
    for (int k = 0; k < kZ; ++k) acc[k] = 0.0f;

    // one view: wait for its window, set up and issue the next view's, then
    // the kZ updates (two Col records alternate, so nothing is copied)
    auto step = [&](int p, const Col& c, Col& next) {
        __builtin_amdgcn_s_waitcnt(0);  // this view's window has landed (nothing else is in flight)
        if (p + 1 < nviews) column(p + 1, next);
        // the bilinear arithmetic sits in each branch: with it after the
        // branch, the two gathers would be merged into generic (flat) loads
        float smp[kZ];
        if (c.staged)
            samples_lds<kZ>(lds, c.lb, c.pv0, c.dpv, c.wu, hv, smp);
        else
            samples_global<kZ>(qbase + (long long)p * QP * 16, rowb, c.colb, c.pv0, c.dpv, c.wu, hv, smp);
#pragma unroll
        for (int k = 0; k < kZ; ++k) acc[k] = fmaf(smp[k], c.wt, acc[k]);
    };
    Col ca, cb;
    if (nviews > 0) column(0, ca);
    int p = 0;
    for (; p + 1 < nviews; p += 2) {
        step(p, ca, cb);
        step(p + 1, cb, ca);
    }
    if (p < nviews) step(p, ca, cb);
    if (ix < nx && iy < ny) store_column<kZ>(acc, vol, ix, iy, z0, nx, ny, nz, scale, ep);
}

template <int kW, int kZ>
__global__ __launch_bounds__(256, 3) void atb_lds_kernel(const float* __restrict__ proj, int nU, int nV, int nviews,
                                                      const BwdView* __restrict__ views, float* vol, int nx,
                                                      int ny, int nz, float scale, AtbUpdate ep) {
    atb_lds_block<kW, kZ>(proj, nU, nV, nviews, views, vol, nx, ny, nz, scale, ep, blockIdx.x, blockIdx.y,
                          blockIdx.z);
}

// Many independent FDK backprojections with the SART update epilogue in one
// launch: job j owns blocks [begin, begin + gx gy gz) of the grid.
template <int kZ>
__global__ __launch_bounds__(256, 3) void atb_lds_multi_kernel(const AtbJob* __restrict__ jobs, int njobs) {
    const int b = blockIdx.x;
    const int j = find_job(jobs, njobs, b);
    const AtbJob& J = jobs[j];
    const int l = b - J.begin;
    AtbUpdate ep;
    ep.x = J.x;
    ep.inv_v = J.inv_v;
    ep.noneg = J.noneg;
    atb_lds_block<0, kZ>(J.quads, J.nU, J.nV, J.nviews, J.views, J.vol, J.nx, J.ny, J.nz, J.scale, ep, l % J.gx,
                         (l / J.gx) % J.gy, l / (J.gx * J.gy));
}

// Quad copy of the projections (see above), view-major, rows of quad_det(nU)
// elements.  Reads either detector layout.  Elements first, first + stride, ...
__device__ __forceinline__ void quad_projections_range(const void* __restrict__ proj, int dtype, int nU, int nV,
                                                       int nviews, int layout, float4* __restrict__ quads,
                                                       long long first, long long stride) {
    const long long QU = quad_det(nU), QV = quad_det(nV);
    const long long total = QU * QV * nviews;
    for (long long q = first; q < total; q += stride) {
        const long long u = q % QU - 1, v = (q / QU) % QV - 1, p = q / (QU * QV);
        const long long base = p * (long long)nU * nV;
        float t[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const long long uu = u + (k & 1), vv = v + (k >> 1);
            float val = 0.0f;
            if (uu >= 0 && uu < nU && vv >= 0 && vv < nV)
                val = load_storage(proj, (layout == TIGRE_LAYOUT_MATLAB) ? base + uu * nV + vv : base + vv * nU + uu,
                                   dtype);
            t[k] = val;
        }
        const float du0 = t[1] - t[0], du1 = t[3] - t[2];
        quads[q] = make_float4(t[0], du0, t[2] - t[0], du1 - du0);
    }
}

__global__ __launch_bounds__(256) void quad_projections_kernel(const void* __restrict__ proj, int dtype, int nU,
                                                               int nV, int nviews, int layout,
                                                               float4* __restrict__ quads) {
    quad_projections_range(proj, dtype, nU, nV, nviews, layout, quads,
                           (long long)blockIdx.x * blockDim.x + threadIdx.x, (long long)gridDim.x * blockDim.x);
}

// Quad copies of many FP32 projection sets (Python layout) in one launch.
__global__ __launch_bounds__(256) void quad_projections_multi_kernel(const QuadJob* __restrict__ jobs, int njobs) {
    const int b = blockIdx.x;
    const int j = find_job(jobs, njobs, b);
    const QuadJob& J = jobs[j];
    quad_projections_range(J.proj, TIGRE_DTYPE_F32, J.nU, J.nV, J.nviews, TIGRE_LAYOUT_PYTHON,
                           reinterpret_cast<float4*>(J.quads), (long long)(b - J.begin) * 256 + threadIdx.x,
                           (long long)J.blocks * 256);
}

int quad_job_blocks(int nU, int nV, int nviews) {
    const long long total = quad_det(nU) * quad_det(nV) * (long long)nviews;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    return (int)(blocks < 1 ? 1 : blocks);
}

int launch_quad_multi(const QuadJob* d_jobs, int njobs, long long blocks, hipStream_t stream) {
    TIGRE_REQUIRE(njobs > 0 && blocks > 0 && blocks < (1LL << 31), "bad multi-job launch");
    KernelTimerScope timer(TIGRE_KERNEL_RELAYOUT, stream);
    hipLaunchKernelGGL(quad_projections_multi_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, d_jobs, njobs);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

int launch_atb_multi(const AtbJob* d_jobs, int njobs, long long blocks, int zper, hipStream_t stream) {
    TIGRE_REQUIRE(njobs > 0 && blocks > 0 && blocks < (1LL << 31), "bad multi-job launch");
    TIGRE_REQUIRE(zper == 4 || zper == 8 || zper == 16, "slices per lane must be 4, 8 or 16");
    KernelTimerScope timer(TIGRE_KERNEL_ATB_FDK, stream);
    if (zper == 16)
        hipLaunchKernelGGL(atb_lds_multi_kernel<16>, dim3((unsigned)blocks), dim3(256), 0, stream, d_jobs, njobs);
    else if (zper == 8)
        hipLaunchKernelGGL(atb_lds_multi_kernel<8>, dim3((unsigned)blocks), dim3(256), 0, stream, d_jobs, njobs);
    else
        hipLaunchKernelGGL(atb_lds_multi_kernel<4>, dim3((unsigned)blocks), dim3(256), 0, stream, d_jobs, njobs);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

int launch_quad_projections(const void* proj, int dtype, int nU, int nV, int nviews, int layout, float* quads,
                            hipStream_t stream) {
    const long long total = quad_det(nU) * quad_det(nV) * (long long)nviews;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    KernelTimerScope timer(TIGRE_KERNEL_RELAYOUT, stream);
    hipLaunchKernelGGL(quad_projections_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, proj, dtype, nU, nV,
                       nviews, layout, reinterpret_cast<float4*>(quads));
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

int launch_atb(const float* quads, const HostGeometry& g, const BwdView* d_views, bool matched, bool axial,
               float* vol, int zper, const AtbUpdate& ep, hipStream_t stream) {
    // the parallel-beam backprojector serves both 'FDK' and 'matched' (_Atb.pyx:56-76)
    const int kind = g.parallel ? 2 : (matched ? 1 : 0);
    TIGRE_REQUIRE(zper == 4 || zper == 8 || zper == 16, "slices per lane must be 4, 8 or 16");
    // per-view byte offsets are 32-bit and the row stride enters a 24-bit multiply
    TIGRE_REQUIRE(quad_det(g.nDetec[0]) * 16 < (1LL << 24) &&
                      quad_det(g.nDetec[0]) * quad_det(g.nDetec[1]) * 16 < (1LL << 32),
                  "detector too large for one backprojection launch");
    const int nx = g.nVoxel[0], ny = g.nVoxel[1], nz = g.nVoxel[2];
    dim3 grid((nx + 15) / 16, (ny + 15) / 16, (nz + zper - 1) / zper);
    TIGRE_REQUIRE(grid.z <= 65535u && grid.y <= 65535u, "volume too large for one launch");
    // matched: final scaling by dVx dVy dVz / (dU dV) (voxel_backprojection2.cu:585)
    const float scale =
        kind == 1 ? (float)(g.dVoxel[0] * g.dVoxel[1] * g.dVoxel[2] / (g.dDetec[0] * g.dDetec[1])) : 1.0f;
    KernelTimerScope timer(kind == 1 ? TIGRE_KERNEL_ATB_MATCHED : TIGRE_KERNEL_ATB_FDK, stream);
#define TIGRE_ATB_LAUNCH(M, A, Z)                                                                                \
    hipLaunchKernelGGL((atb_kernel<M, A, Z>), grid, dim3(256), 0, stream, quads, g.nDetec[0], g.nDetec[1],     \
                       g.nangles, d_views, vol, nx, ny, nz, scale, ep)
#define TIGRE_ATB_Z(M, A)                                  \
    do {                                                   \
        if (zper == 16) TIGRE_ATB_LAUNCH(M, A, 16);        \
        else if (zper == 8) TIGRE_ATB_LAUNCH(M, A, 8);     \
        else TIGRE_ATB_LAUNCH(M, A, 4);                    \
    } while (0)
#define TIGRE_ATB_LDS_LAUNCH(M, Z)                                                                            \
    hipLaunchKernelGGL((atb_lds_kernel<M, Z>), grid, dim3(256), 0, stream, quads, g.nDetec[0], g.nDetec[1], \
                       g.nangles, d_views, vol, nx, ny, nz, scale, ep)
#define TIGRE_ATB_LDS_Z(M)                                 \
    do {                                                   \
        if (zper == 16) TIGRE_ATB_LDS_LAUNCH(M, 16);       \
        else if (zper == 8) TIGRE_ATB_LDS_LAUNCH(M, 8);    \
        else TIGRE_ATB_LDS_LAUNCH(M, 4);                   \
    } while (0)
    // the staged kernel packs two row indices into 16 bits each
    const bool staged = axial && kind != 1 && g.nDetec[1] + 1 < 65536;
    if (kind == 1) {
        if (axial) TIGRE_ATB_Z(1, true);
        else TIGRE_ATB_Z(1, false);
    } else if (kind == 0) {
        if (staged) TIGRE_ATB_LDS_Z(0);
        else if (axial) TIGRE_ATB_Z(0, true);
        else TIGRE_ATB_Z(0, false);
    } else {
        if (staged) TIGRE_ATB_LDS_Z(2);
        else if (axial) TIGRE_ATB_Z(2, true);
        else TIGRE_ATB_Z(2, false);
    }
#undef TIGRE_ATB_LDS_Z
#undef TIGRE_ATB_LDS_LAUNCH
#undef TIGRE_ATB_Z
#undef TIGRE_ATB_LAUNCH
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

}  // namespace tigre
