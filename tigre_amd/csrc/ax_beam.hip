// Beam-staged Siddon forward projector for gfx950.
//
// A 1024-thread workgroup traces the rays of a 16 x 16 detector tile in up to
// four views that share a principal axis P (the horizontal axis, x or y, along
// which their rays advance fastest) and its direction.  Their rays form a thin
// beam.  The workgroup sweeps the volume along P in slabs of a few planes; for
// each slab it copies the box of voxels the beam crosses into LDS (16-byte
// loads of rows of a re-laid-out copy of the volume), and every ray walks its
// Siddon crossings inside the slab reading voxels from LDS.  The gathers of
// the walk thus never reach the texture/L1 path, which bounds a gather-driven
// walk; the slab copies are coalesced and are shared by the 1024 rays.
//
// The staging copy ("stage") of the volume is [p][z][a], a = the other
// horizontal axis fastest, with a zero border: p in -1 .. nP + 1, z in -1 ..
// nz, a in -4 .. round_up(nA, 4) + 3.  Every voxel a walk can reach (indices
// -1 .. n) therefore exists, rows start on 16-byte boundaries, and so does the
// upper interpolation tap of a sample clamped to the far face (index n + 1:
// row nz + 1 of plane p is row -1 of plane p + 1, and the extra zero plane
// nP + 1 ends the copy), as do the whole rows of the last plane a slab box
// copies.
//
// Planning: every wavefront evaluates the 16 corner rays of the tile (four
// corners in each view) and computes the box origins of a share of the slabs;
// the box extents, maxed over the slabs, are the same for the whole sweep, so
// a ray's LDS offset carries from one slab to the next by a uniform shift.
// Slabs are double-buffered: slab s + 1 streams into LDS by direct
// global-to-LDS loads while the rays walk slab s, and one barrier ends each
// slab.  A tile whose box does not fit the LDS buffer even with one plane per
// slab walks the staging copy in global memory instead (same arithmetic).
#include <cmath>

#include "common.hpp"
#include "ray.hpp"

namespace tigre {

namespace {

constexpr int kBeamTile = 16;       // detector pixels per tile edge
constexpr int kBeamViews = 4;       // views per workgroup
constexpr int kBeamThreads = kBeamViews * kBeamTile * kBeamTile;
constexpr int kSlabCap = 9216;      // floats per LDS slab buffer (two per workgroup)
constexpr int kSlabPlanesMax = 8;   // principal planes per slab, at most (the tunable limit is <= this)

// Minimum / maximum over lanes 0..15 (the 16 corner rays; the other rows of
// the wavefront hold copies), by four DPP row shifts, returned as a uniform.
template <int kCtrl>
__device__ __forceinline__ float row_shr(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), kCtrl, 0xF, 0xF, false));
}
__device__ __forceinline__ float corner_min(float x) {
    x = fminf(x, row_shr<0x111>(x));
    x = fminf(x, row_shr<0x112>(x));
    x = fminf(x, row_shr<0x114>(x));
    x = fminf(x, row_shr<0x118>(x));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 15));
}
__device__ __forceinline__ float corner_max(float x) {
    x = fmaxf(x, row_shr<0x111>(x));
    x = fmaxf(x, row_shr<0x112>(x));
    x = fmaxf(x, row_shr<0x114>(x));
    x = fmaxf(x, row_shr<0x118>(x));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 15));
}

// Extent of the beam over principal planes lo .. hi: the beam's lines are
// A(p) = cA + kA p, z(p) = cz + kz p (corner ray c in lane c; the other rows of
// the wavefront hold copies), the voxels of the planes span P coordinates
// [lo, hi + 1], and a margin of one voxel covers rounding; `tap` = 1 adds the
// upper interpolation tap.  Clipped to -1 .. n + tap (no walk or tap leaves
// it: a sample clamped to the face n reads the zero tap n + 1 with weight 0);
// a0 is aligned down to 4 for 16-byte rows.
struct BoxExtent {
    int a0, a1, z0, z1;
};
__device__ inline BoxExtent beam_extent(int lo, int hi, float cA, float kA, float cz, float kz, int nA, int nz,
                                        int tap) {
    const float p0 = (float)lo, p1 = (float)(hi + 1);
    const float A0 = fmaf(kA, p0, cA), A1 = fmaf(kA, p1, cA);
    const float Z0 = fmaf(kz, p0, cz), Z1 = fmaf(kz, p1, cz);
    float amin = corner_min(fminf(A0, A1)), amax = corner_max(fmaxf(A0, A1));
    float zmin = corner_min(fminf(Z0, Z1)), zmax = corner_max(fmaxf(Z0, Z1));
    amin = fminf(fmaxf(amin, -2.0f), (float)nA + 1.0f);
    amax = fminf(fmaxf(amax, -2.0f), (float)nA + 1.0f);
    zmin = fminf(fmaxf(zmin, -2.0f), (float)nz + 1.0f);
    zmax = fminf(fmaxf(zmax, -2.0f), (float)nz + 1.0f);
    BoxExtent e;
    e.a0 = max((int)floorf(amin) - 1, -1) & ~3;
    e.a1 = min((int)floorf(amax) + 1 + tap, nA + tap);
    e.z0 = max((int)floorf(zmin) - 1, -1);
    e.z1 = min((int)floorf(zmax) + 1 + tap, nz + tap);
    return e;
}

// Slab s of a sweep with `planes` planes per slab along direction dir over
// principal indices -1 .. nP.
__device__ __forceinline__ void slab_range(int s, int planes, int dir, int nP, int& lo, int& hi) {
    if (dir > 0) {
        lo = -1 + s * planes;
        hi = min(lo + planes - 1, nP);
    } else {
        hi = nP - s * planes;
        lo = max(hi - planes + 1, -1);
    }
}

struct StageGeom {
    int NZ, NA;  // bordered extents of the staging copy (z, a)
};

__device__ __forceinline__ unsigned long long stage_index(const StageGeom& s, int p, int z, int a) {
    return ((unsigned long long)(p + 1) * (unsigned)s.NZ + (unsigned)(z + 1)) * (unsigned)s.NA + (unsigned)(a + 4);
}

// Per-ray Siddon state (Siddon_projection.cu:110-268 restated: exact path
// through the voxel grid, crossings taken in increasing parameter order).
// The crossing counts k per axis give the next crossing on each axis,
// t0 + k dt (recomputed exactly, never accumulated).
struct RayWalk {
    float tx, ty, tz, t0x, t0y, t0z, dtx, dty, dtz;
    float kx, ky, kz;      // crossings so far per axis (exact small integers)
    unsigned off;          // voxel offset in the current source (LDS box or staging copy)
    unsigned dX, dY, dZ;   // offset steps of one crossing on each axis
    int kexit;             // principal crossing count at which the ray leaves the slab
    float T, tc, sum;
};

// Walks crossings until the ray leaves the volume (returns true) or its
// principal crossing count reaches kexit (returns false).  Voxel reads are
// at offset (off + slot) from `base`, in bytes for the LDS slabs (base is the
// shared buffer, a constant address folded into the load; slot selects the
// half of the double buffer) and in elements for the staging copy in global
// memory.  The axis of each crossing is selected without branches, the
// principal axis P first on ties, then the other horizontal axis A, then z;
// every step reads one voxel, issued one step before its value is needed;
// the steps are unrolled four to a guard count.
// Leaving the slab and leaving the volume are one test, tm >= Te with
// Te = min(T, time of the slab's last principal crossing): that time is the
// same fma as the crossing's own, and P wins ties, so the test can only
// pass on the principal crossing itself (or at the volume's exit).  Segments
// end at the crossing times themselves: at a slab's exit the crossing time is
// Te exactly, and the volume's exit T is itself a crossing, whose computed
// time differs from T by rounding only (a ray that stops short of it reads
// the zero border for its next segment), so no segment is clamped to Te.
template <int kP, bool kLds>
__device__ inline bool walk(RayWalk& r, const float* __restrict__ base, unsigned slot, int guard) {
    // locals only, named by role (P principal, A other horizontal, z): a
    // conditional between struct members would be compiled as a select of
    // their addresses and force the struct into scratch memory
    float tp = kP == 0 ? r.tx : r.ty, ta = kP == 0 ? r.ty : r.tx, tz = r.tz;
    float kp = kP == 0 ? r.kx : r.ky, ka = kP == 0 ? r.ky : r.kx, kz = r.kz;
    const float t0p = kP == 0 ? r.t0x : r.t0y, t0a = kP == 0 ? r.t0y : r.t0x, t0z = r.t0z;
    const float dtp = kP == 0 ? r.dtx : r.dty, dta = kP == 0 ? r.dty : r.dtx, dtz = r.dtz;
    // LDS: byte offsets; global staging copy: element offsets (its byte offsets can pass 2^32)
    constexpr int kSh = kLds ? 2 : 0;
    const unsigned dP = (kP == 0 ? r.dX : r.dY) << kSh, dA = (kP == 0 ? r.dY : r.dX) << kSh, dZ = r.dZ << kSh;
    unsigned off = (r.off + slot) << kSh;
    float tc = r.tc, sum = r.sum;
    const float T = r.T;
    const float Te = r.kexit >= 1 ? fminf(T, fmaf((float)(r.kexit - 1), dtp, t0p)) : T;
    auto read = [&](unsigned o) {
        return kLds ? *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + o) : base[o];
    };
    float tm = tc;
    float va = read(off), vb = 0.0f;
#define TIGRE_BEAM_STEP(CUR, NEXT)                                       \
    {                                                                    \
        tm = fminf(tp, fminf(ta, tz));                                   \
        sum = fmaf(tm - tc, CUR, sum);                                   \
        tc = tm;                                                         \
        const bool bp = tp == tm, ba = ta == tm;                         \
        kp += bp ? 1.0f : 0.0f;                                          \
        ka += (ba && !bp) ? 1.0f : 0.0f;                                 \
        kz += (bp || ba) ? 0.0f : 1.0f;                                  \
        off += bp ? dP : (ba ? dA : dZ);                                 \
        tp = fmaf(kp, dtp, t0p);                                         \
        ta = fmaf(ka, dta, t0a);                                         \
        tz = fmaf(kz, dtz, t0z);                                         \
        if (tm >= Te) break;                                             \
        NEXT = read(off);                                                \
    }
    for (; guard > 0; --guard) {
        TIGRE_BEAM_STEP(va, vb)
        TIGRE_BEAM_STEP(vb, va)
        TIGRE_BEAM_STEP(va, vb)
        TIGRE_BEAM_STEP(vb, va)
    }
#undef TIGRE_BEAM_STEP
    r.tx = kP == 0 ? tp : ta;
    r.ty = kP == 0 ? ta : tp;
    r.tz = tz;
    r.kx = kP == 0 ? kp : ka;
    r.ky = kP == 0 ? ka : kp;
    r.kz = kz;
    r.off = (off >> kSh) - slot;
    r.tc = tc;
    r.sum = sum;
    return tm >= T;
}

// Workgroup table entry: views [first, first + count) of the ordered view
// records share principal axis `principal` and direction `dir`.
struct BeamGroup {
    int first, count, principal, dir;
};

constexpr int kMaxSlabs = 272;

struct SweepShared {
    float buf[2 * kSlabCap];
    int2 origin[kMaxSlabs];  // (a0, z0) of each slab box
    int ext[2];              // box extents na (multiple of 4) and nzb, maxed over the slabs
};

// Per-ray state of the interpolated projector (ray_interpolated_projection.cu
// :117-206 restated): samples m = 0 .. count - 1 at q + m f (voxel-centre index
// frame), already clipped to the ones inside the zero-bordered volume.
struct RaySamples {
    float qP, qA, qz, fP, fA, fz;
    int m, count;
    float sum;
};

// Trilinear (kMode 1) or nearest (kMode 2) sample of a box whose fastest axis
// is A: `src` holds planes P, rows z, columns A with strides SP and SZ; (pc, zc,
// ac) is the lower tap relative to the box origin.
template <int kMode>
__device__ __forceinline__ float box_sample(const float* __restrict__ src, int SP, int SZ, float P, float A, float Z,
                                            int P0, int Z0, int A0) {
    // tap indices relative to the box origin are >= 0 and < 2^24, strides < 2^24
    if (kMode == 2) {
        const unsigned pc = (unsigned)((int)floorf(P + 0.5f) - P0), zc = (unsigned)((int)floorf(Z + 0.5f) - Z0),
                       ac = (unsigned)((int)floorf(A + 0.5f) - A0);
        return src[__umul24(pc, (unsigned)SP) + __umul24(zc, (unsigned)SZ) + ac];
    }
    const float flP = floorf(P), flA = floorf(A), flZ = floorf(Z);
    const float wP = P - flP, wA = A - flA, wZ = Z - flZ;
    const unsigned o = __umul24((unsigned)((int)flP - P0), (unsigned)SP) +
                       __umul24((unsigned)((int)flZ - Z0), (unsigned)SZ) + (unsigned)((int)flA - A0);
    const float c00 = fmaf(wA, src[o + 1] - src[o], src[o]);
    const float c10 = fmaf(wA, src[o + SZ + 1] - src[o + SZ], src[o + SZ]);
    const float c01 = fmaf(wA, src[o + SP + 1] - src[o + SP], src[o + SP]);
    const float c11 = fmaf(wA, src[o + SP + SZ + 1] - src[o + SP + SZ], src[o + SP + SZ]);
    const float c0 = fmaf(wZ, c10 - c00, c00), c1 = fmaf(wZ, c11 - c01, c01);
    return fmaf(wP, c1 - c0, c0);
}

// Samples of one slab (cells lo .. hi along P, box origin (a0, z0), LDS or
// global `src`); stops at the first sample whose cell lies beyond the slab.
// Returns true when the ray has no samples left.  The samples of the slab are
// counted once (P advances monotonically with the sample index), so the loop
// tests one bound per sample.  kLds: offsets inside a box are below the slab
// capacity (far below 2^24) and are formed from box-relative cell indices in
// float arithmetic, exactly; the staging copy in global memory needs the
// integer form.
template <int kMode, bool kLds>
__device__ inline bool sample_slab(RaySamples& r, const float* __restrict__ base, unsigned slot, int SP, int SZ,
                                   int P0, int Z0, int A0, int lo, int hi, int dir, float hP, float hA, float hz) {
    int m = r.m;
    const int count = r.count;
    const float qP = r.qP, qA = r.qA, qz = r.qz, fP = r.fP, fA = r.fA, fz = r.fz;
    float sum = r.sum;
    // positions stay in [-1, n] (rounding at the clipped ends), so the taps
    // stay in the zero-bordered range
    auto posP = [&](int k) { return __builtin_amdgcn_fmed3f(fmaf((float)k, fP, qP), -1.0f, hP); };
    auto cell = [&](int k) { return (int)floorf(kMode == 2 ? posP(k) + 0.5f : posP(k)); };
    auto inside = [&](int k) { return dir > 0 ? cell(k) <= hi : cell(k) >= lo; };
    // A sample clamped to the far face P = nP lies in cell nP, beyond every
    // slab (both its taps are on the zero border, so it contributes exactly
    // 0).  A sweep towards +P meets it last and stops before it; a sweep
    // towards -P meets it first, and would take it into the first slab with
    // its upper tap outside the box: skip it.
    if (dir < 0)
        while (m < count && cell(m) > hi) ++m;
    // first sample beyond the slab: an estimate, then corrected with the
    // per-sample test itself
    int me = count;
    if (fP != 0.0f) {
        const float lim = (dir > 0 ? (float)hi + 1.0f : (float)lo) - (kMode == 2 ? 0.5f : 0.0f);
        const float est = ceilf((lim - qP) / fP);
        me = est < (float)m ? m : (est > (float)count ? count : (int)est);
        while (me > m && !inside(me - 1)) --me;
        while (me < count && inside(me)) ++me;
    }
    const bool done = me >= count;
    // trilinear: the 8 taps of the last cell and their four differences along
    // A stay in registers; consecutive samples (half a voxel apart at the
    // default accuracy) mostly share a cell, so only the lanes that moved to a
    // new cell read again.  The sample index runs as an exact float.
    const float fP0 = (float)P0, fSP = (float)SP, fSZ = (float)SZ;
    // A0 + Z0 SZ - slot, so that fma(flZ, SZ, flA - cza) = (flZ - Z0) SZ + flA - A0 + slot
    // (exact: every term is an integer of magnitude below 2^24)
    const float cza = (float)(A0 + Z0 * SZ - (int)slot);
    unsigned cached = 0xffffffffu;
    float c000 = 0.f, c010 = 0.f, c100 = 0.f, c110 = 0.f, d00 = 0.f, d10 = 0.f, d01 = 0.f, d11 = 0.f;
    const float fme = (float)me;
    float fm = (float)m;
    for (; fm < fme; fm += 1.0f) {
        const float P = __builtin_amdgcn_fmed3f(fmaf(fm, fP, qP), -1.0f, hP);
        const float A = __builtin_amdgcn_fmed3f(fmaf(fm, fA, qA), -1.0f, hA);
        const float Z = __builtin_amdgcn_fmed3f(fmaf(fm, fz, qz), -1.0f, hz);
        if (kMode == 2) {
            sum += box_sample<kMode>(base + slot, SP, SZ, P, A, Z, P0, Z0, A0);
        } else {
            const float flP = floorf(P), flA = floorf(A), flZ = floorf(Z);
            const float wP = P - flP, wA = A - flA, wZ = Z - flZ;
            const unsigned o =
                kLds ? (unsigned)fmaf(flP - fP0, fSP, fmaf(flZ, fSZ, flA - cza))
                     : __umul24((unsigned)((int)flP - P0), (unsigned)SP) +
                           __umul24((unsigned)((int)flZ - Z0), (unsigned)SZ) + (unsigned)((int)flA - A0);
            if (o != cached) {
                const float* t = base + o;
                c000 = t[0];
                c010 = t[SZ];
                c100 = t[SP];
                c110 = t[SP + SZ];
                d00 = t[1] - c000;
                d10 = t[SZ + 1] - c010;
                d01 = t[SP + 1] - c100;
                d11 = t[SP + SZ + 1] - c110;
                cached = o;
            }
            // the arithmetic of box_sample, tap for tap
            const float a00 = fmaf(wA, d00, c000);
            const float a10 = fmaf(wA, d10, c010);
            const float a01 = fmaf(wA, d01, c100);
            const float a11 = fmaf(wA, d11, c110);
            const float b0 = fmaf(wZ, a10 - a00, a00), b1 = fmaf(wZ, a11 - a01, a01);
            sum += fmaf(wP, b1 - b0, b0);
        }
    }
    r.m = (int)fm;
    r.sum = sum;
    return done;
}

// One workgroup: the rays of a 16 x 16 detector tile in up to four views.
// kMode 0: Siddon; 1: trilinear interpolation; 2: nearest-voxel sampling.
template <int kP, int kMode>
__device__ inline void beam_sweep(const float* __restrict__ stage, const StageGeom& sg, int nx, int ny, int nz,
                                  const FwdView* __restrict__ views, const BeamGroup& grp, int tile_u, int tile_v,
                                  int stack, int nU, int nV, long long su, long long sv, long long sview,
                                  float* __restrict__ out, double acc, double dvx, double dvy, double dvz,
                                  int max_planes, SweepShared& sh) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int nP = kP == 0 ? nx : ny, nA = kP == 0 ? ny : nx;
    const int dir = grp.dir;
    // Siddon visits principal voxels -1 .. nP; samples have lower taps in cells
    // -1 .. nP - 1 and upper taps one plane further
    const int last_cell = kMode == 0 ? nP : nP - 1;
    const int tap = kMode == 0 ? 0 : 1;

    // ---- the lane's ray (FP64 set-up) ----
    // Slot s (four wavefronts) traces view first + s of the group in tile
    // (tile_u, tile_v).  Stacked, the slots trace `stack` tiles of each of the
    // 4 / stack views a group may hold: two (groups of at most two views), the
    // tiles (2 tile_u + (s >> 1), tile_v) of view s & 1; four (groups of one
    // view), the 2 x 2 tiles (2 tile_u + (s & 1), 2 tile_v + (s >> 1)).
    const int slot = wave >> 2;
    const int per = 4 / stack;  // views per workgroup
    auto slot_view = [&](int s) { return grp.first + min(s % per, grp.count - 1); };
    auto slot_tu = [&](int s) { return stack == 1 ? tile_u : 2 * tile_u + (stack == 2 ? s >> 1 : s & 1); };
    auto slot_tv = [&](int s) { return stack == 4 ? 2 * tile_v + (s >> 1) : tile_v; };
    const int u = slot_tu(slot) * kBeamTile + (lane & 15);
    const int v = slot_tv(slot) * kBeamTile + (wave & 3) * 4 + (lane >> 4);
    const bool active = slot % per < grp.count && u < nU && v < nV;
    const FwdView& w = views[slot_view(slot)];
    RayWalk r;
    r.sum = 0.0f;
    RaySamples rs;
    rs.sum = 0.0f;
    int state = 0;  // 0: no ray or finished, 1: waiting for the slab of its entry, 2: walking
    int ip0 = 0, ia0 = 0, iz0 = 0, sx = 0, sy = 0, sz = 0;
    double len = 0.0;
    if (active) {
        const double pv = (double)(nV - 1 - v);  // detector row flip, Siddon_projection.cu:141
        double sp[3];
        ray_source(w, u, pv, sp);
        double rx = ray_dir(w, u, pv, 0);
        double ry = ray_dir(w, u, pv, 1);
        double rz = ray_dir(w, u, pv, 2);
        if (kMode == 0) {
            if (w.pad[1] != 0.0) parallel_siddon_height(sp, rx, ry, rz, nx, ny);
            // components below 1e-3 voxel are treated as exactly zero (:155-159)
            if (fabs(rx) < 1e-3) rx = 0.0;
            if (fabs(ry) < 1e-3) ry = 0.0;
            if (fabs(rz) < 1e-3) rz = 0.0;
            len = sqrt(rx * rx * dvx * dvx + ry * ry * dvy * dvy + rz * rz * dvz * dvz);  // mm (:241, :267)
            double am = -INFINITY, aM = INFINITY;
            bool hit = slab(sp[0], rx, nx, am, aM);
            hit = hit && slab(sp[1], ry, ny, am, aM);
            hit = hit && slab(sp[2], rz, nz, am, aM);
            if (hit && am < aM) {  // a ray that misses the volume projects to 0 (:179-180)
                const Axis ax = axis_setup(sp[0], rx, am);
                const Axis ay = axis_setup(sp[1], ry, am);
                const Axis az = axis_setup(sp[2], rz, am);
                r.tx = r.t0x = ax.t0;
                r.ty = r.t0y = ay.t0;
                r.tz = r.t0z = az.t0;
                r.dtx = ax.dt;
                r.dty = ay.dt;
                r.dtz = az.dt;
                r.kx = r.ky = r.kz = 0.0f;
                r.T = (float)(aM - am);
                r.tc = 0.0f;
                ip0 = kP == 0 ? ax.idx : ay.idx;
                ia0 = kP == 0 ? ay.idx : ax.idx;
                iz0 = az.idx;
                sx = ax.step;
                sy = ay.step;
                sz = az.step;
                state = 1;
            }
        } else {
            // samples per ray and step (ray_interpolated_projection.cu:164-169)
            const double n = sample_count(w, sqrt(rx * rx + ry * ry + rz * rz), acc);
            if (n >= 1.0 && n < 1e9) {  // else: degenerate ray (pixel at the source), 0
                const double st[3] = {rx / n, ry / n, rz / n};
                // sample range [first, last] of the reference (:182-193), clipped to
                // the samples strictly inside (-1, n) on every axis (the others are
                // exactly zero)
                const double last = (w.cap < n) ? ceil(w.cap) : n;
                double lo = -INFINITY, hi = INFINITY;
                const int nn[3] = {nx, ny, nz};
                bool empty = false;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    if (st[a] == 0.0) {
                        if (!(sp[a] > -1.0 && sp[a] < (double)nn[a])) empty = true;
                    } else {
                        const double a0 = (-1.0 - sp[a]) / st[a], a1 = ((double)nn[a] - sp[a]) / st[a];
                        lo = fmax(lo, fmin(a0, a1));
                        hi = fmin(hi, fmax(a0, a1));
                    }
                }
                const double i0 = fmax(w.first, floor(lo));
                const double i1 = fmin(last, ceil(hi));
                len = sqrt(st[0] * dvx * st[0] * dvx + st[1] * dvy * st[1] * dvy + st[2] * dvz * st[2] * dvz);
                if (!empty && i0 <= i1) {
                    // positions relative to the first sample keep the FP32 loop accurate
                    rs.qP = (float)(sp[kP] + i0 * st[kP]);
                    rs.qA = (float)(sp[1 - kP] + i0 * st[1 - kP]);
                    rs.qz = (float)(sp[2] + i0 * st[2]);
                    rs.fP = (float)st[kP];
                    rs.fA = (float)st[1 - kP];
                    rs.fz = (float)st[2];
                    rs.m = 0;
                    rs.count = (int)(i1 - i0) + 1;
                    state = 2;
                }
            }
        }
    }

    // ---- corner rays of the tile (lanes 0..15 of every wavefront) ----
    float cA, kA, cz, kz;
    {
        const int c = lane & 15;
        const int cs = c >> 2;
        const FwdView& cw = views[slot_view(cs)];
        const int cu = min(slot_tu(cs) * kBeamTile + ((c & 1) ? kBeamTile - 1 : 0), nU - 1);
        const int cvv = min(slot_tv(cs) * kBeamTile + ((c & 2) ? kBeamTile - 1 : 0), nV - 1);
        const double pv = (double)(nV - 1 - cvv);
        double csrc[3];
        ray_source(cw, cu, pv, csrc);
        const double dP = cw.org[kP] + cu * cw.du[kP] + pv * cw.dv[kP] - csrc[kP];
        const double dA = cw.org[1 - kP] + cu * cw.du[1 - kP] + pv * cw.dv[1 - kP] - csrc[1 - kP];
        const double dZ = cw.org[2] + cu * cw.du[2] + pv * cw.dv[2] - csrc[2];
        const double ka = dA / dP, kzz = dZ / dP;
        kA = (float)ka;
        cA = (float)(csrc[1 - kP] - csrc[kP] * ka);
        kz = (float)kzz;
        cz = (float)(csrc[2] - csrc[kP] * kzz);
    }

    // ---- plan: box origins of every slab (in parallel over the wavefronts)
    // and the extents shared by all of them; fewer planes per slab until the
    // box fits an LDS buffer ----
    int planes = 0, nslabs = 0, na = 0, nzb = 0;
    bool staged = false;
    for (int attempt = 0; attempt < 5 && !staged; ++attempt) {
        const int pl_try = attempt == 0 ? max_planes
                                        : min(max_planes, attempt == 1 ? 6 : (attempt == 2 ? 4 : (attempt == 3 ? 2 : 1)));
        const int ns = (last_cell + 2 + pl_try - 1) / pl_try;
        if (ns > kMaxSlabs) continue;
        if (tid == 0) sh.ext[0] = sh.ext[1] = 0;
        __syncthreads();
        for (int s = wave; s < ns; s += kBeamThreads / 64) {
            int lo, hi;
            slab_range(s, pl_try, dir, last_cell, lo, hi);
            const BoxExtent e = beam_extent(lo, hi, cA, kA, cz, kz, nA, nz, tap);
            if (lane == 0) {
                sh.origin[s] = make_int2(e.a0, e.z0);
                atomicMax(&sh.ext[0], (e.a1 - e.a0 + 4) & ~3);
                atomicMax(&sh.ext[1], e.z1 - e.z0 + 1);
            }
        }
        __syncthreads();
        na = sh.ext[0];
        nzb = sh.ext[1];
        if (na * nzb * (pl_try + tap) <= kSlabCap) {
            staged = true;
            planes = pl_try;
            nslabs = ns;
        }
        __syncthreads();  // ext is reset by the next attempt
    }

    if (!staged) {
        // The beam's box does not fit an LDS buffer: read the staging copy.
        if (kMode == 0) {
            if (state == 1) {
                r.off = (unsigned)stage_index(sg, ip0, iz0, ia0);
                r.dX = (unsigned)(sx * (kP == 0 ? sg.NZ * sg.NA : 1));
                r.dY = (unsigned)(sy * (kP == 1 ? sg.NZ * sg.NA : 1));
                r.dZ = (unsigned)(sz * sg.NA);
                r.kexit = -1;  // never reached
                walk<kP, false>(r, stage, 0u, (nx + ny + nz + 10) / 4);
            }
        } else if (state == 2) {
            // box origin (-1, -1, -4): the bordered staging copy itself
            sample_slab<kMode, false>(rs, stage, 0u, sg.NZ * sg.NA, sg.NA, -1, -1, -4, -1, last_cell, dir,
                                      (float)nP, (float)nA, (float)nz);
        }
    } else {
        const int SP = nzb * na;
        const int splanes = planes + tap;  // staged planes per slab
        if (kMode == 0) {
            // offset steps inside a box
            r.dX = (unsigned)(sx * (kP == 0 ? SP : 1));
            r.dY = (unsigned)(sy * (kP == 1 ? SP : 1));
            r.dZ = (unsigned)(sz * na);
        }
        // this thread's quads of a box: row (plane * nzb + z) and 4-float column
        const int nq4 = na >> 2, nq = splanes * nzb * nq4;
        unsigned long long qoff[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int q = min(tid + j * kBeamThreads, nq - 1);
            const int row = q / nq4, aq = q - row * nq4;
            const int pl = row / nzb, zr = row - pl * nzb;
            qoff[j] = ((unsigned long long)pl * (unsigned)sg.NZ + (unsigned)zr) * (unsigned)sg.NA + 4u * aq;
        }
        const int wave_first = tid & ~63;
        // Direct global-to-LDS copies (16 bytes per lane) of slab s: quad q lands
        // at buf[4 q]; the LDS base of each instruction is uniform per wavefront.
        auto fetch = [&](int s, float* dst) {
            int lo, hi;
            slab_range(s, planes, dir, last_cell, lo, hi);
            const int2 o = sh.origin[s];
            const unsigned long long base = stage_index(sg, lo, o.y, o.x);
            const int nqs = (hi - lo + 1 + tap) * nzb * nq4;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                if (wave_first + j * kBeamThreads < nqs) {
                    if (tid + j * kBeamThreads < nqs)
                        __builtin_amdgcn_global_load_lds(stage + base + qoff[j], dst + 4 * (wave_first + j * kBeamThreads),
                                                         16, 0, 0);
                }
            }
        };
        fetch(0, sh.buf);
        __builtin_amdgcn_s_waitcnt(0);  // the direct LDS loads have landed
        __syncthreads();
        int plo = 0, pa0 = 0, pz0 = 0;
        for (int s = 0; s < nslabs; ++s) {
            if (s + 1 < nslabs) fetch(s + 1, sh.buf + ((s + 1) & 1) * kSlabCap);
            int lo, hi;
            slab_range(s, planes, dir, last_cell, lo, hi);
            const int2 o = sh.origin[s];
            const unsigned slot = (unsigned)((s & 1) * kSlabCap);
            if (kMode == 0) {
                if (state == 2) {
                    // paused at the first plane of this slab: re-base the offset
                    r.off += (unsigned)(-((lo - plo) * SP + (o.y - pz0) * na + (o.x - pa0)));
                    r.kexit += planes;
                } else if (state == 1 && ip0 >= lo && ip0 <= hi) {
                    r.off = (unsigned)(((ip0 - lo) * nzb + (iz0 - o.y)) * na + (ia0 - o.x));
                    r.kexit = dir > 0 ? hi + 1 - ip0 : ip0 - lo + 1;
                    state = 2;
                }
                if (state == 2) {
                    // crossings inside one slab: its planes plus the box's a and z planes
                    if (walk<kP, true>(r, sh.buf, slot, (planes + na + nzb + 7) / 4)) state = 0;
                }
            } else if (state == 2) {
                if (sample_slab<kMode, true>(rs, sh.buf, slot, SP, na, lo, o.y, o.x, lo, hi, dir, (float)nP,
                                             (float)nA, (float)nz))
                    state = 0;
            }
            plo = lo;
            pa0 = o.x;
            pz0 = o.y;
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
        }
    }

    if (active) {
        float* dst = out + (long long)w.pad[0] * sview + (long long)u * su + (long long)v * sv;
        *dst = (float)((double)(kMode == 0 ? r.sum : rs.sum) * len);
    }
}

// Workgroup L of one projection: tile L % tiles of group L / tiles (stacked
// by two: tiles of 2 x 1 detector tiles; by four: 2 x 2).
template <int kMode>
__device__ __forceinline__ void beam_block(const float* __restrict__ stage_x, const float* __restrict__ stage_y,
                                           int nx, int ny, int nz, int na_x, int na_y,
                                           const FwdView* __restrict__ views, const BeamGroup* __restrict__ groups,
                                           int ngroups, int nU, int nV, long long su, long long sv, long long sview,
                                           float* __restrict__ out, double acc, double dvx, double dvy, double dvz,
                                           int max_planes, int stack, int L, SweepShared& sh) {
    const int edge_u = stack >= 2 ? 2 * kBeamTile : kBeamTile, edge_v = stack == 4 ? 2 * kBeamTile : kBeamTile;
    const int tilesU = (nU + edge_u - 1) / edge_u, tilesV = (nV + edge_v - 1) / edge_v;
    const int tiles = tilesU * tilesV;
    const int gi = L / tiles, tile = L % tiles;
    if (gi >= ngroups) return;  // uniform per workgroup
    const BeamGroup grp = groups[gi];
    StageGeom sg;
    sg.NZ = nz + 2;
    if (grp.principal == 0) {
        sg.NA = na_x + 8;
        beam_sweep<0, kMode>(stage_x, sg, nx, ny, nz, views, grp, tile % tilesU, tile / tilesU, stack, nU, nV, su, sv,
                             sview, out, acc, dvx, dvy, dvz, max_planes, sh);
    } else {
        sg.NA = na_y + 8;
        beam_sweep<1, kMode>(stage_y, sg, nx, ny, nz, views, grp, tile % tilesU, tile / tilesU, stack, nU, nV, su, sv,
                             sview, out, acc, dvx, dvy, dvz, max_planes, sh);
    }
}

// Two workgroups per CU (8 wavefronts per SIMD): the register budget is held
// to that occupancy, which the slab barriers need to keep the SIMDs busy.
template <int kMode>
__global__ __launch_bounds__(kBeamThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void ax_beam_kernel(
    const float* __restrict__ stage_x, const float* __restrict__ stage_y, int nx, int ny, int nz, int na_x, int na_y,
    const FwdView* __restrict__ views, const BeamGroup* __restrict__ groups, int ngroups, int nU, int nV,
    long long su, long long sv, long long sview, float* __restrict__ out, double acc, double dvx, double dvy,
    double dvz, int max_planes, int stack) {
    __shared__ SweepShared sh;
    beam_block<kMode>(stage_x, stage_y, nx, ny, nz, na_x, na_y, views, groups, ngroups, nU, nV, su, sv, sview, out,
                      acc, dvx, dvy, dvz, max_planes, stack, xcd_logical_block(blockIdx.x, gridDim.x), sh);
}

// Siddon projections of many independent problems (Python detector layout) in
// one launch; the XCD-ordered block index runs over the whole grid, so the
// consecutive workgroups of one group still share an XCD.
__global__ __launch_bounds__(kBeamThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void ax_beam_multi_kernel(
    const BeamJob* __restrict__ jobs, int njobs) {
    __shared__ SweepShared sh;
    const int L = xcd_logical_block(blockIdx.x, gridDim.x);
    const int j = find_job(jobs, njobs, L);
    const BeamJob& J = jobs[j];
    beam_block<0>(J.stage_x, J.stage_y, J.nx, J.ny, J.nz, J.na_x, J.na_y, J.views,
                  reinterpret_cast<const BeamGroup*>(J.groups), J.ngroups, J.nU, J.nV, 1, J.nU,
                  (long long)J.nU * J.nV, J.out, J.acc, J.dvx, J.dvy, J.dvz, J.max_planes, J.stack, L - J.begin, sh);
}

// Staging copy for principal axis P: out[p][z][a] over the bordered ranges
// p -1 .. nP + 1, z -1 .. nz, a -4 .. na_pad + 3 (zero outside the volume), from the
// x-fastest volume.  64 x 64 tiles of (p, a) per z row go through LDS, so both
// the reads (x fastest) and the writes (a fastest) are coalesced whichever of
// p, a is x.
__device__ __forceinline__ void stage_volume_block(const void* __restrict__ vol, int dtype, int nx, int ny, int nz,
                                                   int principal, int na_pad, float* __restrict__ out, int bx, int by,
                                                   int bz) {
    __shared__ float tile[64][65];
    const int NP = (principal == 0 ? nx : ny) + 3, NZ = nz + 2, NA = na_pad + 8;
    const int a_base = bx * 64 - 4, p_base = by * 64 - 1, z = bz - 1;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    // read: x fastest across threads
    for (int r = ty; r < 64; r += 4) {
        int x, y;
        if (principal == 0) {  // p = x, a = y: rows of the tile are a, columns p
            x = p_base + tx;
            y = a_base + r;
        } else {               // p = y, a = x
            x = a_base + tx;
            y = p_base + r;
        }
        float val = 0.0f;
        if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz)
            val = load_storage(vol, ((long long)z * ny + y) * nx + x, dtype);
        if (principal == 0)
            tile[r][tx] = val;   // tile[a][p]
        else
            tile[tx][r] = val;   // tile[a][p]
    }
    __syncthreads();
    // write: a fastest across threads
    for (int r = ty; r < 64; r += 4) {
        const int p = p_base + r, a = a_base + tx;
        if (p + 1 < NP && a + 4 < NA)
            out[((long long)(p + 1) * NZ + (z + 1)) * NA + (a + 4)] = tile[tx][r];
    }
}

__global__ __launch_bounds__(256) void stage_volume_kernel(const void* __restrict__ vol, int dtype, int nx, int ny,
                                                           int nz, int principal, int na_pad,
                                                           float* __restrict__ out) {
    stage_volume_block(vol, dtype, nx, ny, nz, principal, na_pad, out, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Staging copies of many FP32 volumes in one launch.
__global__ __launch_bounds__(256) void stage_volume_multi_kernel(const StageJob* __restrict__ jobs, int njobs) {
    const int b = blockIdx.x;
    const int j = find_job(jobs, njobs, b);
    const StageJob& J = jobs[j];
    const int l = b - J.begin;
    stage_volume_block(J.vol, TIGRE_DTYPE_F32, J.nx, J.ny, J.nz, J.principal, J.na_pad, J.out, l % J.gx,
                       (l / J.gx) % J.gy, l / (J.gx * J.gy));
}

}  // namespace

int stage_pad(int n) { return (n + 3) & ~3; }

long long stage_elements(const HostGeometry& g, int principal) {
    const int nP = principal == 0 ? g.nVoxel[0] : g.nVoxel[1];
    const int nA = principal == 0 ? g.nVoxel[1] : g.nVoxel[0];
    // planes -1 .. nP + 1 written, plus one unwritten plane of slack: the
    // last rows a slab box copies may run past the copy by less than a plane
    return (long long)(nP + 4) * (g.nVoxel[2] + 2) * (stage_pad(nA) + 8);
}

int stage_job_blocks(const HostGeometry& g, int principal, int* gx, int* gy) {
    const int nP = principal == 0 ? g.nVoxel[0] : g.nVoxel[1];
    const int nA = principal == 0 ? g.nVoxel[1] : g.nVoxel[0];
    *gx = (stage_pad(nA) + 8 + 63) / 64;
    *gy = (nP + 3 + 63) / 64;
    return *gx * *gy * (g.nVoxel[2] + 2);
}

long long beam_job_blocks(const HostGeometry& g, int ngroups, int stack) {
    const int edge_u = stack >= 2 ? 2 * kBeamTile : kBeamTile, edge_v = stack == 4 ? 2 * kBeamTile : kBeamTile;
    const long long tiles = ((g.nDetec[0] + edge_u - 1) / edge_u) * ((g.nDetec[1] + edge_v - 1) / edge_v);
    return tiles * ngroups;
}

int launch_stage_multi(const StageJob* d_jobs, int njobs, long long blocks, hipStream_t stream) {
    TIGRE_REQUIRE(njobs > 0 && blocks > 0 && blocks < (1LL << 31), "bad multi-job launch");
    KernelTimerScope timer(TIGRE_KERNEL_RELAYOUT, stream);
    hipLaunchKernelGGL(stage_volume_multi_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, d_jobs, njobs);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

int launch_ax_beam_multi(const BeamJob* d_jobs, int njobs, long long blocks, hipStream_t stream) {
    TIGRE_REQUIRE(njobs > 0 && blocks > 0 && blocks < (1LL << 31), "bad multi-job launch");
    KernelTimerScope timer(TIGRE_KERNEL_AX_SIDDON, stream);
    hipLaunchKernelGGL(ax_beam_multi_kernel, dim3((unsigned)blocks), dim3(kBeamThreads), 0, stream, d_jobs, njobs);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

int launch_stage_volume(const void* vol, int dtype, const HostGeometry& g, int principal, float* stage,
                        hipStream_t stream) {
    const int nP = principal == 0 ? g.nVoxel[0] : g.nVoxel[1];
    const int nA = principal == 0 ? g.nVoxel[1] : g.nVoxel[0];
    const int NA = stage_pad(nA) + 8;
    dim3 grid((unsigned)((NA + 63) / 64), (unsigned)((nP + 3 + 63) / 64), (unsigned)(g.nVoxel[2] + 2));
    TIGRE_REQUIRE(grid.y <= 65535u && grid.z <= 65535u, "volume too large for the staging copy");
    KernelTimerScope timer(TIGRE_KERNEL_RELAYOUT, stream);
    hipLaunchKernelGGL(stage_volume_kernel, grid, dim3(256), 0, stream, vol, dtype, g.nVoxel[0], g.nVoxel[1],
                       g.nVoxel[2],
                       principal, stage_pad(nA), stage);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

int launch_ax_beam(const float* stage_x, const float* stage_y, const HostGeometry& g, const FwdView* d_views,
                   const int* d_groups, int ngroups, int stack, int projection_type, int layout, float* proj,
                   int max_planes, hipStream_t stream) {
    TIGRE_REQUIRE(max_planes >= 1 && max_planes <= kSlabPlanesMax, "slab planes out of range");
    const long long nU = g.nDetec[0], nV = g.nDetec[1];
    long long su = 1, sv = nU;
    if (layout == TIGRE_LAYOUT_MATLAB) {
        su = nV;
        sv = 1;
    }
    const long long blocks = beam_job_blocks(g, ngroups, stack);
    TIGRE_REQUIRE(blocks < (1LL << 31), "too many detector tiles for one launch");
    // LDS offsets are 32-bit; global element offsets of the staging copy too
    TIGRE_REQUIRE(stage_elements(g, 0) < (1LL << 32) && stage_elements(g, 1) < (1LL << 32),
                  "volume too large for 32-bit staging offsets");
    const bool interp = projection_type == TIGRE_PROJ_INTERPOLATED;
    // point filtering when accuracy > 1 (ray_interpolated_projection.cu:589-595)
    auto kern = !interp ? ax_beam_kernel<0> : (g.accuracy > 1.0 ? ax_beam_kernel<2> : ax_beam_kernel<1>);
    KernelTimerScope timer(interp ? TIGRE_KERNEL_AX_INTERPOLATED : TIGRE_KERNEL_AX_SIDDON, stream);
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kBeamThreads), 0, stream, stage_x, stage_y, g.nVoxel[0],
                       g.nVoxel[1], g.nVoxel[2], stage_pad(g.nVoxel[1]), stage_pad(g.nVoxel[0]), d_views,
                       reinterpret_cast<const BeamGroup*>(d_groups), ngroups, g.nDetec[0], g.nDetec[1], su, sv,
                       nU * nV, proj, g.accuracy, g.dVoxel[0], g.dVoxel[1], g.dVoxel[2], max_planes, stack);
    TIGRE_HIP_TRY(hipGetLastError());
    return TIGRE_OK;
}

}  // namespace tigre
