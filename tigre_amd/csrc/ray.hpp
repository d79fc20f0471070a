// Device helpers shared by the ray-driven forward projectors.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace tigre {

// Workgroups are dispatched round-robin over the 8 XCDs; remap the hardware
// id so that each XCD runs a contiguous range of logical blocks (neighbouring
// tiles of the same views share that XCD's L2).
__device__ __forceinline__ int xcd_logical_block(int b, int nblk) {
    const int q = nblk / 8, r = nblk % 8;
    const int xcd = b % 8, local = b / 8;
    return xcd * q + min(xcd, r) + local;
}

// Slab [0, n] limits of the parametric ray s + a r (IEEE semantics of
// Siddon_projection.cu:168-176: a ray parallel to the slab and strictly inside
// is unconstrained, otherwise it misses).
__device__ __forceinline__ bool slab(double s, double r, int n, double& am, double& aM) {
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
__device__ __forceinline__ Axis axis_setup(double s, double r, double am) {
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

// Parallel-beam Siddon (Siddon_projection_parallel.cu:173-227): entry and exit
// come from the x and y planes only, and the z index stays at the midpoint of
// the first segment, so a ray that moves in z (a detector offset tilts the
// rays) is summed as the horizontal ray at that height, with the length of its
// x-y part.  Applied to the ray start s and direction before the 3D set-up.
__device__ inline void parallel_siddon_height(double s[3], double rx, double ry, double& rz, int nx, int ny) {
    double am = -INFINITY, aM = INFINITY;
    if (!slab(s[0], rx, nx, am, aM) || !slab(s[1], ry, ny, am, aM) || !(am < aM)) {
        rz = 0.0;
        return;
    }
    const Axis ax = axis_setup(s[0], rx, am), ay = axis_setup(s[1], ry, am);
    double first = am + fmin((double)ax.t0, (double)ay.t0);
    if (!(first <= aM)) first = aM;
    s[2] += 0.5 * (am + first) * rz;
    rz = 0.0;
}

}  // namespace tigre
