/*
 * TEST INFRASTRUCTURE ONLY -- the FP64 CPU oracle (see tigre_oracle.h).
 *
 * Every function below cites the reference file:line it restates.  The
 * reference is CUDA and cannot be built here; this file is an independent
 * FP64 restatement of its arithmetic.  Deliberate, documented differences:
 *
 *  - Texture filtering: the reference samples through CUDA texture units whose
 *    linear filter uses 8-bit fixed-point fractional weights.  The oracle uses
 *    exact trilinear/bilinear weights with the same texel-centre convention
 *    and the same zero border.
 *  - Siddon: the reference walks the ray incrementally (Jacobs) and takes the
 *    first voxel from the midpoint of the first segment
 *    (Siddon_projection.cu:229-231).  The oracle uses the original Siddon
 *    formulation: merge the sorted plane-crossing parameters of the three
 *    axes and take every segment's voxel from that segment's own midpoint.
 *    Both compute the same radiological path; this form has no index drift.
 */
#include "tigre_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double x, y, z; } d3;

void og_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int og_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Rotations                                                                 */
/* ------------------------------------------------------------------------ */

/* Forward ZYZ Euler rotation, Siddon_projection.cu:799-818. */
static d3 euler_zyz(double a, double t, double p, d3 v) {
    double ca = cos(a), sa = sin(a), ct = cos(t), st = sin(t), cp = cos(p), sp = sin(p);
    d3 r;
    r.x = (ca * ct * cp - sa * sp) * v.x + (-ca * ct * sp - sa * cp) * v.y + ca * st * v.z;
    r.y = (sa * ct * cp + ca * sp) * v.x + (-sa * ct * sp + ca * cp) * v.y + sa * st * v.z;
    r.z = -st * cp * v.x + st * sp * v.y + ct * v.z;
    return r;
}

/* Detector roll/pitch/yaw, projection.cpp:18-35. */
static d3 roll_pitch_yaw(double roll, double pitch, double yaw, d3 v) {
    double cr = cos(roll), sr = sin(roll), cp = cos(pitch), sp = sin(pitch), cw = cos(yaw), sw = sin(yaw);
    d3 r;
    r.x = cr * cp * v.x + (cr * sp * sw - sr * cw) * v.y + (cr * sp * cw + sr * sw) * v.z;
    r.y = sr * cp * v.x + (sr * sp * sw + cr * cw) * v.y + (sr * sp * cw - cr * sw) * v.z;
    r.z = -sp * v.x + cp * sw * v.y + cp * cw * v.z;
    return r;
}

/* Transposed ZYZ Euler used by the backprojectors, voxel_backprojection.cu:762-787
 * (called with the negated angles). */
static d3 euler_zyz_t(double a, double t, double p, d3 v) {
    double ca = cos(a), sa = sin(a), ct = cos(t), st = sin(t), cp = cos(p), sp = sin(p);
    d3 r;
    r.x = v.x * (cp * ct * ca - sp * sa) + v.y * (-cp * ct * sa - sp * ca) + v.z * cp * st;
    r.y = v.x * (sp * ct * ca + cp * sa) + v.y * (-sp * ct * sa + cp * ca) + v.z * sp * st;
    r.z = -v.x * st * ca + v.y * st * sa + v.z * ct;
    return r;
}

/* Transposed roll/pitch/yaw, voxel_backprojection.cu:789-816. */
static d3 roll_pitch_yaw_t(double roll, double pitch, double yaw, d3 v) {
    double cr = cos(roll), sr = sin(roll), cp = cos(pitch), sp = sin(pitch), cw = cos(yaw), sw = sin(yaw);
    d3 r;
    r.x = cr * cp * v.x + sr * cp * v.y - sp * v.z;
    r.y = (cr * sp * sw - sr * cw) * v.x + (sr * sp * sw + cr * cw) * v.y + cp * sw * v.z;
    r.z = (cr * sp * cw + sr * sw) * v.x + (sr * sp * cw - cr * sw) * v.y + cp * cw * v.z;
    return r;
}

/* ------------------------------------------------------------------------ */
/* Per-view frames                                                           */
/* ------------------------------------------------------------------------ */

/* Forward frame: Siddon_projection.cu:666-763 (voxel_centre_origin = 0, the
 * volume corner at 0) and ray_interpolated_projection.cu:632-741
 * (voxel_centre_origin = 1, voxel centres at integers).  Output, in voxel
 * index units: pixel(0,0) position, per-u step, per-v step, source. */
void og_forward_frame(const og_geo* g, int i, int voxel_centre_origin, double out[12]) {
    double a = g->angles[3 * i], t = g->angles[3 * i + 1], p = g->angles[3 * i + 2];
    double DSO = g->DSO[i], DSD = g->DSD[i];
    d3 S = {DSO, 0.0, 0.0};
    d3 P = {0.0, g->dDetecU * (-((double)g->nDetecU / 2.0) + 0.5),
            g->dDetecV * (((double)g->nDetecV / 2.0) - 0.5)};
    d3 Pu = {0.0, g->dDetecU, 0.0};
    d3 Pv = {0.0, 0.0, -g->dDetecV};
    /* roll/pitch/yaw about the detector centre */
    P = roll_pitch_yaw(g->dRoll[i], g->dPitch[i], g->dYaw[i], P);
    Pu = roll_pitch_yaw(g->dRoll[i], g->dPitch[i], g->dYaw[i], Pu);
    Pv = roll_pitch_yaw(g->dRoll[i], g->dPitch[i], g->dYaw[i], Pv);
    P.x -= (DSD - DSO);
    P.y += g->offDetecU[i];
    P.z += g->offDetecV[i];
    P = euler_zyz(a, t, p, P);
    Pu = euler_zyz(a, t, p, Pu);
    Pv = euler_zyz(a, t, p, Pv);
    S = euler_zyz(a, t, p, S);
    /* image offset, corner (or first voxel centre) at 0, voxel units */
    double shx = g->sVoxelX / 2.0 - (voxel_centre_origin ? g->dVoxelX / 2.0 : 0.0);
    double shy = g->sVoxelY / 2.0 - (voxel_centre_origin ? g->dVoxelY / 2.0 : 0.0);
    double shz = g->sVoxelZ / 2.0 - (voxel_centre_origin ? g->dVoxelZ / 2.0 : 0.0);
    P.x = (P.x - g->offOrigX[i] + shx) / g->dVoxelX;
    P.y = (P.y - g->offOrigY[i] + shy) / g->dVoxelY;
    P.z = (P.z - g->offOrigZ[i] + shz) / g->dVoxelZ;
    S.x = (S.x - g->offOrigX[i] + shx) / g->dVoxelX;
    S.y = (S.y - g->offOrigY[i] + shy) / g->dVoxelY;
    S.z = (S.z - g->offOrigZ[i] + shz) / g->dVoxelZ;
    Pu.x /= g->dVoxelX; Pu.y /= g->dVoxelY; Pu.z /= g->dVoxelZ;
    Pv.x /= g->dVoxelX; Pv.y /= g->dVoxelY; Pv.z /= g->dVoxelZ;
    /* centre-of-rotation shift (Siddon_projection.cu:752-756) */
    double corx = -g->COR[i] * sin(a) / g->dVoxelX;
    double cory = g->COR[i] * cos(a) / g->dVoxelY;
    P.x += corx; P.y += cory;
    S.x += corx; S.y += cory;
    out[0] = P.x;  out[1] = P.y;  out[2] = P.z;
    out[3] = Pu.x; out[4] = Pu.y; out[5] = Pu.z;
    out[6] = Pv.x; out[7] = Pv.y; out[8] = Pv.z;
    out[9] = S.x;  out[10] = S.y; out[11] = S.z;
}

/* Safe skip distance before sampling starts, ray_interpolated_projection.cu:743-768.
 * nVoxel/2 is integer division in the reference and stays so here. */
double og_maxdistance_cuboid(const og_geo* g, int i) {
    double mx = (double)(g->nVoxelX / 2) + fabs(g->offOrigX[i]) / g->dVoxelX;
    double my = (double)(g->nVoxelY / 2) + fabs(g->offOrigY[i]) / g->dVoxelY;
    double mz = (double)(g->nVoxelZ / 2) + fabs(g->offOrigZ[i]) / g->dVoxelZ;
    double a = g->DSO[i] / g->dVoxelX, b = g->DSO[i] / g->dVoxelY;
    double al = g->angles[3 * i], th = g->angles[3 * i + 1], ps = g->angles[3 * i + 2];
    if (th == 0.0 && ps == 0.0) {
        double s = sin(al), c = cos(al);
        double v = a * b / sqrt(a * a * s * s + b * b * c * c) - sqrt(mx * mx + my * my);
        return v > 0.0 ? v : 0.0;
    }
    double dmax = fmax(fmax(g->dVoxelX, g->dVoxelY), g->dVoxelZ);
    double v = g->DSO[i] / dmax - sqrt(mx * mx + my * my + mz * mz);
    return v > 0.0 ? v : 0.0;
}

/* Backward frame, computeDeltasCube voxel_backprojection.cu:826-898, with the
 * angles negated as the host loop does (voxel_backprojection.cu:510-512).
 * x stays in mm; y and z are in detector-pixel units.
 * out: origin(3), dX(3), dY(3), dZ(3), source(3). */
void og_backward_frame(const og_geo* g, int i, double out[15]) {
    double a = -g->angles[3 * i], t = -g->angles[3 * i + 1], p = -g->angles[3 * i + 2];
    double DSO = g->DSO[i], DSD = g->DSD[i];
    d3 P = {-(g->sVoxelX / 2.0 - g->dVoxelX / 2.0) + g->offOrigX[i],
            -(g->sVoxelY / 2.0 - g->dVoxelY / 2.0) + g->offOrigY[i],
            -(g->sVoxelZ / 2.0 - g->dVoxelZ / 2.0) + g->offOrigZ[i]};
    d3 Q[4];
    Q[0] = P;
    Q[1] = P; Q[1].x += g->dVoxelX;
    Q[2] = P; Q[2].y += g->dVoxelY;
    Q[3] = P; Q[3].z += g->dVoxelZ;
    for (int k = 0; k < 4; ++k) {
        Q[k] = euler_zyz_t(a, t, p, Q[k]);
        Q[k].z -= g->offDetecV[i];
        Q[k].y -= g->offDetecU[i];
        Q[k].x += (DSD - DSO);
        Q[k] = roll_pitch_yaw_t(g->dRoll[i], g->dPitch[i], g->dYaw[i], Q[k]);
        Q[k].x -= (DSD - DSO);
        Q[k].z /= g->dDetecV;
        Q[k].y /= g->dDetecU;
    }
    d3 S = {DSD, -g->offDetecU[i], -g->offDetecV[i]};
    S = roll_pitch_yaw_t(g->dRoll[i], g->dPitch[i], g->dYaw[i], S);
    S.x -= (DSD - DSO);
    S.z /= g->dDetecV;
    S.y /= g->dDetecU;
    out[0] = Q[0].x; out[1] = Q[0].y; out[2] = Q[0].z;
    for (int k = 1; k < 4; ++k) {
        out[3 * k + 0] = Q[k].x - Q[0].x;
        out[3 * k + 1] = Q[k].y - Q[0].y;
        out[3 * k + 2] = Q[k].z - Q[0].z;
    }
    out[12] = S.x; out[13] = S.y; out[14] = S.z;
}

/* ------------------------------------------------------------------------ */
/* Forward projectors                                                        */
/* ------------------------------------------------------------------------ */

static inline size_t proj_index(const og_geo* g, int view, int u, int v, int layout_matlab) {
    size_t P = (size_t)g->nDetecU * (size_t)g->nDetecV;
    /* Siddon_projection.cu:125-129 */
    if (layout_matlab) return (size_t)view * P + (size_t)u * g->nDetecV + v;
    return (size_t)view * P + (size_t)v * g->nDetecU + u;
}

/* Axis limits of the parametric ray S + a*r against the slab [0, n]
 * (Siddon_projection.cu:168-176, IEEE semantics for r == 0: inside the open
 * slab no constraint, otherwise a miss). Returns 0 on a miss. */
static int slab_limits(double s, double r, int n, double* lo, double* hi) {
    if (r == 0.0) {
        if (s > 0.0 && s < (double)n) { *lo = -INFINITY; *hi = INFINITY; return 1; }
        return 0;
    }
    double a0 = (0.0 - s) / r, a1 = ((double)n - s) / r;
    *lo = fmin(a0, a1);
    *hi = fmax(a0, a1);
    return 1;
}

/* Plane-crossing parameters of one axis inside (am, aM), increasing. */
static int axis_crossings(double s, double r, int n, double am, double aM, double* out) {
    int cnt = 0;
    if (r == 0.0) return 0;
    if (r > 0.0) {
        for (int p = 0; p <= n; ++p) {
            double a = ((double)p - s) / r;
            if (a > am && a < aM) out[cnt++] = a;
        }
    } else {
        for (int p = n; p >= 0; --p) {
            double a = ((double)p - s) / r;
            if (a > am && a < aM) out[cnt++] = a;
        }
    }
    return cnt;
}

/* One Siddon ray, Siddon_projection.cu:131-268. */
static double siddon_ray(const float* img, const og_geo* g, d3 S, d3 P, double* work) {
    const int nx = g->nVoxelX, ny = g->nVoxelY, nz = g->nVoxelZ;
    d3 r = {P.x - S.x, P.y - S.y, P.z - S.z};
    const double eps = 0.001;                         /* :155-159 */
    if (fabs(r.x) < eps) r.x = 0.0;
    if (fabs(r.y) < eps) r.y = 0.0;
    if (fabs(r.z) < eps) r.z = 0.0;
    double lo, hi, am = -INFINITY, aM = INFINITY;
    if (!slab_limits(S.x, r.x, nx, &lo, &hi)) return 0.0;
    am = fmax(am, lo); aM = fmin(aM, hi);
    if (!slab_limits(S.y, r.y, ny, &lo, &hi)) return 0.0;
    am = fmax(am, lo); aM = fmin(aM, hi);
    if (!slab_limits(S.z, r.z, nz, &lo, &hi)) return 0.0;
    am = fmax(am, lo); aM = fmin(aM, hi);
    if (!(am < aM)) return 0.0;                       /* miss => 0 (:179-180) */

    double* ax = work;
    double* ay = ax + (nx + 1);
    double* az = ay + (ny + 1);
    int cx = axis_crossings(S.x, r.x, nx, am, aM, ax);
    int cy = axis_crossings(S.y, r.y, ny, am, aM, ay);
    int cz = axis_crossings(S.z, r.z, nz, am, aM, az);
    int ix = 0, iy = 0, iz = 0;
    double prev = am, sum = 0.0;
    for (;;) {
        double nxt = aM;
        int which = -1;
        if (ix < cx && ax[ix] < nxt) { nxt = ax[ix]; which = 0; }
        if (iy < cy && ay[iy] < nxt) { nxt = ay[iy]; which = 1; }
        if (iz < cz && az[iz] < nxt) { nxt = az[iz]; which = 2; }
        if (nxt > prev) {
            double mid = 0.5 * (prev + nxt);
            long i = (long)floor(S.x + mid * r.x);
            long j = (long)floor(S.y + mid * r.y);
            long k = (long)floor(S.z + mid * r.z);
            if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz)
                sum += (nxt - prev) * (double)img[((size_t)k * ny + j) * nx + i];
            prev = nxt;
        }
        if (which == 0) ++ix;
        else if (which == 1) ++iy;
        else if (which == 2) ++iz;
        else break;
    }
    /* ray length in mm (:241,:267) */
    double len = sqrt(r.x * r.x * g->dVoxelX * g->dVoxelX + r.y * r.y * g->dVoxelY * g->dVoxelY +
                      r.z * r.z * g->dVoxelZ * g->dVoxelZ);
    return sum * len;
}

void og_ax_siddon(const float* img, const og_geo* g, int layout_matlab, double* out) {
    const int nU = g->nDetecU, nV = g->nDetecV;
    for (int view = 0; view < g->nangles; ++view) {
        double f[12];
        og_forward_frame(g, view, 0, f);
        d3 O = {f[0], f[1], f[2]}, DU = {f[3], f[4], f[5]}, DV = {f[6], f[7], f[8]}, S = {f[9], f[10], f[11]};
#pragma omp parallel
        {
            double* work = (double*)malloc(sizeof(double) * (size_t)(g->nVoxelX + g->nVoxelY + g->nVoxelZ + 3));
#pragma omp for schedule(dynamic, 4)
            for (int v = 0; v < nV; ++v) {
                double pv = (double)(nV - v - 1);     /* row flip, :141 */
                for (int u = 0; u < nU; ++u) {
                    d3 P = {O.x + u * DU.x + pv * DV.x, O.y + u * DU.y + pv * DV.y, O.z + u * DU.z + pv * DV.z};
                    out[proj_index(g, view, u, v, layout_matlab)] = siddon_ray(img, g, S, P, work);
                }
            }
            free(work);
        }
    }
}

/* Zero-border trilinear sample at a continuous index position whose voxel
 * centres sit at integers (tex3D(p + 0.5) with linear filtering). */
static inline double vox(const float* img, int nx, int ny, int nz, long i, long j, long k) {
    if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return 0.0;
    return (double)img[((size_t)k * ny + j) * nx + i];
}

static double trilinear0(const float* img, int nx, int ny, int nz, double x, double y, double z) {
    double fx = floor(x), fy = floor(y), fz = floor(z);
    long i = (long)fx, j = (long)fy, k = (long)fz;
    double wx = x - fx, wy = y - fy, wz = z - fz;
    double c00 = vox(img, nx, ny, nz, i, j, k) * (1 - wx) + vox(img, nx, ny, nz, i + 1, j, k) * wx;
    double c10 = vox(img, nx, ny, nz, i, j + 1, k) * (1 - wx) + vox(img, nx, ny, nz, i + 1, j + 1, k) * wx;
    double c01 = vox(img, nx, ny, nz, i, j, k + 1) * (1 - wx) + vox(img, nx, ny, nz, i + 1, j, k + 1) * wx;
    double c11 = vox(img, nx, ny, nz, i, j + 1, k + 1) * (1 - wx) + vox(img, nx, ny, nz, i + 1, j + 1, k + 1) * wx;
    double c0 = c00 * (1 - wy) + c10 * wy;
    double c1 = c01 * (1 - wy) + c11 * wy;
    return c0 * (1 - wz) + c1 * wz;
}

/* Nearest-texel sample (point filtering, used when accuracy > 1,
 * ray_interpolated_projection.cu:589-595). */
static double nearest0(const float* img, int nx, int ny, int nz, double x, double y, double z) {
    return vox(img, nx, ny, nz, (long)floor(x + 0.5), (long)floor(y + 0.5), (long)floor(z + 0.5));
}

/* Parameter range (in sample units) where S + i*step lies inside (-1, n) on
 * one axis; samples outside contribute exactly 0 to the zero-border sum. */
static void sample_window(double s, double st, int n, double* lo, double* hi) {
    if (st == 0.0) {
        if (s > -1.0 && s < (double)n) { *lo = -INFINITY; *hi = INFINITY; }
        else { *lo = INFINITY; *hi = -INFINITY; }
        return;
    }
    double a0 = (-1.0 - s) / st, a1 = ((double)n - s) / st;
    *lo = fmin(a0, a1);
    *hi = fmax(a0, a1);
}

/* Interpolated projector, ray_interpolated_projection.cu:117-206 and the
 * host loop :353-392 (crop distance, spherical flag). */
void og_ax_interp(const float* img, const og_geo* g, int layout_matlab, double* out) {
    const int nU = g->nDetecU, nV = g->nDetecV;
    const int nx = g->nVoxelX, ny = g->nVoxelY, nz = g->nVoxelZ;
    const double acc = g->accuracy;
    const int point = acc > 1.0;
    for (int view = 0; view < g->nangles; ++view) {
        double f[12];
        og_forward_frame(g, view, 1, f);
        d3 O = {f[0], f[1], f[2]}, DU = {f[3], f[4], f[5]}, DV = {f[6], f[7], f[8]}, S = {f[9], f[10], f[11]};
        double crop = floor(og_maxdistance_cuboid(g, view));
        /* the flag sums |theta| + |psi| over the view's batch of PROJ_PER_BLOCK = 9
         * consecutive views of the call (ray_interpolated_projection.cu:75, :352-365) */
        int spherical = 0;
        for (int j = (view / 9) * 9; j < (view / 9) * 9 + 9 && j < g->nangles; ++j)
            if (g->angles[3 * j + 1] != 0.0 || g->angles[3 * j + 2] != 0.0) spherical = 1;
        double dmin = spherical ? fmin(fmin(g->dVoxelX, g->dVoxelY), g->dVoxelZ) : fmin(g->dVoxelX, g->dVoxelY);
        double cap = (2.0 * g->DSO[view] / dmin + crop) / acc;
#pragma omp parallel for schedule(dynamic, 4)
        for (int v = 0; v < nV; ++v) {
            double pv = (double)(nV - v - 1);
            for (int u = 0; u < nU; ++u) {
                d3 P = {O.x + u * DU.x + pv * DV.x, O.y + u * DU.y + pv * DV.y, O.z + u * DU.z + pv * DV.z};
                d3 d = {P.x - S.x, P.y - S.y, P.z - S.z};
                double len = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
                double n = ceil(len / acc);
                d3 st = {d.x / n, d.y / n, d.z / n};
                double last = (cap < n) ? ceil(cap) : n;
                double first = floor(crop / acc);
                double lo, hi, wlo = -INFINITY, whi = INFINITY;
                sample_window(S.x, st.x, nx, &lo, &hi); wlo = fmax(wlo, lo); whi = fmin(whi, hi);
                sample_window(S.y, st.y, ny, &lo, &hi); wlo = fmax(wlo, lo); whi = fmin(whi, hi);
                sample_window(S.z, st.z, nz, &lo, &hi); wlo = fmax(wlo, lo); whi = fmin(whi, hi);
                double i0 = fmax(first, floor(wlo) - 1.0);
                double i1 = fmin(last, ceil(whi) + 1.0);
                double sum = 0.0;
                for (double i = i0; i <= i1; i += 1.0) {
                    double x = S.x + i * st.x, y = S.y + i * st.y, z = S.z + i * st.z;
                    sum += point ? nearest0(img, nx, ny, nz, x, y, z) : trilinear0(img, nx, ny, nz, x, y, z);
                }
                double dl = sqrt(st.x * g->dVoxelX * st.x * g->dVoxelX + st.y * g->dVoxelY * st.y * g->dVoxelY +
                                 st.z * g->dVoxelZ * st.z * g->dVoxelZ);
                out[proj_index(g, view, u, v, layout_matlab)] = sum * dl;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Parallel beam                                                             */
/* ------------------------------------------------------------------------ */

/* Parallel-beam frames.  Siddon (voxel_centre_origin = 0):
 * computeDeltas_Siddon_parallel, Siddon_projection_parallel.cu:388-470 --
 * detector points with x = 0 go through roll/pitch/yaw, x -= DSD - DSO, the
 * detector offset is added once, then a rotation by alpha alone.  Interpolated
 * (voxel_centre_origin = 1): computeDeltas_parallel,
 * ray_interpolated_projection_parallel.cu:313-390 -- the detector offset is
 * added twice and the full ZYZ rotation is applied.  The source is pixel
 * (0, 0)'s detector point at x = 0 through roll/pitch/yaw, then x += DSO, with
 * no detector offset; every pixel's source is this source plus the pixel's
 * u / v steps.  Output as og_forward_frame. */
void og_forward_frame_parallel(const og_geo* g, int i, int voxel_centre_origin, double out[12]) {
    double a = g->angles[3 * i], t = g->angles[3 * i + 1], p = g->angles[3 * i + 2];
    double DSO = g->DSO[i], DSD = g->DSD[i];
    d3 P = {0.0, g->dDetecU * (-((double)g->nDetecU / 2.0) + 0.5),
            g->dDetecV * (((double)g->nDetecV / 2.0) - 0.5)};
    d3 S = P;
    d3 Pu = {0.0, g->dDetecU, 0.0};
    d3 Pv = {0.0, 0.0, -g->dDetecV};
    P = roll_pitch_yaw(g->dRoll[i], g->dPitch[i], g->dYaw[i], P);
    Pu = roll_pitch_yaw(g->dRoll[i], g->dPitch[i], g->dYaw[i], Pu);
    Pv = roll_pitch_yaw(g->dRoll[i], g->dPitch[i], g->dYaw[i], Pv);
    S = roll_pitch_yaw(g->dRoll[i], g->dPitch[i], g->dYaw[i], S);
    P.x -= (DSD - DSO);
    S.x += DSO;
    int times = voxel_centre_origin ? 2 : 1;
    P.y += times * g->offDetecU[i];
    P.z += times * g->offDetecV[i];
    if (voxel_centre_origin) {
        P = euler_zyz(a, t, p, P);
        Pu = euler_zyz(a, t, p, Pu);
        Pv = euler_zyz(a, t, p, Pv);
        S = euler_zyz(a, t, p, S);
    } else {
        P = euler_zyz(a, 0.0, 0.0, P);
        Pu = euler_zyz(a, 0.0, 0.0, Pu);
        Pv = euler_zyz(a, 0.0, 0.0, Pv);
        S = euler_zyz(a, 0.0, 0.0, S);
    }
    double shx = g->sVoxelX / 2.0 - (voxel_centre_origin ? g->dVoxelX / 2.0 : 0.0);
    double shy = g->sVoxelY / 2.0 - (voxel_centre_origin ? g->dVoxelY / 2.0 : 0.0);
    double shz = g->sVoxelZ / 2.0 - (voxel_centre_origin ? g->dVoxelZ / 2.0 : 0.0);
    P.x = (P.x - g->offOrigX[i] + shx) / g->dVoxelX;
    P.y = (P.y - g->offOrigY[i] + shy) / g->dVoxelY;
    P.z = (P.z - g->offOrigZ[i] + shz) / g->dVoxelZ;
    S.x = (S.x - g->offOrigX[i] + shx) / g->dVoxelX;
    S.y = (S.y - g->offOrigY[i] + shy) / g->dVoxelY;
    S.z = (S.z - g->offOrigZ[i] + shz) / g->dVoxelZ;
    Pu.x /= g->dVoxelX; Pu.y /= g->dVoxelY; Pu.z /= g->dVoxelZ;
    Pv.x /= g->dVoxelX; Pv.y /= g->dVoxelY; Pv.z /= g->dVoxelZ;
    double corx = -g->COR[i] * sin(a) / g->dVoxelX;
    double cory = g->COR[i] * cos(a) / g->dVoxelY;
    P.x += corx; P.y += cory;
    S.x += corx; S.y += cory;
    out[0] = P.x;  out[1] = P.y;  out[2] = P.z;
    out[3] = Pu.x; out[4] = Pu.y; out[5] = Pu.z;
    out[6] = Pv.x; out[7] = Pv.y; out[8] = Pv.z;
    out[9] = S.x;  out[10] = S.y; out[11] = S.z;
}

/* One parallel-beam Siddon ray, kernelPixelDetector_parallel
 * (Siddon_projection_parallel.cu:150-268): crossings of the x and y planes
 * only, the z index fixed at the first segment's midpoint, the ray length
 * without its z component.  For the usual horizontal rays (no detector pitch
 * or roll) this is the full 3D Siddon sum. */
static double siddon_ray_parallel(const float* img, const og_geo* g, d3 S, d3 P, double* work) {
    const int nx = g->nVoxelX, ny = g->nVoxelY, nz = g->nVoxelZ;
    d3 r = {P.x - S.x, P.y - S.y, P.z - S.z};
    double lo, hi, am = -INFINITY, aM = INFINITY;
    if (!slab_limits(S.x, r.x, nx, &lo, &hi)) return 0.0;
    am = fmax(am, lo); aM = fmin(aM, hi);
    if (!slab_limits(S.y, r.y, ny, &lo, &hi)) return 0.0;
    am = fmax(am, lo); aM = fmin(aM, hi);
    if (!(am < aM)) return 0.0;
    double* ax = work;
    double* ay = ax + (nx + 1);
    int cx = axis_crossings(S.x, r.x, nx, am, aM, ax);
    int cy = axis_crossings(S.y, r.y, ny, am, aM, ay);
    double first = aM;
    if (cx > 0 && ax[0] < first) first = ax[0];
    if (cy > 0 && ay[0] < first) first = ay[0];
    long k = (long)floor(S.z + 0.5 * (am + first) * r.z);
    if (k < 0 || k >= nz) return 0.0;
    int ix = 0, iy = 0;
    double prev = am, sum = 0.0;
    for (;;) {
        double nxt = aM;
        int which = -1;
        if (ix < cx && ax[ix] < nxt) { nxt = ax[ix]; which = 0; }
        if (iy < cy && ay[iy] < nxt) { nxt = ay[iy]; which = 1; }
        if (nxt > prev) {
            double mid = 0.5 * (prev + nxt);
            long i = (long)floor(S.x + mid * r.x);
            long j = (long)floor(S.y + mid * r.y);
            if (i >= 0 && i < nx && j >= 0 && j < ny)
                sum += (nxt - prev) * (double)img[((size_t)k * ny + j) * nx + i];
            prev = nxt;
        }
        if (which == 0) ++ix;
        else if (which == 1) ++iy;
        else break;
    }
    return sum * sqrt(r.x * r.x * g->dVoxelX * g->dVoxelX + r.y * r.y * g->dVoxelY * g->dVoxelY);
}

void og_ax_siddon_parallel(const float* img, const og_geo* g, int layout_matlab, double* out) {
    const int nU = g->nDetecU, nV = g->nDetecV;
    for (int view = 0; view < g->nangles; ++view) {
        double f[12];
        og_forward_frame_parallel(g, view, 0, f);
        d3 O = {f[0], f[1], f[2]}, DU = {f[3], f[4], f[5]}, DV = {f[6], f[7], f[8]}, S0 = {f[9], f[10], f[11]};
#pragma omp parallel
        {
            double* work = (double*)malloc(sizeof(double) * (size_t)(g->nVoxelX + g->nVoxelY + 2));
#pragma omp for schedule(dynamic, 4)
            for (int v = 0; v < nV; ++v) {
                double pv = (double)(nV - v - 1);
                for (int u = 0; u < nU; ++u) {
                    d3 sh = {u * DU.x + pv * DV.x, u * DU.y + pv * DV.y, u * DU.z + pv * DV.z};
                    d3 P = {O.x + sh.x, O.y + sh.y, O.z + sh.z};
                    d3 S = {S0.x + sh.x, S0.y + sh.y, S0.z + sh.z};
                    out[proj_index(g, view, u, v, layout_matlab)] = siddon_ray_parallel(img, g, S, P, work);
                }
            }
            free(work);
        }
    }
}

/* kernelPixelDetector_parallel of ray_interpolated_projection_parallel.cu
 * :117-191: samples from the pixel's own source, the first at
 * floor(crop / accuracy), the last at min(n, ceil((2 DSO / dVoxelX + crop) /
 * accuracy)). */
void og_ax_interp_parallel(const float* img, const og_geo* g, int layout_matlab, double* out) {
    const int nU = g->nDetecU, nV = g->nDetecV;
    const int nx = g->nVoxelX, ny = g->nVoxelY, nz = g->nVoxelZ;
    const double acc = g->accuracy;
    const int point = acc > 1.0;
    for (int view = 0; view < g->nangles; ++view) {
        double f[12];
        og_forward_frame_parallel(g, view, 1, f);
        d3 O = {f[0], f[1], f[2]}, DU = {f[3], f[4], f[5]}, DV = {f[6], f[7], f[8]}, S0 = {f[9], f[10], f[11]};
        double crop = floor(og_maxdistance_cuboid(g, view));
        double cap = (2.0 * g->DSO[view] / g->dVoxelX + crop) / acc;
#pragma omp parallel for schedule(dynamic, 4)
        for (int v = 0; v < nV; ++v) {
            double pv = (double)(nV - v - 1);
            for (int u = 0; u < nU; ++u) {
                d3 sh = {u * DU.x + pv * DV.x, u * DU.y + pv * DV.y, u * DU.z + pv * DV.z};
                d3 S = {S0.x + sh.x, S0.y + sh.y, S0.z + sh.z};
                /* the pixel shifts of source and pixel cancel: every ray is O - S0 */
                d3 d = {O.x - S0.x, O.y - S0.y, O.z - S0.z};
                /* |d| / acc is often an exact integer here (|d| = DSD / dVoxel); values
                 * within 1e-9 of an integer count as that integer */
                double n = ceil(sqrt(d.x * d.x + d.y * d.y + d.z * d.z) / acc - 1e-9);
                d3 st = {d.x / n, d.y / n, d.z / n};
                double last = (cap < n) ? ceil(cap) : n;
                double first = floor(crop / acc);
                double lo, hi, wlo = -INFINITY, whi = INFINITY;
                sample_window(S.x, st.x, nx, &lo, &hi); wlo = fmax(wlo, lo); whi = fmin(whi, hi);
                sample_window(S.y, st.y, ny, &lo, &hi); wlo = fmax(wlo, lo); whi = fmin(whi, hi);
                sample_window(S.z, st.z, nz, &lo, &hi); wlo = fmax(wlo, lo); whi = fmin(whi, hi);
                double i0 = fmax(first, floor(wlo) - 1.0);
                double i1 = fmin(last, ceil(whi) + 1.0);
                double sum = 0.0;
                for (double i = i0; i <= i1; i += 1.0) {
                    double x = S.x + i * st.x, y = S.y + i * st.y, z = S.z + i * st.z;
                    sum += point ? nearest0(img, nx, ny, nz, x, y, z) : trilinear0(img, nx, ny, nz, x, y, z);
                }
                double dl = sqrt(st.x * g->dVoxelX * st.x * g->dVoxelX + st.y * g->dVoxelY * st.y * g->dVoxelY +
                                 st.z * g->dVoxelZ * st.z * g->dVoxelZ);
                out[proj_index(g, view, u, v, layout_matlab)] = sum * dl;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Backprojectors                                                            */
/* ------------------------------------------------------------------------ */

/* Zero-border bilinear sample of one projection at texture coordinates
 * (tu, tv): texel (i, j) has its centre at (i + 0.5, j + 0.5). */
static double bilinear0(const float* pj, const og_geo* g, int layout_matlab, double tu, double tv) {
    const int nU = g->nDetecU, nV = g->nDetecV;
    double x = tu - 0.5, y = tv - 0.5;
    double fx = floor(x), fy = floor(y);
    long i = (long)fx, j = (long)fy;
    double wx = x - fx, wy = y - fy;
    double t[2][2];
    for (int b = 0; b < 2; ++b)
        for (int a = 0; a < 2; ++a) {
            long ii = i + a, jj = j + b;
            if (ii < 0 || jj < 0 || ii >= nU || jj >= nV) t[b][a] = 0.0;
            else t[b][a] = layout_matlab ? (double)pj[(size_t)ii * nV + jj] : (double)pj[(size_t)jj * nU + ii];
        }
    return (t[0][0] * (1 - wx) + t[0][1] * wx) * (1 - wy) + (t[1][0] * (1 - wx) + t[1][1] * wx) * wy;
}

/* Shared voxel-driven loop.  matched = 0: kernelPixelBackprojectionFDK
 * (voxel_backprojection.cu:137-281); matched = 1: kernelPixelBackprojection
 * (voxel_backprojection2.cu:142-306) with the final constant scaling (:585). */
static void atb_common(const float* proj, const og_geo* g, int layout_matlab, double* out, int matched) {
    const int nx = g->nVoxelX, ny = g->nVoxelY, nz = g->nVoxelZ;
    const size_t nvox = (size_t)nx * ny * nz;
    const size_t P = (size_t)g->nDetecU * g->nDetecV;
    memset(out, 0, nvox * sizeof(double));
    for (int view = 0; view < g->nangles; ++view) {
        double f[15];
        og_backward_frame(g, view, f);
        const double al = -g->angles[3 * view];
        const double sa = sin(al), ca = cos(al);
        const double COR = g->COR[view], DSD = g->DSD[view], DSO = g->DSO[view];
        const float* pj = proj + (size_t)view * P;
        const double ox = g->offOrigX[view], oy = g->offOrigY[view], oz = g->offOrigZ[view];
#pragma omp parallel for schedule(static)
        for (long kz = 0; kz < nz; ++kz) {
            for (long jy = 0; jy < ny; ++jy) {
                for (long ix = 0; ix < nx; ++ix) {
                    double Px = f[0] + ix * f[3] + jy * f[6] + kz * f[9];
                    double Py = f[1] + ix * f[4] + jy * f[7] + kz * f[10] - COR / g->dDetecU;
                    double Pz = f[2] + ix * f[5] + jy * f[8] + kz * f[11];
                    double vx = Px - f[12], vy = Py - f[13], vz = Pz - f[14];
                    double t = (DSO - DSD - f[12]) / vx;
                    double yy = vy * t + f[13], zz = vz * t + f[14];
                    double u = yy + (double)g->nDetecU * 0.5;
                    double v = zz + (double)g->nDetecV * 0.5;
                    double sample = bilinear0(pj, g, layout_matlab, u, v);
                    double w;
                    if (!matched) {
                        /* :243-249 */
                        double realx = -(g->sVoxelX - g->dVoxelX) * 0.5 + ix * g->dVoxelX + ox;
                        double realy = -(g->sVoxelY - g->dVoxelY) * 0.5 + jy * g->dVoxelY + oy + COR;
                        double ww = (DSO + realy * sa - realx * ca) / DSO;
                        w = 1.0 / (ww * ww);
                    } else {
                        /* voxel_backprojection2.cu:201-279 */
                        double Sx = DSO * ca, Sy = -DSO * sa;
                        double rvx = -g->sVoxelX / 2 + g->dVoxelX / 2 + ox + ix * g->dVoxelX;
                        double rvy = -g->sVoxelY / 2 + g->dVoxelY / 2 + oy + jy * g->dVoxelY;
                        double rvz = -g->sVoxelZ / 2 + g->dVoxelZ / 2 + oz + kz * g->dVoxelZ;
                        double dax = -(DSD - DSO);
                        double day = (-g->sDetecU + g->dDetecU) * 0.5 + u * g->dDetecU + g->offDetecU[view];
                        double Dz = (-g->sDetecV + g->dDetecV) * 0.5 + v * g->dDetecV + g->offDetecV[view];
                        double Dx = dax * ca + day * sa;
                        double Dy = -dax * sa + day * ca;
                        double L = sqrt((Sx - Dx) * (Sx - Dx) + (Sy - Dy) * (Sy - Dy) + Dz * Dz);
                        double lsq = (Sx - rvx) * (Sx - rvx) + (Sy - rvy) * (Sy - rvy) + rvz * rvz;
                        w = L * L * L / (DSD * lsq);
                    }
                    out[((size_t)kz * ny + jy) * nx + ix] += sample * w;
                }
            }
        }
    }
    if (matched) {
        double c = g->dVoxelX * g->dVoxelY * g->dVoxelZ / (g->dDetecU * g->dDetecV);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < (long)nvox; ++q) out[q] *= c;
    }
}

/* kernelPixelBackprojection_parallel, voxel_backprojection_parallel.cu:138-281:
 * the voxel (frame of computeDeltasCubeParallel, :510-585, the same as the
 * cone backprojectors' frame) projects orthogonally, texture coordinates
 * (y + nU / 2, z + nV / 2) with y shifted by -COR / dDetecU; weight 1, no
 * final scaling. */
void og_atb_parallel(const float* proj, const og_geo* g, int layout_matlab, double* out) {
    const int nx = g->nVoxelX, ny = g->nVoxelY, nz = g->nVoxelZ;
    const size_t nvox = (size_t)nx * ny * nz;
    const size_t P = (size_t)g->nDetecU * g->nDetecV;
    memset(out, 0, nvox * sizeof(double));
    for (int view = 0; view < g->nangles; ++view) {
        double f[15];
        og_backward_frame(g, view, f);
        const double COR = g->COR[view];
        const float* pj = proj + (size_t)view * P;
#pragma omp parallel for schedule(static)
        for (long kz = 0; kz < nz; ++kz)
            for (long jy = 0; jy < ny; ++jy)
                for (long ix = 0; ix < nx; ++ix) {
                    double Py = f[1] + ix * f[4] + jy * f[7] + kz * f[10] - COR / g->dDetecU;
                    double Pz = f[2] + ix * f[5] + jy * f[8] + kz * f[11];
                    out[((size_t)kz * ny + jy) * nx + ix] +=
                        bilinear0(pj, g, layout_matlab, Py + (double)g->nDetecU * 0.5, Pz + (double)g->nDetecV * 0.5);
                }
    }
}

void og_atb_fdk(const float* proj, const og_geo* g, int layout_matlab, double* out) {
    atb_common(proj, g, layout_matlab, out, 0);
}

void og_atb_matched(const float* proj, const og_geo* g, int layout_matlab, double* out) {
    atb_common(proj, g, layout_matlab, out, 1);
}

/* ------------------------------------------------------------------------ */
/* TV steepest descent, GD_TV.cu                                             */
/* ------------------------------------------------------------------------ */

/* f with the single-GPU z padding of pocs_tv: the image occupies slices
 * [0, nz) and every slice outside reads as zero (GD_TV.cu:367-368,456-465). */
static inline double fz(const double* f, long nx, long ny, long nz, long z, long y, long x) {
    if (z < 0 || z >= nz) return 0.0;
    return f[(z * ny + y) * nx + x];
}

/* Backward differences at (z,y,x), GD_TV.cu:95-115: the y and x terms are
 * dropped (left 0) at the low edge; z always has a (zero) neighbour. */
static inline void bdiff(const double* f, long nx, long ny, long nz, long z, long y, long x, double d[3]) {
    double c = fz(f, nx, ny, nz, z, y, x);
    d[0] = c - fz(f, nx, ny, nz, z - 1, y, x);
    d[1] = (y >= 1) ? c - fz(f, nx, ny, nz, z, y - 1, x) : 0.0;
    d[2] = (x >= 1) ? c - fz(f, nx, ny, nz, z, y, x - 1) : 0.0;
}

/* gradientTV, GD_TV.cu:117-143. */
void og_tv_gradient(const double* f, double* g, long nx, long ny, long nz) {
    const double eps = 1e-8;
#pragma omp parallel for schedule(static)
    for (long z = 0; z < nz; ++z)
        for (long y = 0; y < ny; ++y)
            for (long x = 0; x < nx; ++x) {
                double df[3], di[3] = {0, 0, 0}, dj[3] = {0, 0, 0}, dk[3];
                bdiff(f, nx, ny, nz, z, y, x, df);
                if (x + 1 < nx) bdiff(f, nx, ny, nz, z, y, x + 1, di);
                if (y + 1 < ny) bdiff(f, nx, ny, nz, z, y + 1, x, dj);
                bdiff(f, nx, ny, nz, z + 1, y, x, dk);
                double v = (df[0] + df[1] + df[2]) / (sqrt(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]) + eps);
                v -= di[2] / (sqrt(di[0] * di[0] + di[1] * di[1] + di[2] * di[2]) + eps);
                v -= dj[1] / (sqrt(dj[0] * dj[0] + dj[1] * dj[1] + dj[2] * dj[2]) + eps);
                v -= dk[0] / (sqrt(dk[0] * dk[0] + dk[1] * dk[1] + dk[2] * dk[2]) + eps);
                g[(z * ny + y) * nx + x] = v;
            }
}

/* Backward differences without z padding (PICCS.cu:109-128): the z term is
 * dropped at the low edge like the y and x terms; (z, y, x) inside the volume. */
static inline void bdiff_unpadded(const double* f, long nx, long ny, long z, long y, long x, double d[3]) {
    const long i = (z * ny + y) * nx + x;
    d[0] = (z >= 1) ? f[i] - f[i - nx * ny] : 0.0;
    d[1] = (y >= 1) ? f[i] - f[i - nx] : 0.0;
    d[2] = (x >= 1) ? f[i] - f[i - 1] : 0.0;
}

/* Generalised sub-gradient: pad_z selects the GD_TV / GD_AwTV z padding (else
 * PICCS's unpadded volume); inv_delta != 0 weights each difference d by
 * exp(-(d inv_delta)^2) (GD_AwTV.cu:117-170); eps as in each reference kernel. */
void og_tv_gradient_ex(const double* f, double* g, long nx, long ny, long nz, int pad_z, double inv_delta,
                       double eps) {
#pragma omp parallel for schedule(static)
    for (long z = 0; z < nz; ++z)
        for (long y = 0; y < ny; ++y)
            for (long x = 0; x < nx; ++x) {
                double df[3], di[3] = {0, 0, 0}, dj[3] = {0, 0, 0}, dk[3] = {0, 0, 0};
                if (pad_z) {
                    bdiff(f, nx, ny, nz, z, y, x, df);
                    if (x + 1 < nx) bdiff(f, nx, ny, nz, z, y, x + 1, di);
                    if (y + 1 < ny) bdiff(f, nx, ny, nz, z, y + 1, x, dj);
                    bdiff(f, nx, ny, nz, z + 1, y, x, dk);
                } else {
                    bdiff_unpadded(f, nx, ny, z, y, x, df);
                    if (x + 1 < nx) bdiff_unpadded(f, nx, ny, z, y, x + 1, di);
                    if (y + 1 < ny) bdiff_unpadded(f, nx, ny, z, y + 1, x, dj);
                    if (z + 1 < nz) bdiff_unpadded(f, nx, ny, z + 1, y, x, dk);
                }
                double w[12];
                const double* all[4] = {df, di, dj, dk};
                for (int a = 0; a < 4; ++a)
                    for (int c = 0; c < 3; ++c) {
                        const double t = all[a][c] * inv_delta;
                        w[3 * a + c] = inv_delta != 0.0 ? exp(-t * t) : 1.0;
                    }
                double v = (w[0] * df[0] + w[1] * df[1] + w[2] * df[2]) /
                           (sqrt(w[0] * df[0] * df[0] + w[1] * df[1] * df[1] + w[2] * df[2] * df[2]) + eps);
                v -= w[5] * di[2] / (sqrt(w[3] * di[0] * di[0] + w[4] * di[1] * di[1] + w[5] * di[2] * di[2]) + eps);
                v -= w[7] * dj[1] / (sqrt(w[6] * dj[0] * dj[0] + w[7] * dj[1] * dj[1] + w[8] * dj[2] * dj[2]) + eps);
                v -= w[9] * dk[0] / (sqrt(w[9] * dk[0] * dk[0] + w[10] * dk[1] * dk[1] + w[11] * dk[2] * dk[2]) + eps);
                g[(z * ny + y) * nx + x] = v;
            }
}

/* aw_pocs_tv (GD_AwTV.cu:276-): pocs_tv with the weighted gradient. */
void og_aw_pocs_tv(const double* img, double* dst, double alpha, long nx, long ny, long nz, int maxiter,
                   double delta) {
    size_t n = (size_t)nx * ny * nz;
    double* g = (double*)malloc(n * sizeof(double));
    memcpy(dst, img, n * sizeof(double));
    for (int it = 0; it < maxiter; ++it) {
        og_tv_gradient_ex(dst, g, nx, ny, nz, 1, 1.0 / delta, 1e-8);
        double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
        for (long q = 0; q < (long)n; ++q) s += g[q] * g[q];
        double scale = alpha / sqrt(s + 1e-7);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < (long)n; ++q) dst[q] -= g[q] * scale;
    }
    free(g);
}

/* piccs_tv (PICCS.cu:274-380): g = (1 - ratio) grad TV(img) + ratio grad TV(img -
 * prior), both unpadded with eps 1e-6; img -= alpha g / sqrt(sum g^2). */
void og_piccs_tv(const double* img, const double* prior, double* dst, double alpha, double ratio, long nx, long ny,
                 long nz, int maxiter) {
    size_t n = (size_t)nx * ny * nz;
    double* g = (double*)malloc(n * sizeof(double));
    double* h = (double*)malloc(n * sizeof(double));
    double* d = (double*)malloc(n * sizeof(double));
    memcpy(dst, img, n * sizeof(double));
    for (int it = 0; it < maxiter; ++it) {
        og_tv_gradient_ex(dst, g, nx, ny, nz, 0, 0.0, 1e-6);
        for (size_t q = 0; q < n; ++q) d[q] = dst[q] - prior[q];
        og_tv_gradient_ex(d, h, nx, ny, nz, 0, 0.0, 1e-6);
        double s = 0.0;
        for (size_t q = 0; q < n; ++q) {
            g[q] = (1.0 - ratio) * g[q] + ratio * h[q];
            s += g[q] * g[q];
        }
        double scale = alpha / sqrt(s);
        for (size_t q = 0; q < n; ++q) dst[q] -= g[q] * scale;
    }
    free(g);
    free(h);
    free(d);
}

/* pocs_tv single-device iteration, GD_TV.cu:489-597:
 * g = grad TV(img); img -= alpha * g / sqrt(sum(g^2) + 1e-7). */
void og_pocs_tv(const double* img, double* dst, double alpha, long nx, long ny, long nz, int maxiter) {
    size_t n = (size_t)nx * ny * nz;
    double* g = (double*)malloc(n * sizeof(double));
    memcpy(dst, img, n * sizeof(double));
    for (int it = 0; it < maxiter; ++it) {
        og_tv_gradient(dst, g, nx, ny, nz);
        double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
        for (long q = 0; q < (long)n; ++q) s += g[q] * g[q];
        s += 1e-7;
        double scale = alpha / sqrt(s);
#pragma omp parallel for schedule(static)
        for (long q = 0; q < (long)n; ++q) dst[q] -= g[q] * scale;
    }
    free(g);
}

/* ------------------------------------------------------------------------ */
/* ROF primal-dual TV denoise, tv_proximal.cu                                */
/* ------------------------------------------------------------------------ */

/* Single-device domain (tv_proximal.cu:436-470 with cpy_from_host :672-693):
 * nz + 2 slices; slice 0 mirrors image slice 1, slices 1..nz hold the image,
 * slice nz + 1 is zero.  All slices, padding included, evolve; the result is
 * slices 1..nz of u. */
void og_tvdenoise(const double* src, double* dst, double lambda, const double spacing[3], long nx, long ny,
                  long nz, int maxiter) {
    const long D = nz + 2, sl = nx * ny;
    const size_t n = (size_t)D * sl;
    const double dx = spacing[0], dy = spacing[1], dz = spacing[2];
    double* f = (double*)calloc(n, sizeof(double));
    double* u = (double*)calloc(n, sizeof(double));
    double* px = (double*)calloc(n, sizeof(double));
    double* py = (double*)calloc(n, sizeof(double));
    double* pz = (double*)calloc(n, sizeof(double));
    if (nz > 1) memcpy(f, src + sl, sl * sizeof(double));          /* slice 0 <- image slice 1 */
    memcpy(f + sl, src, (size_t)nz * sl * sizeof(double));
    memcpy(u, f, n * sizeof(double));
    for (int it = 0; it < maxiter; ++it) {
        double tau2 = 0.3 + 0.02 * it;                                /* :478-479 */
        double tau1 = (1.0 / tau2) * ((1.0 / 6.0) - (5.0 / (15.0 + it)));
        /* update_u, :136-155 */
#pragma omp parallel for schedule(static)
        for (long z = 0; z < D; ++z)
            for (long y = 0; y < ny; ++y)
                for (long x = 0; x < nx; ++x) {
                    long idx = z * sl + y * nx + x;
                    double dv = 0.0;
                    dv += (z >= 1) ? (pz[idx] - pz[idx - sl]) / dz : pz[idx];
                    dv += (y >= 1) ? (py[idx] - py[idx - nx]) / dy : py[idx];
                    dv += (x >= 1) ? (px[idx] - px[idx - 1]) / dx : px[idx];
                    u[idx] = u[idx] * (1.0 - tau1) + tau1 * (f[idx] + (1.0 / lambda) * dv);
                }
        /* update_p, :158-181 */
#pragma omp parallel for schedule(static)
        for (long z = 0; z < D; ++z)
            for (long y = 0; y < ny; ++y)
                for (long x = 0; x < nx; ++x) {
                    long idx = z * sl + y * nx + x;
                    double gz = (z + 1 < D) ? (u[idx + sl] - u[idx]) / dz : 0.0;
                    double gy = (y + 1 < ny) ? (u[idx + nx] - u[idx]) / dy : 0.0;
                    double gx = (x + 1 < nx) ? (u[idx + 1] - u[idx]) / dx : 0.0;
                    double qz = pz[idx] + tau2 * gz, qy = py[idx] + tau2 * gy, qx = px[idx] + tau2 * gx;
                    double nrm = fmax(1.0, sqrt(qz * qz + qy * qy + qx * qx));
                    pz[idx] = qz / nrm;
                    py[idx] = qy / nrm;
                    px[idx] = qx / nrm;
                }
    }
    memcpy(dst, u + sl, (size_t)nz * sl * sizeof(double));
    free(f); free(u); free(px); free(py); free(pz);
}
