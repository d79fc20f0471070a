// Host-side geometry core: imports the caller's float32 per-projection
// geometry into FP64 and builds the per-view records the kernels consume.
//
// The frames follow the reference's conventions:
//   forward  - computeDeltas_Siddon (Common/CUDA/Siddon_projection.cu:666-763)
//              and computeDeltas (Common/CUDA/ray_interpolated_projection.cu:632-741);
//   backward - computeDeltasCube (Common/CUDA/voxel_backprojection.cu:826-898)
//              with the angles negated (voxel_backprojection.cu:510-512).
// Here every rotation is built once per view as a 3x3 matrix in FP64 and
// applied to the detector corner/steps, instead of rotating points one by one.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.hpp"

namespace tigre {

namespace {

struct M3 {
    double a[3][3];
};
struct V3 {
    double x, y, z;
};

inline V3 mul(const M3& m, V3 v) {
    return {m.a[0][0] * v.x + m.a[0][1] * v.y + m.a[0][2] * v.z,
            m.a[1][0] * v.x + m.a[1][1] * v.y + m.a[1][2] * v.z,
            m.a[2][0] * v.x + m.a[2][1] * v.y + m.a[2][2] * v.z};
}
inline M3 mul(const M3& p, const M3& q) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.a[i][j] = p.a[i][0] * q.a[0][j] + p.a[i][1] * q.a[1][j] + p.a[i][2] * q.a[2][j];
    return r;
}
inline M3 transpose(const M3& m) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.a[i][j] = m.a[j][i];
    return r;
}

// Z(a) * Y(t) * Z(p): the ZYZ Euler rotation of Siddon_projection.cu:799-818.
M3 euler_zyz(double a, double t, double p) {
    const double ca = std::cos(a), sa = std::sin(a), ct = std::cos(t), st = std::sin(t);
    const double cp = std::cos(p), sp = std::sin(p);
    M3 z1{{{ca, -sa, 0}, {sa, ca, 0}, {0, 0, 1}}};
    M3 y{{{ct, 0, st}, {0, 1, 0}, {-st, 0, ct}}};
    M3 z2{{{cp, -sp, 0}, {sp, cp, 0}, {0, 0, 1}}};
    return mul(mul(z1, y), z2);
}

// Z(roll) * Y(pitch) * X(yaw): the detector rotation of projection.cpp:18-35.
M3 roll_pitch_yaw(double roll, double pitch, double yaw) {
    const double cr = std::cos(roll), sr = std::sin(roll), cp = std::cos(pitch), sp = std::sin(pitch);
    const double cw = std::cos(yaw), sw = std::sin(yaw);
    M3 z{{{cr, -sr, 0}, {sr, cr, 0}, {0, 0, 1}}};
    M3 y{{{cp, 0, sp}, {0, 1, 0}, {-sp, 0, cp}}};
    M3 x{{{1, 0, 0}, {0, cw, -sw}, {0, sw, cw}}};
    return mul(mul(z, y), x);
}

// eulerZYZT (voxel_backprojection.cu:762-787) evaluated at the negated
// angles is Z(-p) * Y(-t) * Z(-a)... expressed through the same builders:
// its matrix equals Z(p') Y(t') Z(a') with primed = negated angles, i.e. the
// ZYZ product with the roles of alpha and psi exchanged.
M3 euler_zyz_t(double a, double t, double p) { return euler_zyz(p, t, a); }

}  // namespace

int import_geometry(const TigreGeometry* geo, const float* angles, int nangles, HostGeometry* g) {
    TIGRE_REQUIRE(geo != nullptr, "geometry pointer is null");
    TIGRE_REQUIRE(angles != nullptr, "angles pointer is null");
    TIGRE_REQUIRE(nangles > 0, "nangles must be positive");
    TIGRE_REQUIRE(geo->nVoxelX > 0 && geo->nVoxelY > 0 && geo->nVoxelZ > 0, "nVoxel must be positive");
    TIGRE_REQUIRE(geo->nDetecU > 0 && geo->nDetecV > 0, "nDetector must be positive");
    TIGRE_REQUIRE(geo->dVoxelX > 0 && geo->dVoxelY > 0 && geo->dVoxelZ > 0, "dVoxel must be positive");
    TIGRE_REQUIRE(geo->dDetecU > 0 && geo->dDetecV > 0, "dDetector must be positive");
    TIGRE_REQUIRE(geo->DSO && geo->DSD && geo->offOrigX && geo->offOrigY && geo->offOrigZ && geo->offDetecU &&
                      geo->offDetecV && geo->dRoll && geo->dPitch && geo->dYaw && geo->COR,
                  "geometry per-view arrays must all be set");
    g->nVoxel[0] = geo->nVoxelX;
    g->nVoxel[1] = geo->nVoxelY;
    g->nVoxel[2] = geo->nVoxelZ;
    g->sVoxel[0] = geo->sVoxelX;
    g->sVoxel[1] = geo->sVoxelY;
    g->sVoxel[2] = geo->sVoxelZ;
    g->dVoxel[0] = geo->dVoxelX;
    g->dVoxel[1] = geo->dVoxelY;
    g->dVoxel[2] = geo->dVoxelZ;
    g->nDetec[0] = geo->nDetecU;
    g->nDetec[1] = geo->nDetecV;
    g->sDetec[0] = geo->sDetecU;
    g->sDetec[1] = geo->sDetecV;
    g->dDetec[0] = geo->dDetecU;
    g->dDetec[1] = geo->dDetecV;
    g->accuracy = geo->accuracy;
    TIGRE_REQUIRE(g->accuracy > 0, "accuracy must be positive");
    g->nangles = nangles;
    g->angles.assign(angles, angles + 3 * (size_t)nangles);
    g->DSO.resize(nangles);
    g->DSD.resize(nangles);
    g->COR.resize(nangles);
    g->offOrig.resize(3 * (size_t)nangles);
    g->offDetec.resize(2 * (size_t)nangles);
    g->rot.resize(3 * (size_t)nangles);
    for (int i = 0; i < nangles; ++i) {
        g->DSO[i] = geo->DSO[i];
        g->DSD[i] = geo->DSD[i];
        g->COR[i] = geo->COR[i];
        g->offOrig[3 * i + 0] = geo->offOrigX[i];
        g->offOrig[3 * i + 1] = geo->offOrigY[i];
        g->offOrig[3 * i + 2] = geo->offOrigZ[i];
        g->offDetec[2 * i + 0] = geo->offDetecU[i];
        g->offDetec[2 * i + 1] = geo->offDetecV[i];
        g->rot[3 * i + 0] = geo->dRoll[i];
        g->rot[3 * i + 1] = geo->dPitch[i];
        g->rot[3 * i + 2] = geo->dYaw[i];
        TIGRE_REQUIRE(std::isfinite(g->DSO[i]) && std::isfinite(g->DSD[i]), "DSO and DSD must be finite");
    }
    return TIGRE_OK;
}

void forward_views(const HostGeometry& g, bool centre, std::vector<FwdView>* out) {
    out->resize(g.nangles);
    const double dU = g.dDetec[0], dV = g.dDetec[1];
    const int nU = g.nDetec[0], nV = g.nDetec[1];
    for (int i = 0; i < g.nangles; ++i) {
        const double al = g.angles[3 * i], th = g.angles[3 * i + 1], ps = g.angles[3 * i + 2];
        const double DSO = g.DSO[i], DSD = g.DSD[i];
        const M3 E = euler_zyz(al, th, ps);
        const M3 D = roll_pitch_yaw(g.rot[3 * i], g.rot[3 * i + 1], g.rot[3 * i + 2]);
        const M3 ED = mul(E, D);
        // Detector pixel (0,0) about the detector centre, then moved to the
        // detector plane and offset, then rotated with the gantry.
        V3 corner = mul(D, V3{0.0, dU * (0.5 - nU / 2.0), dV * (nV / 2.0 - 0.5)});
        corner.x -= (DSD - DSO);
        corner.y += g.offDetec[2 * i + 0];
        corner.z += g.offDetec[2 * i + 1];
        V3 P = mul(E, corner);
        V3 S = mul(E, V3{DSO, 0.0, 0.0});
        V3 u = mul(ED, V3{0.0, dU, 0.0});
        V3 v = mul(ED, V3{0.0, 0.0, -dV});
        if (g.parallel) {
            // Parallel beam: the source of pixel (0, 0) is that pixel's detector
            // point moved to x = DSO before the offsets (no detector offset on
            // the source), and every pixel carries its own source.  Siddon
            // rotates by alpha alone (Siddon_projection_parallel.cu:388-470);
            // the interpolated projector applies the full ZYZ rotation and the
            // detector offset twice (ray_interpolated_projection_parallel.cu
            // :313-390), both kept as the reference does.
            V3 s0 = mul(D, V3{0.0, dU * (0.5 - nU / 2.0), dV * (nV / 2.0 - 0.5)});
            s0.x += DSO;
            const M3 R = centre ? E : euler_zyz(al, 0.0, 0.0);
            if (centre) {
                corner.y += g.offDetec[2 * i + 0];
                corner.z += g.offDetec[2 * i + 1];
            }
            P = mul(R, corner);
            S = mul(R, s0);
            const M3 RD = mul(R, D);
            u = mul(RD, V3{0.0, dU, 0.0});
            v = mul(RD, V3{0.0, 0.0, -dV});
        }
        FwdView& r = (*out)[i];
        const double corx = -g.COR[i] * std::sin(al), cory = g.COR[i] * std::cos(al);
        const double* off = &g.offOrig[3 * i];
        double pos[2][3] = {{P.x, P.y, P.z}, {S.x, S.y, S.z}};
        for (int k = 0; k < 2; ++k) {
            for (int c = 0; c < 3; ++c) {
                double shift = g.sVoxel[c] / 2.0 - (centre ? g.dVoxel[c] / 2.0 : 0.0);
                pos[k][c] = (pos[k][c] - off[c] + shift) / g.dVoxel[c];
            }
            pos[k][0] += corx / g.dVoxel[0];
            pos[k][1] += cory / g.dVoxel[1];
        }
        const double uu[3] = {u.x, u.y, u.z}, vv[3] = {v.x, v.y, v.z};
        for (int c = 0; c < 3; ++c) {
            r.org[c] = pos[0][c];
            r.src[c] = pos[1][c];
            r.du[c] = uu[c] / g.dVoxel[c];
            r.dv[c] = vv[c] / g.dVoxel[c];
        }
        // Interpolated projector's sample window (ray_interpolated_projection.cu:365-392,
        // :743-768): crop = floor(safe skip distance), first = floor(crop/acc),
        // and the cap on the last sample.
        const double mx = (double)(g.nVoxel[0] / 2) + std::fabs(off[0]) / g.dVoxel[0];
        const double my = (double)(g.nVoxel[1] / 2) + std::fabs(off[1]) / g.dVoxel[1];
        const double mz = (double)(g.nVoxel[2] / 2) + std::fabs(off[2]) / g.dVoxel[2];
        double maxdist;
        const bool planar = (th == 0.0 && ps == 0.0);
        if (planar) {
            const double a = DSO / g.dVoxel[0], b = DSO / g.dVoxel[1];
            const double s = std::sin(al), c = std::cos(al);
            maxdist = a * b / std::sqrt(a * a * s * s + b * b * c * c) - std::sqrt(mx * mx + my * my);
        } else {
            const double dmax = std::fmax(std::fmax(g.dVoxel[0], g.dVoxel[1]), g.dVoxel[2]);
            maxdist = DSO / dmax - std::sqrt(mx * mx + my * my + mz * mz);
        }
        if (maxdist < 0.0) maxdist = 0.0;
        const double crop = std::floor(maxdist);
        // The cap's voxel size comes from a flag of the whole launch batch: the
        // reference projects 9 views per kernel and takes the three-axis minimum
        // for all of them when any view of the batch has theta or psi nonzero
        // (ray_interpolated_projection.cu:352-392, PROJ_PER_BLOCK = 9); crop
        // above is per view.
        bool batch_spherical = false;
        const int b0 = (i / 9) * 9, b1 = std::min(b0 + 9, g.nangles);
        for (int j = b0; j < b1; ++j)
            if (g.angles[3 * j + 1] != 0.0 || g.angles[3 * j + 2] != 0.0) batch_spherical = true;
        const double dmin = batch_spherical ? std::fmin(std::fmin(g.dVoxel[0], g.dVoxel[1]), g.dVoxel[2])
                                            : std::fmin(g.dVoxel[0], g.dVoxel[1]);
        r.first = std::floor(crop / g.accuracy);
        // parallel beam: the cap uses dVoxelX (ray_interpolated_projection_parallel.cu:175-176)
        r.cap = (2.0 * DSO / (g.parallel ? g.dVoxel[0] : dmin) + crop) / g.accuracy;
        r.pad[0] = (double)i;
        r.pad[1] = g.parallel ? 1.0 : 0.0;
    }
}

void backward_views(const HostGeometry& g, std::vector<BwdView>* out) {
    out->resize(g.nangles);
    const double dU = g.dDetec[0], dV = g.dDetec[1];
    for (int i = 0; i < g.nangles; ++i) {
        const double al = -g.angles[3 * i], th = -g.angles[3 * i + 1], ps = -g.angles[3 * i + 2];
        const double DSO = g.DSO[i], DSD = g.DSD[i];
        const double offU = g.offDetec[2 * i + 0], offV = g.offDetec[2 * i + 1];
        const M3 R = euler_zyz_t(al, th, ps);
        const M3 DT = transpose(roll_pitch_yaw(g.rot[3 * i], g.rot[3 * i + 1], g.rot[3 * i + 2]));
        const M3 A = mul(DT, R);
        const double* off = &g.offOrig[3 * i];
        const V3 P0{-(g.sVoxel[0] / 2.0 - g.dVoxel[0] / 2.0) + off[0],
                    -(g.sVoxel[1] / 2.0 - g.dVoxel[1] / 2.0) + off[1],
                    -(g.sVoxel[2] / 2.0 - g.dVoxel[2] / 2.0) + off[2]};
        // Point map: detector-frame position in (mm, pixel, pixel) units.
        auto place = [&](V3 p) {
            V3 q = mul(R, p);
            q.z -= offV;
            q.y -= offU;
            q.x += (DSD - DSO);
            q = mul(DT, q);
            q.x -= (DSD - DSO);
            return V3{q.x, q.y / dU, q.z / dV};
        };
        auto dir = [&](V3 d) {
            V3 q = mul(A, d);
            return V3{q.x, q.y / dU, q.z / dV};
        };
        V3 O = place(P0);
        V3 dx = dir(V3{g.dVoxel[0], 0, 0});
        V3 dy = dir(V3{0, g.dVoxel[1], 0});
        V3 dz = dir(V3{0, 0, g.dVoxel[2]});
        V3 S = mul(DT, V3{DSD, -offU, -offV});
        S.x -= (DSD - DSO);
        S.y /= dU;
        S.z /= dV;
        BwdView& r = (*out)[i];
        const double COR = g.COR[i];
        r.org[0] = (float)(O.x - S.x);
        r.org[1] = (float)(O.y - COR / dU - S.y);
        r.org[2] = (float)(O.z - S.z);
        r.dX[0] = (float)dx.x; r.dX[1] = (float)dx.y; r.dX[2] = (float)dx.z;
        r.dY[0] = (float)dy.x; r.dY[1] = (float)dy.y; r.dY[2] = (float)dy.z;
        r.dZ[0] = (float)dz.x; r.dZ[1] = (float)dz.y; r.dZ[2] = (float)dz.z;
        r.src[0] = (float)S.x; r.src[1] = (float)S.y; r.src[2] = (float)S.z;
        r.tnum = (float)(DSO - DSD - S.x);
        // detector coordinates in the quad copy's storage frame: pixel index + 1
        r.uoff = (float)(S.y + g.nDetec[0] * 0.5 - 0.5 + 1.0);
        r.voff = (float)(S.z + g.nDetec[1] * 0.5 - 0.5 + 1.0);
        const double sa = std::sin(al), ca = std::cos(al);
        r.DSD = (float)DSD;
        // real position of voxel (ix, iy): rx = realx0 + ix dVx, ry = realy0 + iy dVy
        const double realx0 = -(g.sVoxel[0] - g.dVoxel[0]) * 0.5 + off[0];
        const double realy0 = -(g.sVoxel[1] - g.dVoxel[1]) * 0.5 + off[1] + COR;
        r.wfdk[0] = (float)((DSO + realy0 * sa - realx0 * ca) / DSO);
        r.wfdk[1] = (float)(-g.dVoxel[0] * ca / DSO);
        r.wfdk[2] = (float)(g.dVoxel[1] * sa / DSO);
        for (int c = 0; c < 3; ++c) {
            r.mrv0[c] = (float)((-g.sVoxel[c] / 2 + g.dVoxel[c] / 2 + off[c]) / DSD);
            r.mdv[c] = (float)(g.dVoxel[c] / DSD);
        }
        r.mS[0] = (float)(DSO * ca / DSD);
        r.mS[1] = (float)(-DSO * sa / DSD);
        // realD uses the texture coordinate u = pixel + 0.5 (voxel_backprojection2.cu:259-260);
        // the kernel's coordinate is pixel + 1, hence the further - dU, - dV
        r.mdet[0] = (float)(dU / DSD);
        r.mdet[1] = (float)(dV / DSD);
        r.mdet[2] = (float)(((-g.sDetec[0] + dU) * 0.5 + offU + 0.5 * dU - dU) / DSD);
        r.mdet[3] = (float)(((-g.sDetec[1] + dV) * 0.5 + offV + 0.5 * dV - dV) / DSD);
        for (float& x : r.pad) x = 0.0f;
    }
}

bool backward_is_axial(const HostGeometry&, const std::vector<BwdView>& v) {
    for (const BwdView& r : v)
        if (r.dZ[0] != 0.0f || r.dZ[1] != 0.0f) return false;
    return true;
}

}  // namespace tigre
