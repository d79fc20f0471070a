"""Pins of the oracle's backprojectors.

A projection that is linear in the pixel coordinates, p(u, v) = a u + b v + c,
is reproduced exactly by bilinear sampling, so a voxel whose footprint stays
inside the detector receives  sum_views W(view, voxel) (a u + b v + c)  with
(u, v) the voxel centre's projection.  The projection is computed here from the
world-frame model in tests/refgeom.py and the weights from their formulas:
  FDK     W = 1 / U^2, U = (DSO - x cos(a) - y sin(a)) / DSO
          (voxel_backprojection.cu:241-249),
  matched W = L^3 / (DSD |S - X|^2) * dVx dVy dVz / (dU dV)
          (voxel_backprojection2.cu:201-279, :585).
"""
import numpy as np
import pytest

import oracle
from tests import refgeom
from tigre_amd.geometry import cone_geometry, prepared


def ramp(g, n, a, b, c):
    nV, nU = int(g.nDetector[0]), int(g.nDetector[1])
    vv, uu = np.meshgrid(np.arange(nV, dtype=np.float64), np.arange(nU, dtype=np.float64), indexing="ij")
    return np.broadcast_to((a * uu + b * vv + c), (n, nV, nU)).astype(np.float32).copy()


def geo_with_views(tilt):
    g = cone_geometry((6, 7, 8), (6.0, 7.7, 7.2), (40, 44), (1.0, 1.1), 120.0, 200.0)
    n = 5
    rng = np.random.default_rng(1)
    if tilt:
        g.offOrigin = rng.uniform(-0.5, 0.5, size=(n, 3))
        g.offDetector = rng.uniform(-2, 2, size=(n, 2))
        g.rotDetector = rng.uniform(-0.05, 0.05, size=(n, 3))
        ang = np.stack([rng.uniform(0, 2 * np.pi, n), rng.uniform(-0.2, 0.2, n), rng.uniform(-0.2, 0.2, n)], 1)
    else:
        ang = np.stack([rng.uniform(0, 2 * np.pi, n), np.zeros(n), np.zeros(n)], 1)
    return g, ang


@pytest.mark.parametrize("tilt", [False, True])
def test_fdk_linear_projection(tilt):
    g, ang = geo_with_views(tilt)
    gp = prepared(g, ang)
    a, b, c = 0.375, -0.25, 5.0   # dyadic: the float32 ramp is exact
    proj = ramp(gp, len(ang), a, b, c)
    got = oracle.atb(proj, gp, "FDK")
    nz, ny, nx = [int(x) for x in gp.nVoxel]
    for iz in range(nz):
        for iy in range(ny):
            for ix in range(nx):
                ref = 0.0
                for i in range(len(ang)):
                    X = refgeom.voxel_centre(gp, i, iz, iy, ix)
                    u, v = refgeom.project_point(gp, i, X)
                    al = float(gp.angles[i][0])
                    DSO = float(np.asarray(gp.DSO).reshape(-1)[i])
                    U = (DSO - X[0] * np.cos(al) - X[1] * np.sin(al)) / DSO
                    ref += (a * u + b * v + c) / U ** 2
                assert got[iz, iy, ix] == pytest.approx(ref, rel=1e-9)


def test_matched_linear_projection():
    g, ang = geo_with_views(False)
    gp = prepared(g, ang)
    a, b, c = -0.125, 0.3125, 7.0
    proj = ramp(gp, len(ang), a, b, c)
    got = oracle.atb(proj, gp, "matched")
    nz, ny, nx = [int(x) for x in gp.nVoxel]
    nV, nU = int(gp.nDetector[0]), int(gp.nDetector[1])
    dV, dU = float(gp.dDetector[0]), float(gp.dDetector[1])
    dvx, dvy, dvz = [float(x) for x in np.asarray(gp.dVoxel)[::-1]]
    for iz in range(0, nz, 2):
        for iy in range(ny):
            for ix in range(nx):
                ref = 0.0
                for i in range(len(ang)):
                    X = refgeom.voxel_centre(gp, i, iz, iy, ix)
                    u, v = refgeom.project_point(gp, i, X)
                    al = float(gp.angles[i][0])
                    DSO = float(np.asarray(gp.DSO).reshape(-1)[i])
                    DSD = float(np.asarray(gp.DSD).reshape(-1)[i])
                    S = np.array([DSO * np.cos(al), DSO * np.sin(al), 0.0])
                    # detector point of texture coordinate (u + 1/2, v + 1/2), rotated by alpha only
                    y = (-nU * dU + dU) / 2 + (u + 0.5) * dU
                    z = (-nV * dV + dV) / 2 + (v + 0.5) * dV
                    D = refgeom.rz(al) @ np.array([-(DSD - DSO), y, 0.0])
                    L = np.sqrt((S[0] - D[0]) ** 2 + (S[1] - D[1]) ** 2 + z ** 2)
                    W = L ** 3 / (DSD * np.sum((S - X) ** 2))
                    ref += W * (a * u + b * v + c)
                ref *= dvx * dvy * dvz / (dU * dV)
                assert got[iz, iy, ix] == pytest.approx(ref, rel=1e-9)


def test_fdk_border_is_zero():
    """Voxels projecting outside the detector receive nothing (zero-border sampling)."""
    g = cone_geometry((4, 4, 4), (40.0, 40.0, 40.0), (4, 4), (0.5, 0.5), 300.0, 400.0)
    gp = prepared(g, np.array([0.0]))
    got = oracle.atb(np.ones((1, 4, 4), dtype=np.float32), gp, "FDK")
    assert got[0, 0, 0] == 0.0 and got[3, 3, 3] == 0.0


def test_matched_is_near_adjoint_of_siddon():
    """'matched' approximates the transpose of Siddon: <Ax, y> ~ <x, A^T y> (gap recorded, not zero)."""
    g = cone_geometry((24, 24, 24), (24.0, 24.0, 24.0), (32, 32), (1.2, 1.2), 200.0, 300.0)
    ang = np.linspace(0, 2 * np.pi, 12, endpoint=False)
    gp = prepared(g, ang)
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 1, size=(24, 24, 24)).astype(np.float32)
    y = rng.uniform(0, 1, size=(12, 32, 32)).astype(np.float32)
    lhs = float((oracle.ax(x, gp) * y).sum())
    rhs = float((x * oracle.atb(y, gp, "matched")).sum())
    assert abs(lhs - rhs) / abs(lhs) < 0.05
