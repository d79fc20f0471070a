"""Pins of the oracle's forward projectors against independent statements.

- Siddon (Siddon_projection.cu:110-268) against the exact ray/voxel-box
  intersection lengths of tests/refgeom.py (world frame, brute force), and a
  constant volume against the ray/box chord length.
- Interpolated (ray_interpolated_projection.cu:117-206) against scipy's
  trilinear map_coordinates(mode='grid-constant') at the reference's uniform
  sample positions, computed in the world frame.
- The voxelised ellipsoid phantom against its closed-form line integrals.
"""
import numpy as np
import pytest
from scipy.ndimage import map_coordinates

import oracle
from tests import refgeom
from tigre_amd.geometry import cone_geometry, prepared
from tigre_amd.phantom import ellipsoid_line_integral, ellipsoid_params, ellipsoid_phantom


def small_geo(n=(4, 5, 6), s=(6.0, 5.5, 7.2), det=(6, 7), ddet=(2.0, 1.8), DSO=60.0, DSD=100.0):
    return cone_geometry(n, s, det, ddet, DSO, DSD)


def tilted(g, nviews, seed):
    rng = np.random.default_rng(seed)
    g.offOrigin = rng.uniform(-0.6, 0.6, size=(nviews, 3))
    g.offDetector = rng.uniform(-1.0, 1.0, size=(nviews, 2))
    g.rotDetector = rng.uniform(-0.1, 0.1, size=(nviews, 3))
    g.COR = rng.uniform(-0.4, 0.4, size=nviews)
    ang = np.stack([rng.uniform(0, 2 * np.pi, nviews), rng.uniform(-0.3, 0.3, nviews),
                    rng.uniform(-0.3, 0.3, nviews)], axis=1)
    return g, ang


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_siddon_brute_force_tiny_volume(seed):
    g = small_geo()
    nviews = 3
    g, ang = tilted(g, nviews, seed)
    gp = prepared(g, ang)
    img = np.random.default_rng(seed + 10).uniform(0.1, 1.0, size=tuple(g.nVoxel)).astype(np.float32)
    got = oracle.ax(img, gp, "Siddon")
    nV, nU = int(g.nDetector[0]), int(g.nDetector[1])
    for i in range(nviews):
        for v in range(nV):
            for u in range(nU):
                w = refgeom.voxel_path_lengths(gp, i, u, v)
                ref = float((w * img).sum())
                assert got[i, v, u] == pytest.approx(ref, rel=1e-9, abs=1e-9), (i, v, u)


def test_siddon_constant_volume_is_chord_length():
    g = cone_geometry((32, 40, 36), (32.0, 36.0, 40.0), (48, 52), (1.6, 1.4), 500.0, 800.0)
    g, ang = tilted(g, 6, 7)
    gp = prepared(g, ang)
    img = np.full(tuple(g.nVoxel), 2.5, dtype=np.float32)
    got = oracle.ax(img, gp, "Siddon")
    nV, nU = int(g.nDetector[0]), int(g.nDetector[1])
    for i in range(6):
        lo, hi = refgeom.box(gp, i)
        for v in range(0, nV, 5):
            for u in range(0, nU, 5):
                S, P = refgeom.source_and_pixel(gp, i, u, v)
                assert got[i, v, u] == pytest.approx(2.5 * refgeom.segment_box_length(S, P, lo, hi),
                                                     rel=1e-9, abs=1e-9)


def test_siddon_row_orientation():
    """Only the top detector rows see a slab placed in the upper half of the volume."""
    g = cone_geometry((16, 16, 16), (16.0, 16.0, 16.0), (20, 20), (1.6, 1.6), 200.0, 300.0)
    gp = prepared(g, np.array([0.3]))
    img = np.zeros((16, 16, 16), dtype=np.float32)
    img[12:, :, :] = 1.0      # z index 12..15: positive z
    p = oracle.ax(img, gp)[0]
    assert p[:10].sum() == 0.0 and p[10:].sum() > 0.0


def _interp_reference(img, gp, i, u, v):
    """Sum of zero-border trilinear samples at S + k (P - S)/n, k = 0..n, times the step length."""
    S, P = refgeom.source_and_pixel(gp, i, u, v)
    lo, _ = refgeom.box(gp, i)
    d = refgeom.voxel_size(gp)
    Si = (S - lo) / d - 0.5          # voxel-centre index coordinates (x, y, z)
    Pi = (P - lo) / d - 0.5
    acc = float(gp.accuracy)
    n = np.ceil(np.linalg.norm(Pi - Si) / acc)
    k = np.arange(0, n + 1)
    pts = Si[None, :] + k[:, None] * ((Pi - Si) / n)[None, :]
    vals = map_coordinates(img.astype(np.float64), [pts[:, 2], pts[:, 1], pts[:, 0]], order=1,
                           mode="grid-constant", cval=0.0)
    step_mm = np.linalg.norm((Pi - Si) / n * d)
    return vals.sum() * step_mm


@pytest.mark.parametrize("seed", [3, 4])
def test_interpolated_matches_scipy_trilinear(seed):
    g = cone_geometry((10, 12, 14), (10.0, 13.2, 12.6), (14, 16), (1.5, 1.5), 80.0, 130.0)
    g, ang = tilted(g, 3, seed)
    gp = prepared(g, ang)
    img = np.random.default_rng(seed).uniform(0, 1, size=tuple(g.nVoxel)).astype(np.float32)
    got = oracle.ax(img, gp, "interpolated")
    for i in range(3):
        for v in range(0, 14, 3):
            for u in range(0, 16, 3):
                ref = _interp_reference(img, gp, i, u, v)
                assert got[i, v, u] == pytest.approx(ref, rel=1e-9, abs=1e-9), (i, v, u)


def test_interpolated_reproduces_linear_field_inside():
    """Trilinear sampling is exact for a linear field: ray sum = step * sum of the field at the samples."""
    # 4 x 4 detector: no pixel sits on the central ray, whose length DSD / accuracy
    # would be an exact integer and make the sample count ceil(len / acc) ambiguous
    g = cone_geometry((24, 24, 24), (24.0, 24.0, 24.0), (4, 4), (0.5, 0.5), 300.0, 400.0)
    gp = prepared(g, np.array([0.4]))
    z, y, x = np.meshgrid(*[np.arange(24.0)] * 3, indexing="ij")
    img = (0.5 + 0.01 * x + 0.02 * y + 0.03 * z).astype(np.float32)
    got = oracle.ax(img, gp, "interpolated")[0]
    for v, u in [(1, 1), (1, 2), (2, 2), (0, 3)]:
        assert got[v, u] == pytest.approx(_interp_reference(img, gp, 0, u, v), rel=1e-9)


def _phantom_error(n, params):
    g = cone_geometry((n, n, n), (80.0, 80.0, 80.0), (24, 24), (2.5, 2.5), 400.0, 600.0)
    ang = np.array([0.0, 1.1, 2.3])
    gp = prepared(g, ang)
    img = ellipsoid_phantom(g.nVoxel, g.sVoxel, len(params), seed=0, params=params)
    got = oracle.ax(img, gp, "Siddon")
    num, den = 0.0, 0.0
    for i in range(len(ang)):
        for v in range(24):
            for u in range(24):
                S, P = refgeom.source_and_pixel(gp, i, u, v)
                ref = ellipsoid_line_integral(params, S[::-1], (P - S)[::-1])
                num += (got[i, v, u] - ref) ** 2
                den += ref ** 2
    return np.sqrt(num / den)


def test_phantom_line_integrals_converge():
    """Siddon of the voxelised phantom converges to the analytic line integrals
    of the continuous phantom as the voxels shrink (a flipped or rotated frame
    would not converge)."""
    params = ellipsoid_params(6, (80.0, 80.0, 80.0), seed=5, axes_mm=(5.0, 12.0))  # fully inside the box
    e80, e160 = _phantom_error(80, params), _phantom_error(160, params)
    assert e160 < 0.07 and e160 < 0.6 * e80


def test_frames_closed_form():
    """Forward frame of an unrotated view: source on +x, detector on -x, voxel units."""
    g = cone_geometry((8, 10, 12), (8.0, 15.0, 24.0), (6, 4), (2.0, 3.0), 100.0, 160.0)
    gp = prepared(g, np.array([0.0]))
    f = oracle.forward_frame(gp, 0, 0)
    # sVoxel (x,y,z) = (24, 15, 8) mm; dVoxel = (2, 1.5, 1) mm
    assert np.allclose(f[3], [(100.0 + 12.0) / 2.0, 7.5 / 1.5, 4.0 / 1.0])          # source
    # nDetector (v, u) = (6, 4), dDetector (v, u) = (2, 3) mm
    assert np.allclose(f[0], [(-60.0 + 12.0) / 2.0, (3.0 * (0.5 - 2.0) + 7.5) / 1.5,
                              (2.0 * (3.0 - 0.5) + 4.0) / 1.0])                         # pixel (0, top row)
    assert np.allclose(f[1], [0.0, 3.0 / 1.5, 0.0]) and np.allclose(f[2], [0.0, 0.0, -2.0])
    fc = oracle.forward_frame(gp, 0, 1)
    assert np.allclose(fc[3] - f[3], [-0.5, -0.5, -0.5])


def _batch_geometry():
    """Voxels 400x finer in z than in x/y: the planar sample cap
    (2 DSO / min(dVx, dVy) + crop) / accuracy ends inside the volume for the
    steep rows, the three-axis cap (min over dVz as well) beyond it."""
    g = cone_geometry((400, 8, 8), (2.0, 16.0, 16.0), (16, 8), (0.2, 3.0), 100.0, 150.0)
    ang = np.zeros((9, 3))
    ang[:, 0] = np.linspace(0.1, 1.7, 9)
    ang[0, 1] = 0.02   # only the first view of the batch leaves the circular orbit
    return g, ang


def test_interpolated_cap_flag_is_per_batch():
    """The interpolated projector picks its sample cap with one flag per launch
    batch of 9 views (ray_interpolated_projection.cu:352-392): planar views that
    share a batch with a tilted view get the three-axis cap, which here lies
    beyond the volume, so they equal the uncapped ray sums; the same views
    projected without the tilted view get the planar cap, which cuts the steep
    rows short."""
    g, ang = _batch_geometry()
    img = np.random.default_rng(9).uniform(0.5, 1.0, size=tuple(g.nVoxel)).astype(np.float32)
    with_tilted = oracle.ax(img, prepared(g, ang), "interpolated")
    alone = oracle.ax(img, prepared(g, ang[1:]), "interpolated")
    gp = prepared(g, ang)
    for i in (1, 4, 8):
        for v in (0, 3, 8, 15):
            for u in (0, 3, 7):
                ref = _interp_reference(img, gp, i, u, v)
                assert with_tilted[i, v, u] == pytest.approx(ref, rel=1e-9), (i, v, u)
    assert np.all(alone <= with_tilted[1:] * (1 + 1e-12))
    cut = (with_tilted[1:] - alone) / with_tilted[1:]
    assert cut.max() > 0.05          # steep rows lose samples inside the volume
    assert cut[:, 7:9, :].max() < 1e-12  # the central rows keep all of theirs
