"""CPU: pins of the oracle's parallel-beam restatements (oracle/tigre_oracle.c).

Closed forms on the reference's default parallel geometry (1 mm voxels and
pixels, geometry.py:274-300): at alpha = 0 every ray runs along x through
voxel centres, so Siddon and the interpolated projector both give the sums of
the image along x; at alpha = pi / 2 the sums along y, mirrored in u.  The
backprojector at alpha = 0 copies each pixel to its row of voxels, which makes
it the exact adjoint of the Siddon projector there.
"""
import numpy as np
import pytest

import oracle
from tigre_amd.geometry import geometry, prepared

SHAPE = (6, 8, 10)


@pytest.fixture(scope="module")
def setup():
    geo = geometry(mode="parallel", nVoxel=np.array(SHAPE))
    img = np.random.default_rng(0).random(SHAPE).astype(np.float32)
    return geo, img


@pytest.mark.parametrize("kind", ["Siddon", "interpolated"])
def test_axis_aligned_views_are_sums(setup, kind):
    geo, img = setup
    p = oracle.ax(img, prepared(geo, np.array([0.0, np.pi / 2])), kind)
    s0 = img.astype(np.float64).sum(axis=2)  # (z, y)
    s1 = img.astype(np.float64).sum(axis=1)  # (z, x)
    # detector 10 wide, volume 8 wide in y: pixels 1..8 see voxel rows 0..7
    np.testing.assert_allclose(p[0][:, 1:9], s0, rtol=1e-6, atol=1e-6)
    assert not p[0][:, [0, 9]].any()
    np.testing.assert_allclose(p[1], s1[:, ::-1], rtol=1e-6, atol=1e-6)


def test_backprojection_axis_aligned_is_adjoint(setup):
    geo, img = setup
    g = prepared(geo, np.array([0.0]))
    proj = np.random.default_rng(1).random((1,) + tuple(int(v) for v in geo.nDetector)).astype(np.float32)
    back = oracle.atb(proj, g, "FDK")
    # every voxel (z, y, x) lands on pixel centre (z, y + 1)
    np.testing.assert_allclose(back, np.broadcast_to(proj[0][:, 1:9, None], SHAPE), rtol=1e-12, atol=1e-12)
    lhs = np.sum(oracle.ax(img, g, "Siddon") * proj)
    assert np.isclose(lhs, np.sum(img * back), rtol=1e-9)
    # one backprojector serves 'FDK' and 'matched'
    assert np.array_equal(back, oracle.atb(proj, g, "matched"))


def test_rays_are_parallel(setup):
    geo, _ = setup
    g = prepared(geo, np.array([0.7]))
    f = oracle.forward_frame(g, 0, 0)  # pixel (0, 0), u step, v step, source
    d = f[0] - f[3]
    # the ray direction does not depend on the pixel: source and pixel move together
    dsd = float(np.asarray(g.DSD).reshape(-1)[0])
    assert np.allclose(d, [-dsd * np.cos(0.7), -dsd * np.sin(0.7), 0.0], atol=1e-5)
