"""Pins of the oracle's algorithm restatements (oracle/algorithms.py).

The reference's own algorithm test thresholds a normalised error of the
reconstruction against the phantom (Python/tests/algorithm_test.py:62-65);
here the relative L2 error against a seeded ellipsoid phantom must fall below
fixed bars, plus closed-form and monotonicity properties.
"""
import numpy as np
import pytest

import oracle
import oracle.algorithms as oalg
from tigre_amd.geometry import cone_geometry, prepared
from tigre_amd.phantom import ellipsoid_phantom


def nrmse(x, ref):
    """Relative L2 error ||x - ref|| / ||ref||."""
    return float(np.linalg.norm((x - ref).ravel()) / np.linalg.norm(ref.ravel()))


@pytest.fixture(scope="module")
def problem():
    geo = cone_geometry((40, 40, 40), (40.0, 40.0, 40.0), (56, 64), (1.0, 1.0), 300.0, 450.0)
    angles = np.linspace(0, 2 * np.pi, 60, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 6, seed=2, )
    proj = oracle.ax(img, prepared(geo, angles))
    return geo, angles, img, proj


def test_ramp_filter_is_abs_frequency():
    """The ram-lak response approximates |w| / pi on [0, pi] (the ideal ramp, sampled)."""
    filt, L = oalg.ramp_filter(cone_geometry((4, 4, 4), (4.0,) * 3, (4, 100), (1.0, 1.0), 10.0, 20.0))
    k = np.arange(8, L // 2)
    assert np.allclose(filt[k], k / (L / 2), atol=2e-3)
    assert filt[L // 2 + 1:].tolist() == filt[1:L // 2][::-1].tolist()


def test_fdk_reconstructs_phantom(problem):
    geo, angles, img, proj = problem
    rec, _ = oalg.fdk(proj, geo, angles)
    assert nrmse(rec, img) < 0.3


def test_os_sart_converges(problem):
    geo, angles, img, proj = problem
    e = [nrmse(oalg.os_sart(proj, geo, angles, n, blocksize=20), img) for n in (1, 3)]
    assert e[1] < e[0] and e[1] < 0.2


def test_cgls_residual_decreases(problem):
    geo, angles, img, proj = problem
    x, r = oalg.cgls(proj, geo, angles, 6)
    assert all(b < a for a, b in zip(r, r[1:]))
    assert nrmse(x, img) < 0.2


def test_wang_weights_redundancy():
    """Detector-offset weights: a ray measured twice (positions t and -t of the
    central plane inside the overlap) gets total weight 1; beyond the overlap the
    short side gets 0 and the long side 1; the product and the oracle agree."""
    from oracle import algorithms as oalg
    from tigre_amd.algorithms import wang_weights
    from tigre_amd.geometry import cone_geometry

    geo = cone_geometry((32, 32, 32), (32.0, 32.0, 32.0), (16, 64), (1.0, 1.0), 500.0, 1000.0)
    for off in (10.0, -13.0):
        geo.offDetector = np.array([0.0, off])
        proj = np.zeros((2, 16, 64))
        _, zgeo, theta = oalg.wang_pad(proj, geo)
        w = oalg.wang_weight(zgeo, theta)
        n = int(zgeo.nDetector[1])
        DSO, DSD = 500.0, 1000.0
        offset = float(zgeo.offDetector[1])
        t = ((np.arange(n) - n / 2 + 0.5) * 1.0 + abs(offset)) * DSO / DSD
        lim = abs(theta) * DSO / DSD
        ww = w[::-1] if theta < 0 else w
        inside = np.abs(t) <= lim
        # mirror pairs t_k = -t_j inside the overlap
        pairs = 0
        for k in np.nonzero(inside)[0]:
            j = np.argmin(np.abs(t + t[k]))
            if abs(t[j] + t[k]) < 1e-9:
                assert abs(ww[k] + ww[j] - 1.0) < 1e-12
                pairs += 1
        assert pairs >= 8
        assert np.all(ww[t < -lim] == 0.0) and np.all(ww[t > lim] == 1.0)
        assert np.allclose(wang_weights(zgeo, theta), w, rtol=0, atol=1e-12)
