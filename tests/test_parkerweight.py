"""Parker short-scan weights (tigre_amd.utilities.parkerweight, restating
Python/tigre/utilities/parkerweight.py:23-78) against the property that defines
them: over a short scan every ray measured twice has weights that sum to one.

In the fan-angle convention of the reference (alpha = -atan(u / DSD) for
detector position u), the ray (beta, alpha) is measured again at gantry angle
beta + pi + 2 alpha by the mirrored detector position (-alpha).  The
weights are smooth in beta, so the conjugate weight is read by linear
interpolation on a fine angle grid (error ~1e-6 here)."""
import warnings

import numpy as np
import pytest

from tigre_amd.geometry import cone_geometry
from tigre_amd.utilities.parkerweight import parker_weights, parkerweight


def _fan(geo):
    n = int(round(geo.sDetector[0] / geo.dDetector[0]))
    pos = -geo.sDetector[0] / 2 + geo.dDetector[0] / 2 + geo.dDetector[0] * np.arange(n)
    alpha = -np.arctan(pos / np.mean(geo.DSD))
    return n, alpha, abs(alpha[0] - alpha[-1]) / 2


@pytest.mark.parametrize("overscan,q", [(0.3, 0.1), (0.6, 0.5), (0.1, 1.0)])
def test_conjugate_rays_sum_to_one(overscan, q):
    geo = cone_geometry((64, 64, 64), (64.0,) * 3, (64, 64), (1.6, 1.6), 1000.0, 1536.0)
    n, alpha, delta = _fan(geo)
    total = np.pi + 2 * delta + overscan
    angles = np.linspace(0.0, total, 20001)
    w = parker_weights(geo, angles, q)
    assert np.all(w >= -1e-12) and np.all(w <= 1 + 1e-12)
    worst, count = 0.0, 0
    for ii in range(0, len(angles), 97):
        for j in range(n):
            bc = angles[ii] + np.pi + 2 * alpha[j]
            if bc < 0.0 or bc > total:
                continue
            wc = np.interp(bc, angles, w[:, n - 1 - j])
            worst = max(worst, abs(w[ii, j] + wc - 1.0))
            count += 1
    assert count > 300
    assert worst < 1e-4


def test_rays_measured_once_have_weight_one():
    """A ray whose conjugate lies outside the scan is measured once: weight 1
    away from the smooth transitions at both ends of the scan."""
    geo = cone_geometry((64, 64, 64), (64.0,) * 3, (64, 64), (1.6, 1.6), 1000.0, 1536.0)
    n, alpha, delta = _fan(geo)
    total = np.pi + 2 * delta + 0.3
    angles = np.linspace(0.0, total, 4001)
    w = parker_weights(geo, angles, 0.1)
    mid = (angles > 0.5 * total - 0.3) & (angles < 0.5 * total + 0.3)
    assert np.allclose(w[mid], 1.0, atol=1e-12)


def test_parkerweight_scales_each_view_along_its_last_axis():
    """As the reference (parkerweight.py:75), the weight row is tiled over the
    rows of each view, so it varies along the last (u) axis; the fan angles
    come from the first detector axis's size and spacing, which is why the
    reference (and this restatement) needs a square detector."""
    geo = cone_geometry((32, 32, 32), (32.0,) * 3, (32, 32), (1.6, 1.6), 1000.0, 1536.0)
    _, _, delta = _fan(geo)
    angles = np.linspace(0.0, np.pi + 2 * delta + 0.4, 50)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        w = parker_weights(geo, angles, 0.1)
        proj = np.random.default_rng(0).random((50, 32, 32)).astype(np.float32)
        out = parkerweight(proj.copy(), geo, angles, 0.1)
    assert np.allclose(out, proj * w[:, None, :].astype(np.float32), rtol=1e-6)
    assert not np.allclose(w[:, 0], w[:, -1])
