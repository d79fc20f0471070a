"""Pins of the oracle's TV steps.

POCS-TV (GD_TV.cu): its sub-gradient is the derivative of the discrete TV
functional  TV(f) = sum_p |D^- f(p)|  (backward differences; the image sits
between zero slices in z, the y/x terms vanish at the low edges), checked by
central finite differences; one step moves the image by alpha * |g| /
sqrt(|g|^2 + 1e-7) in L2; a constant image has gradient +1 on both z faces and
0 elsewhere.

ROF denoise (tv_proximal.cu): with lambda -> infinity the iteration leaves f
unchanged; on a tiny volume many iterations reach the minimiser of
sum |D^+ u| + lambda/2 |u - f_ext|^2 over the extended domain (slice 0 =
image slice 1, last slice 0).
"""
import numpy as np
import pytest

import oracle


def tv_functional(f):
    nz, ny, nx = f.shape
    F = np.zeros((nz + 2, ny, nx))
    F[1:-1] = f
    total = 0.0
    for z in range(1, nz + 2):           # image slices and the zero slice above
        dz = F[z] - F[z - 1]
        dy = np.zeros((ny, nx))
        dy[1:] = F[z][1:] - F[z][:-1]
        dx = np.zeros((ny, nx))
        dx[:, 1:] = F[z][:, 1:] - F[z][:, :-1]
        total += np.sqrt(dz ** 2 + dy ** 2 + dx ** 2).sum()
    return total


def test_tv_gradient_is_derivative_of_tv():
    rng = np.random.default_rng(0)
    f = rng.uniform(0, 1, size=(4, 5, 6))
    g = oracle.tv_gradient(f)
    h = 1e-6
    for idx in [(0, 0, 0), (1, 2, 3), (3, 4, 5), (2, 0, 5), (0, 4, 0), (3, 1, 1)]:
        fp, fm = f.copy(), f.copy()
        fp[idx] += h
        fm[idx] -= h
        num = (tv_functional(fp) - tv_functional(fm)) / (2 * h)
        assert g[idx] == pytest.approx(num, rel=1e-5, abs=1e-6), idx


def test_tv_constant_image():
    g = oracle.tv_gradient(np.full((5, 4, 3), 2.0))
    assert np.allclose(g[0], 1.0) and np.allclose(g[-1], 1.0) and np.allclose(g[1:-1], 0.0)


def test_pocs_tv_step_length():
    rng = np.random.default_rng(1)
    f = rng.uniform(0, 1, size=(6, 7, 8))
    alpha = 0.3
    g = oracle.tv_gradient(f)
    n = np.sqrt((g ** 2).sum())
    out = oracle.pocs_tv(f, alpha, 1)
    assert np.sqrt(((out - f) ** 2).sum()) == pytest.approx(alpha * n / np.sqrt(n ** 2 + 1e-7), rel=1e-12)
    two = oracle.pocs_tv(f, alpha, 2)
    assert np.sqrt(((two - out) ** 2).sum()) == pytest.approx(alpha, rel=1e-6)


def test_tvdenoise_infinite_lambda_is_identity():
    f = np.random.default_rng(2).uniform(0, 1, size=(4, 5, 6))
    assert np.allclose(oracle.tvdenoise(f, 1e30, 20), f, atol=1e-12)


def _grad(u):
    gz = np.zeros_like(u)
    gz[:-1] = u[1:] - u[:-1]
    gy = np.zeros_like(u)
    gy[:, :-1] = u[:, 1:] - u[:, :-1]
    gx = np.zeros_like(u)
    gx[:, :, :-1] = u[:, :, 1:] - u[:, :, :-1]
    return gz, gy, gx


def _div(pz, py, px):
    """Negative adjoint of _grad."""
    d = np.zeros_like(pz)
    d[0] += pz[0]
    d[1:-1] += pz[1:-1] - pz[:-2]
    d[-1] -= pz[-2]
    d[:, 0] += py[:, 0]
    d[:, 1:-1] += py[:, 1:-1] - py[:, :-2]
    d[:, -1] -= py[:, -2]
    d[:, :, 0] += px[:, :, 0]
    d[:, :, 1:-1] += px[:, :, 1:-1] - px[:, :, :-2]
    d[:, :, -1] -= px[:, :, -2]
    return d


def rof_chambolle(f, lam, iters=60000, tau=1.0 / 12.0):
    """ROF minimiser by Chambolle's dual projection algorithm (J. Math. Imaging Vis. 20, 2004),
    an algorithm independent of the primal-dual iteration under test."""
    pz, py, px = np.zeros_like(f), np.zeros_like(f), np.zeros_like(f)
    for _ in range(iters):
        gz, gy, gx = _grad(_div(pz, py, px) - lam * f)
        nrm = 1.0 + tau * np.sqrt(gz ** 2 + gy ** 2 + gx ** 2)
        pz, py, px = (pz + tau * gz) / nrm, (py + tau * gy) / nrm, (px + tau * gx) / nrm
    return f - _div(pz, py, px) / lam


def test_tvdenoise_reaches_rof_minimiser():
    rng = np.random.default_rng(3)
    nz, ny, nx = 3, 3, 4
    f = rng.uniform(0, 1, size=(nz, ny, nx))
    lam = 4.0
    fe = np.zeros((nz + 2, ny, nx))       # extended domain: mirror slice, image, zero slice
    fe[0] = f[1]
    fe[1:-1] = f
    u_star = rof_chambolle(fe, lam)[1:-1]
    got = oracle.tvdenoise(f, lam, 3000)
    assert np.max(np.abs(got - u_star)) < 1e-4


def _tv_functional(f, pad_z, w=None):
    """Sum over voxels of sqrt(sum_k w_k d_k^2) with the reference's backward
    differences (x/y dropped at the low edge; z padded with zeros or dropped)."""
    nz, ny, nx = f.shape
    if pad_z:
        g = np.zeros((nz + 2, ny, nx))
        g[1:-1] = f
        zs = range(1, nz + 2)
    else:
        g = f
        zs = range(nz)
    total = 0.0
    for z in zs:
        for y in range(ny):
            for x in range(nx):
                d0 = g[z, y, x] - g[z - 1, y, x] if (z >= 1) else 0.0
                d1 = g[z, y, x] - g[z, y - 1, x] if y >= 1 else 0.0
                d2 = g[z, y, x] - g[z, y, x - 1] if x >= 1 else 0.0
                ww = (1, 1, 1) if w is None else w(z, y, x)
                total += np.sqrt(ww[0] * d0 * d0 + ww[1] * d1 * d1 + ww[2] * d2 * d2)
    return total


def test_aw_gradient_is_derivative_with_frozen_weights():
    """GD_AwTV's gradient is the derivative of sum sqrt(sum_k w_k d_k^2) with the
    weights w_k = exp(-(d_k/delta)^2) held at their current values."""
    rng = np.random.default_rng(3)
    f = rng.normal(0, 1, (4, 5, 6))
    delta = 0.9
    g = oracle.tv_gradient_ex(f, pad_z=True, delta=delta, eps=1e-12)
    nz, ny, nx = f.shape
    fp = np.zeros((nz + 2, ny, nx))
    fp[1:-1] = f
    W = {}
    for z in range(1, nz + 2):
        for y in range(ny):
            for x in range(nx):
                d = [fp[z, y, x] - fp[z - 1, y, x], fp[z, y, x] - fp[z, y - 1, x] if y else 0.0,
                     fp[z, y, x] - fp[z, y, x - 1] if x else 0.0]
                W[(z, y, x)] = [np.exp(-(v / delta) ** 2) for v in d]
    h = 1e-6
    for (z, y, x) in [(0, 0, 0), (1, 2, 3), (3, 4, 5), (2, 0, 4)]:
        fa, fb = f.copy(), f.copy()
        fa[z, y, x] += h
        fb[z, y, x] -= h
        num = (_tv_functional(fa, True, lambda *k: W[k]) - _tv_functional(fb, True, lambda *k: W[k])) / (2 * h)
        assert abs(num - g[z, y, x]) < 1e-6


def test_aw_tends_to_tv_for_large_delta():
    rng = np.random.default_rng(4)
    f = rng.normal(0, 1, (5, 6, 7))
    a = oracle.aw_pocs_tv(f, 0.1, 4, 1e9)
    b = oracle.pocs_tv(f, 0.1, 4)
    assert np.max(np.abs(a - b)) < 1e-12


def test_piccs_gradient_unpadded_and_prior_cases():
    rng = np.random.default_rng(5)
    f = rng.normal(0, 1, (4, 5, 6))
    g = oracle.tv_gradient_ex(f, pad_z=False, eps=1e-12)
    h = 1e-6
    for (z, y, x) in [(0, 0, 0), (3, 4, 5), (1, 2, 3)]:
        fa, fb = f.copy(), f.copy()
        fa[z, y, x] += h
        fb[z, y, x] -= h
        num = (_tv_functional(fa, False) - _tv_functional(fb, False)) / (2 * h)
        assert abs(num - g[z, y, x]) < 1e-6
    # prior equal to the image: the prior term vanishes and the normalised step
    # is the plain one
    a = oracle.piccs_tv(f, f, 0.05, 0.5, 1)
    b = oracle.piccs_tv(f, rng.normal(0, 1, f.shape), 0.05, 0.0, 1)
    assert np.max(np.abs(a - b)) < 1e-12
    # every step moves the image by alpha in L2
    c = oracle.piccs_tv(f, rng.normal(0, 1, f.shape), 0.05, 0.3, 1)
    assert abs(np.linalg.norm(c - f) - 0.05) < 1e-12
