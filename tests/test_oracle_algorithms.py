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


# -- OS-SART weights ------------------------------------------------------------------
def _views_with_offsets(n, seed):
    rng = np.random.default_rng(seed)
    ang = np.stack([rng.uniform(0, 2 * np.pi, n), np.zeros(n), np.zeros(n)], 1)
    return rng, ang


def test_weights_w_is_inverse_chord():
    """W = 1 / (Siddon projection of ones over the 2^3 enlarged volume)
    (iterative_recon_alg.py:214-230): per ray, the inverse of its chord through
    the box of size (max(sDetector_v, s_z), 1.1 s_y, 1.1 s_x) centred on the
    view's offOrigin, and 0 (1 / inf) where the chord is at most min(dVoxel) / 2.
    Chords from the independent world-frame ray/box model."""
    from tests import refgeom

    geo = cone_geometry((10, 12, 14), (10.0, 12.0, 14.0), (20, 24), (1.0, 1.0), 100.0, 160.0)
    rng, ang = _views_with_offsets(6, 3)
    geo.offOrigin = rng.uniform(-1, 1, size=(6, 3))
    geo.offDetector = rng.uniform(-2, 2, size=(6, 2))
    gp = prepared(geo, ang)
    W = oalg.weights_w(gp, gp.angles)
    s = np.asarray(gp.sVoxel, dtype=np.float64)
    big = np.array([max(float(gp.sDetector[0]), s[0]), 1.1 * s[1], 1.1 * s[2]])[::-1]  # (x, y, z)
    floor = min(np.asarray(gp.dVoxel, dtype=np.float64)) / 2
    nV, nU = W.shape[1:]
    checked = zeroed = 0
    for i in range(len(ang)):
        off = np.asarray(gp.offOrigin, dtype=np.float64).reshape(-1, 3)[i][::-1]
        lo, hi = off - big / 2, off + big / 2
        for v in range(0, nV, 3):
            for u in range(0, nU, 3):
                S, P = refgeom.source_and_pixel(gp, i, u, v)
                chord = refgeom.segment_box_length(S, P, lo, hi)
                expect = 1.0 / chord if chord > floor else 0.0
                assert W[i, v, u] == pytest.approx(expect, rel=1e-6, abs=1e-12), (i, v, u)
                checked += 1
                zeroed += expect == 0.0
    assert checked > 200 and zeroed < checked


def test_weights_v_brute_force():
    """V[block] = mean over z of the FDK backprojection of all-ones projections
    over the volume shrunk by 0.9 max(s_y, s_x) / |(s_y, s_x)| on every axis
    (iterative_recon_alg.py:232-270): per voxel, the sum over the block's views
    of 1/U^2 times the bilinear sample of ones (zero border) at the voxel
    centre's detector position, from the world-frame model."""
    from tests import refgeom

    geo = cone_geometry((5, 6, 7), (10.0, 12.0, 15.0), (14, 16), (1.0, 1.0), 60.0, 100.0)
    _, ang = _views_with_offsets(6, 5)
    blocks = [np.arange(0, 3), np.arange(3, 6)]
    gp = prepared(geo, ang)
    V = oalg.weights_v(gp, gp.angles, blocks)
    s = np.asarray(gp.sVoxel, dtype=np.float64)
    k = 0.9 * max(s[1], s[2]) / np.hypot(s[1], s[2])
    gs = prepared(geo, ang)
    gs.sVoxel = (s * k).astype(np.float32)
    gs.dVoxel = (gs.sVoxel / np.asarray(gs.nVoxel)).astype(np.float32)
    nz, ny, nx = [int(x) for x in gs.nVoxel]
    nV, nU = int(gs.nDetector[0]), int(gs.nDetector[1])

    def edge(t, n):
        return min(max(min(t + 1.0, n - t), 0.0), 1.0)

    for b, idx in enumerate(blocks):
        for iy in range(ny):
            for ix in range(nx):
                acc = 0.0
                for iz in range(nz):
                    for i in idx:
                        X = refgeom.voxel_centre(gs, i, iz, iy, ix)
                        u, v = refgeom.project_point(gs, i, X)
                        al = float(gs.angles[i][0])
                        DSO = float(np.asarray(gs.DSO).reshape(-1)[i])
                        U = (DSO - X[0] * np.cos(al) - X[1] * np.sin(al)) / DSO
                        acc += edge(u, nU) * edge(v, nV) / U ** 2
                assert V[b, iy, ix] == pytest.approx(acc / nz, rel=1e-6), (b, iy, ix)


def test_weights_v_parallel_is_block_size():
    """Parallel beam: V = the number of views of the block (iterative_recon_alg.py:265-267)."""
    from tigre_amd.geometry import geometry

    g = geometry(mode="parallel", nVoxel=np.array([4, 6, 8]))
    ang = np.linspace(0, np.pi, 7, endpoint=False)
    V = oalg.weights_v(g, ang, [np.arange(0, 5), np.arange(5, 7)])
    assert V.shape == (2, 6, 8) and np.all(V[0] == 5.0) and np.all(V[1] == 2.0)


# -- ASD-POCS ----------------------------------------------------------------------
@pytest.fixture(scope="module")
def small_problem():
    geo = cone_geometry((16, 16, 16), (16.0, 16.0, 16.0), (24, 24), (1.0, 1.0), 200.0, 300.0)
    angles = np.linspace(0, 2 * np.pi, 12, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 4, seed=5)
    proj = oracle.ax(img, prepared(geo, angles))
    return geo, angles, img, proj


def test_asd_pocs_without_tv_is_sart(small_problem):
    """With no TV step (tviter = 0, or alpha = 0 so dtvg = 0), ASD-POCS is plain
    SART: OS-SART with one view per subset, the same number of sweeps."""
    geo, angles, _, proj = small_problem
    sart = oalg.os_sart(proj, geo, angles, 3, blocksize=1)
    for kw in ({"tviter": 0}, {"alpha": 0.0}):
        trace = []
        res = oalg.asd_pocs(proj, geo, angles, 3, trace=trace, **kw)
        assert len(trace) == 3
        assert np.array_equal(res, sart), kw


def test_asd_pocs_tv_step_length(small_problem):
    """Each TV step moves the image by dtvg |g| / sqrt(|g|^2 + 1e-7) = dtvg in L2
    (GD_TV.cu:561-595), so with one TV step per iteration dg equals the step the
    call used, and the first step is alpha * dp (pocs_algorithms.py:144-145);
    with five steps dg is at most 5 dtvg."""
    geo, angles, _, proj = small_problem
    alpha = 0.05
    trace = []
    oalg.asd_pocs(proj, geo, angles, 4, alpha=alpha, tviter=1, trace=trace)
    assert trace[0][5] == pytest.approx(alpha * trace[0][2], rel=1e-12)
    for (_, _, _, dg, _, step) in trace:
        assert dg == pytest.approx(step, rel=1e-6)
    trace5 = []
    oalg.asd_pocs(proj, geo, angles, 3, alpha=alpha, tviter=5, trace=trace5)
    for (_, _, _, dg, _, step) in trace5:
        assert step < dg <= 5 * step * (1 + 1e-9)


@pytest.mark.parametrize("rmax,maxl2err,reduced", [(0.0, 0.0, True), (10.0, 0.0, False), (0.0, 1e30, False)])
def test_asd_pocs_step_reduction_rule(small_problem, rmax, maxl2err, reduced):
    """dtvg shrinks by alpha_red after an iteration exactly when dg > rmax dp and
    dd > epsilon (pocs_algorithms.py:152-153).  The step of each iteration is
    read back from dg (one TV step per iteration moves the image by the step);
    dd and dp are recomputed from the iterates."""
    geo, angles, _, proj = small_problem
    trace, snaps = [], []
    alpha_red = 0.7
    oalg.asd_pocs(proj, geo, angles, 4, alpha=0.05, alpha_red=alpha_red, rmax=rmax, tviter=1,
                  epsilon=maxl2err, trace=trace, snapshots=snaps)
    gfull = prepared(geo, angles)
    prev_post = np.zeros_like(snaps[0][0])
    for k, ((_, dd, dp, dg, c, step), (pre, post)) in enumerate(zip(trace, snaps)):
        assert dd == pytest.approx(np.linalg.norm(oracle.ax(pre.astype(np.float32), gfull) - proj), rel=1e-9)
        assert dp == pytest.approx(np.linalg.norm(pre - prev_post), rel=1e-9)
        assert dg == pytest.approx(np.linalg.norm(post - pre), rel=1e-9)
        assert -1.0 - 1e-9 <= c <= 1.0 + 1e-9
        prev_post = post
        if k + 1 < len(trace):
            ratio = trace[k + 1][3] / dg
            assert ratio == pytest.approx(alpha_red if reduced else 1.0, rel=1e-5), (k, ratio)


def test_asd_pocs_beta_stop(small_problem):
    """beta = lmbda is multiplied by lmbda_red every iteration and the loop stops
    once beta < 0.005 (pocs_algorithms.py:155-158): with lmbda_red = 0.5 that is
    after 8 iterations (0.5^8 < 0.005 < 0.5^7), whatever niter allows; beta
    never scales the SART update, so the first iterations equal those of a run
    with lmbda_red = 1."""
    geo, angles, _, proj = small_problem
    t_half, t_one = [], []
    oalg.asd_pocs(proj, geo, angles, 20, tviter=2, lmbda_red=0.5, trace=t_half)
    oalg.asd_pocs(proj, geo, angles, 3, tviter=2, lmbda_red=1.0, trace=t_one)
    assert len(t_half) == 8 and len(t_one) == 3
    # (the oracle's OpenMP norm sum may round differently with another thread count)
    assert np.allclose(np.array(t_half[:3], dtype=np.float64), np.array(t_one, dtype=np.float64), rtol=1e-12, atol=0)


# -- SART-TV, OS-SART-TV, FISTA, ISTA (the ROF denoiser's callers) ----------------
def test_im3ddenoise_identity_and_equivariance():
    """im3ddenoise normalises its input to [0, 1] before the ROF denoiser and
    undoes it after (im_3d_denoise.py:7-28): with no iterations it returns
    the image, and it commutes with a positive scaling and a shift."""
    x = np.random.default_rng(3).random((10, 11, 12))
    assert np.abs(oalg.im3ddenoise(x, 0, 50.0) - x).max() <= 1e-14
    a, c = 3.7, -0.4
    y1 = oalg.im3ddenoise(a * x + c, 15, 30.0)
    y2 = a * oalg.im3ddenoise(x, 15, 30.0) + c
    assert np.abs(y1 - y2).max() <= 1e-12 * np.abs(y2).max()
    # and, at the reference's defaults (50 iterations, lambda 15), it denoises:
    # a noisy phantom loses total variation and comes closer to the clean one
    img = ellipsoid_phantom((16, 16, 16), (16.0, 16.0, 16.0), 4, seed=5)
    noisy = img + np.random.default_rng(1).normal(0, 0.05, img.shape)
    tv = lambda z: sum(np.abs(np.diff(z, axis=k)).sum() for k in range(3))  # noqa: E731
    den = oalg.im3ddenoise(noisy, 50, 15.0)
    assert tv(den) < 0.8 * tv(noisy)
    assert nrmse(den, img) < nrmse(noisy, img)


@pytest.mark.parametrize("blocksize", [1, 4])
def test_sart_tv_without_denoise_is_os_sart(small_problem, blocksize):
    """With no denoising iterations SART-TV / OS-SART-TV are the SART / OS-SART
    sweeps (art_family_algorithms.py:86-108, :131-153)."""
    geo, angles, _, proj = small_problem
    ref = oalg.os_sart(proj, geo, angles, 3, blocksize=blocksize)
    got = oalg.sart_tv(proj, geo, angles, 3, blocksize=blocksize, tviter=0)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def _dense_operators(geo, angles):
    """Interpolated Ax and matched Atb of a tiny problem as dense matrices,
    column by column from unit vectors."""
    g = prepared(geo, angles)
    nvox = int(np.prod(geo.nVoxel))
    shape_v = tuple(int(v) for v in geo.nVoxel)
    A = np.stack([oracle.ax(np.eye(nvox, 1, -i).reshape(shape_v).astype(np.float32), g, "interpolated").ravel()
                  for i in range(nvox)], axis=1)
    shape_p = (len(angles),) + tuple(int(v) for v in geo.nDetector)
    npix = int(np.prod(shape_p))
    B = np.stack([oracle.atb(np.eye(npix, 1, -i).reshape(shape_p).astype(np.float32), g, "matched").ravel()
                  for i in range(npix)], axis=1)
    return A, B


def test_fista_data_step_and_acceleration():
    """FISTA's data step is res + (2 / hyper) Atb_matched(b - Ax_interpolated(res))
    (ista_algorithms.py:128-153): with no denoising and no clipping its first
    iterate equals that step computed with the dense operators; and its
    extrapolation accelerates the iteration: after 25 iterations its data
    residual is below that of the plain gradient iteration with the same step."""
    geo = cone_geometry((6, 6, 6), (6.0, 6.0, 6.0), (8, 8), (1.0, 1.0), 60.0, 90.0)
    angles = np.linspace(0, 2 * np.pi, 5, endpoint=False)
    A, B = _dense_operators(geo, angles)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 2, seed=8)
    b = A @ img.ravel().astype(np.float64)
    proj = b.reshape((5, 8, 8))
    # FISTA's step 2 / hyper must not exceed 1 / L (L the largest eigenvalue of
    # Atb Ax): hyper = 2.4 L
    hyper = 2.4 * float(np.max(np.abs(np.linalg.eigvals(B @ A))))
    first = oalg.fista(proj, geo, angles, 1, hyper=hyper, tviter=0, noneg=False)
    step = 2.0 / hyper * (B @ b)
    assert np.abs(first.ravel() - step).max() <= 1e-5 * np.abs(step).max()
    fast = oalg.fista(proj, geo, angles, 25, hyper=hyper, tviter=0, noneg=False)
    x = np.zeros(A.shape[1])
    for _ in range(25):
        x = x + 2.0 / hyper * (B @ (b - A @ x))
    r_fista = np.linalg.norm(A @ fast.ravel() - b)
    r_gd = np.linalg.norm(A @ x - b)
    assert r_fista < 0.7 * r_gd, (r_fista, r_gd)


# -- LSQR and MLEM -------------------------------------------------------------------
def _dense_siddon_matched(geo, angles):
    """Siddon Ax and matched Atb of a tiny problem as dense matrices."""
    g = prepared(geo, angles)
    nvox = int(np.prod(geo.nVoxel))
    shape_v = tuple(int(v) for v in geo.nVoxel)
    A = np.stack([oracle.ax(np.eye(nvox, 1, -i).reshape(shape_v).astype(np.float32), g, "Siddon").ravel()
                  for i in range(nvox)], axis=1)
    shape_p = (len(angles),) + tuple(int(v) for v in geo.nDetector)
    npix = int(np.prod(shape_p))
    B = np.stack([oracle.atb(np.eye(npix, 1, -i).reshape(shape_p).astype(np.float32), g, "matched").ravel()
                  for i in range(npix)], axis=1)
    return A, B


def test_lsqr_matches_scipy_lsqr():
    """The LSQR restatement follows Paige and Saunders' recurrences: on a
    tiny problem it equals scipy's lsqr run on the same operators (Siddon
    forward, matched backward as dense matrices) for the same iterations."""
    from scipy.sparse.linalg import LinearOperator, lsqr as scipy_lsqr

    geo = cone_geometry((6, 6, 6), (6.0, 6.0, 6.0), (8, 8), (1.0, 1.0), 60.0, 90.0)
    angles = np.linspace(0, 2 * np.pi, 5, endpoint=False)
    A, B = _dense_siddon_matched(geo, angles)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 2, seed=8)
    b = A @ img.ravel().astype(np.float64) + np.random.default_rng(2).normal(0, 0.01, A.shape[0])
    op = LinearOperator(A.shape, matvec=lambda x: A @ x, rmatvec=lambda y: B @ y, dtype=np.float64)
    for n in (1, 3, 6):
        got, norms = oalg.lsqr(b.reshape((5, 8, 8)), geo, angles, n)
        if len(norms) < n:  # the residual rose: the restatement stopped early
            break
        ref = scipy_lsqr(op, b, atol=0, btol=0, conlim=0, iter_lim=n)[0]
        assert np.abs(got.ravel() - ref).max() <= 1e-4 * np.abs(ref).max(), n


def test_mlem_keeps_an_exact_solution():
    """MLEM's update res * Atb(proj / Ax res) / Atb(1) leaves an image whose
    forward projection is the data unchanged (the ratio is 1 everywhere), and
    from a flat start it raises the Poisson log-likelihood of the data."""
    geo = cone_geometry((16, 16, 16), (16.0, 16.0, 16.0), (24, 24), (1.0, 1.0), 200.0, 300.0)
    angles = np.linspace(0, 2 * np.pi, 12, endpoint=False)
    g = prepared(geo, angles)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 4, seed=5) + 0.1
    proj = oracle.ax(img.astype(np.float32), g, "interpolated")
    out = oalg.mlem(proj, geo, angles, 2, init=img)
    assert np.abs(out - img).max() <= 1e-5 * np.abs(img).max()

    def loglik(x):
        p = oracle.ax(x.astype(np.float32), g, "interpolated")
        m = p > 0
        return float(np.sum(proj[m] * np.log(p[m]) - p[m]))

    one = oalg.mlem(proj, geo, angles, 1)
    three = oalg.mlem(proj, geo, angles, 3)
    assert loglik(np.ones_like(img)) < loglik(one) < loglik(three)


@pytest.mark.parametrize("damp", [1.0, 0.0])
def test_lsmr_matches_scipy_lsmr(damp):
    """The LSMR restatement follows Fong and Saunders' recurrences with the
    damping as the reference passes it: on a tiny problem it equals scipy's
    lsmr(damp=...) on the same operators as dense matrices."""
    from scipy.sparse.linalg import LinearOperator, lsmr as scipy_lsmr

    geo = cone_geometry((6, 6, 6), (6.0, 6.0, 6.0), (8, 8), (1.0, 1.0), 60.0, 90.0)
    angles = np.linspace(0, 2 * np.pi, 5, endpoint=False)
    A, B = _dense_siddon_matched(geo, angles)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 2, seed=8)
    b = A @ img.ravel().astype(np.float64) + np.random.default_rng(2).normal(0, 0.01, A.shape[0])
    op = LinearOperator(A.shape, matvec=lambda x: A @ x, rmatvec=lambda y: B @ y, dtype=np.float64)
    compared = 0
    for n in (1, 3, 6):
        got, norms = oalg.lsmr(b.reshape((5, 8, 8)), geo, angles, n, lmbda=damp)
        if len(norms) < n:  # the residual rose: the restatement stopped early
            break
        ref = scipy_lsmr(op, b, damp=damp, atol=0, btol=0, conlim=0, maxiter=n)[0]
        assert np.abs(got.ravel() - ref).max() <= 1e-4 * np.abs(ref).max(), n
        compared += 1
    assert compared >= 2


@pytest.mark.parametrize("variant", ["ba", "ab"])
def test_gmres_matches_scipy_gmres(variant):
    """AB- / BA-GMRES restated: on a tiny problem the k-th iterate equals
    scipy's gmres after k Arnoldi steps on B A x = B b (BA) or A B u = b with
    x = B u (AB), the operators as dense matrices (Siddon forward, matched
    backward)."""
    from scipy.sparse.linalg import gmres as scipy_gmres

    geo = cone_geometry((6, 6, 6), (6.0, 6.0, 6.0), (8, 8), (1.0, 1.0), 60.0, 90.0)
    angles = np.linspace(0, 2 * np.pi, 5, endpoint=False)
    A, B = _dense_siddon_matched(geo, angles)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 2, seed=8)
    b = A @ img.ravel().astype(np.float64) + np.random.default_rng(2).normal(0, 0.01, A.shape[0])
    compared = 0
    for n in (1, 2, 4):
        got, norms = oalg.gmres(b.reshape((5, 8, 8)), geo, angles, n, variant=variant, restart=False)
        if len(norms) < n:
            break
        if variant == "ba":
            ref = scipy_gmres(B @ A, B @ b, x0=np.zeros(A.shape[1]), rtol=0.0, atol=0.0, restart=n, maxiter=1)[0]
        else:
            u = scipy_gmres(A @ B, b, x0=np.zeros(A.shape[0]), rtol=0.0, atol=0.0, restart=n, maxiter=1)[0]
            ref = B @ u
        assert np.abs(got.ravel() - ref).max() <= 1e-4 * np.abs(ref).max(), n
        compared += 1
    assert compared >= 2


@pytest.mark.parametrize("damp", [1.0, 0.0])
def test_hybrid_lsqr_is_damped_lsqr(damp):
    """Hybrid LSQR with a fixed Tikhonov parameter solves the damped least-squares
    problem over the Krylov subspace of the bidiagonalisation, which is what
    LSQR with that damping computes: on a tiny problem it equals scipy's
    lsqr(damp=...) on the same operators as dense matrices: the reference's
    form for one and two iterations (its elementwise update against earlier
    vectors starts at the third), the textbook form for four."""
    from scipy.sparse.linalg import LinearOperator, lsqr as scipy_lsqr

    geo = cone_geometry((6, 6, 6), (6.0, 6.0, 6.0), (8, 8), (1.0, 1.0), 60.0, 90.0)
    angles = np.linspace(0, 2 * np.pi, 5, endpoint=False)
    A, B = _dense_siddon_matched(geo, angles)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 2, seed=8)
    b = A @ img.ravel().astype(np.float64) + np.random.default_rng(2).normal(0, 0.01, A.shape[0])
    op = LinearOperator(A.shape, matvec=lambda x: A @ x, rmatvec=lambda y: B @ y, dtype=np.float64)
    compared = 0
    for n in (1, 2, 4):
        got, norms = oalg.hybrid_lsqr(b.reshape((5, 8, 8)), geo, angles, n, lmbda=damp, restart=False,
                                      elementwise_reorth=n <= 2)
        if len(norms) < n:
            break
        ref = scipy_lsqr(op, b, damp=damp, atol=0, btol=0, conlim=0, iter_lim=n)[0]
        assert np.abs(got.ravel() - ref).max() <= 1e-4 * np.abs(ref).max(), n
        compared += 1
    assert compared >= 2


# -- PCSD ----------------------------------------------------------------------------
def test_pcsd_without_tv_is_sart(small_problem):
    """With no TV iterations and a data bound every image exceeds, PCSD is
    SART sweeps, niter + 1 of them (the reference stops at n_iter > niter)."""
    geo, angles, _, proj = small_problem
    sart = oalg.os_sart(proj, geo, angles, 3, blocksize=1)
    res = oalg.pcsd(proj, geo, angles, 2, tviter=0, epsilon=1e-12)
    assert np.array_equal(res, sart)


def test_pcsd_tv_step_rule(small_problem):
    """The TV step size is 1 in the first iteration and d_p / d_p1 after it:
    recomputed here from the data residuals of the images themselves; with one
    TV iteration each step moves the image by that size in L2 (one POCS-TV
    step moves it by alpha |g| / sqrt(|g|^2 + 1e-7))."""
    geo, angles, _, proj = small_problem
    g = prepared(geo, angles)
    trace = []
    oalg.pcsd(proj, geo, angles, 3, tviter=1, epsilon=1e-12, trace=trace)
    assert len(trace) == 4
    assert trace[0][4] == 1.0
    # d_p1 = residual of the first pre-TV image = dd of the first iteration
    d_p1 = trace[0][2]
    for n_iter, d_p, dd, dp, dtvg, dg, c in trace[1:]:
        assert abs(dtvg - d_p / d_p1) <= 1e-12 * dtvg
    for row in trace:
        assert abs(row[5] - row[4]) <= 1e-6 * row[4]
    # d_p of each iteration is the residual of the previous post-TV image
    res = oalg.pcsd(proj, geo, angles, 0, tviter=1, epsilon=1e-12)
    assert abs(np.linalg.norm(oracle.ax(res.astype(np.float32), g) - proj) - trace[1][1]) <= 1e-9 * trace[1][1]


def test_fbp_reconstructs_parallel_phantom():
    """Parallel-beam FBP (single_pass_algorithms.py:208-224) reconstructs a
    seeded phantom from 180 views over pi to the FDK bar, and is linear."""
    from tigre_amd.geometry import geometry

    geo = geometry(mode="parallel", nVoxel=np.array([32, 40, 40]))
    angles = np.linspace(0, np.pi, 180, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 5, seed=12)
    proj = oracle.ax(img, prepared(geo, angles))
    rec = oalg.fbp(proj, geo, angles)
    assert nrmse(rec, img) < 0.3
    rec2 = oalg.fbp(2.5 * proj, geo, angles)
    assert np.abs(rec2 - 2.5 * rec).max() <= 1e-5 * np.abs(rec).max()


def test_irn_differences_closed_form():
    """On a ramp x = 2 i along one axis the differences are -2 except the last
    two entries, which keep x; the other axes' differences vanish except
    there.  irn_ltx equals the transpose of irn_lx except on the last two rows
    of each axis (the reference's form), checked on the dense matrices."""
    n = (5, 4, 6)
    x = np.broadcast_to(2.0 * np.arange(n[0])[:, None, None], n).copy()
    d = oalg.irn_diffs(x)
    assert np.all(d[0][:-2] == -2.0) and np.all(d[0][-2:] == x[-2:])
    assert np.all(d[1][:, :-2] == 0.0) and np.all(d[1][:, -2:] == x[:, -2:])
    rng = np.random.default_rng(3)
    W = rng.uniform(0.5, 2.0, n)
    N = int(np.prod(n))
    E = np.eye(N)
    L = np.stack([np.concatenate([t.ravel() for t in oalg.irn_lx(W, E[k].reshape(n))]) for k in range(N)], 1)
    E3 = np.eye(3 * N)
    Lt = np.stack([oalg.irn_ltx(W, [E3[k][j * N:(j + 1) * N].reshape(n) for j in range(3)]).ravel()
                   for k in range(3 * N)], 1)
    diff = np.abs(Lt - L.T).reshape(n + (3 * N,))
    keep = np.ones(n, bool)
    keep[-2:, :, :] = keep[:, -2:, :] = keep[:, :, -2:] = False
    assert diff[keep].max() == 0.0 and diff[~keep].max() > 0.1


@pytest.mark.parametrize("start", ["zero", "random"])
def test_irn_cycle_is_lsqr_on_stacked_system(start):
    """One outer cycle of the IRN loop is CGLS on [A; sqrt(lmbda) W D] x =
    [b; 0]: with an exact transpose pair its k-th iterate equals scipy's lsqr
    (k steps, no stopping tests) on the stacked dense matrix from the same
    start; with a second cycle the restart is from the image before the last
    update, with weights from the last one."""
    from scipy.sparse.linalg import lsqr as sp_lsqr

    rng = np.random.default_rng(11)
    n = (3, 3, 4)
    N = int(np.prod(n))
    A = rng.standard_normal((30, N))
    b = rng.standard_normal(30)
    Ds = [rng.standard_normal((N, N)) for _ in range(3)]
    W = rng.uniform(0.5, 1.5, N)
    lmbda = 0.7
    x0 = np.zeros(n) if start == "zero" else rng.standard_normal(n)
    fA = lambda x: A @ np.ravel(x)
    fAt = lambda y: (A.T @ y).reshape(n)
    weights = lambda x: W.reshape(n)
    lx = lambda Wt, x: [Wt.ravel() * (D @ np.ravel(x)) for D in Ds]
    ltx = lambda Wt, y: sum(D.T @ (Wt.ravel() * yy) for D, yy in zip(Ds, y)).reshape(n)
    K = np.vstack([A] + [np.sqrt(lmbda) * W[:, None] * D for D in Ds])
    rhs = np.concatenate([b, np.zeros(3 * N)])
    for k in (1, 3, 6):
        x, norms = oalg.irn_cycles(b, x0, fA, fAt, weights, lx, ltx, k, 1, lmbda)
        ref = x0.ravel() + sp_lsqr(K, rhs - K @ x0.ravel(), atol=0, btol=0, conlim=0, iter_lim=k)[0]
        assert np.abs(x.ravel() - ref).max() <= 1e-8 * np.abs(ref).max()
        assert abs(norms[-1] - np.linalg.norm(b - A @ x.ravel())) <= 1e-12 * norms[-1]
    # two cycles = one cycle of 4, then one from the iterate after 3 steps
    two, _ = oalg.irn_cycles(b, x0, fA, fAt, weights, lx, ltx, 4, 2, lmbda)
    x3, _ = oalg.irn_cycles(b, x0, fA, fAt, weights, lx, ltx, 3, 1, lmbda)
    again, _ = oalg.irn_cycles(b, x3, fA, fAt, weights, lx, ltx, 4, 1, lmbda)
    assert np.abs(two - again).max() <= 1e-9 * np.abs(again).max()


def test_irn_tv_cgls_weights_follow_the_image():
    """On the tomography problem the first cycle (W constant from a zero
    image) lowers the data residual every step, and later cycles reweight:
    the result differs from a single cycle of the same total length."""
    geo = cone_geometry((6, 6, 6), (6.0, 6.0, 6.0), (8, 8), (1.5, 1.5), 60.0, 90.0)
    angles = np.linspace(0, 2 * np.pi, 12, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 3, seed=5)
    proj = oracle.ax(img, prepared(geo, angles))
    x, norms = oalg.irn_tv_cgls(proj, geo, angles, 5, niter_outer=1)
    assert all(b <= a * (1 + 1e-12) for a, b in zip(norms, norms[1:]))
    x2, _ = oalg.irn_tv_cgls(proj, geo, angles, 5, niter_outer=2)
    assert np.abs(x2 - x).max() > 1e-6 * np.abs(x).max()


@pytest.mark.parametrize("case", ["least_squares", "compatible", "non_adjoint"])
def test_device_lsqr_matches_scipy_lsqr(case):
    """The LSQR used for the hybrid fLSQR-TV preconditioner (torch, run here on
    CPU tensors) equals scipy's lsqr with its default stopping tests, at
    iteration limits 3 and 50, on an inconsistent system, a compatible one
    (stops early on the residual test) and a non-adjoint pair of operators
    like the reference's."""
    import torch
    from scipy.sparse.linalg import LinearOperator, lsqr as sp_lsqr

    from tigre_amd.algorithms import _lsqr_solve

    rng = np.random.default_rng(21)
    K = rng.standard_normal((40, 25))
    Kt = K.T + (0.05 * rng.standard_normal((25, 40)) if case == "non_adjoint" else 0.0)
    b = K @ rng.standard_normal(25) if case == "compatible" else rng.standard_normal(40)
    op = LinearOperator(K.shape, matvec=lambda x: K @ x, rmatvec=lambda y: Kt @ y, dtype=np.float64)
    Kd, Ktd = torch.tensor(K), torch.tensor(Kt)
    for lim in (3, 50):
        ref = sp_lsqr(op, b, iter_lim=lim)
        got = _lsqr_solve(lambda x: Kd @ x, lambda y: Ktd @ y, torch.tensor(b), lim).numpy()
        # near the Krylov dimension rounding grows (3e-11 at 20 steps, 1e-8 at 24); a
        # different stopping step would differ by the step size, 1e-5 here
        tol = 1e-10 if lim == 3 else 1e-6
        assert np.abs(got - ref[0]).max() <= tol * np.abs(ref[0]).max(), (lim, ref[1], ref[2])
    if case == "compatible":
        assert sp_lsqr(op, b, iter_lim=50)[2] < 50


@pytest.mark.parametrize("lmbda", [0.0, 0.7])
def test_hybrid_flsqr_tv_solves_the_projected_problem(lmbda):
    """The image after k steps is start + xA0 + Z y with y = argmin
    ||A Z y - r0||^2 + lmbda^2 ||W D Z y||^2 over the k preconditioned vectors
    (solved here by dense least squares on A Z and W D Z), so the U / M
    recurrence and the QR factor represent the problem exactly; xA0 is the
    constant image whose projection best fits the data (r0 is orthogonal to
    the projection of a constant image); the entry point's image is the start."""
    geo = cone_geometry((6, 6, 6), (6.0, 6.0, 6.0), (8, 8), (1.5, 1.5), 60.0, 90.0)
    angles = np.linspace(0, 2 * np.pi, 12, endpoint=False)
    # not a constant image: its data would be fitted by xA0 alone (r0 = 0)
    img = np.random.default_rng(5).random((6, 6, 6)).astype(np.float32)
    g = prepared(geo, angles)
    proj = oracle.ax(img, g)
    tr = {}
    x, norms, start = oalg.hybrid_flsqr_tv(proj, geo, angles, 3, lmbda=lmbda, trace=tr)
    assert np.all(start == 0.0) and len(norms) == 3
    ones = oracle.ax(np.ones(img.shape, np.float32), g)
    # xA0 reaches the projector rounded to float32: orthogonal to 1e-7
    assert abs(np.sum(ones * tr["r0"])) <= 1e-6 * np.linalg.norm(ones) * np.linalg.norm(tr["r0"])
    AZ = np.stack([oracle.ax(z.astype(np.float32), g).ravel() for z in tr["Z"]], 1)
    LZ = np.stack([np.concatenate([t.ravel() for t in oalg.irn_lx(tr["W"], z)]) for z in tr["Z"]], 1)
    y = np.linalg.lstsq(np.vstack([AZ, lmbda * LZ]), np.concatenate([tr["r0"].ravel(), np.zeros(LZ.shape[0])]),
                        rcond=None)[0]
    want = tr["xA0"] + sum(yj * zj for yj, zj in zip(y, tr["Z"]))
    assert np.abs(x - want).max() <= 1e-6 * np.abs(want).max()
