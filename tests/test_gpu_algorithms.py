"""GPU: device-resident algorithms against the oracle's FP64 restatements.

Tolerances: FDK is linear, so FP32 storage/accumulation must stay within a
relative L2 error of 1e-5 of the FP64 result; the iterative methods feed FP32
results back into the next step, so a few iterations are compared at 1e-4.
"""
import numpy as np
import pytest

import oracle
import oracle.algorithms as oalg
from tigre_amd import algorithms as algs
from tigre_amd.geometry import cone_geometry, prepared
from tigre_amd.phantom import ellipsoid_phantom

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem():
    geo = cone_geometry((40, 40, 40), (40.0, 40.0, 40.0), (56, 64), (1.0, 1.0), 300.0, 450.0)
    angles = np.linspace(0, 2 * np.pi, 60, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 6, seed=2)
    proj = oracle.ax(img, prepared(geo, angles)).astype(np.float32)
    return geo, angles, img, proj


@pytest.mark.parametrize("filt", [None, "shepp_logan", "hann"])
def test_fdk(cuda, problem, filt):
    geo, angles, img, proj = problem
    got = algs.fdk(proj, geo, angles, filter=filt)
    ref, _ = oalg.fdk(proj, geo, angles, filt)
    err = oracle.rel_l2(got, ref)
    print("FDK %s rel-L2 %.3e" % (filt, err))
    assert err <= 1e-5


def test_ossart(cuda, problem):
    geo, angles, img, proj = problem
    got = algs.ossart(proj, geo, angles, 2, blocksize=20)
    ref = oalg.os_sart(proj, geo, angles, 2, blocksize=20)
    err = oracle.rel_l2(got, ref)
    print("OS-SART rel-L2 %.3e" % err)
    assert err <= 1e-4


def test_cgls(cuda, problem):
    geo, angles, img, proj = problem
    got = algs.cgls(proj, geo, angles, 3)
    ref, _ = oalg.cgls(proj, geo, angles, 3)
    err = oracle.rel_l2(got, ref)
    print("CGLS rel-L2 %.3e" % err)
    assert err <= 1e-4


def test_asd_pocs(cuda):
    geo = cone_geometry((24, 24, 24), (24.0, 24.0, 24.0), (32, 32), (1.0, 1.0), 300.0, 450.0)
    angles = np.linspace(0, 2 * np.pi, 20, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 4, seed=3)
    proj = oracle.ax(img, prepared(geo, angles)).astype(np.float32)
    got = algs.asd_pocs(proj, geo, angles, 2)
    ref = oalg.asd_pocs(proj, geo, angles, 2)
    err = oracle.rel_l2(got, ref)
    print("ASD-POCS rel-L2 %.3e" % err)
    assert err <= 1e-3


def test_fdk_offset_detector(cuda):
    """FDK with a laterally offset detector (Wang redundancy weights) against
    the oracle restatement, and closer to the phantom than without the weights."""
    from oracle import algorithms as oalg

    geo = cone_geometry((48, 48, 48), (48.0, 48.0, 48.0), (48, 56), (1.6, 1.6), 1000.0, 1536.0)
    geo.offDetector = np.array([0.0, 12.0])
    angles = np.linspace(0, 2 * np.pi, 90, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 6, seed=41)
    proj = oracle.ax(img, prepared(geo, angles)).astype(np.float32)
    got = algs.fdk(proj, geo, angles)
    ref = oalg.fdk(proj, geo, angles)[0]
    err = oracle.rel_l2(got, ref)
    plain = algs.fdk(proj, geo, angles, dowang=False)
    print("FDK offset detector rel-L2 vs oracle %.3e; image error %.3f (Wang) vs %.3f (none)" % (
        err, oracle.rel_l2(got, img), oracle.rel_l2(plain, img)))
    assert err <= 1e-5
    assert oracle.rel_l2(got, img) < oracle.rel_l2(plain, img)


@pytest.mark.parametrize("variant", ["awasd_pocs", "piccs", "os_asd_pocs"])
def test_asd_pocs_variants_run(cuda, variant):
    """The ASD-POCS family: each variant reduces the data residual well below
    that of the zero image (the TV step follows the non-negativity clamp, so
    small negative values are expected)."""
    geo = cone_geometry((32, 32, 32), (32.0, 32.0, 32.0), (40, 40), (1.6, 1.6), 1000.0, 1536.0)
    angles = np.linspace(0, 2 * np.pi, 40, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 5, seed=51)
    gp = prepared(geo, angles)
    proj = oracle.ax(img, gp).astype(np.float32)
    kw = {"prior": img * 0.9} if variant == "piccs" else {}
    if variant == "os_asd_pocs":
        kw["blocksize"] = 10
    rec = getattr(algs, variant)(proj, geo, angles, 3, **kw)
    r0 = np.linalg.norm(proj)
    r = np.linalg.norm(oracle.ax(rec.astype(np.float32), gp) - proj)
    print("%s residual %.3e of %.3e" % (variant, r, r0))
    assert r < 0.5 * r0


@pytest.mark.parametrize("blocksize", [1, 20])
def test_sart_tv(cuda, problem, blocksize):
    """SART-TV (one view per subset) and OS-SART-TV (20) with the reference's
    denoising defaults (50 iterations, lambda 50) against the FP64 restatement."""
    geo, angles, img, proj = problem
    if blocksize == 1:
        got = algs.sart_tv(proj, geo, angles, 2)
    else:
        got = algs.ossart_tv(proj, geo, angles, 2, blocksize=blocksize)
    ref = oalg.sart_tv(proj, geo, angles, 2, blocksize=blocksize, tviter=50, tvlambda=50.0)
    err = oracle.rel_l2(got, ref)
    print("SART-TV blocksize %d rel-L2 %.3e" % (blocksize, err))
    assert err <= 1e-4


def _largest_eigenvalue(geo, angles):
    """Largest eigenvalue of Atb_matched Ax_interpolated by power iteration on the device."""
    import torch

    from tigre_amd import _native

    vs = algs.ViewSets(geo, angles, [np.arange(len(angles))])
    x = torch.rand(vs.vol_shape, device="cuda", generator=torch.Generator(device="cuda").manual_seed(4))
    lam = 0.0
    for _ in range(30):
        y = vs.atb(vs.ax(x, 0, _native.PROJ_INTERPOLATED), 0, _native.BACKPROJ_MATCHED)
        lam = float(torch.linalg.vector_norm(y.double()) / torch.linalg.vector_norm(x.double()))
        x = y / torch.linalg.vector_norm(y)
    return lam


@pytest.mark.parametrize("variant", ["fista", "ista"])
def test_fista_ista(cuda, problem, variant):
    """FISTA / ISTA with the step 2 / hyper = 1 / (1.2 L) (L the largest
    eigenvalue of Atb Ax) and, for FISTA, TV denoising of strength 30 after
    every data step, against the FP64 restatement."""
    geo, angles, img, proj = problem
    hyper = 2.4 * _largest_eigenvalue(geo, angles)
    if variant == "fista":
        got = algs.fista(proj, geo, angles, 4, hyper=hyper, tvlambda=hyper / 60.0)
        ref = oalg.fista(proj, geo, angles, 4, hyper=hyper, tvlambda=hyper / 60.0)
    else:
        got = algs.ista(proj, geo, angles, 4, hyper=hyper)
        ref = oalg.fista(proj, geo, angles, 4, hyper=hyper, momentum=False)
    err = oracle.rel_l2(got, ref)
    print("%s rel-L2 %.3e (vs phantom %.3f)" % (variant, err, oracle.rel_l2(got, img)))
    assert err <= 1e-4


def test_lsqr(cuda, problem):
    geo, angles, img, proj = problem
    got, l2 = algs.lsqr(proj, geo, angles, 4, computel2=True)
    ref, norms = oalg.lsqr(proj, geo, angles, 4)
    err = oracle.rel_l2(got, ref)
    print("LSQR rel-L2 %.3e, residuals %s" % (err, l2))
    assert err <= 1e-4
    assert np.allclose(l2[0], norms, rtol=1e-4)


def test_mlem(cuda, problem):
    geo, angles, img, proj = problem
    got = algs.mlem(proj, geo, angles, 3)
    ref = oalg.mlem(proj, geo, angles, 3)
    err = oracle.rel_l2(got, ref)
    print("MLEM rel-L2 %.3e" % err)
    assert err <= 1e-4


def test_lsmr(cuda, problem):
    geo, angles, img, proj = problem
    got, l2 = algs.lsmr(proj, geo, angles, 4, computel2=True)
    ref, norms = oalg.lsmr(proj, geo, angles, 4)
    err = oracle.rel_l2(got, ref)
    print("LSMR rel-L2 %.3e, residuals %s" % (err, l2))
    assert err <= 1e-4
    assert np.allclose(l2[0], norms, rtol=1e-4)


@pytest.mark.parametrize("variant", ["ab", "ba"])
def test_gmres(cuda, problem, variant):
    geo, angles, img, proj = problem
    fn = algs.ab_gmres if variant == "ab" else algs.ba_gmres
    got, l2 = fn(proj, geo, angles, 4, computel2=True)
    ref, norms = oalg.gmres(proj, geo, angles, 4, variant=variant)
    err = oracle.rel_l2(got, ref)
    print("%s-GMRES rel-L2 %.3e, residuals %s" % (variant.upper(), err, l2))
    assert err <= 1e-4
    assert np.allclose(l2[0], norms, rtol=1e-4)


def test_hybrid_lsqr(cuda, problem):
    geo, angles, img, proj = problem
    got, l2 = algs.hybrid_lsqr(proj, geo, angles, 4, computel2=True)
    ref, norms = oalg.hybrid_lsqr(proj, geo, angles, 4)
    err = oracle.rel_l2(got, ref)
    print("hybrid LSQR rel-L2 %.3e, residuals %s" % (err, l2))
    assert err <= 1e-4
    assert np.allclose(l2[0], norms, rtol=1e-4)


@pytest.mark.parametrize("variant", ["pcsd", "aw_pcsd", "os_pcsd", "os_aw_pcsd"])
def test_pcsd(cuda, variant):
    """PCSD and its forms against the FP64 restatement.  The adaptive-weighted
    forms use delta = -0.5: at the default -0.005 the weights exp(-(d / delta)^2)
    vanish at every edge of the piecewise-constant phantom and the step is
    driven by differences at the rounding level of FP32, which FP64 does not
    share."""
    geo = cone_geometry((24, 24, 24), (24.0, 24.0, 24.0), (32, 32), (1.0, 1.0), 300.0, 450.0)
    angles = np.linspace(0, 2 * np.pi, 20, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 4, seed=3)
    proj = oracle.ax(img, prepared(geo, angles)).astype(np.float32)
    fn = getattr(algs, variant)
    kw = {"blocksize": 5} if variant.startswith("os") else {}
    if "aw" in variant:
        kw["delta"] = -0.5
    got = fn(proj, geo, angles, 2, **kw)
    ref = oalg.pcsd(proj, geo, angles, 2, blocksize=kw.get("blocksize", 1), delta=kw.get("delta"))
    err = oracle.rel_l2(got, ref)
    print("%s rel-L2 %.3e" % (variant, err))
    assert err <= 1e-3


def test_fbp_parallel(cuda):
    """Parallel-beam FBP against the FP64 restatement (FBP is linear: 1e-5);
    a cone geometry is refused as in the reference; the filter keyword lands
    on the geometry for the next call, as the reference does."""
    from tigre_amd.geometry import geometry

    geo = geometry(mode="parallel", nVoxel=np.array([32, 40, 40]))
    angles = np.linspace(0, np.pi, 90, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 5, seed=12)
    proj = oracle.ax(img, prepared(geo, angles)).astype(np.float32)
    got = algs.fbp(proj, geo, angles, filter="hann")
    err = oracle.rel_l2(got, oalg.fbp(proj, geo, angles))
    print("FBP rel-L2 %.3e" % err)
    assert err <= 1e-5
    assert geo.filter == "hann"
    again = algs.fbp(proj, geo, angles)
    assert oracle.rel_l2(again, oalg.fbp(proj, geo, angles, "hann")) <= 1e-5
    with pytest.raises(ValueError):
        algs.fbp(proj, cone_geometry((8, 8, 8), (8.0, 8.0, 8.0), (8, 8), (1.0, 1.0), 100.0, 150.0), angles)


def test_irn_tv_cgls(cuda, problem):
    """IRN-TV-CGLS (two cycles of three steps) against the FP64 restatement;
    the device differences / Ltx against the oracle's on a ragged volume."""
    import torch

    from tigre_amd.algorithms import _irn_diffs, _irn_ltx

    rng = np.random.default_rng(4)
    x = rng.standard_normal((5, 7, 4))
    y = rng.standard_normal((3, 5, 7, 4))
    W = rng.uniform(0.5, 2.0, (5, 7, 4))
    xt, yt, Wt = (torch.tensor(v, dtype=torch.float64, device="cuda") for v in (x, y, W))
    assert np.abs(_irn_diffs(xt).cpu().numpy() - np.stack(oalg.irn_diffs(x))).max() <= 1e-12
    assert np.abs(_irn_ltx(Wt, yt).cpu().numpy() - oalg.irn_ltx(W, list(y))).max() <= 1e-12
    geo, angles, img, proj = problem
    got, l2 = algs.irn_tv_cgls(proj, geo, angles, 3, niter_outer=2, computel2=True)
    ref, norms = oalg.irn_tv_cgls(proj, geo, angles, 3, niter_outer=2)
    err = oracle.rel_l2(got, ref)
    print("IRN-TV-CGLS rel-L2 %.3e, residuals %s" % (err, l2))
    assert err <= 1e-4
    assert np.allclose(l2[0], norms, rtol=1e-4)


def test_hybrid_flsqr_tv(cuda, problem):
    """Hybrid fLSQR-TV (three steps): the reconstruction (return_x) and the
    residual history against the FP64 restatement; the entry point's default
    result is the starting image, as the reference's."""
    geo, angles, img, proj = problem
    got, l2 = algs.hybrid_flsqr_tv(proj, geo, angles, 3, computel2=True, return_x=True)
    ref, norms, start = oalg.hybrid_flsqr_tv(proj, geo, angles, 3)
    err = oracle.rel_l2(got, ref)
    print("hybrid fLSQR-TV rel-L2 %.3e, residuals %s vs %s" % (err, l2, norms))
    assert err <= 1e-4
    assert np.allclose(l2[0], norms, rtol=1e-4)
    assert np.all(algs.hybrid_flsqr_tv(proj, geo, angles, 1) == 0.0)


def test_iterativereconalg_is_ossart(cuda, problem):
    """The base class's entry point runs the OS-SART loop with its defaults."""
    geo, angles, img, proj = problem
    assert np.array_equal(algs.iterativereconalg(proj, geo, angles, 2), algs.ossart(proj, geo, angles, 2))
