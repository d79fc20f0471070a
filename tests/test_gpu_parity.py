"""GPU parity: the gfx950 operators against the FP64 oracle on seeded inputs.

Tolerance (FP32 storage, FP32 accumulation): relative L2 error <= 1e-5 against
the FP64 oracle, per result.
"""
import ctypes

import numpy as np
import pytest

import oracle
import tigre_amd
from tigre_amd import _native
from tigre_amd.geometry import cone_geometry, prepared
from tigre_amd.phantom import ellipsoid_phantom

pytestmark = pytest.mark.gpu
TOL = 1e-5


def small_cone_geometry():
    # DSO 1000 / DSD 1536 mm, 128^3 at 1 mm, 128^2 detector at 1.6 mm
    return cone_geometry((128, 128, 128), (128.0, 128.0, 128.0), (128, 128), (1.6, 1.6), 1000.0, 1536.0)


@pytest.fixture(scope="module")
def small_cone():
    geo = small_cone_geometry()
    angles = np.linspace(0, 2 * np.pi, 128, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 10, seed=20261203)
    return geo, angles, img, prepared(geo, angles)


def tilted_geometry(n=40, nviews=24, seed=5):
    """Offsets, COR, detector roll/pitch/yaw and non-planar Euler angles."""
    rng = np.random.default_rng(seed)
    geo = cone_geometry((n, n + 4, n - 4), (n * 1.0, (n + 4) * 0.9, (n - 4) * 1.1), (n + 8, n + 12), (1.5, 1.4),
                        900.0, 1400.0)
    geo.offOrigin = rng.uniform(-3, 3, size=(nviews, 3))
    geo.offDetector = rng.uniform(-2, 2, size=(nviews, 2))
    geo.rotDetector = rng.uniform(-0.05, 0.05, size=(nviews, 3))
    geo.COR = rng.uniform(-1, 1, size=nviews)
    angles = np.stack([np.linspace(0, 2 * np.pi, nviews, endpoint=False), rng.uniform(-0.2, 0.2, nviews),
                       rng.uniform(-0.2, 0.2, nviews)], axis=1)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 6, seed=seed)
    return geo, angles, img


@pytest.mark.parametrize("ptype", ["Siddon", "interpolated"])
def test_ax_small_cone(cuda, small_cone, ptype):
    import torch

    geo, angles, img, gp = small_cone
    got = tigre_amd.Ax(torch.from_numpy(img).to(cuda), geo, angles, ptype)
    torch.cuda.synchronize()
    ref = oracle.ax(img, gp, ptype)
    err = oracle.rel_l2(got.cpu().numpy(), ref)
    print("Ax %s small_cone rel-L2 %.3e" % (ptype, err))
    assert err <= TOL


@pytest.mark.parametrize("btype", ["FDK", "matched"])
def test_atb_small_cone(cuda, small_cone, btype):
    import torch

    geo, angles, img, gp = small_cone
    proj = oracle.ax(img, gp).astype(np.float32)
    got = tigre_amd.Atb(torch.from_numpy(proj).to(cuda), geo, angles, btype)
    torch.cuda.synchronize()
    ref = oracle.atb(proj, gp, btype)
    err = oracle.rel_l2(got.cpu().numpy(), ref)
    print("Atb %s small_cone rel-L2 %.3e" % (btype, err))
    assert err <= TOL


def test_host_contract_matches_device(cuda, small_cone):
    """NumPy in/out (reference contract) equals the resident device path bit for bit."""
    import torch

    geo, angles, img, gp = small_cone
    sub = angles[:16]
    a = tigre_amd.Ax(img, geo, sub)
    b = tigre_amd.Ax(torch.from_numpy(img).to(cuda), geo, sub).cpu().numpy()
    assert np.array_equal(a, b)
    c = tigre_amd.Atb(a, geo, sub, "matched")
    d = tigre_amd.Atb(torch.from_numpy(a).to(cuda), geo, sub, "matched").cpu().numpy()
    assert np.array_equal(c, d)


@pytest.mark.parametrize("ptype", ["Siddon", "interpolated"])
def test_ax_tilted(cuda, ptype):
    import torch

    geo, angles, img = tilted_geometry()
    gp = prepared(geo, angles)
    got = tigre_amd.Ax(torch.from_numpy(img).to(cuda), geo, angles, ptype).cpu().numpy()
    err = oracle.rel_l2(got, oracle.ax(img, gp, ptype))
    print("Ax %s tilted rel-L2 %.3e" % (ptype, err))
    assert err <= TOL


@pytest.mark.parametrize("btype", ["FDK", "matched"])
def test_atb_tilted(cuda, btype):
    import torch

    geo, angles, img = tilted_geometry()
    gp = prepared(geo, angles)
    proj = oracle.ax(img, gp).astype(np.float32)
    got = tigre_amd.Atb(torch.from_numpy(proj).to(cuda), geo, angles, btype).cpu().numpy()
    err = oracle.rel_l2(got, oracle.atb(proj, gp, btype))
    print("Atb %s tilted rel-L2 %.3e" % (btype, err))
    assert err <= TOL


def _siddon_case(name):
    """Geometries that drive each Siddon code path: the beam-staged cell walk and
    crossing walk (small_cone views away from / near 45 degrees), a slab box too large
    for LDS (coarse detector: walk from global memory), and views whose rays
    barely advance along any horizontal axis (wide fan: gather kernel)."""
    if name == "coarse":
        geo = cone_geometry((64, 64, 64), (64.0, 64.0, 64.0), (8, 8), (24.0, 24.0), 1000.0, 1536.0)
        angles = np.linspace(0, 2 * np.pi, 20, endpoint=False)
    else:
        geo = cone_geometry((32, 32, 32), (32.0, 32.0, 32.0), (16, 48), (4.0, 4.0), 60.0, 100.0)
        angles = np.linspace(0, 2 * np.pi, 24, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 6, seed=21)
    return geo, angles, img


@pytest.mark.parametrize("name", ["coarse", "widefan"])
def test_siddon_paths(cuda, name):
    import torch

    geo, angles, img = _siddon_case(name)
    got = tigre_amd.Ax(torch.from_numpy(img).to(cuda), geo, angles, "Siddon").cpu().numpy()
    ref = oracle.ax(img, prepared(geo, angles), "Siddon")
    err = oracle.rel_l2(got, ref)
    worst = float(np.max(np.abs(got - ref))) / float(np.max(np.abs(ref)))
    print("Ax Siddon %s rel-L2 %.3e max-rel %.3e" % (name, err, worst))
    assert err <= TOL and worst <= 1e-4


def test_siddon_small_cone_max_error(cuda, small_cone):
    """Every ray, not only the L2 aggregate: a wrong voxel on a few rays shows here."""
    import torch

    geo, angles, img, gp = small_cone
    got = tigre_amd.Ax(torch.from_numpy(img).to(cuda), geo, angles, "Siddon").cpu().numpy()
    ref = oracle.ax(img, gp, "Siddon")
    worst = float(np.max(np.abs(got - ref))) / float(np.max(np.abs(ref)))
    print("Ax Siddon small_cone max-rel %.3e" % worst)
    assert worst <= 1e-4


def test_interp_point_filter(cuda):
    """accuracy > 1 selects nearest-voxel sampling (ray_interpolated_projection.cu:589-595)."""
    import torch

    geo = cone_geometry((40, 40, 40), (40.0, 40.0, 40.0), (48, 48), (1.6, 1.6), 1000.0, 1536.0, accuracy=1.5)
    angles = np.linspace(0, 2 * np.pi, 12, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 5, seed=11)
    got = tigre_amd.Ax(torch.from_numpy(img).to(cuda), geo, angles, "interpolated").cpu().numpy()
    err = oracle.rel_l2(got, oracle.ax(img, prepared(geo, angles), "interpolated"))
    assert err <= TOL


@pytest.mark.parametrize("kind", ["ax", "atb"])
def test_matlab_layout(cuda, kind):
    """The detector layout flag (v fastest) through the host C entry points."""
    geo = cone_geometry((24, 28, 32), (24.0, 28.0, 32.0), (20, 36), (1.6, 1.6), 1000.0, 1536.0)
    angles = np.linspace(0, 2 * np.pi, 10, endpoint=False)
    gp = prepared(geo, angles)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 4, seed=3)
    n, nv, nu = 10, 20, 36
    cg = _native.CGeometry(gp, n)
    ang = np.ascontiguousarray(gp.angles, dtype=np.float32)
    L = _native.lib()
    if kind == "ax":
        out = np.empty((n, nu, nv), dtype=np.float32)
        ptrs = (ctypes.POINTER(ctypes.c_float) * n)(
            *[ctypes.cast(out.ctypes.data + 4 * nu * nv * i, ctypes.POINTER(ctypes.c_float)) for i in range(n)])
        st = L.tigre_ax(img.ctypes.data_as(ctypes.c_void_p), cg.ptr(), ptrs, ang.ctypes.data_as(ctypes.c_void_p), n,
                        0, _native.LAYOUT_MATLAB, None, 0)
        assert st == 0, _native.last_error()
        ref = oracle.ax(img, gp, "Siddon", layout_matlab=True)
        assert oracle.rel_l2(out, ref) <= TOL
    else:
        proj = oracle.ax(img, gp, "Siddon", layout_matlab=True).astype(np.float32)
        out = np.empty((24, 28, 32), dtype=np.float32)  # (nZ, nY, nX)
        st = L.tigre_atb(proj.ctypes.data_as(ctypes.c_void_p), cg.ptr(), out.ctypes.data_as(ctypes.c_void_p),
                         ang.ctypes.data_as(ctypes.c_void_p), n, 0, _native.LAYOUT_MATLAB, None, 0)
        assert st == 0, _native.last_error()
        ref = oracle.atb(proj, gp, "FDK", layout_matlab=True)
        assert oracle.rel_l2(out, ref) <= TOL


@pytest.mark.parametrize("shape", [(40, 48, 56), (19, 13, 130), (2, 17, 67), (1, 9, 300)])
def test_minTV(cuda, shape):
    """minTV against the oracle; the second shape has several x tiles with a
    ragged last one, a ragged y tile and a 3-plane z chunk, the third fewer
    planes than the kernel prefetches, the fourth a single plane across three
    x tiles."""
    import torch

    img = ellipsoid_phantom(shape, tuple(float(v) for v in shape), 6, seed=9)
    img += np.random.default_rng(1).normal(0, 0.05, img.shape).astype(np.float32)
    got = tigre_amd.minTV(torch.from_numpy(img).to(cuda), 0.02, 20).cpu().numpy()
    err = oracle.rel_l2(got, oracle.pocs_tv(img, 0.02, 20))
    print("minTV rel-L2 %.3e" % err)
    assert err <= TOL
    host = tigre_amd.minTV(img, 0.02, 20)
    assert np.array_equal(host, got)


def test_tvdenoise(cuda):
    import torch

    img = ellipsoid_phantom((30, 34, 38), (30.0, 34.0, 38.0), 5, seed=4)
    img += np.random.default_rng(2).normal(0, 0.05, img.shape).astype(np.float32)
    got = tigre_amd.tvdenoise(torch.from_numpy(img).to(cuda), 25, 15.0).cpu().numpy()
    err = oracle.rel_l2(got, oracle.tvdenoise(img, 15.0, 25))
    print("tvdenoise rel-L2 %.3e" % err)
    assert err <= TOL


def test_aw_minTV(cuda):
    import torch

    img = ellipsoid_phantom((36, 40, 44), (36.0, 40.0, 44.0), 6, seed=12)
    img += np.random.default_rng(6).normal(0, 0.05, img.shape).astype(np.float32)
    got = tigre_amd.AwminTV(torch.from_numpy(img).to(cuda), 0.02, 15, delta=-0.05).cpu().numpy()
    err = oracle.rel_l2(got, oracle.aw_pocs_tv(img, 0.02, 15, -0.05))
    print("AwminTV rel-L2 %.3e" % err)
    assert err <= TOL
    assert np.array_equal(tigre_amd.AwminTV(img, 0.02, 15, delta=-0.05), got)


@pytest.mark.parametrize("shape", [(36, 40, 44), (19, 13, 130)])
def test_minPICCS(cuda, shape):
    import torch

    img = ellipsoid_phantom(shape, tuple(float(v) for v in shape), 6, seed=13)
    prior = ellipsoid_phantom(shape, tuple(float(v) for v in shape), 6, seed=14)
    img += np.random.default_rng(7).normal(0, 0.05, img.shape).astype(np.float32)
    got = tigre_amd.minPICCS(torch.from_numpy(img).to(cuda), torch.from_numpy(prior).to(cuda), 0.02, 15,
                             ratio=0.6).cpu().numpy()
    err = oracle.rel_l2(got, oracle.piccs_tv(img, prior, 0.02, 0.6, 15))
    print("minPICCS rel-L2 %.3e" % err)
    assert err <= TOL
    assert np.array_equal(tigre_amd.minPICCS(img, prior, 0.02, 15, ratio=0.6), got)


@pytest.mark.parametrize("ptype", ["Siddon", "interpolated"])
@pytest.mark.parametrize("nviews,pixel", [(24, 0.2), (48, 0.3)])
def test_ax_sparse_views_stacked_tiles(cuda, ptype, nviews, pixel):
    """Sparse views of a fine detector: a group of views spans at most the
    angle over which the beam of a 16-pixel tile moves 2.5 tile widths at the
    volume's edge (45 mm from the axis): 6.6 degrees for 0.2 mm pixels, less
    than the 15-degree spacing of 24 views, so every view forms a group of its
    own and each workgroup traces 2 x 2 tiles of one view (the 100 x 90
    detector has 4 x 3 tiles of 32 x 32 pixels, 288 >= 256 workgroups); 9.9
    degrees for 0.3 mm pixels, between one and two spacings of 48 views, so
    views go in pairs and each workgroup traces two tiles of each (4 x 6 tiles
    of 32 x 16 pixels, 24 pairs).  The last column and row of tiles lie partly
    off the detector."""
    import torch

    geo = cone_geometry((64, 64, 64), (64.0, 64.0, 64.0), (100, 90), (pixel, pixel), 1000.0, 1536.0)
    angles = np.linspace(0, 2 * np.pi, nviews, endpoint=False) + 0.1
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 8, seed=33)
    cap = 2.5 * (16 * pixel * 1000.0 / 1536.0) / (0.5 * np.hypot(64.0, 64.0))
    spacing = 2 * np.pi / nviews
    if nviews == 24:
        assert cap < spacing and ((100 + 31) // 32) * ((90 + 31) // 32) * 24 >= 256
    else:
        assert spacing < cap < 2 * spacing and ((100 + 31) // 32) * ((90 + 15) // 16) * 24 >= 256
    d_img = torch.from_numpy(img).to(cuda)
    got = tigre_amd.Ax(d_img, geo, angles, ptype).cpu().numpy()
    ref = oracle.ax(img, prepared(geo, angles), ptype)
    err = oracle.rel_l2(got, ref)
    worst = float(np.max(np.abs(got - ref))) / float(np.max(np.abs(ref)))
    print("Ax %s sparse views rel-L2 %.3e max-rel %.3e" % (ptype, err, worst))
    assert err <= TOL and worst <= 1e-4
    # the same views one call each (one group, one tile per workgroup): same bits
    one = np.concatenate([tigre_amd.Ax(d_img, geo, angles[i:i + 1], ptype).cpu().numpy() for i in range(nviews)])
    assert np.array_equal(one, got)


def test_siddon_exact_crossing_ties(cuda):
    """Views at odd multiples of 45 degrees whose central detector column passes
    through voxel corners: x and y crossings coincide exactly at every step,
    also at the planes where the beam kernel moves a ray to its next LDS slab
    (regression: a tie at that plane must not end the slab before the
    principal crossing)."""
    import torch

    geo = cone_geometry((40, 48, 48), (40.0, 48.0, 48.0), (40, 64), (1.0, 1.0), 300.0, 450.0)
    geo.offDetector = np.array([0.0, 0.5])  # a pixel centre on the central ray
    angles = np.array([np.pi / 4, 3 * np.pi / 4, 5 * np.pi / 4, 7 * np.pi / 4, 0.3, 1.9])
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 6, seed=17) + 0.1
    got = tigre_amd.Ax(torch.from_numpy(img).to(cuda), geo, angles, "Siddon").cpu().numpy()
    ref = oracle.ax(img, prepared(geo, angles), "Siddon")
    assert np.all(np.isfinite(got))
    err = oracle.rel_l2(got, ref)
    print("Siddon with exact crossing ties rel-L2 %.3e" % err)
    assert err <= TOL


def _window_case(name):
    """Geometries for the LDS-staged FDK backprojector's window logic: voxels
    projecting to several pixels (windows wider than the staged 16 x 24 quads:
    every view falls back to the global gather), per-view source distances that
    make some views fit and others not, and a volume wider than the detector's
    field of view (windows clamped at the detector edges)."""
    if name == "wide":
        geo = cone_geometry((24, 24, 24), (96.0, 96.0, 96.0), (64, 64), (1.0, 1.0), 300.0, 450.0)
        nviews = 12
    elif name == "mixed":
        nviews = 16
        geo = cone_geometry((40, 40, 40), (40.0, 40.0, 40.0), (80, 80), (1.0, 1.0), 400.0, 600.0)
        geo.DSO = np.linspace(250.0, 600.0, nviews)
        geo.DSD = geo.DSO * np.linspace(1.3, 3.0, nviews)
    else:  # "edge"
        geo = cone_geometry((64, 64, 64), (64.0, 64.0, 64.0), (32, 40), (1.0, 1.0), 500.0, 750.0)
        nviews = 20
    angles = np.linspace(0.1, 2 * np.pi + 0.1, nviews, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 5, seed=31)
    return geo, angles, img


@pytest.mark.parametrize("name", ["wide", "mixed", "edge"])
def test_atb_fdk_windows(cuda, name):
    import torch

    geo, angles, img = _window_case(name)
    gp = prepared(geo, angles)
    proj = oracle.ax(img, gp).astype(np.float32)
    got = tigre_amd.Atb(torch.from_numpy(proj).to(cuda), geo, angles, "FDK").cpu().numpy()
    err = oracle.rel_l2(got, oracle.atb(proj, gp, "FDK"))
    print("Atb FDK windows %s rel-L2 %.3e" % (name, err))
    assert err <= TOL


def test_interpolated_cap_flag_is_per_batch(cuda):
    """The interpolated sample cap follows the reference's 9-view launch batch
    (ray_interpolated_projection.cu:352-392): planar views batched with a tilted
    one take the three-axis cap.  On voxels 400x finer in z than in x/y the two
    caps give different ray sums (up to the whole ray); the HIP path matches the
    oracle with and without the tilted view in the call."""
    import torch

    from tests.test_oracle_forward import _batch_geometry

    geo, ang = _batch_geometry()
    img = np.random.default_rng(9).uniform(0.5, 1.0, size=tuple(geo.nVoxel)).astype(np.float32)
    d_img = torch.from_numpy(img).to(cuda)
    for views in (ang, ang[1:]):
        print("interpolated batch cap: %d views" % len(views), flush=True)
        got = tigre_amd.Ax(d_img, geo, views, "interpolated")
        torch.cuda.synchronize()
        ref = oracle.ax(img, prepared(geo, views), "interpolated")
        err = oracle.rel_l2(got.cpu().numpy(), ref)
        print("Ax interpolated batch cap (%d views) rel-L2 %.3e" % (len(views), err))
        assert err <= TOL
