"""GPU: parallel-beam operators against the FP64 oracle (rel-L2 <= 1e-5).

Geometries: the reference's default parallel geometry with generic angles,
and one with a detector offset and a centre-of-rotation shift (the
interpolated projector's doubled detector offset and the Siddon projector's
undisplaced source are restated in the oracle as the reference computes
them).  The host contract and the device path must agree bit for bit.
"""
import numpy as np
import pytest

import oracle
import tigre_amd
from tigre_amd.geometry import geometry, prepared
from tigre_amd.phantom import ellipsoid_phantom

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _geometries():
    g1 = geometry(mode="parallel", nVoxel=np.array([40, 48, 56]))
    g2 = geometry(mode="parallel", nVoxel=np.array([36, 44, 40]))
    g2.offDetector = np.array([1.5, -2.25])
    g2.COR = 0.8
    g2.nDetector = np.array([40, 50])
    g2.sDetector = g2.nDetector * np.asarray(g2.dDetector)
    return {"default": g1, "offsets": g2}


@pytest.mark.parametrize("name", ["default", "offsets"])
@pytest.mark.parametrize("kind", ["Siddon", "interpolated"])
def test_parallel_ax(cuda, name, kind):
    import torch

    geo = _geometries()[name]
    angles = np.linspace(0.05, np.pi, 33, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 6, seed=23)
    got = tigre_amd.Ax(torch.from_numpy(img).to(cuda), geo, angles, kind).cpu().numpy()
    ref = oracle.ax(img, prepared(geo, angles), kind)
    err = oracle.rel_l2(got, ref)
    print("parallel Ax %s %s rel-L2 %.3e" % (kind, name, err))
    assert err <= TOL
    assert np.array_equal(tigre_amd.Ax(img, geo, angles, kind), got)


@pytest.mark.parametrize("name", ["default", "offsets"])
def test_parallel_atb(cuda, name):
    import torch

    geo = _geometries()[name]
    angles = np.linspace(0.05, np.pi, 33, endpoint=False)
    rng = np.random.default_rng(5)
    proj = rng.random((len(angles),) + tuple(int(v) for v in geo.nDetector)).astype(np.float32)
    got = tigre_amd.Atb(torch.from_numpy(proj).to(cuda), geo, angles, "FDK").cpu().numpy()
    ref = oracle.atb(proj, prepared(geo, angles), "FDK")
    err = oracle.rel_l2(got, ref)
    print("parallel Atb %s rel-L2 %.3e" % (name, err))
    assert err <= TOL
    assert np.array_equal(tigre_amd.Atb(proj, geo, angles, "matched"), got)


def test_interpolated_ignores_stale_lds(cuda):
    """Regression: a sample clamped to the far face (P = nP) of a sweep towards -P
    lies beyond every slab; it must be skipped, not read with its upper tap
    outside the slab box.  That read had weight 0, so it only showed when the
    LDS still held NaN from an earlier kernel: dirty the LDS first (weight
    gradients of NaN activations stage them there), then project."""
    import ctypes

    import torch

    from tigre_amd import _native
    from tigre_amd.operators import _stream_handle

    lib = _native.lib()
    shape = (2, 32, 32, 64)
    n = ctypes.c_size_t()
    _native.check(lib.tigre_cnn_wgrad_workspace_size(*shape, ctypes.byref(n)), "ws")
    ws = torch.empty(n.value, dtype=torch.uint8, device="cuda")
    nan = torch.full(shape + (32,), float("nan"), dtype=torch.bfloat16, device="cuda")
    out = torch.empty((28, 32, 32), device="cuda")
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    for _ in range(3):
        _native.check(lib.tigre_cnn_wgrad(p(nan), _native.CNN_IN_BF16X32, p(nan), _native.CNN_IN_BF16X32, p(out),
                                          *shape, p(ws), n.value, _stream_handle()), "wgrad")
    geo = _geometries()["default"]
    angles = np.linspace(0.05, np.pi, 33, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 6, seed=23)
    got = tigre_amd.Ax(torch.from_numpy(img).to(cuda), geo, angles, "interpolated").cpu().numpy()
    assert np.isfinite(got).all()
    ref = oracle.ax(img, prepared(geo, angles), "interpolated")
    assert oracle.rel_l2(got, ref) <= TOL


@pytest.mark.parametrize("kind", ["Siddon", "interpolated"])
def test_parallel_ax_sparse_views_stacked_tiles(cuda, kind):
    """Sparse parallel views of a fine detector (0.2 mm pixels, 64 mm volume):
    a group of views spans at most the angle over which a 16-pixel tile's beam
    moves 2.5 tile widths at the volume's edge, 2.5 * 3.2 / 45.3 rad = 10
    degrees, less than the 15-degree spacing of 24 views over 2 pi, so each
    view is a group of its own and every workgroup traces 2 x 2 tiles of one
    view (the 100 x 90 detector has 4 x 3 tiles of 32 x 32 pixels, 288 >= 256
    workgroups over the 24 views).  Against the oracle and one call per view."""
    import torch

    geo = geometry(mode="parallel", nVoxel=np.array([64, 64, 64]))
    geo.nDetector = np.array([100, 90])
    geo.dDetector = np.array([0.2, 0.2])
    geo.sDetector = geo.nDetector * geo.dDetector
    angles = np.linspace(0, 2 * np.pi, 24, endpoint=False) + 0.1
    assert 2.5 * 16 * 0.2 / (0.5 * np.hypot(64.0, 64.0)) < 2 * np.pi / 24
    assert ((100 + 31) // 32) * ((90 + 31) // 32) * 24 >= 256
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 8, seed=41)
    d_img = torch.from_numpy(img).to(cuda)
    got = tigre_amd.Ax(d_img, geo, angles, kind).cpu().numpy()
    ref = oracle.ax(img, prepared(geo, angles), kind)
    err = oracle.rel_l2(got, ref)
    print("parallel Ax %s sparse views rel-L2 %.3e" % (kind, err))
    assert err <= TOL
    one = np.concatenate([tigre_amd.Ax(d_img, geo, angles[i:i + 1], kind).cpu().numpy() for i in range(24)])
    assert np.array_equal(one, got)


def test_parallel_ossart_cgls(cuda):
    """OS-SART and CGLS on a parallel-beam geometry go through the parallel
    operators (V = block size) and match the FP64 restatements."""
    import oracle.algorithms as oalg
    import tigre_amd.algorithms as algs

    geo = geometry(mode="parallel", nVoxel=np.array([24, 32, 32]))
    angles = np.linspace(0, np.pi, 40, endpoint=False)
    img = np.random.default_rng(9).random((24, 32, 32)).astype(np.float32)
    proj = oracle.ax(img, prepared(geo, angles)).astype(np.float32)
    err_os = oracle.rel_l2(algs.ossart(proj, geo, angles, 2, blocksize=10, verbose=False),
                           oalg.os_sart(proj, geo, angles, 2, blocksize=10))
    err_cg = oracle.rel_l2(algs.cgls(proj, geo, angles, 4), oalg.cgls(proj, geo, angles, 4)[0])
    print("parallel OS-SART rel-L2 %.3e, CGLS %.3e" % (err_os, err_cg))
    assert err_os <= 1e-4 and err_cg <= 1e-4
