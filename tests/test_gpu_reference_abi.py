"""GPU: the reference's C++ entry points (siddon_ray_projection & co., the
symbols a gateway compiled against the reference headers calls) against the
FP64 oracle, called through their mangled names with the geometry passed by
value and the device list by reference, as the Cython gateways call them
(_Ax.pyx:78-87, _Atb.pyx:68-76, _minTV.pyx:60, _tv_proximal.pyx:66).

Tolerance: relative L2 <= 1e-5 (FP32 storage and accumulation)."""
import ctypes

import numpy as np
import pytest

import oracle
from tigre_amd import _native
from tigre_amd.geometry import cone_geometry, prepared
from tigre_amd.phantom import ellipsoid_phantom

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def case():
    geo = cone_geometry((40, 44, 48), (40.0, 44.0, 48.0), (52, 60), (1.4, 1.4), 700.0, 1100.0)
    rng = np.random.default_rng(17)
    n = 12
    geo.offDetector = rng.uniform(-2, 2, size=(n, 2))
    angles = np.stack([np.linspace(0, 2 * np.pi, n, endpoint=False), np.zeros(n), np.zeros(n)], axis=1)
    gp = prepared(geo, angles)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 6, seed=17)
    return gp, img


def _device_list():
    ids = (ctypes.c_int * 1)(0)
    g = _native.GpuIds()
    g.m_piDeviceIds = ctypes.cast(ids, ctypes.POINTER(ctypes.c_int))
    g.m_iCount = 1
    return g, ids


def _forward(abi, name, img, gp, matlab=False):
    n = gp.angles.shape[0]
    nV, nU = int(gp.nDetector[0]), int(gp.nDetector[1])
    out = np.full((n, nU, nV) if matlab else (n, nV, nU), np.nan, dtype=np.float32)
    ptrs = (ctypes.c_void_p * n)(*[out.ctypes.data + k * nU * nV * 4 for k in range(n)])
    ang = np.ascontiguousarray(gp.angles, dtype=np.float32)
    cg = _native.CGeometry(gp, n)
    ids, keep = _device_list()
    st = abi[name](np.ascontiguousarray(img).ctypes.data, cg.struct, ptrs, ang.ctypes.data, n, ctypes.byref(ids))
    assert st == 0, _native.last_error()
    return out


@pytest.mark.parametrize("name,ptype", [("siddon_ray_projection", "Siddon"),
                                        ("interpolation_projection", "interpolated")])
def test_forward_entry_points(cuda, case, name, ptype):
    gp, img = case
    got = _forward(_native.reference_abi(), name, img, gp)
    err = oracle.rel_l2(got, oracle.ax(img, gp, ptype))
    print("%s rel-L2 %.3e" % (name, err))
    assert err <= TOL


def test_forward_entry_point_matlab_build(cuda, case):
    """The MATLAB-layout build of the same symbols writes v fastest."""
    gp, img = case
    got = _forward(_native.reference_abi(_native.LIB_PATH_MATLAB), "siddon_ray_projection", img, gp, matlab=True)
    err = oracle.rel_l2(got, oracle.ax(img, gp, "Siddon", layout_matlab=True))
    print("siddon_ray_projection (MATLAB layout) rel-L2 %.3e" % err)
    assert err <= TOL


@pytest.mark.parametrize("name,btype", [("voxel_backprojection", "FDK"), ("voxel_backprojection2", "matched")])
def test_backward_entry_points(cuda, case, name, btype):
    gp, img = case
    abi = _native.reference_abi()
    proj = oracle.ax(img, gp).astype(np.float32)
    n = gp.angles.shape[0]
    out = np.full(img.shape, np.nan, dtype=np.float32)
    ang = np.ascontiguousarray(gp.angles, dtype=np.float32)
    cg = _native.CGeometry(gp, n)
    ids, keep = _device_list()
    st = abi[name](proj.ctypes.data, cg.struct, out.ctypes.data, ang.ctypes.data, n, ctypes.byref(ids))
    assert st == 0, _native.last_error()
    err = oracle.rel_l2(out, oracle.atb(proj, gp, btype))
    print("%s rel-L2 %.3e" % (name, err))
    assert err <= TOL


def test_tv_entry_points(cuda, case):
    _, img = case
    abi = _native.reference_abi()
    ids, keep = _device_list()
    nz, ny, nx = img.shape
    size = (ctypes.c_long * 3)(nx, ny, nz)
    src = np.ascontiguousarray(img, dtype=np.float32)
    out = np.full_like(src, np.nan)
    abi["pocs_tv"](src.ctypes.data, out.ctypes.data, 0.05, size, 6, ctypes.byref(ids))
    err = oracle.rel_l2(out, oracle.pocs_tv(img, 0.05, 6))
    print("pocs_tv rel-L2 %.3e" % err)
    assert err <= TOL
    spacing = np.ones(3, dtype=np.float32)
    abi["tvdenoising"](src.ctypes.data, out.ctypes.data, 20.0, spacing.ctypes.data, size, 10, ctypes.byref(ids))
    err = oracle.rel_l2(out, oracle.tvdenoise(img, 20.0, 10))
    print("tvdenoising rel-L2 %.3e" % err)
    assert err <= TOL


def test_device_queries(cuda):
    abi = _native.reference_abi()
    n = abi["GetGpuCount"]()
    assert n >= 1
    name = ctypes.create_string_buffer(128)
    abi["GetGpuName"](0, name)
    assert len(name.value) > 0
    ids = (ctypes.c_int * n)()
    msg = ctypes.create_string_buffer(65535)
    assert abi["GetGpuIdArray"](b"", ids, n, msg) == n and list(ids) == list(range(n))
    assert abi["GetGpuIdArray"](name.value, ids, n, msg) >= 1 and ids[0] == 0
    assert abi["GetGpuIdArray"](b"no such device", ids, n, msg) == 0
    assert b"Found GPUs" in msg.value
