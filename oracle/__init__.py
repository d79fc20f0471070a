"""TEST INFRASTRUCTURE ONLY: Python access to the FP64 CPU oracle.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
package, and only as the checker.  The product (tigre_amd) never imports it.

Geometry objects passed here must already be expanded by ``check_geo`` and
rounded by ``cast_to_single`` (the Python front end's preparation,
Python/tigre/utilities/geometry.py:24-140); the oracle reads the same float32
values as doubles.  Axis handling restates convert_to_c_geometry
(Python/tigre/utilities/cuda_interface/_types.pxd:69-166).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)


class OgGeo(ctypes.Structure):
    _fields_ = [
        ("nVoxelX", ctypes.c_int), ("nVoxelY", ctypes.c_int), ("nVoxelZ", ctypes.c_int),
        ("sVoxelX", ctypes.c_double), ("sVoxelY", ctypes.c_double), ("sVoxelZ", ctypes.c_double),
        ("dVoxelX", ctypes.c_double), ("dVoxelY", ctypes.c_double), ("dVoxelZ", ctypes.c_double),
        ("nDetecU", ctypes.c_int), ("nDetecV", ctypes.c_int),
        ("sDetecU", ctypes.c_double), ("sDetecV", ctypes.c_double),
        ("dDetecU", ctypes.c_double), ("dDetecV", ctypes.c_double),
        ("accuracy", ctypes.c_double),
        ("nangles", ctypes.c_int),
        ("angles", _dp), ("DSO", _dp), ("DSD", _dp), ("COR", _dp),
        ("offOrigX", _dp), ("offOrigY", _dp), ("offOrigZ", _dp),
        ("offDetecU", _dp), ("offDetecV", _dp),
        ("dRoll", _dp), ("dPitch", _dp), ("dYaw", _dp),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        g = ctypes.POINTER(OgGeo)
        for name in ("og_ax_siddon", "og_ax_interp", "og_atb_fdk", "og_atb_matched", "og_ax_siddon_parallel",
                     "og_ax_interp_parallel", "og_atb_parallel"):
            fn = getattr(L, name)
            fn.argtypes = [vp, g, ctypes.c_int, vp]
            fn.restype = None
        L.og_forward_frame.argtypes = [g, ctypes.c_int, ctypes.c_int, vp]
        L.og_forward_frame_parallel.argtypes = [g, ctypes.c_int, ctypes.c_int, vp]
        L.og_backward_frame.argtypes = [g, ctypes.c_int, vp]
        L.og_maxdistance_cuboid.argtypes = [g, ctypes.c_int]
        L.og_maxdistance_cuboid.restype = ctypes.c_double
        L.og_pocs_tv.argtypes = [vp, vp, ctypes.c_double, ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_int]
        L.og_tv_gradient.argtypes = [vp, vp, ctypes.c_long, ctypes.c_long, ctypes.c_long]
        L.og_tv_gradient_ex.argtypes = [vp, vp, ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double]
        L.og_aw_pocs_tv.argtypes = [vp, vp, ctypes.c_double, ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                    ctypes.c_int, ctypes.c_double]
        L.og_piccs_tv.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_long, ctypes.c_long,
                                  ctypes.c_long, ctypes.c_int]
        L.og_tvdenoise.argtypes = [vp, vp, ctypes.c_double, vp, ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                   ctypes.c_int]
        L.og_set_threads.argtypes = [ctypes.c_int]
        L.og_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def set_threads(n):
    lib().og_set_threads(int(n))


def get_threads():
    return lib().og_get_threads()


class _Geo:
    def __init__(self, geo):
        n = geo.angles.shape[0]
        d = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))  # noqa: E731
        off = np.asarray(geo.offOrigin, dtype=np.float64).reshape(n, 3)
        offd = np.asarray(geo.offDetector, dtype=np.float64).reshape(n, 2)
        rot = np.asarray(geo.rotDetector, dtype=np.float64).reshape(n, 3)
        self.keep = {
            "angles": d(geo.angles),
            "DSO": d(np.broadcast_to(np.asarray(geo.DSO, dtype=np.float64), (n,))),
            "DSD": d(np.broadcast_to(np.asarray(geo.DSD, dtype=np.float64), (n,))),
            "COR": d(np.broadcast_to(np.asarray(geo.COR, dtype=np.float64), (n,))),
            "offOrigX": d(off[:, 2]), "offOrigY": d(off[:, 1]), "offOrigZ": d(off[:, 0]),
            "offDetecU": d(offd[:, 1]), "offDetecV": d(offd[:, 0]),
            "dRoll": d(rot[:, 2]), "dPitch": d(rot[:, 1]), "dYaw": d(rot[:, 0]),
        }
        s = OgGeo()
        nv = [int(v) for v in np.asarray(geo.nVoxel)]
        sv = np.asarray(geo.sVoxel, dtype=np.float64)
        dv = np.asarray(geo.dVoxel, dtype=np.float64)
        nd = [int(v) for v in np.asarray(geo.nDetector)]
        sd = np.asarray(geo.sDetector, dtype=np.float64)
        dd = np.asarray(geo.dDetector, dtype=np.float64)
        s.nVoxelX, s.nVoxelY, s.nVoxelZ = nv[2], nv[1], nv[0]
        s.sVoxelX, s.sVoxelY, s.sVoxelZ = sv[2], sv[1], sv[0]
        s.dVoxelX, s.dVoxelY, s.dVoxelZ = dv[2], dv[1], dv[0]
        s.nDetecU, s.nDetecV = nd[1], nd[0]
        s.sDetecU, s.sDetecV = sd[1], sd[0]
        s.dDetecU, s.dDetecV = dd[1], dd[0]
        s.accuracy = float(geo.accuracy)
        s.nangles = n
        for k, v in self.keep.items():
            setattr(s, k, v.ctypes.data_as(_dp))
        self.s = s
        self.n = n
        self.nv, self.nd = nv, nd


def _vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _parallel(geo):
    return getattr(geo, "mode", "cone") == "parallel"


def ax(img, geo, projection_type="Siddon", layout_matlab=False):
    """FP64 forward projection; returns (n, nV, nU) (or (n, nU, nV) for MATLAB layout).
    geo.mode 'parallel' selects the parallel-beam restatements."""
    G = _Geo(geo)
    img = np.ascontiguousarray(img, dtype=np.float32)
    shape = (G.n, G.nd[1], G.nd[0]) if layout_matlab else (G.n, G.nd[0], G.nd[1])
    out = np.empty(shape, dtype=np.float64)
    if _parallel(geo):
        fn = lib().og_ax_interp_parallel if projection_type == "interpolated" else lib().og_ax_siddon_parallel
    else:
        fn = lib().og_ax_interp if projection_type == "interpolated" else lib().og_ax_siddon
    fn(_vp(img), ctypes.byref(G.s), int(layout_matlab), _vp(out))
    return out


def atb(proj, geo, backprojection_type="FDK", layout_matlab=False):
    """FP64 backprojection; returns (nZ, nY, nX)."""
    G = _Geo(geo)
    proj = np.ascontiguousarray(proj, dtype=np.float32)
    out = np.empty((G.nv[0], G.nv[1], G.nv[2]), dtype=np.float64)
    if _parallel(geo):
        fn = lib().og_atb_parallel  # one parallel-beam backprojector for 'FDK' and 'matched'
    else:
        fn = lib().og_atb_matched if backprojection_type == "matched" else lib().og_atb_fdk
    fn(_vp(proj), ctypes.byref(G.s), int(layout_matlab), _vp(out))
    return out


def forward_frame(geo, view, voxel_centre_origin):
    G = _Geo(geo)
    out = np.empty(12)
    fn = lib().og_forward_frame_parallel if _parallel(geo) else lib().og_forward_frame
    fn(ctypes.byref(G.s), int(view), int(voxel_centre_origin), _vp(out))
    return out.reshape(4, 3)


def backward_frame(geo, view):
    G = _Geo(geo)
    out = np.empty(15)
    lib().og_backward_frame(ctypes.byref(G.s), int(view), _vp(out))
    return out.reshape(5, 3)


def maxdistance_cuboid(geo, view):
    G = _Geo(geo)
    return lib().og_maxdistance_cuboid(ctypes.byref(G.s), int(view))


def pocs_tv(img, alpha, maxiter):
    img = np.ascontiguousarray(img, dtype=np.float64)
    out = np.empty_like(img)
    nz, ny, nx = img.shape
    lib().og_pocs_tv(_vp(img), _vp(out), float(alpha), nx, ny, nz, int(maxiter))
    return out


def tv_gradient(img):
    img = np.ascontiguousarray(img, dtype=np.float64)
    out = np.empty_like(img)
    nz, ny, nx = img.shape
    lib().og_tv_gradient(_vp(img), _vp(out), nx, ny, nz)
    return out


def tv_gradient_ex(img, pad_z=True, delta=None, eps=1e-8):
    img = np.ascontiguousarray(img, dtype=np.float64)
    out = np.empty_like(img)
    nz, ny, nx = img.shape
    lib().og_tv_gradient_ex(_vp(img), _vp(out), nx, ny, nz, int(bool(pad_z)),
                            0.0 if delta is None else 1.0 / float(delta), float(eps))
    return out


def aw_pocs_tv(img, alpha, maxiter, delta):
    img = np.ascontiguousarray(img, dtype=np.float64)
    out = np.empty_like(img)
    nz, ny, nx = img.shape
    lib().og_aw_pocs_tv(_vp(img), _vp(out), float(alpha), nx, ny, nz, int(maxiter), float(delta))
    return out


def piccs_tv(img, prior, alpha, ratio, maxiter):
    img = np.ascontiguousarray(img, dtype=np.float64)
    prior = np.ascontiguousarray(prior, dtype=np.float64)
    out = np.empty_like(img)
    nz, ny, nx = img.shape
    lib().og_piccs_tv(_vp(img), _vp(prior), _vp(out), float(alpha), float(ratio), nx, ny, nz, int(maxiter))
    return out


def tvdenoise(img, lmbda, maxiter, spacing=(1.0, 1.0, 1.0)):
    img = np.ascontiguousarray(img, dtype=np.float64)
    out = np.empty_like(img)
    nz, ny, nx = img.shape
    sp = np.asarray(spacing, dtype=np.float64)
    lib().og_tvdenoise(_vp(img), _vp(out), float(lmbda), _vp(sp), nx, ny, nz, int(maxiter))
    return out


def rel_l2(a, b):
    """||a - b|| / ||b|| in FP64."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))
