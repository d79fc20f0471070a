"""ctypes binding of the gfx950 library ``tigre_amd/lib/libtigre_amd.so``.

The shared library is the product: every operator in this package calls it.
There is no CPU fallback; importing an operator without the built library
raises ``ImportError`` with the build command.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libtigre_amd.so")

LAYOUT_PYTHON = 0
LAYOUT_MATLAB = 1
PROJ_SIDDON = 0
PROJ_INTERPOLATED = 1
BACKPROJ_FDK = 0
BACKPROJ_MATCHED = 1
DTYPE_F32 = 0
DTYPE_BF16 = 1
DTYPE_F16 = 2
CNN_IN_BF16X32 = 0
CNN_IN_F32X1 = 1
CNN_EPI_BIAS_RELU = 0
CNN_EPI_RELU_MASK = 1
CNN_EPI_RESIDUAL = 2
KERNEL_AX_SIDDON = 0
KERNEL_AX_INTERPOLATED = 1
KERNEL_ATB_FDK = 2
KERNEL_ATB_MATCHED = 3
KERNEL_RELAYOUT = 4
KERNEL_POCS_TV = 5
KERNEL_TVDENOISE = 6
KERNEL_PROBE = 7

_fp = ctypes.POINTER(ctypes.c_float)


class TigreGeometry(ctypes.Structure):
    """Mirror of ``TigreGeometry`` (include/tigre_amd.h), itself byte-compatible
    with the reference ``Geometry`` (Common/CUDA/types_TIGRE.hpp:48-85)."""

    _fields_ = [
        ("nVoxelX", ctypes.c_int), ("nVoxelY", ctypes.c_int), ("nVoxelZ", ctypes.c_int),
        ("sVoxelX", ctypes.c_float), ("sVoxelY", ctypes.c_float), ("sVoxelZ", ctypes.c_float),
        ("dVoxelX", ctypes.c_float), ("dVoxelY", ctypes.c_float), ("dVoxelZ", ctypes.c_float),
        ("offOrigX", _fp), ("offOrigY", _fp), ("offOrigZ", _fp),
        ("DSO", _fp),
        ("nDetecU", ctypes.c_int), ("nDetecV", ctypes.c_int),
        ("sDetecU", ctypes.c_float), ("sDetecV", ctypes.c_float),
        ("dDetecU", ctypes.c_float), ("dDetecV", ctypes.c_float),
        ("offDetecU", _fp), ("offDetecV", _fp),
        ("DSD", _fp),
        ("dRoll", _fp), ("dPitch", _fp), ("dYaw", _fp),
        ("unitX", ctypes.c_float), ("unitY", ctypes.c_float), ("unitZ", ctypes.c_float),
        ("alpha", ctypes.c_float), ("theta", ctypes.c_float), ("psi", ctypes.c_float),
        ("COR", _fp),
        ("maxLength", ctypes.c_float),
        ("accuracy", ctypes.c_float),
    ]


# Every symbol include/tigre_amd.h declares, with its ctypes signature.
SIGNATURES = {
    "tigre_ax": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TigreGeometry), ctypes.POINTER(_fp),
                                ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_void_p, ctypes.c_int]),
    "tigre_atb": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TigreGeometry), ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.c_int]),
    "tigre_ax_parallel": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TigreGeometry), ctypes.POINTER(_fp),
                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_int]),
    "tigre_atb_parallel": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TigreGeometry), ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]),
    "tigre_ax_parallel_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TigreGeometry), ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_void_p]),
    "tigre_atb_parallel_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TigreGeometry), ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_void_p]),
    "tigre_pocs_tv": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p,
                                     ctypes.c_int, ctypes.c_void_p, ctypes.c_int]),
    "tigre_aw_pocs_tv": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p,
                                        ctypes.c_int, ctypes.c_float, ctypes.c_void_p, ctypes.c_int]),
    "tigre_piccs_tv": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]),
    "tigre_tvdenoising": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]),
    "tigre_ax_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TigreGeometry), ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p]),
    "tigre_atb_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TigreGeometry), ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p]),
    "tigre_ax_device_typed": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TigreGeometry), ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_void_p]),
    "tigre_atb_device_typed": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TigreGeometry), ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_void_p]),
    "tigre_pocs_tv_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                            ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "tigre_aw_pocs_tv_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p,
                                               ctypes.c_int, ctypes.c_float, ctypes.c_void_p]),
    "tigre_piccs_tv_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                             ctypes.c_float, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "tigre_tvdenoising_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                ctypes.c_void_p]),
    "tigre_plan_create": (ctypes.c_int, [ctypes.POINTER(TigreGeometry), ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "tigre_plan_workspace_size": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]),
    "tigre_plan_ax": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "tigre_plan_atb": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "tigre_plan_sart_update": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                              ctypes.c_size_t, ctypes.c_void_p]),
    "tigre_sart_residual": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
                                           ctypes.c_void_p]),
    "tigre_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "tigre_batch_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "tigre_batch_sart": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "tigre_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "tigre_cnn_conv": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_void_p]),
    "tigre_cnn_wgrad_workspace_size": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                      ctypes.POINTER(ctypes.c_size_t)]),
    "tigre_cnn_wgrad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_size_t, ctypes.c_void_p]),
    "tigre_cnn_wgrad_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                               ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_void_p]),
    "tigre_cnn_wgrad_noalloc_limit": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t)]),
    "tigre_valu_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "tigre_get_gpu_count": (ctypes.c_int, []),
    "tigre_get_gpu_name": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
    "tigre_kernel_timing": (ctypes.c_int, [ctypes.c_int]),
    "tigre_kernel_time": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_int)]),
    "tigre_last_error": (ctypes.c_char_p, []),
    "tigre_version": (ctypes.c_char_p, []),
}

# The reference's C++ entry points (include/tigre_reference_abi.hpp), by their
# Itanium-mangled names: what a gateway compiled against the reference headers
# links to.  Geometry travels by value (a TigreGeometry), GpuIds by reference.
REFERENCE_ABI = {
    "siddon_ray_projection": "_Z21siddon_ray_projectionPf8GeometryPS_PKfiRK6GpuIds",
    "interpolation_projection": "_Z24interpolation_projectionPf8GeometryPS_PKfiRK6GpuIds",
    "siddon_ray_projection_parallel": "_Z30siddon_ray_projection_parallelPf8GeometryPS_PKfiRK6GpuIds",
    "interpolation_projection_parallel": "_Z33interpolation_projection_parallelPf8GeometryPS_PKfiRK6GpuIds",
    "voxel_backprojection": "_Z20voxel_backprojectionPf8GeometryS_PKfiRK6GpuIds",
    "voxel_backprojection2": "_Z21voxel_backprojection2Pf8GeometryS_PKfiRK6GpuIds",
    "voxel_backprojection_parallel": "_Z29voxel_backprojection_parallelPf8GeometryS_PKfiRK6GpuIds",
    "pocs_tv": "_Z7pocs_tvPfS_fPKliRK6GpuIds",
    "aw_pocs_tv": "_Z10aw_pocs_tvPfS_fPKlifRK6GpuIds",
    "piccs_tv": "_Z8piccs_tvPKfS0_PfffPKliRK6GpuIds",
    "tvdenoising": "_Z11tvdenoisingPfS_fPKfPKliRK6GpuIds",
    "GetGpuCount": "_Z11GetGpuCountv",
    "GetGpuIdArray": "_Z13GetGpuIdArrayPKcPiiPc",
    "GetGpuName": "_Z10GetGpuNameiPc",
    "GpuIds::GpuIds": "_ZN6GpuIdsC1Ev",
    "GpuIds::~GpuIds": "_ZN6GpuIdsD1Ev",
    "GpuIds::SetIds": "_ZN6GpuIds6SetIdsEiPi",
    "GpuIds::GetLength": "_ZNK6GpuIds9GetLengthEv",
    "GpuIds::SetAllGpus": "_ZN6GpuIds10SetAllGpusEi",
    "GpuIds::operator[]": "_ZN6GpuIdsixEi",
    "GpuIds::operator[] const": "_ZNK6GpuIdsixEi",
    "GpuIds::AreEqualDevices": "_ZNK6GpuIds15AreEqualDevicesEv",
}
# The same entry points with the MATLAB detector layout (v fastest), as the
# reference's MEX build compiles them.
LIB_PATH_MATLAB = os.path.join(os.path.dirname(LIB_PATH), "libtigre_amd_matlab.so")


class TigreBatchStep(ctypes.Structure):
    """Mirror of ``TigreBatchStep`` (include/tigre_amd.h)."""

    _fields_ = [("plan", ctypes.c_void_p), ("d_x", ctypes.c_void_p), ("d_proj", ctypes.c_void_p),
                ("d_w", ctypes.c_void_p), ("d_ax", ctypes.c_void_p), ("d_inv_v", ctypes.c_void_p),
                ("d_workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_size_t), ("noneg", ctypes.c_int)]


class GpuIds(ctypes.Structure):
    """Mirror of the reference's struct GpuIds (GpuIds.hpp:3-14)."""

    _fields_ = [("m_piDeviceIds", ctypes.POINTER(ctypes.c_int)), ("m_iCount", ctypes.c_int)]


def reference_abi(path=None):
    """The reference-signature entry points of the library (or of the MATLAB-layout
    build), as ctypes functions keyed by their C++ names."""
    L = ctypes.CDLL(path or LIB_PATH)
    vp, ip = ctypes.c_void_p, ctypes.c_int
    geo, ids = TigreGeometry, ctypes.POINTER(GpuIds)
    sig = {
        "siddon_ray_projection": (ip, [vp, geo, vp, vp, ip, ids]),
        "interpolation_projection": (ip, [vp, geo, vp, vp, ip, ids]),
        "siddon_ray_projection_parallel": (ip, [vp, geo, vp, vp, ip, ids]),
        "interpolation_projection_parallel": (ip, [vp, geo, vp, vp, ip, ids]),
        "voxel_backprojection": (ip, [vp, geo, vp, vp, ip, ids]),
        "voxel_backprojection2": (ip, [vp, geo, vp, vp, ip, ids]),
        "voxel_backprojection_parallel": (ip, [vp, geo, vp, vp, ip, ids]),
        "pocs_tv": (None, [vp, vp, ctypes.c_float, vp, ip, ids]),
        "aw_pocs_tv": (None, [vp, vp, ctypes.c_float, vp, ip, ctypes.c_float, ids]),
        "piccs_tv": (None, [vp, vp, vp, ctypes.c_float, ctypes.c_float, vp, ip, ids]),
        "tvdenoising": (None, [vp, vp, ctypes.c_float, vp, vp, ip, ids]),
        "GetGpuCount": (ip, []),
        "GetGpuIdArray": (ip, [ctypes.c_char_p, vp, ip, vp]),
        "GetGpuName": (None, [ip, vp]),
        "GpuIds::GpuIds": (None, [ids]),
        "GpuIds::~GpuIds": (None, [ids]),
        "GpuIds::SetIds": (None, [ids, ip, vp]),
        "GpuIds::GetLength": (ip, [ids]),
        "GpuIds::SetAllGpus": (None, [ids, ip]),
        "GpuIds::operator[]": (ctypes.POINTER(ctypes.c_int), [ids, ip]),
        "GpuIds::operator[] const": (ip, [ids, ip]),
        "GpuIds::AreEqualDevices": (ctypes.c_bool, [ids]),
    }
    out = {}
    for name, mangled in REFERENCE_ABI.items():
        fn = getattr(L, mangled)
        fn.restype, fn.argtypes = sig[name]
        out[name] = fn
    return out


_lib = None


def lib():
    """Load (once) and return the library; raise ImportError when it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "tigre_amd: native library not built (%s). Build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C tigre_amd/csrc`." % LIB_PATH)
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error() -> str:
    msg = lib().tigre_last_error()
    return msg.decode() if msg else ""


def check(status: int, where: str) -> None:
    """Turn a library status code into the reference's exception types."""
    if status == 0:
        return
    from .errors import TigreCudaCallError

    msg = last_error()
    if status == 6:
        raise ValueError("%s %s" % (where, msg))
    raise TigreCudaCallError(where, status, msg)


class CGeometry:
    """A filled ``TigreGeometry`` plus the float32 arrays it points into.

    ``geo`` must already be expanded by ``check_geo`` and rounded by
    ``cast_to_single``.  Axis conventions follow convert_to_c_geometry
    (Python/tigre/utilities/cuda_interface/_types.pxd:69-166): Python arrays
    are (z, y, x) / (v, u), the C struct is (X, Y, Z) / (U, V), and
    rotDetector[i] = (yaw, pitch, roll)."""

    def __init__(self, geo, nangles: int):
        f32 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))  # noqa: E731
        off = np.asarray(geo.offOrigin, dtype=np.float32).reshape(nangles, 3)
        offd = np.asarray(geo.offDetector, dtype=np.float32).reshape(nangles, 2)
        rot = np.asarray(geo.rotDetector, dtype=np.float32).reshape(nangles, 3)
        self._keep = {
            "offOrigX": f32(off[:, 2]), "offOrigY": f32(off[:, 1]), "offOrigZ": f32(off[:, 0]),
            "DSO": f32(np.broadcast_to(np.asarray(geo.DSO, dtype=np.float32), (nangles,))),
            "DSD": f32(np.broadcast_to(np.asarray(geo.DSD, dtype=np.float32), (nangles,))),
            "offDetecU": f32(offd[:, 1]), "offDetecV": f32(offd[:, 0]),
            "dRoll": f32(rot[:, 2]), "dPitch": f32(rot[:, 1]), "dYaw": f32(rot[:, 0]),
            "COR": f32(np.broadcast_to(np.asarray(geo.COR, dtype=np.float32), (nangles,))),
        }
        s = TigreGeometry()
        nv = [int(x) for x in np.asarray(geo.nVoxel).reshape(3)]
        sv = np.asarray(geo.sVoxel, dtype=np.float32).reshape(3)
        dv = np.asarray(geo.dVoxel, dtype=np.float32).reshape(3)
        nd = [int(x) for x in np.asarray(geo.nDetector).reshape(2)]
        sd = np.asarray(geo.sDetector, dtype=np.float32).reshape(2)
        dd = np.asarray(geo.dDetector, dtype=np.float32).reshape(2)
        s.nVoxelX, s.nVoxelY, s.nVoxelZ = nv[2], nv[1], nv[0]
        s.sVoxelX, s.sVoxelY, s.sVoxelZ = float(sv[2]), float(sv[1]), float(sv[0])
        s.dVoxelX, s.dVoxelY, s.dVoxelZ = float(dv[2]), float(dv[1]), float(dv[0])
        s.nDetecU, s.nDetecV = nd[1], nd[0]
        s.sDetecU, s.sDetecV = float(sd[1]), float(sd[0])
        s.dDetecU, s.dDetecV = float(dd[1]), float(dd[0])
        s.unitX = s.unitY = s.unitZ = 1.0
        s.accuracy = float(np.float32(geo.accuracy))
        for name, arr in self._keep.items():
            setattr(s, name, arr.ctypes.data_as(_fp))
        self.struct = s

    def ptr(self):
        return ctypes.byref(self.struct)
