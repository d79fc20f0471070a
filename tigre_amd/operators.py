"""``Ax`` / ``Atb``: the cone-beam projector pair.

Same call signatures, checks and defaults as Python/tigre/utilities/Ax.py:29-59
and Atb.py:9-40 (via _Ax.pyx:32-101 and _Atb.pyx:31-90):

    Ax(img, geo, angles, projection_type="Siddon", gpuids=None)
    Atb(proj, geo, angles, backprojection_type="FDK", gpuids=None)

``img`` is (nZ, nY, nX) float32, projections are (nangles, nV, nU) float32.
NumPy inputs follow the reference contract (host in, host out).  torch
tensors that live on a GPU stay there: the operator runs on the current
stream and returns a device tensor, with no host round trip.  Device tensors
may also be stored in bfloat16 or float16: the result has the input's storage
type, and all arithmetic is float32.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .geometry import prepared

__all__ = ["Ax", "Atb"]


def _is_device_tensor(x):
    try:
        import torch
    except ImportError:  # pragma: no cover - torch is part of the image
        return False
    return isinstance(x, torch.Tensor) and x.is_cuda


def _gpu_ids(gpuids):
    if gpuids is None:
        return None, 0
    devs = list(getattr(gpuids, "devices", gpuids))
    if not devs:
        return None, 0
    arr = (ctypes.c_int * len(devs))(*[int(d) for d in devs])
    return arr, len(devs)


def _projection_kind(projection_type):
    if projection_type in ("Siddon", "ray-voxel"):
        return _native.PROJ_SIDDON
    if projection_type == "interpolated":
        return _native.PROJ_INTERPOLATED
    print("Warning: Unknown projection_type, using default Siddon")
    return _native.PROJ_SIDDON


def _backprojection_kind(backprojection_type):
    if backprojection_type == "FDK":
        return _native.BACKPROJ_FDK
    if backprojection_type == "matched":
        return _native.BACKPROJ_MATCHED
    print("Warning: Unknown backprojector, using default matched")
    return _native.BACKPROJ_MATCHED


def _is_parallel(geo):
    """Parallel or cone beam; an unknown mode warns and means cone (_Ax.pyx:67-73)."""
    if geo.mode == "parallel":
        return True
    if geo.mode not in ("cone", None):
        print("Warning: Unknown mode, using default cone beam")
    return False


def _storage_code(t):
    """Storage type code of a device tensor: float32, bfloat16 or float16."""
    import torch

    codes = {torch.float32: _native.DTYPE_F32, torch.bfloat16: _native.DTYPE_BF16, torch.float16: _native.DTYPE_F16}
    if t.dtype not in codes:
        raise TypeError("Input data should be float32 (or, on the GPU, bfloat16 / float16), not " + str(t.dtype))
    return codes[t.dtype]


def _stream_handle():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def Ax(img, geo, angles, projection_type="Siddon", **kwargs):
    """Forward projection (Ax.py:29-59)."""
    lib = _native.lib()
    device = _is_device_tensor(img)
    if device:
        import torch

        code = _storage_code(img)
        shape = tuple(img.shape)
    else:
        if img.dtype != np.float32:
            raise TypeError("Input data should be float32, not " + str(img.dtype))
        if not np.isreal(img).all():
            raise ValueError("Complex types not compatible for projection.")
        shape = img.shape
    if tuple(shape) != tuple(int(v) for v in np.asarray(geo.nVoxel)):
        raise ValueError("Input data should be of shape geo.nVoxel: " + str(geo.nVoxel) + " not:" + str(shape))
    g = prepared(geo, angles)
    parallel = _is_parallel(g)
    kind = _projection_kind(projection_type)
    n = g.angles.shape[0]
    cg = _native.CGeometry(g, n)
    ang = np.ascontiguousarray(g.angles, dtype=np.float32)
    nv, nu = int(g.nDetector[0]), int(g.nDetector[1])
    if device:
        import torch

        src = img.contiguous()
        out = torch.empty((n, nv, nu), dtype=img.dtype, device=img.device)
        fn = lib.tigre_ax_parallel_device if parallel else lib.tigre_ax_device_typed
        with torch.cuda.device(img.device):
            st = fn(ctypes.c_void_p(src.data_ptr()), cg.ptr(), ctypes.c_void_p(out.data_ptr()),
                    ang.ctypes.data_as(ctypes.c_void_p), n, kind, _native.LAYOUT_PYTHON, code, _stream_handle())
        _native.check(st, "TIGRE: Call to Ax failed")
        return out
    src = np.ascontiguousarray(img)
    out = np.empty((n, nv, nu), dtype=np.float32)
    P = nv * nu
    base = out.ctypes.data
    ptrs = (ctypes.POINTER(ctypes.c_float) * n)(
        *[ctypes.cast(base + 4 * P * i, ctypes.POINTER(ctypes.c_float)) for i in range(n)])
    ids, nids = _gpu_ids(kwargs.get("gpuids"))
    fn = lib.tigre_ax_parallel if parallel else lib.tigre_ax
    st = fn(src.ctypes.data_as(ctypes.c_void_p), cg.ptr(), ptrs, ang.ctypes.data_as(ctypes.c_void_p), n, kind,
            _native.LAYOUT_PYTHON, ids, nids)
    _native.check(st, "TIGRE: Call to Ax failed")
    return out


def Atb(projections, geo, angles, backprojection_type="FDK", **kwargs):
    """Backprojection (Atb.py:9-40)."""
    lib = _native.lib()
    device = _is_device_tensor(projections)
    if device:
        import torch

        code = _storage_code(projections)
    else:
        if projections.dtype != np.float32:
            raise TypeError("Input data should be float32, not " + str(projections.dtype))
        if not np.isreal(projections).all():
            raise ValueError("Complex types not compatible for back projection.")
    g = prepared(geo, angles)
    parallel = _is_parallel(g)
    expect = (g.angles.shape[0], int(g.nDetector[0]), int(g.nDetector[1]))
    if tuple(projections.shape) != expect:
        raise ValueError("Expected data shape: " + str(expect) + " not compatible with: "
                         + str(tuple(projections.shape)))
    kind = _backprojection_kind(backprojection_type)
    n = g.angles.shape[0]
    cg = _native.CGeometry(g, n)
    ang = np.ascontiguousarray(g.angles, dtype=np.float32)
    vshape = tuple(int(v) for v in np.asarray(g.nVoxel))
    if device:
        import torch

        src = projections.contiguous()
        out = torch.empty(vshape, dtype=projections.dtype, device=projections.device)
        with torch.cuda.device(projections.device):
            if parallel:
                st = lib.tigre_atb_parallel_device(ctypes.c_void_p(src.data_ptr()), cg.ptr(),
                                                   ctypes.c_void_p(out.data_ptr()),
                                                   ang.ctypes.data_as(ctypes.c_void_p), n, _native.LAYOUT_PYTHON,
                                                   code, _stream_handle())
            else:
                st = lib.tigre_atb_device_typed(ctypes.c_void_p(src.data_ptr()), cg.ptr(),
                                                ctypes.c_void_p(out.data_ptr()), ang.ctypes.data_as(ctypes.c_void_p),
                                                n, kind, _native.LAYOUT_PYTHON, code, _stream_handle())
        _native.check(st, "Atb:")
        return out
    src = np.ascontiguousarray(projections)
    out = np.empty(vshape, dtype=np.float32)
    ids, nids = _gpu_ids(kwargs.get("gpuids"))
    if parallel:
        st = lib.tigre_atb_parallel(src.ctypes.data_as(ctypes.c_void_p), cg.ptr(), out.ctypes.data_as(ctypes.c_void_p),
                                    ang.ctypes.data_as(ctypes.c_void_p), n, _native.LAYOUT_PYTHON, ids, nids)
    else:
        st = lib.tigre_atb(src.ctypes.data_as(ctypes.c_void_p), cg.ptr(), out.ctypes.data_as(ctypes.c_void_p),
                           ang.ctypes.data_as(ctypes.c_void_p), n, kind, _native.LAYOUT_PYTHON, ids, nids)
    _native.check(st, "Atb:")
    return out
