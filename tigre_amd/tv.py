"""TV steps: ``minTV`` / ``minimizeTV`` (POCS-TV steepest descent), its
adaptive-weighted and PICCS variants ``AwminTV`` / ``minPICCS``, and
``tvdenoise`` / ``im3ddenoise`` (ROF primal-dual denoise).

Signatures follow Python/tigre/utilities/cuda_interface/_minTV.pyx:39-64,
_tv_proximal.pyx:39-71, im_3d_denoise.py:7-28 and MATLAB/Utilities/minimizeTV.m.
NumPy in, NumPy out (host contract); GPU torch tensors stay on the device.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .operators import _gpu_ids, _is_device_tensor, _stream_handle

__all__ = ["minTV", "minimizeTV", "AwminTV", "minPICCS", "tvdenoise", "im3ddenoise"]


def _sizes(shape):
    # the native side takes {X, Y, Z}: the reversed (z, y, x) shape (_minTV.pyx:51-54)
    return (ctypes.c_long * 3)(int(shape[2]), int(shape[1]), int(shape[0]))


def minTV(src, alpha=15.0, maxiter=100, gpuids=None):
    """``maxiter`` steps of img -= alpha * g / ||g|| with g the TV sub-gradient."""
    lib = _native.lib()
    if _is_device_tensor(src):
        import torch

        if src.dtype != torch.float32 or src.dim() != 3:
            raise TypeError("minTV expects a 3D float32 tensor")
        x = src.contiguous()
        out = torch.empty_like(x)
        with torch.cuda.device(x.device):
            st = lib.tigre_pocs_tv_device(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                          float(alpha), _sizes(x.shape), int(maxiter), _stream_handle())
        _native.check(st, "minimizeTV:POCS_TV:")
        return out
    src = np.ascontiguousarray(src, dtype=np.float32)
    if src.ndim != 3:
        raise ValueError("minTV expects a 3D array")
    out = np.empty_like(src)
    ids, nids = _gpu_ids(gpuids)
    st = lib.tigre_pocs_tv(src.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), float(alpha),
                           _sizes(src.shape), int(maxiter), ids, nids)
    _native.check(st, "minimizeTV:POCS_TV:")
    return out


def minimizeTV(img, dtvg=1.0, ng=20, gpuids=None):
    """MATLAB-style name (MATLAB/Utilities/minimizeTV.m:23-39): ng steps of size dtvg."""
    return minTV(img, dtvg, ng, gpuids=gpuids)


def AwminTV(src, alpha=15.0, maxiter=100, delta=-0.005, gpuids=None):
    """Adaptive-weighted TV steps (_AwminTV.pyx:34-59, GD_AwTV.cu): as minTV with
    each difference d weighted by exp(-(d / delta)^2)."""
    lib = _native.lib()
    if _is_device_tensor(src):
        import torch

        if src.dtype != torch.float32 or src.dim() != 3:
            raise TypeError("AwminTV expects a 3D float32 tensor")
        x = src.contiguous()
        out = torch.empty_like(x)
        with torch.cuda.device(x.device):
            st = lib.tigre_aw_pocs_tv_device(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                             float(alpha), _sizes(x.shape), int(maxiter), float(delta),
                                             _stream_handle())
        _native.check(st, "TIGRE:Call to aw_pocs_tv failed")
        return out
    src = np.ascontiguousarray(src, dtype=np.float32)
    out = np.empty_like(src)
    ids, nids = _gpu_ids(gpuids)
    st = lib.tigre_aw_pocs_tv(src.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), float(alpha),
                              _sizes(src.shape), int(maxiter), float(delta), ids, nids)
    _native.check(st, "TIGRE:Call to aw_pocs_tv failed")
    return out


def minPICCS(src, prior, alpha=15.0, maxiter=100, ratio=0.95, gpuids=None):
    """PICCS steps (_minPICCS.pyx:34-66, PICCS.cu): descent along
    (1 - ratio) grad TV(img) + ratio grad TV(img - prior)."""
    lib = _native.lib()
    if _is_device_tensor(src):
        import torch

        x, p = src.contiguous(), prior.contiguous()
        if x.dtype != torch.float32 or p.dtype != torch.float32 or x.shape != p.shape or x.dim() != 3:
            raise TypeError("minPICCS expects two 3D float32 tensors of one shape")
        out = torch.empty_like(x)
        with torch.cuda.device(x.device):
            st = lib.tigre_piccs_tv_device(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(p.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()), float(alpha), float(ratio),
                                           _sizes(x.shape), int(maxiter), _stream_handle())
        _native.check(st, "minimizeTV:PICCS_TV:")
        return out
    src = np.ascontiguousarray(src, dtype=np.float32)
    prior = np.ascontiguousarray(prior, dtype=np.float32)
    if src.shape != prior.shape or src.ndim != 3:
        raise ValueError("minPICCS expects two 3D arrays of one shape")
    out = np.empty_like(src)
    ids, nids = _gpu_ids(gpuids)
    st = lib.tigre_piccs_tv(src.ctypes.data_as(ctypes.c_void_p), prior.ctypes.data_as(ctypes.c_void_p),
                            out.ctypes.data_as(ctypes.c_void_p), float(alpha), float(ratio), _sizes(src.shape),
                            int(maxiter), ids, nids)
    _native.check(st, "minimizeTV:PICCS_TV:")
    return out


def tvdenoise(src, maxiter=100, lamda=15.0, gpuids=None, spacing=(1.0, 1.0, 1.0)):
    """ROF TV denoise (_tv_proximal.pyx:39-71); spacing is (dx, dy, dz), 1 in the reference."""
    lib = _native.lib()
    sp = (ctypes.c_float * 3)(*[float(s) for s in spacing])
    if _is_device_tensor(src):
        import torch

        x = src.contiguous()
        out = torch.empty_like(x)
        with torch.cuda.device(x.device):
            st = lib.tigre_tvdenoising_device(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                              float(lamda), sp, _sizes(x.shape), int(maxiter), _stream_handle())
        _native.check(st, "tvdenoising:")
        return out
    src = np.ascontiguousarray(src, dtype=np.float32)
    out = np.empty_like(src)
    ids, nids = _gpu_ids(gpuids)
    st = lib.tigre_tvdenoising(src.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p),
                               float(lamda), sp, _sizes(src.shape), int(maxiter), ids, nids)
    _native.check(st, "tvdenoising:")
    return out


def im3ddenoise(img, iter=50, lmbda=15.0, gpuids=None):
    """Normalise to [0, 1], TV-denoise, undo the normalisation (im_3d_denoise.py:7-28)."""
    if img.ndim != 3:
        raise ValueError("Expected a 3D array (nZ, nY, nX), got shape %s." % (img.shape,))
    imgmin = np.amin(img.ravel())
    img = img - imgmin
    imgmax = np.amax(img.ravel())
    img = img / imgmax
    img = tvdenoise(img.astype(np.float32), iter, lmbda, gpuids)
    img *= imgmax
    img += imgmin
    return img
