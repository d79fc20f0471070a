"""Reconstruction algorithms on device-resident data.

Same entry points and defaults as Python/tigre/algorithms (__init__.py:5-60):
``fdk``/``FDK``, ``fbp``, ``ossart``, ``iterativereconalg``, ``sart``, ``sirt``, ``sart_tv``, ``ossart_tv``,
``fista``, ``ista``, ``cgls``, ``lsqr``, ``hybrid_lsqr``, ``lsmr``, ``irn_tv_cgls``, ``hybrid_flsqr_tv``,
``ab_gmres``, ``ba_gmres``, ``mlem``, ``asd_pocs``, ``os_asd_pocs``,
``awasd_pocs``, ``os_awasd_pocs``, ``piccs``, ``os_piccs``, ``pcsd``,
``aw_pcsd``, ``os_pcsd``, ``os_aw_pcsd``.
The volume, the projections and every intermediate stay in HBM as torch
tensors; each projector call goes straight to the gfx950 entry points with
per-subset geometry records prepared once (no per-call host round trip,
no re-validation of the geometry).  Norms and inner products are FP64.

Inputs may be NumPy arrays (copied to the GPU once) or GPU tensors; results
are NumPy arrays unless ``return_tensor=True``.
"""
from __future__ import annotations

import copy
import ctypes
import time

import numpy as np

from .. import _native
from ..geometry import prepared
from ..operators import _stream_handle
from ..utilities.filtering import filtering
from ..utilities.order_subsets import order_subsets

__all__ = ["fdk", "FDK", "fbp", "ossart", "iterativereconalg", "sart", "sirt", "sart_tv", "ossart_tv", "fista", "ista", "cgls", "lsqr",
           "hybrid_lsqr", "lsmr", "irn_tv_cgls",
           "hybrid_flsqr_tv", "ab_gmres", "ba_gmres", "mlem", "asd_pocs", "os_asd_pocs", "awasd_pocs",
           "os_awasd_pocs", "piccs", "os_piccs", "pcsd", "aw_pcsd", "os_pcsd", "os_aw_pcsd", "ViewSets",
           "BatchedOSSART"]


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("tigre_amd.algorithms run on the GPU; no HIP device is visible")
    return torch


def _sub_geometry(geo, idx):
    """Geometry restricted to views idx (iterative_recon_alg.py:335-341)."""
    g = copy.deepcopy(geo)
    for a in ("DSD", "DSO", "COR"):
        setattr(g, a, np.asarray(getattr(geo, a))[idx])
    for a in ("offOrigin", "offDetector", "rotDetector"):
        setattr(g, a, np.asarray(getattr(geo, a))[idx])
    return g


class ViewSets:
    """Projector calls on fixed subsets of views of one geometry.

    The per-subset geometry (checked, rounded to float32, converted to the C
    struct) is prepared once; ``ax``/``atb`` then call the device entry points
    directly on the current stream."""

    def __init__(self, geo, angles, blocks):
        self.geo = copy.deepcopy(geo)
        self.geo.check_geo(angles)
        self.angles = self.geo.angles
        # the reference's Ax / Atb pick the parallel-beam operators from
        # geo.mode (_Ax.pyx:67-87, _Atb.pyx:56-76)
        self.parallel = getattr(self.geo, "mode", "cone") == "parallel"
        self.sets = []
        for idx in blocks:
            idx = np.asarray(idx, dtype=np.int64).reshape(-1)
            g = prepared(_sub_geometry(self.geo, idx), self.angles[idx])
            cg = _native.CGeometry(g, len(idx))
            ang = np.ascontiguousarray(g.angles, dtype=np.float32)
            self.sets.append((idx, g, cg, ang))
        self.lib = _native.lib()
        self.vol_shape = tuple(int(v) for v in np.asarray(self.geo.nVoxel))
        self.det_shape = (int(self.geo.nDetector[0]), int(self.geo.nDetector[1]))

    def __len__(self):
        return len(self.sets)

    def index(self, j):
        return self.sets[j][0]

    def ax(self, img, j, kind=_native.PROJ_SIDDON, out=None):
        """Projections of subset j, stored like ``img`` (float32, bfloat16 or float16)."""
        torch = _torch()
        idx, g, cg, ang = self.sets[j]
        if out is None:
            out = torch.empty((len(idx),) + self.det_shape, dtype=img.dtype, device=img.device)
        fn = self.lib.tigre_ax_parallel_device if self.parallel else self.lib.tigre_ax_device_typed
        st = fn(ctypes.c_void_p(img.data_ptr()), cg.ptr(), ctypes.c_void_p(out.data_ptr()),
                ang.ctypes.data_as(ctypes.c_void_p), len(idx), kind, _native.LAYOUT_PYTHON, _storage_code(img),
                _stream_handle())
        _native.check(st, "TIGRE: Call to Ax failed")
        return out

    def atb(self, proj, j, kind=_native.BACKPROJ_FDK, out=None):
        """Backprojection of subset j, stored like ``proj``."""
        torch = _torch()
        idx, g, cg, ang = self.sets[j]
        if out is None:
            out = torch.empty(self.vol_shape, dtype=proj.dtype, device=proj.device)
        if self.parallel:
            # one parallel backprojector serves 'FDK' and 'matched' (_Atb.pyx:56-76)
            st = self.lib.tigre_atb_parallel_device(ctypes.c_void_p(proj.data_ptr()), cg.ptr(),
                                                    ctypes.c_void_p(out.data_ptr()),
                                                    ang.ctypes.data_as(ctypes.c_void_p), len(idx),
                                                    _native.LAYOUT_PYTHON, _storage_code(proj), _stream_handle())
        else:
            st = self.lib.tigre_atb_device_typed(ctypes.c_void_p(proj.data_ptr()), cg.ptr(),
                                                 ctypes.c_void_p(out.data_ptr()), ang.ctypes.data_as(ctypes.c_void_p),
                                                 len(idx), kind, _native.LAYOUT_PYTHON, _storage_code(proj),
                                                 _stream_handle())
        _native.check(st, "Atb:")
        return out


def _storage_code(t):
    torch = _torch()
    codes = {torch.float32: _native.DTYPE_F32, torch.bfloat16: _native.DTYPE_BF16, torch.float16: _native.DTYPE_F16}
    if t.dtype not in codes:
        raise TypeError("volumes and projections are stored as float32, bfloat16 or float16, not " + str(t.dtype))
    return codes[t.dtype]


def _to_device(x):
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return x.to(device="cuda", dtype=torch.float32).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float32), device="cuda")


def _result(x, return_tensor):
    return x if return_tensor else x.float().cpu().numpy()


def _norm(x):
    torch = _torch()
    return float(torch.linalg.vector_norm(x, dtype=torch.float64))


# ---------------------------------------------------------------------------
# FDK
# ---------------------------------------------------------------------------
def _wang_padding(p, geo):
    """Zero padding for an offset detector (single_pass_algorithms.py:78-111):
    2 |offset| / dU columns of zeros on the side facing away from the offset,
    detector size and offset updated to match.  Returns (padded, geometry,
    theta), theta = (sU / 2 - |offset|) sign(offset) = the overlap half-width."""
    torch = _torch()
    zgeo = copy.deepcopy(geo)
    zgeo.offDetector = np.array(zgeo.offDetector, dtype=np.float64)
    zgeo.nDetector = np.array(zgeo.nDetector)
    zgeo.sDetector = np.array(zgeo.sDetector, dtype=np.float64)
    # the offset the reference reads: [1] of a per-scan pair, [1, 0] of a per-view table
    off = float(zgeo.offDetector[1, 0] if zgeo.offDetector.ndim == 2 else zgeo.offDetector[1])
    dU = float(geo.dDetector[1])
    pad = int(2 * off / dU)
    zgeo.offDetector[1] = off - pad / 2 * dU
    zgeo.nDetector[1] = abs(pad) + int(geo.nDetector[1])
    zgeo.sDetector[1] = zgeo.nDetector[1] * dU
    theta = (float(geo.sDetector[1]) / 2 - abs(off)) * np.sign(off)
    zeros = torch.zeros(p.shape[0], p.shape[1], abs(pad), dtype=p.dtype, device=p.device)
    padded = torch.cat((zeros, p), dim=2) if off > 0 else torch.cat((p, zeros), dim=2)
    return padded.contiguous(), zgeo, theta


def wang_weights(geo, theta):
    """Detector-offset redundancy weights over the u axis (single_pass_algorithms.py:113-163):
    positions t = (u_k + |offset + (DSD / DSO) COR|) DSO / DSD on the central plane;
    w = (sin(pi / 2 atan(t / DSO) / atan(|theta| DSO / DSD / DSO)) + 1) / 2 inside the
    overlap |t| <= |theta| DSO / DSD, 0 beyond it on the short side, 1 on the long side;
    mirrored when theta < 0.  The projections are then scaled by 2 w."""
    off = np.asarray(geo.offDetector, dtype=np.float64)
    offset = float(off[1] if off.ndim == 1 else off[1, 0])
    DSO = float(np.asarray(geo.DSO).reshape(-1)[0])
    DSD = float(np.asarray(geo.DSD).reshape(-1)[0])
    COR = float(np.asarray(getattr(geo, "COR", 0.0)).reshape(-1)[0])
    offset = offset + (DSD / DSO) * COR
    nU = int(geo.nDetector[1])
    us = (np.arange(-nU / 2 + 0.5, nU / 2 - 0.5 + 1) * float(geo.dDetector[1]) + abs(offset)) * DSO / DSD
    abstheta = abs(theta * DSO / DSD)
    w = np.ones(nU)
    inside = np.abs(us) <= abstheta
    w[inside] = 0.5 * (np.sin((np.pi / 2) * np.arctan(us[inside] / DSO) / np.arctan(abstheta / DSO)) + 1)
    w[us < -abstheta] = 0
    return w[::-1].copy() if theta < 0 else w


def fdk(proj, geo, angles, filter=None, verbose=False, return_tensor=False, **kwargs):
    """Feldkamp-Davis-Kress, single_pass_algorithms.py:12-202: detector-offset
    (Wang) weights when the detector is offset, cosine weight, ramp
    filtering, FDK-weighted backprojection."""
    torch = _torch()
    p = _to_device(proj)
    dowang = kwargs.get("dowang", True) and bool(np.any(np.asarray(getattr(geo, "offDetector", 0))))
    if dowang:
        if verbose:
            print("FDK: applying detector offset weights")
        p, geo, theta = _wang_padding(p, geo)
        w = torch.as_tensor((2.0 * wang_weights(geo, theta)).astype(np.float32), device=p.device)
        p = p * w
    g = copy.deepcopy(geo)
    g.check_geo(angles)
    g.filter = filter
    nV, nU = int(g.nDetector[0]), int(g.nDetector[1])
    dV, dU = float(g.dDetector[0]), float(g.dDetector[1])
    DSD0 = float(np.asarray(g.DSD).reshape(-1)[0])
    xv = (np.arange(nU) - nU / 2 + 0.5) * dU
    yv = (np.arange(nV) - nV / 2 + 0.5) * dV
    xx, yy = np.meshgrid(xv, yv)
    w = torch.as_tensor((DSD0 / np.sqrt(DSD0 ** 2 + xx ** 2 + yy ** 2)).astype(np.float32), device=p.device)
    filt = filtering(p * w, g, g.angles, parker=False, verbose=verbose)
    vs = ViewSets(g, g.angles, [np.arange(len(g.angles))])
    return _result(vs.atb(filt, 0, _native.BACKPROJ_FDK), return_tensor)


FDK = fdk


def fbp(proj, geo, angles, filter=None, verbose=False, return_tensor=False, **kwargs):
    """Filtered backprojection for a parallel beam (single_pass_algorithms.py
    :208-224): the projections ramp-filtered as in FDK (no cosine weight),
    backprojected, scaled by DSO / DSD of the first view in float32.  As in the
    reference, the filter applied is the one ``geo`` holds on entry (ram-lak
    when none) and ``filter`` is stored on ``geo`` for later calls."""
    if getattr(geo, "mode", "cone") != "parallel":
        raise ValueError("Only use FBP for parallel beam. Check geo.mode.")
    g = copy.deepcopy(geo)
    g.check_geo(angles)
    geo.filter = filter
    p = _to_device(proj)
    filt = filtering(p, g, g.angles, parker=False, verbose=verbose)
    weight = float(np.float32(float(np.asarray(g.DSO).reshape(-1)[0]) / float(np.asarray(g.DSD).reshape(-1)[0])))
    vs = ViewSets(g, g.angles, [np.arange(len(g.angles))])
    out = vs.atb(filt, 0, _native.BACKPROJ_MATCHED)
    out *= weight
    return _result(out, return_tensor)


# ---------------------------------------------------------------------------
# SART family (iterative_recon_alg.py:131-401, art_family_algorithms.py)
# ---------------------------------------------------------------------------
class _SartState:
    def __init__(self, proj, geo, angles, blocksize, order=None, init=None, lmbda=1.0, noneg=True, verbose=False,
                 use_plans=True):
        torch = _torch()
        self.geo = copy.deepcopy(geo)
        self.geo.check_geo(angles)
        self.angles = self.geo.angles
        _, index = order_subsets(self.angles, blocksize, order)
        self.blocks = [np.atleast_1d(np.asarray(i)) for i in index]
        self.views = ViewSets(self.geo, self.angles, self.blocks)
        self.proj = _to_device(proj)
        self.lmbda = lmbda
        self.noneg = noneg
        self.verbose = verbose
        self.W = self._weights_w()
        self.V = self._weights_v()
        shape = self.views.vol_shape
        if init is None:
            self.res = torch.zeros(shape, dtype=torch.float32, device="cuda")
        elif isinstance(init, str) and init == "FDK":
            self.res = torch.clamp(fdk(self.proj, self.geo, self.angles, return_tensor=True), min=0)
        else:
            self.res = _to_device(init).reshape(shape).clone()
        self.proj_blocks = [self.proj[torch.as_tensor(i, device="cuda")].contiguous() for i in self.blocks]
        self.W_blocks = [self.W[torch.as_tensor(i, device="cuda")].contiguous() for i in self.blocks]
        # Cone beam: one projection plan per subset (records on the device once)
        # and the fused residual / update kernels, which keep the rounding
        # steps of the tensor expressions of the plain loop below (same bits)
        # and make a subset a handful of launches with no host set-up.
        self.plans = None
        if use_plans and getattr(self.geo, "mode", "cone") == "cone" and self.res.dtype == torch.float32:
            from ..plan import Plan

            self.plans = [Plan(_sub_geometry(self.geo, idx), self.angles[idx]) for idx in self.blocks]
            self.ws = torch.empty(max(max(p.workspace_bytes() for p in self.plans), 256), dtype=torch.uint8,
                                  device="cuda")
            nV, nU = self.plans[0].proj_shape[1:]
            self.axbuf = torch.empty((max(len(b) for b in self.blocks), nV, nU), dtype=torch.float32,
                                     device="cuda")
            self._inv_v = (None, None)  # (lambda, [lambda / V_j])

    def _weights_w(self):
        """W = 1 / Ax(ones(2^3)) on an enlarged volume (iterative_recon_alg.py:214-230)."""
        torch = _torch()
        g = copy.deepcopy(self.geo)
        g.sVoxel = np.asarray(g.sVoxel, dtype=np.float64).copy()
        g.sVoxel[1:] = g.sVoxel[1:] * 1.1
        g.sVoxel[0] = max(g.sDetector[0], g.sVoxel[0])
        g.nVoxel = np.array([2, 2, 2])
        g.dVoxel = g.sVoxel / g.nVoxel
        vs = ViewSets(g, self.angles, [np.arange(len(self.angles))])
        W = vs.ax(torch.ones((2, 2, 2), dtype=torch.float32, device="cuda"), 0)
        W = torch.where(W <= float(min(np.asarray(self.geo.dVoxel) / 2)), torch.full_like(W, float("inf")), W)
        return 1.0 / W

    def _weights_v(self):
        """V[block] = mean over z of Atb(ones) on a shrunk volume (iterative_recon_alg.py:232-270);
        for a parallel beam V[block] = the number of views in the block (:265-267)."""
        torch = _torch()
        if self.views.parallel:
            nY, nX = self.views.vol_shape[1:]
            return torch.stack([torch.full((nY, nX), float(len(idx)), dtype=torch.float32, device="cuda")
                                for idx in self.blocks])
        g = copy.deepcopy(self.geo)
        s = np.asarray(g.sVoxel, dtype=np.float64)
        g.sVoxel = s * np.max(s[1:] / np.linalg.norm(s[1:])) * 0.9
        g.dVoxel = g.sVoxel / np.asarray(g.nVoxel)
        vs = ViewSets(g, self.angles, self.blocks)
        out = []
        nV, nU = vs.det_shape
        for j, idx in enumerate(self.blocks):
            ones = torch.ones((len(idx), nV, nU), dtype=torch.float32, device="cuda")
            out.append(vs.atb(ones, j).mean(dim=0))
        V = torch.stack(out)
        return torch.where(V == 0, torch.full_like(V, float("inf")), V)

    def sweep(self):
        """One pass over all subsets (art_data_minimizing, iterative_recon_alg.py:326-345)."""
        torch = _torch()
        if self.plans is not None:
            from ..plan import sart_residual

            if self._inv_v[0] != self.lmbda:
                self._inv_v = (self.lmbda, [(self.lmbda / self.V[j]).contiguous() for j in range(len(self.blocks))])
            for j, p in enumerate(self.plans):
                ax = self.axbuf[: p.nviews]
                p.ax(self.res, ax, self.ws, _native.PROJ_SIDDON)
                sart_residual(self.proj_blocks[j], self.W_blocks[j], ax)
                p.sart_update(ax, self.res, self._inv_v[1][j], self.ws, self.noneg, _native.BACKPROJ_FDK)
            return
        for j in range(len(self.blocks)):
            r = self.W_blocks[j] * (self.proj_blocks[j] - self.views.ax(self.res, j))
            self.res += self.lmbda / self.V[j] * self.views.atb(r, j)
            if self.noneg:
                torch.clamp_(self.res, min=0)

    def residual_norm(self):
        full = ViewSets(self.geo, self.angles, [np.arange(len(self.angles))])
        return _norm(self.proj - full.ax(self.res, 0))


def _sart_family(proj, geo, angles, niter, blocksize, **kwargs):
    st = _SartState(proj, geo, angles, blocksize, order=kwargs.get("OrderStrategy"), init=kwargs.get("init"),
                    lmbda=kwargs.get("lmbda", 1.0), noneg=kwargs.get("noneg", True),
                    verbose=kwargs.get("verbose", False))
    l2 = []
    # lmbda_red is accepted and, as in the reference, not applied: its SART
    # family stores it (iterative_recon_alg.py:143) but every update uses the
    # unreduced lmbda (:389-399)
    for _ in range(niter):
        st.sweep()
        if kwargs.get("computel2", False):
            l2.append(st.residual_norm())
    res = _result(st.res, kwargs.get("return_tensor", False))
    if kwargs.get("computel2", False):
        return res, np.array([l2])
    return res


def ossart(proj, geo, angles, niter, blocksize=20, **kwargs):
    """OS-SART (art_family_algorithms.py:47-62), default 20 views per subset."""
    return _sart_family(proj, geo, angles, niter, blocksize, **kwargs)


def iterativereconalg(proj, geo, angles, niter, **kwargs):
    """The base class's own entry point (iterative_recon_alg.py:487): its
    defaults (blocksize 20, lmbda 1, no reduction, noneg, the ART data step,
    iterative_recon_alg.py:140-156) make its loop the OS-SART one."""
    return ossart(proj, geo, angles, niter, **kwargs)


def sart(proj, geo, angles, niter, **kwargs):
    """SART: one view per subset."""
    kwargs.pop("blocksize", None)
    return _sart_family(proj, geo, angles, niter, 1, **kwargs)


def sirt(proj, geo, angles, niter, **kwargs):
    """SIRT: all views in one subset."""
    kwargs.pop("blocksize", None)
    return _sart_family(proj, geo, angles, niter, len(np.atleast_1d(angles)), **kwargs)


# ---------------------------------------------------------------------------
# The ROF denoiser's callers: SART-TV / OS-SART-TV (art_family_algorithms.py
# :65-153) and FISTA / ISTA (ista_algorithms.py:13-212)
# ---------------------------------------------------------------------------
def _im3ddenoise(img, tviter, lmbda):
    """im3ddenoise (im_3d_denoise.py:7-28) on a device volume: shift to minimum
    0, scale to maximum 1, ROF-denoise, undo; the extrema stay on the device."""
    from ..tv import tvdenoise

    torch = _torch()
    lo = torch.min(img)
    x = img - lo
    hi = torch.max(x)
    out = tvdenoise(x / hi, tviter, lmbda)
    out *= hi
    out += lo
    return out


def _sart_tv(proj, geo, angles, niter, blocksize, tvlambda, tviter, **kwargs):
    st = _SartState(proj, geo, angles, blocksize, order=kwargs.get("OrderStrategy"), init=kwargs.get("init"),
                    lmbda=kwargs.get("lmbda", 1.0), noneg=kwargs.get("noneg", True),
                    verbose=kwargs.get("verbose", False))
    for _ in range(niter):
        st.sweep()
        st.res = _im3ddenoise(st.res, tviter, tvlambda)
    return _result(st.res, kwargs.get("return_tensor", False))


def sart_tv(proj, geo, angles, niter, tvlambda=50, tviter=50, **kwargs):
    """SART with TV regularisation (art_family_algorithms.py:65-108): every
    iteration one SART sweep (one view per subset; a larger blocksize is
    ignored with the reference's warning), then im3ddenoise(res, tviter,
    tvlambda)."""
    if kwargs.pop("blocksize", 1) > 1:
        print("Warning: blocksize is set to 1, please use an OS version of the algorithm for blocksize > 1")
    return _sart_tv(proj, geo, angles, niter, 1, tvlambda, tviter, **kwargs)


def ossart_tv(proj, geo, angles, niter, blocksize=20, tvlambda=50, tviter=50, **kwargs):
    """OS-SART with TV regularisation (art_family_algorithms.py:111-153):
    every iteration one sweep over subsets of ``blocksize`` views, then
    im3ddenoise(res, tviter, tvlambda)."""
    return _sart_tv(proj, geo, angles, niter, blocksize, tvlambda, tviter, **kwargs)


def _ista_family(proj, geo, angles, niter, momentum, hyper=2.0e8, tviter=20, tvlambda=0.1, fista_p=1.0,
                 fista_q=1.0, init=None, noneg=True, verbose=False, return_tensor=False, **kwargs):
    torch = _torch()
    g = copy.deepcopy(geo)
    g.check_geo(angles)
    vs = ViewSets(g, g.angles, [np.arange(len(g.angles))])
    b = _to_device(proj)
    if init is None:
        res = torch.zeros(vs.vol_shape, dtype=torch.float32, device="cuda")
    elif isinstance(init, str) and init == "FDK":
        res = fdk(b, g, g.angles, return_tensor=True)
    else:
        res = _to_device(init).reshape(vs.vol_shape).clone()
    step = 2.0 / float(hyper)  # bm * 2, bm = 1 / hyper (ista_algorithms.py:119-122, :128-153)
    lam = 1.0 / (2.0 / float(hyper) * (float(tvlambda) if momentum else 0.1))
    x_rec = res.clone()
    t = 1.0
    for i in range(niter):
        if verbose:
            print("%s iteration %d of %d" % ("FISTA" if momentum else "ISTA", i + 1, niter))
        r = b - vs.ax(res, 0, _native.PROJ_INTERPOLATED)
        res = res + step * vs.atb(r, 0, _native.BACKPROJ_MATCHED)
        if noneg:
            res.clamp_(min=0)
        if not momentum:
            res = _im3ddenoise(res, 20, lam)
            continue
        x_old = x_rec
        x_rec = _im3ddenoise(res, tviter, lam)
        t_old = t
        t = (fista_p + float(np.sqrt(fista_q + 4 * t ** 2))) / 2
        res = x_rec + (t_old - 1) / t * (x_rec - x_old)
    return _result(res, return_tensor)


def fista(proj, geo, angles, niter, **kwargs):
    """FISTA (ista_algorithms.py:13-182): one block of all views; per
    iteration res += (2 / hyper) Atb_matched(proj - Ax_interpolated(res)),
    clipped at 0, then x = im3ddenoise(res, tviter, 1 / (2 tvlambda / hyper))
    and the extrapolation res = x + (t - 1) / t' (x - x_previous) with
    t' = (fista_p + sqrt(fista_q + 4 t^2)) / 2.  Keywords: hyper (2e8),
    tviter (20), tvlambda (0.1), fista_p (1), fista_q (1), init (None / 'FDK'
    / image), noneg (True), verbose, return_tensor."""
    return _ista_family(proj, geo, angles, niter, True, **kwargs)


def lsqr(proj, geo, angles, niter, restart=True, verbose=False, return_tensor=False, computel2=False, **kwargs):
    """LSQR (krylov_subspace_algorithms.py:101-188), Paige and Saunders'
    bidiagonalisation with Siddon Ax and matched Atb over all views, FP64
    norms.  The data residual is taken after every iteration; when it rises
    the step is undone and the iteration ends, as in the reference (which
    re-initialises and then leaves its loop).  With ``computel2`` the
    residual norms are returned as well."""
    torch = _torch()
    g = copy.deepcopy(geo)
    g.check_geo(angles)
    vs = ViewSets(g, g.angles, [np.arange(len(g.angles))])
    b = _to_device(proj)
    A = lambda x: vs.ax(x, 0, _native.PROJ_SIDDON)  # noqa: E731
    At = lambda y: vs.atb(y, 0, _native.BACKPROJ_MATCHED)  # noqa: E731
    res = torch.zeros(vs.vol_shape, dtype=torch.float32, device="cuda")
    u = b - A(res)
    beta = _norm(u)
    u = u / beta
    phibar = beta
    v = At(u)
    alpha = _norm(v)
    v = v / alpha
    rhobar = alpha
    w = v.clone()
    l2 = []
    for i in range(niter):
        if verbose:
            print("LSQR iteration %d of %d" % (i + 1, niter))
        u = torch.sub(A(v), u, alpha=alpha)
        beta = _norm(u)
        u = u / beta
        v = torch.sub(At(u), v, alpha=beta)
        alpha = _norm(v)
        v = v / alpha
        rho = float(np.sqrt(rhobar ** 2 + beta ** 2))
        c, s = rhobar / rho, beta / rho
        theta = s * alpha
        rhobar = -c * alpha
        phi = c * phibar
        phibar = s * phibar
        res = torch.add(res, w, alpha=phi / rho)
        w_prev = w
        w = torch.sub(v, w, alpha=theta / rho)
        l2.append(_norm(b - A(res)))
        if i > 0 and l2[-1] > l2[-2]:
            res = torch.sub(res, w_prev, alpha=phi / rho)
            if verbose or not restart:
                print("LSQR exited due to divergence.")
            break
    out = _result(res, return_tensor)
    return (out, np.array([l2])) if computel2 else out


def hybrid_lsqr(proj, geo, angles, niter, lmbda=1.0, restart=True, verbose=False, return_tensor=False,
                computel2=False, **kwargs):
    """Hybrid LSQR (krylov_subspace_algorithms.py:190-297): Golub-Kahan
    bidiagonalisation with Siddon Ax and matched Atb, the projected Tikhonov
    problem min |B_k y - beta e1|^2 + lmbda^2 |y|^2 solved through the SVD of
    B_k in FP64 on the host, x = V_k y.  As in the reference, each new vector
    v (and u) has (V_j * v) * V_j subtracted elementwise for the vectors
    j < i - 1 (its reorthogonalisation step as it computes).  The data residual
    of each iterate is taken; a rise at the second iteration (or at any, with
    restart=False) ends the iteration there."""
    torch = _torch()
    g = copy.deepcopy(geo)
    g.check_geo(angles)
    vs = ViewSets(g, g.angles, [np.arange(len(g.angles))])
    b = _to_device(proj)
    A = lambda x: vs.ax(x, 0, _native.PROJ_SIDDON)  # noqa: E731
    At = lambda y: vs.atb(y, 0, _native.BACKPROJ_MATCHED)  # noqa: E731
    u = b - A(torch.zeros(vs.vol_shape, dtype=torch.float32, device="cuda"))
    beta = _norm(u)
    u = u / beta
    U, V = [u], []
    Bm = np.zeros((niter, niter + 1))
    rhs = np.zeros(niter + 1)
    rhs[0] = beta
    l2 = []
    x = None
    for i in range(niter):
        if verbose:
            print("hybrid LSQR iteration %d of %d" % (i + 1, niter))
        v = At(u)
        if i > 0:
            v = torch.sub(v, V[i - 1], alpha=beta)
        for j in range(i - 1):
            v = v - (V[j] * v) * V[j]
        alpha = _norm(v)
        v = v / alpha
        V.append(v)
        u = torch.sub(A(v), u, alpha=alpha)
        for j in range(i - 1):
            u = u - (U[j] * u) * U[j]
        beta = _norm(u)
        u = u / beta
        U.append(u)
        Bm[i, i] = alpha
        Bm[i, i + 1] = beta
        Uk, Sk, Vk = np.linalg.svd(Bm[:i + 1, :i + 2].T)
        y = Vk.T @ (Sk * (Uk.T @ rhs[:i + 2])[:i + 1] / (Sk ** 2 + lmbda ** 2))
        x = torch.zeros(vs.vol_shape, dtype=torch.float32, device="cuda")
        for yj, vj in zip(y, V):
            x = torch.add(x, vj, alpha=float(yj))
        l2.append(_norm(b - A(x)))
        if i > 0 and l2[-1] > l2[-2] and (i == 1 or not restart):
            print("hybrid LSQR exited due to divergence at iteration %d" % i)
            # returned from the loop, never assigned to the image that the
            # reference's entry point returns: the zero start
            # (krylov_subspace_algorithms.py:285-289; iterative_recon_alg.py:470-475)
            x = torch.zeros_like(x)
            break
    out = _result(x, return_tensor)
    return (out, np.array([l2])) if computel2 else out


def lsmr(proj, geo, angles, niter, lmbda=1.0, restart=True, verbose=False, return_tensor=False, computel2=False,
         **kwargs):
    """LSMR (krylov_subspace_algorithms.py:299-439), Fong and Saunders'
    method with damping ``lmbda`` (1 by default, as the reference passes its
    lmbda), Siddon Ax and matched Atb over all views, FP64 scalars and norms.
    The data residual is taken after every iteration; a rise undoes the step
    and ends the iteration, as in LSQR."""
    torch = _torch()
    g = copy.deepcopy(geo)
    g.check_geo(angles)
    vs = ViewSets(g, g.angles, [np.arange(len(g.angles))])
    b = _to_device(proj)
    A = lambda x: vs.ax(x, 0, _native.PROJ_SIDDON)  # noqa: E731
    At = lambda y: vs.atb(y, 0, _native.BACKPROJ_MATCHED)  # noqa: E731
    res = torch.zeros(vs.vol_shape, dtype=torch.float32, device="cuda")
    u = b - A(res)
    beta = _norm(u)
    u = u / beta
    v = At(u)
    alpha = _norm(v)
    v = v / alpha
    alphabar, zetabar = alpha, alpha * beta
    rho, rhobar, cbar, sbar = 1.0, 1.0, 1.0, 0.0
    h, hbar = v.clone(), torch.zeros_like(v)
    l2 = []
    for i in range(niter):
        if verbose:
            print("LSMR iteration %d of %d" % (i + 1, niter))
        u = torch.sub(A(v), u, alpha=alpha)
        beta = _norm(u)
        u = u / beta
        v = torch.sub(At(u), v, alpha=beta)
        alpha = _norm(v)
        v = v / alpha
        alphahat = float(np.sqrt(alphabar ** 2 + lmbda ** 2))
        rhopre = rho
        rho = float(np.sqrt(alphahat ** 2 + beta ** 2))
        c, s = alphahat / rho, beta / rho
        theta = s * alpha
        alphabar = c * alpha
        thetabar = sbar * rho
        rhobarpre = rhobar
        rhobar = float(np.sqrt((cbar * rho) ** 2 + theta ** 2))
        cbar = cbar * rho / rhobar
        sbar = theta / rhobar
        zeta = cbar * zetabar
        zetabar = -sbar * zetabar
        hbar = torch.sub(h, hbar, alpha=thetabar * rho / (rhopre * rhobarpre))
        step = zeta / (rho * rhobar)
        res = torch.add(res, hbar, alpha=step)
        h = torch.sub(v, h, alpha=theta / rho)
        l2.append(_norm(b - A(res)))
        if i > 0 and l2[-1] > l2[-2]:
            res = torch.sub(res, hbar, alpha=step)
            if verbose or not restart:
                print("LSMR exited due to divergence.")
            break
    out = _result(res, return_tensor)
    return (out, np.array([l2])) if computel2 else out


def _irn_diffs(x):
    """Per-axis differences of krylov_subspace_algorithms.py:456-476, stacked:
    d_i = x_i - x_{i+1} for i < n - 2; the last two entries along each axis
    keep x itself, as the reference's slices leave them."""
    torch = _torch()
    out = []
    for a in range(3):
        d = x.clone()
        n = x.shape[a]
        if n > 2:
            d.narrow(a, 0, n - 2).sub_(x.narrow(a, 1, n - 2))
        out.append(d)
    return torch.stack(out)


def _irn_weights(x):
    """(|D x|^2 + 1e-6)^(-1/4) (krylov_subspace_algorithms.py:456-465)."""
    return _irn_diffs(x).square().sum(0).add_(1e-6).pow_(-0.25)


def _irn_ltx(W, y):
    """Ltx (krylov_subspace_algorithms.py:478-496) as the reference computes it:
    per axis w = W y_a, out_0 = w_0, out_i = w_i - w_{i-1} for 0 < i < n - 2,
    out_{n-2} = w_{n-2}, out_{n-1} = -w_{n-2}."""
    w = y * W
    o = w.clone()
    for a in range(3):
        n = w.shape[a + 1]
        if n > 3:
            o[a].narrow(a, 1, n - 3).sub_(w[a].narrow(a, 0, n - 3))
        o[a].select(a, n - 1).copy_(w[a].select(a, n - 2).neg())
    return o.sum(0)


def irn_tv_cgls(proj, geo, angles, niter, lmbda=1.0, niter_outer=4, init=None, verbose=False, return_tensor=False,
                computel2=False, **kwargs):
    """IRN-TV-CGLS (krylov_subspace_algorithms.py:441-567): niter_outer cycles
    of niter CGLS steps on [A; sqrt(lmbda) W D] x = [b; 0], Siddon forward and
    matched backprojection, W = (|D x|^2 + 1e-6)^(-1/4) rebuilt from the image
    at the start of each cycle.  Kept as the reference computes it: the
    differences keep the last two entries of each axis, Ltx is not their exact
    transpose there, and each cycle restarts from the image before the previous
    cycle's last update.  FP64 norms; with ``computel2`` also returns
    ||b - A x|| after every step."""
    torch = _torch()
    g = copy.deepcopy(geo)
    g.check_geo(angles)
    vs = ViewSets(g, g.angles, [np.arange(len(g.angles))])
    b = _to_device(proj)
    if init is None:
        res = torch.zeros(vs.vol_shape, dtype=torch.float32, device="cuda")
    elif isinstance(init, str) and init == "FDK":
        res = fdk(b, g, g.angles, return_tensor=True).clamp_(min=0.0)
    else:
        res = _to_device(init).reshape(vs.vol_shape).clone()
    sl = float(np.sqrt(lmbda))
    res0 = res
    l2 = []
    for outer in range(niter_outer):
        if verbose:
            print("IRN-TV-CGLS outer iteration %d of %d" % (outer + 1, niter_outer))
        W = _irn_weights(res)
        res = res0
        r1 = torch.sub(b, vs.ax(res, 0))
        r2 = _irn_diffs(res).mul_(W).mul_(-sl)
        p = vs.atb(r1, 0, _native.BACKPROJ_MATCHED).add_(_irn_ltx(W, r2), alpha=sl)
        gamma = _norm(p) ** 2
        for _ in range(niter):
            res0 = res
            q1 = vs.ax(p, 0)
            q2 = _irn_diffs(p).mul_(W).mul_(sl)
            alpha = gamma / (_norm(q1) ** 2 + _norm(q2) ** 2)
            res = torch.add(res, p, alpha=alpha)
            if computel2:
                l2.append(_norm(torch.sub(b, vs.ax(res, 0))))
            r1 = torch.sub(r1, q1, alpha=alpha)
            r2 = torch.sub(r2, q2, alpha=alpha)
            s = vs.atb(r1, 0, _native.BACKPROJ_MATCHED).add_(_irn_ltx(W, r2), alpha=sl)
            gamma1 = _norm(s) ** 2
            p = torch.add(s, p, alpha=gamma1 / gamma)
            gamma = gamma1
    out = _result(res, return_tensor)
    return (out, np.array([l2])) if computel2 else out


def _sym_ortho(a, b):
    """Givens rotation (c, s, r) with c a + s b = r >= 0, in the stable form
    of S.-C. Choi's SymOrtho (scipy's LSQR uses the same)."""
    if b == 0.0:
        return float(np.sign(a)), 0.0, abs(a)
    if a == 0.0:
        return 0.0, float(np.sign(b)), abs(b)
    if abs(b) > abs(a):
        tau = a / b
        s = float(np.sign(b)) / np.sqrt(1.0 + tau * tau)
        return s * tau, s, b / s
    tau = b / a
    c = float(np.sign(a)) / np.sqrt(1.0 + tau * tau)
    return c, c * tau, a / c


def _lsqr_solve(matvec, rmatvec, b, iter_lim, atol=1e-6, btol=1e-6, conlim=1e8):
    """min ||K x - b|| by LSQR (Paige & Saunders, ACM TOMS 8 (1982) 43-71)
    from x = 0 with no damping, with the stopping tests of
    scipy.sparse.linalg.lsqr (istop 1-7: compatible system, least-squares
    optimality, condition limit, their machine-precision forms, iteration
    limit), which the reference calls for its preconditioner
    (krylov_subspace_algorithms.py:808-811).  ``matvec`` / ``rmatvec`` play K
    and its transpose; the pair need not be adjoint.  Vectors stay on their
    device; scalars and tests in FP64 on the host."""
    import torch

    def nrm(t):
        return float(torch.linalg.vector_norm(t))

    eps = float(np.finfo(np.float64).eps)
    ctol = 1.0 / conlim if conlim > 0 else 0.0
    anorm = acond = ddnorm = xnorm = xxnorm = z = 0.0
    cs2, sn2 = -1.0, 0.0
    bnorm = beta = nrm(b)
    if beta > 0:
        u = b / beta
        v = rmatvec(u)
        alfa = nrm(v)
    else:
        return torch.zeros_like(rmatvec(b))
    x = torch.zeros_like(v)
    if alfa > 0:
        v = v / alfa
    w = v.clone()
    rhobar, phibar = alfa, beta
    if alfa * beta == 0:
        return x
    itn = 0
    while itn < iter_lim:
        itn += 1
        u = torch.sub(matvec(v), u, alpha=alfa)
        beta = nrm(u)
        if beta > 0:
            u = u / beta
            anorm = np.sqrt(anorm ** 2 + alfa ** 2 + beta ** 2)
            v = torch.sub(rmatvec(u), v, alpha=beta)
            alfa = nrm(v)
            if alfa > 0:
                v = v / alfa
        cs, sn, rho = _sym_ortho(rhobar, beta)
        theta = sn * alfa
        rhobar = -cs * alfa
        phi = cs * phibar
        phibar = sn * phibar
        tau = sn * phi
        ddnorm += (nrm(w) / rho) ** 2
        x = torch.add(x, w, alpha=phi / rho)
        w = torch.add(v, w, alpha=-theta / rho)
        delta = sn2 * rho
        gambar = -cs2 * rho
        rhs = phi - delta * z
        zbar = rhs / gambar
        xnorm = np.sqrt(xxnorm + zbar ** 2)
        gamma = np.sqrt(gambar ** 2 + theta ** 2)
        cs2, sn2 = gambar / gamma, theta / gamma
        z = rhs / gamma
        xxnorm += z ** 2
        acond = anorm * np.sqrt(ddnorm)
        rnorm = abs(phibar)
        arnorm = alfa * abs(tau)
        test1 = rnorm / bnorm
        test2 = arnorm / (anorm * rnorm + eps)
        test3 = 1.0 / (acond + eps)
        t1 = test1 / (1.0 + anorm * xnorm / bnorm)
        rtol = btol + atol * anorm * xnorm / bnorm
        if (itn >= iter_lim or 1.0 + test3 <= 1.0 or 1.0 + test2 <= 1.0 or 1.0 + t1 <= 1.0 or test3 <= ctol
                or test2 <= atol or test1 <= rtol):
            break
    return x


def hybrid_flsqr_tv(proj, geo, angles, niter, lmbda=1.0, init=None, restart=True, verbose=False, return_tensor=False,
                    computel2=False, return_x=False, **kwargs):
    """Hybrid flexible LSQR with TV (krylov_subspace_algorithms.py:698-860):
    the constant image xA0 that best fits the data is split off; each step
    backprojects u (matched), orthogonalises it against the earlier v,
    preconditions it by two LSQR solves of at most 50 iterations with the
    IRN difference operators (Ltx then Lx, as the reference's scipy calls),
    removes <k, z> (k the backprojected constant image, scaled), extends the
    U basis with A z (Siddon) and solves the small problem
    min ||M^T y - beta e1||^2 + lmbda^2 ||WDZ y||^2; the image is
    start + Z y + xA0, and the data residual after each step is kept.

    The reference's entry point returns ``self.res``, which this loop never
    assigns (iterative_recon_alg.py:470-475), so it returns its starting
    image; that is returned here too unless ``return_x`` (no counterpart in
    the reference) asks for the reconstruction.  Preconditioner solves in
    FP64 on the device; the small problem in FP64 on the host, its factor R
    from the Cholesky factor of the Gram matrix (R^T R is all the problem
    uses, the same as the reference's QR up to row signs)."""
    torch = _torch()
    g = copy.deepcopy(geo)
    g.check_geo(angles)
    vs = ViewSets(g, g.angles, [np.arange(len(g.angles))])
    b = _to_device(proj)
    shape = vs.vol_shape
    N = int(np.prod(shape))
    res = (torch.zeros(shape, dtype=torch.float32, device="cuda") if init is None
           else _to_device(init).reshape(shape).clone())
    A = lambda x: vs.ax(x, 0, _native.PROJ_SIDDON)  # noqa: E731
    At = lambda y: vs.atb(y, 0, _native.BACKPROJ_MATCHED)  # noqa: E731
    flat = lambda t: t.reshape(-1).double()  # noqa: E731
    u = torch.sub(b, A(res))
    k = A(torch.full(shape, 1.0 / np.sqrt(N), dtype=torch.float32, device="cuda"))
    c = 1.0 / (np.sqrt(N) * _norm(k) ** 2)
    xA0 = c * float(torch.dot(flat(k), flat(u)))
    kv = flat(At(k)) * c
    u = torch.sub(u, A(torch.full(shape, xA0, dtype=torch.float32, device="cuda")))
    normr = _norm(u)
    u = u / normr
    W = (torch.ones(shape, dtype=torch.float64, device="cuda") if float(res.max()) == 0.0
         else _irn_weights(res).double())
    lx = lambda x: _irn_diffs(x.reshape(shape)).mul_(W).reshape(-1)  # noqa: E731
    ltx = lambda y: _irn_ltx(W, y.reshape((3,) + shape)).reshape(-1)  # noqa: E731
    U, V, Z, WZ = [u], [], [], []
    M = np.zeros((niter, niter + 1))
    rhs = np.zeros(niter + 1)
    rhs[0] = normr
    l2 = []
    x = res
    for i in range(niter):
        if verbose:
            print("hybrid fLSQR-TV iteration %d of %d" % (i + 1, niter))
        v = At(u)
        for vj in V:
            v = torch.sub(v, vj, alpha=float(torch.dot(flat(vj), flat(v))))
        v = v / _norm(v)
        zd = flat(v) - kv * float(flat(v).sum())
        zd = _lsqr_solve(lx, ltx, _lsqr_solve(ltx, lx, zd, 50), 50)
        zd = zd - float(torch.dot(kv, zd))
        V.append(v)
        Z.append(zd.float().reshape(shape))
        WZ.append(lx(flat(Z[-1])))
        u = A(Z[-1])
        for j, uj in enumerate(U):
            M[i, j] = float(torch.dot(flat(uj), flat(u)))
            u = torch.sub(u, uj, alpha=M[i, j])
        M[i, i + 1] = _norm(u)
        u = u / M[i, i + 1]
        U.append(u)
        G = torch.stack(WZ).matmul(torch.stack(WZ).T).cpu().numpy()
        R = np.linalg.cholesky(G).T
        y = np.linalg.lstsq(np.vstack([M[:i + 1, :i + 2].T, lmbda * R]),
                            np.concatenate([rhs[:i + 2], np.zeros(i + 1)]), rcond=None)[0]
        x = torch.full(shape, xA0, dtype=torch.float32, device="cuda").add_(res)
        for yj, zj in zip(y, Z):
            x = torch.add(x, zj, alpha=float(yj))
        l2.append(_norm(torch.sub(b, A(x))))
        if i > 0 and l2[-1] > l2[-2] and (i == 1 or not restart):
            print("hybrid fLSQR-TV exited due to divergence at iteration %d" % i)
            break
    out = _result(x if return_x else res, return_tensor)
    return (out, np.array([l2])) if computel2 else out


def _gmres(proj, geo, angles, niter, left, backprojector="matched", restart=True, verbose=False,
           return_tensor=False, computel2=False, **kwargs):
    """AB-GMRES (left False) / BA-GMRES (left True), krylov_subspace_algorithms.py:570-695."""
    torch = _torch()
    g = copy.deepcopy(geo)
    g.check_geo(angles)
    vs = ViewSets(g, g.angles, [np.arange(len(g.angles))])
    b = _to_device(proj)
    A = lambda x: vs.ax(x, 0, _native.PROJ_SIDDON)  # noqa: E731
    if backprojector == "FDK":
        B = lambda y: fdk(y, g, g.angles, return_tensor=True)  # noqa: E731
    elif backprojector == "matched":
        B = lambda y: vs.atb(y, 0, _native.BACKPROJ_MATCHED)  # noqa: E731
    else:
        raise ValueError("backprojector must be 'matched' or 'FDK'")
    r = B(b) if left else b.clone()
    beta = _norm(r)
    w = [r / beta]
    z = []  # AB: the backprojected basis B(w_i), kept instead of backprojecting it again
    H = np.zeros((niter + 1, niter))
    l2 = []

    def image(y):
        basis = w if left else z
        x = torch.zeros(vs.vol_shape, dtype=torch.float32, device="cuda")
        for yi, bi in zip(y, basis):
            x = torch.add(x, bi, alpha=float(yi))
        return x

    y, x = None, None
    for k in range(niter):
        if verbose:
            print("%s-GMRES iteration %d of %d" % ("BA" if left else "AB", k + 1, niter))
        if left:
            q = B(A(w[k]))
        else:
            z.append(B(w[k]))
            q = A(z[k])
        for i in range(k + 1):
            H[i, k] = float(torch.dot(q.reshape(-1).double(), w[i].reshape(-1).double()))
            q = torch.sub(q, w[i], alpha=H[i, k])
        H[k + 1, k] = _norm(q)
        w.append(q / H[k + 1, k])
        e1 = np.zeros(k + 2)
        e1[0] = beta
        y = np.linalg.lstsq(H[:k + 2, :k + 1], e1, rcond=None)[0]
        x = image(y)
        l2.append(_norm(b - A(x)))
        if k > 0 and l2[-1] > l2[-2] and (k == 1 or not restart):
            print("%s-GMRES exited due to divergence at iteration %d" % ("BA" if left else "AB", k))
            # the reference's loop returns here without assigning its image, and
            # its entry point returns that image: the zero start
            # (krylov_subspace_algorithms.py:623-626, 688-691; iterative_recon_alg.py:470-475)
            x = torch.zeros_like(x)
            break
    out = _result(x, return_tensor)
    return (out, np.array([l2])) if computel2 else out


def ab_gmres(proj, geo, angles, niter, **kwargs):
    """AB-GMRES (krylov_subspace_algorithms.py:570-630): GMRES on A B u = proj
    (Siddon A; B the matched backprojector, or FDK with backprojector='FDK'),
    image x = B u; FP64 inner products and Hessenberg solve.  The data residual
    is taken after every iteration; a rise at the second iteration (or at any,
    with restart=False) ends the iteration there, as in the reference."""
    return _gmres(proj, geo, angles, niter, False, **kwargs)


def ba_gmres(proj, geo, angles, niter, **kwargs):
    """BA-GMRES (krylov_subspace_algorithms.py:634-695): GMRES on B A x = B
    proj over volumes, the same operators and stopping rule as ab_gmres."""
    return _gmres(proj, geo, angles, niter, True, **kwargs)


def mlem(proj, geo, angles, niter, init=None, verbose=False, return_tensor=False, **kwargs):
    """MLEM (statistical_algorithms.py:11-55): res = 1 (or ``init``);
    W = Atb_matched(1), infinite where <= 0; every iteration res *=
    Atb_matched(proj / Ax_interpolated(res)) / W, where a zero forward
    projection divides by infinity; negatives clipped to 0 before and after."""
    torch = _torch()
    g = copy.deepcopy(geo)
    g.check_geo(angles)
    vs = ViewSets(g, g.angles, [np.arange(len(g.angles))])
    b = _to_device(proj)
    if init is None:
        res = torch.ones(vs.vol_shape, dtype=torch.float32, device="cuda")
    elif isinstance(init, str) and init == "FDK":
        res = fdk(b, g, g.angles, return_tensor=True)
    else:
        res = _to_device(init).reshape(vs.vol_shape).clone()
    W = vs.atb(torch.ones_like(b), 0, _native.BACKPROJ_MATCHED)
    W = torch.where(W <= 0.0, torch.full_like(W, float("inf")), W)
    res.clamp_(min=0.0)
    for i in range(niter):
        if verbose:
            print("MLEM iteration %d of %d" % (i + 1, niter))
        den = vs.ax(res, 0, _native.PROJ_INTERPOLATED)
        den = torch.where(den == 0.0, torch.full_like(den, float("inf")), den)
        res = res * (vs.atb(b / den, 0, _native.BACKPROJ_MATCHED) / W)
        res.clamp_(min=0.0)
    return _result(res, return_tensor)


def ista(proj, geo, angles, niter, **kwargs):
    """ISTA (ista_algorithms.py:185-209): FISTA's data step, then
    res = im3ddenoise(res, 20, 1 / (0.2 / hyper)) (its lambda is the fixed 0.1
    and its TV iterations the fixed 20), no extrapolation."""
    return _ista_family(proj, geo, angles, niter, False, **kwargs)


# ---------------------------------------------------------------------------
# CGLS (krylov_subspace_algorithms.py:22-97)
# ---------------------------------------------------------------------------
def cgls(proj, geo, angles, niter, restart=True, verbose=False, return_tensor=False, computel2=False, init=None,
         storage=None, narrow_iterate=False, **kwargs):
    """Conjugate gradient least squares with Siddon forward and matched
    backprojection (krylov_subspace_algorithms.py:22-97); FP64 inner products
    and norms; restarts once on residual growth.

    ``storage`` (a torch dtype: float32 by default, bfloat16 or float16) is the
    type the search direction, its backprojection and the sinograms (data,
    residual, A p) are stored in; the operators and the vector updates compute
    in float32 and round once on store.  The iterate x stays float32 unless
    ``narrow_iterate`` is set: an iterate rounded to a 7-bit significand after
    every update carries that rounding into the result (each of the niter
    updates adds a relative error of up to 2^-8), which a float32 iterate
    avoids at 2 extra bytes per voxel.  With narrow storage the residual is
    the true b - A x of each iteration (computed anyway for the l2 norm)
    instead of the recursive r - alpha q of the reference: the same vector in
    exact arithmetic, without the rounding a narrow recursion accumulates."""
    torch = _torch()
    dt = torch.float32 if storage is None else storage
    xt = dt if narrow_iterate else torch.float32
    g = copy.deepcopy(geo)
    g.check_geo(angles)
    vs = ViewSets(g, g.angles, [np.arange(len(g.angles))])
    b = _to_device(proj).to(dt)
    x = torch.zeros(vs.vol_shape, dtype=xt, device="cuda") if init is None else _to_device(init).to(xt)

    # Narrow storage keeps the search direction normalised, p = n_p p_hat with
    # the scalar n_p in FP64: A p can exceed the float16 range (A^T A has a
    # norm in the thousands), A p_hat does not.  Same iteration as the plain
    # form in exact arithmetic.
    scaled = dt != torch.float32

    def residual():
        # b - A x, rounded once to the storage type
        return torch.sub(b, vs.ax(x, 0)).to(dt)

    def start():
        r = residual()
        p = vs.atb(r, 0, _native.BACKPROJ_MATCHED)
        n_p = _norm(p)
        if scaled:
            p = torch.div(p, n_p)
            return r, p, n_p, n_p ** 2
        return r, p, 1.0, n_p ** 2

    r, p, n_p, gamma = start()
    l2 = np.zeros((1, niter))
    reinit_at = 0
    i = 0
    while i < niter:
        q = vs.ax(p, 0)
        alpha = gamma / ((n_p * _norm(q)) ** 2)
        step = alpha * n_p
        torch.add(x, p, alpha=step, out=x)
        rr = residual()
        l2[0, i] = _norm(rr)
        if i > 0 and l2[0, i] > l2[0, i - 1]:
            torch.sub(x, p, alpha=step, out=x)
            if verbose:
                print("re-initilization of CGLS called at iteration:" + str(i))
            if reinit_at + 1 == i or not restart:
                print("CGLS exited due to divergence.")
                break
            reinit_at = i
            r, p, n_p, gamma = start()
            continue
        if scaled:
            r = rr  # the true residual, rounded once: no drift of a narrow recursive update
        else:
            torch.sub(r, q, alpha=step, out=r)
        s = vs.atb(r, 0, _native.BACKPROJ_MATCHED)
        gamma1 = _norm(s) ** 2
        p = torch.add(s, p, alpha=gamma1 / gamma * n_p)
        if scaled:
            n_p = _norm(p)
            p = torch.div(p, n_p)
        gamma = gamma1
        i += 1
    res = _result(x, return_tensor)
    return (res, l2) if computel2 else res


# ---------------------------------------------------------------------------
# ASD-POCS (pocs_algorithms.py:15-185)
# ---------------------------------------------------------------------------
def asd_pocs(proj, geo, angles, niter, alpha=0.002, alpha_red=0.95, rmax=0.95, tviter=20, maxl2err=None,
             lmbda=1.0, lmbda_red=1.0, OrderStrategy=None, verbose=False, return_tensor=False, regularizer=None,
             blocksize=1, **kwargs):
    """Adaptive steepest descent - projection onto convex sets.

    Each outer iteration: one SART sweep (``blocksize`` views per subset; 1 for
    ASD-POCS, 20 by default for the OS variants) with a non-negativity clamp,
    then ``tviter`` TV steepest-descent steps of size dtvg with the step
    adapted as in the reference; stops on the reference's criteria.
    ``regularizer`` (default: minTV) may replace the TV step with any callable
    (image_tensor, dtvg) -> image_tensor; ``callback(n_iter, image_tensor)``,
    if given, runs after each iteration, outside its timed part."""
    from ..tv import minTV

    st = _SartState(proj, geo, angles, blocksize, order=OrderStrategy, lmbda=lmbda, noneg=True, verbose=verbose)
    full = ViewSets(st.geo, st.angles, [np.arange(len(st.angles))])
    if maxl2err is None:
        epsilon = 0.2 * _norm(full.ax(fdk(st.proj, st.geo, st.angles, return_tensor=True), 0) - st.proj)
    else:
        epsilon = maxl2err
    reg = regularizer if regularizer is not None else (lambda img, step: minTV(img, step, tviter))
    beta, n_iter, dtvg = lmbda, 0, None
    stats = []
    while True:
        t0 = time.perf_counter()
        res_prev = st.res.clone()
        st.sweep()
        n_iter += 1
        dd = _norm(full.ax(st.res, 0) - st.proj)
        dp_vec = st.res - res_prev
        dp = _norm(dp_vec)
        if n_iter == 1:
            dtvg = alpha * dp
        res_prev = st.res.clone()
        st.res = reg(st.res, dtvg)
        dg_vec = st.res - res_prev
        dg = _norm(dg_vec)
        if dg > rmax * dp and dd > epsilon:
            dtvg *= alpha_red
        beta *= lmbda_red
        import torch

        c = float(torch.dot(dg_vec.reshape(-1).double(), dp_vec.reshape(-1).double())) / max(dg * dp, 1e-6)
        stats.append((n_iter, dd, dp, dg, c, time.perf_counter() - t0))
        if kwargs.get("callback") is not None:
            kwargs["callback"](n_iter, st.res)
        if (c < -0.99 and dd <= epsilon) or beta < 0.005 or n_iter >= niter:
            if verbose:
                print("Stop criteria met: c = %s beta = %s iter = %d" % (c, beta, n_iter))
            break
    res = _result(st.res, return_tensor)
    return (res, stats) if kwargs.get("return_stats", False) else res


def os_asd_pocs(proj, geo, angles, niter, blocksize=20, **kwargs):
    """Ordered-subsets ASD-POCS (pocs_algorithms.py OS_ASD_POCS): blocksize 20."""
    return asd_pocs(proj, geo, angles, niter, blocksize=blocksize, **kwargs)


def awasd_pocs(proj, geo, angles, niter, delta=-0.005, tviter=20, blocksize=1, **kwargs):
    """ASD-POCS with the adaptive-weighted (edge preserving) TV step AwminTV
    (pocs_algorithms.py AwASD_POCS; delta controls the smoothing at edges)."""
    from ..tv import AwminTV

    reg = lambda img, step: AwminTV(img, step, tviter, delta)  # noqa: E731
    return asd_pocs(proj, geo, angles, niter, tviter=tviter, regularizer=reg, blocksize=blocksize, **kwargs)


def os_awasd_pocs(proj, geo, angles, niter, blocksize=20, **kwargs):
    """Ordered-subsets AwASD-POCS (pocs_algorithms.py OS_AwASD_POCS)."""
    return awasd_pocs(proj, geo, angles, niter, blocksize=blocksize, **kwargs)


def piccs(proj, geo, angles, niter, prior=None, prior_ratio=0.2, tviter=20, blocksize=1, **kwargs):
    """Prior-image constrained compressed sensing (pocs_algorithms.py PICCS):
    the ASD-POCS loop with the TV step minPICCS(img, prior, dtvg, tviter,
    prior_ratio)."""
    from ..tv import minPICCS

    if prior is None:
        raise ValueError("PICCS needs a prior image (prior=...)")
    pr = _to_device(prior)
    reg = lambda img, step: minPICCS(img, pr, step, tviter, prior_ratio)  # noqa: E731
    return asd_pocs(proj, geo, angles, niter, tviter=tviter, regularizer=reg, blocksize=blocksize, **kwargs)


def os_piccs(proj, geo, angles, niter, blocksize=20, **kwargs):
    """Ordered-subsets PICCS (pocs_algorithms.py OS_PICCS)."""
    return piccs(proj, geo, angles, niter, blocksize=blocksize, **kwargs)


# ---------------------------------------------------------------------------
# PCSD (pocs_algorithms.py:434-646)
# ---------------------------------------------------------------------------
def _pcsd(proj, geo, angles, niter, blocksize, delta=None, tviter=20, maxl2err=None, lmbda=1.0, lmbda_red=1.0,
          OrderStrategy=None, verbose=False, return_tensor=False, **kwargs):
    from ..tv import AwminTV, minTV

    torch = _torch()
    st = _SartState(proj, geo, angles, blocksize, order=OrderStrategy, lmbda=lmbda, noneg=True, verbose=verbose)
    full = ViewSets(st.geo, st.angles, [np.arange(len(st.angles))])
    if maxl2err is None:
        epsilon = 0.2 * _norm(full.ax(fdk(st.proj, st.geo, st.angles, return_tensor=True), 0) - st.proj)
    else:
        epsilon = maxl2err
    beta, n_iter, d_p1 = lmbda, 0, 1.0
    while True:
        res_prev = st.res.clone()
        n_iter += 1
        d_p = _norm(full.ax(st.res, 0) - st.proj)
        if d_p ** 2 > epsilon:  # the reference compares the squared residual norm with epsilon
            st.sweep()
        dd = _norm(full.ax(st.res, 0) - st.proj)
        dp_vec = st.res - res_prev
        dp = _norm(dp_vec)
        dtvg = 1.0 if n_iter == 1 else d_p / d_p1
        res_prev = st.res.clone()
        st.res = minTV(st.res, dtvg, tviter) if delta is None else AwminTV(st.res, dtvg, tviter, delta)
        dg_vec = st.res - res_prev
        dg = _norm(dg_vec)
        if n_iter == 1:
            d_p1 = _norm(full.ax(res_prev, 0) - st.proj)
        beta *= lmbda_red
        c = float(torch.dot(dg_vec.reshape(-1).double(), dp_vec.reshape(-1).double())) / max(dg * dp, 1e-6)
        if verbose:
            print("PCSD iteration %d: data residual %.6g, TV step %.6g" % (n_iter, dd, dtvg))
        if (c < -0.99 and dd <= epsilon) or beta < 0.005 or n_iter > niter:
            break
    return _result(st.res, return_tensor)


def pcsd(proj, geo, angles, niter, **kwargs):
    """Projection-controlled steepest descent (pocs_algorithms.py:434-586):
    every iteration a SART sweep while the squared data residual exceeds
    epsilon (0.2 |Ax(FDK) - proj| unless maxl2err), then ``tviter`` (20) TV
    steps of size 1 in the first iteration and d_p / d_p1 after it; the loop
    ends on the reference's criteria, at the latest after iteration niter + 1."""
    if kwargs.pop("blocksize", 1) > 1:
        print("Warning: blocksize is set to 1, please use an OS version of the algorithm for blocksize > 1")
    return _pcsd(proj, geo, angles, niter, 1, **kwargs)


def aw_pcsd(proj, geo, angles, niter, delta=-0.005, **kwargs):
    """Adaptive-weighted PCSD (pocs_algorithms.py:589-608): the TV step is AwminTV with ``delta``."""
    if kwargs.pop("blocksize", 1) > 1:
        print("Warning: blocksize is set to 1, please use an OS version of the algorithm for blocksize > 1")
    return _pcsd(proj, geo, angles, niter, 1, delta=delta, **kwargs)


def os_pcsd(proj, geo, angles, niter, blocksize=20, **kwargs):
    """Ordered-subsets PCSD (pocs_algorithms.py:611-625), 20 views per subset by default."""
    return _pcsd(proj, geo, angles, niter, blocksize, **kwargs)


def os_aw_pcsd(proj, geo, angles, niter, blocksize=20, delta=-0.005, **kwargs):
    """Ordered-subsets adaptive-weighted PCSD (pocs_algorithms.py:628-646)."""
    return _pcsd(proj, geo, angles, niter, blocksize, delta=delta, **kwargs)


from .batched import BatchedOSSART  # noqa: E402  (uses _SartState above)
