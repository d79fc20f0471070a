"""Projection plans and their autotuner.

A :class:`Plan` prepares the per-view records of one geometry and view list
on the device once (``tigre_plan_create``); ``ax`` / ``atb`` /
``sart_update`` then only enqueue kernels on the current torch stream, with
scratch from a caller workspace, so they can be captured in a CUDA/HIP graph
(``torch.cuda.graph``).  This is the device-side counterpart of the subset
loop of the reference's iterative methods (iterative_recon_alg.py:326-345),
which rebuilds the geometry and allocates on every Ax/Atb call.

The tunables (views per beam workgroup, LDS slab depth, backprojector slices
per lane) are picked by :class:`Autotuner`, which times the candidates on the
device for a representative geometry and remembers the winner per shape.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import _native
from .geometry import prepared
from .operators import _stream_handle

VIEWS_PER_GROUP = (1, 2, 4)
SLAB_PLANES = (2, 4, 8)
SLICES_PER_THREAD = (4, 8, 16)


def _torch():
    import torch

    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Plan:
    """Prepared projector pair for ``geo`` restricted to ``angles``.

    Operands are torch device tensors on the device current at creation:
    volumes (nZ, nY, nX), projections (nviews, nV, nU), float32 (or bfloat16 /
    float16 for ``ax`` / ``atb``)."""

    def __init__(self, geo, angles, views_per_group=4, max_slab_planes=8, slices_per_thread=16):
        g = prepared(geo, angles)
        self.nviews = int(np.asarray(g.angles).shape[0])
        self._cg = _native.CGeometry(g, self.nviews)
        self._ang = np.ascontiguousarray(g.angles, dtype=np.float32)
        self.vol_shape = tuple(int(v) for v in np.asarray(g.nVoxel))
        self.proj_shape = (self.nviews, int(g.nDetector[0]), int(g.nDetector[1]))
        self.params = (int(views_per_group), int(max_slab_planes), int(slices_per_thread))
        self._lib = _native.lib()
        h = ctypes.c_void_p()
        _native.check(self._lib.tigre_plan_create(self._cg.ptr(), self._ang.ctypes.data_as(ctypes.c_void_p),
                                                  self.nviews, self.params[0], self.params[1], self.params[2],
                                                  ctypes.byref(h)), "TIGRE: plan creation failed")
        self._h = h

    def workspace_bytes(self, dtype=None):
        torch = _torch()
        code = _native.DTYPE_F32 if dtype is None else _dtype_code(dtype, torch)
        n = ctypes.c_size_t()
        _native.check(self._lib.tigre_plan_workspace_size(self._h, code, ctypes.byref(n)), "TIGRE: plan workspace")
        return int(n.value)

    def workspace(self, dtype=None, device="cuda"):
        """A workspace tensor (256-byte aligned by the caching allocator)."""
        torch = _torch()
        return torch.empty(max(self.workspace_bytes(dtype), 256), dtype=torch.uint8, device=device)

    def ax(self, img, out, ws, kind=_native.PROJ_SIDDON, layout=_native.LAYOUT_PYTHON):
        torch = _torch()
        _native.check(self._lib.tigre_plan_ax(self._h, _ptr(img), _ptr(out), kind, layout,
                                              _dtype_code(img.dtype, torch), _ptr(ws), ws.numel(),
                                              _stream_handle()), "TIGRE: Call to Ax failed")
        return out

    def atb(self, proj, out, ws, kind=_native.BACKPROJ_FDK, layout=_native.LAYOUT_PYTHON):
        torch = _torch()
        _native.check(self._lib.tigre_plan_atb(self._h, _ptr(proj), _ptr(out), kind, layout,
                                               _dtype_code(proj.dtype, torch), _ptr(ws), ws.numel(),
                                               _stream_handle()), "Atb:")
        return out

    def sart_update(self, residual, x, inv_v, ws, noneg=True, kind=_native.BACKPROJ_FDK,
                    layout=_native.LAYOUT_PYTHON):
        """x <- x + inv_v * Atb(residual), clipped at 0 when noneg (float32)."""
        _native.check(self._lib.tigre_plan_sart_update(self._h, _ptr(residual), _ptr(x), _ptr(inv_v), kind, layout,
                                                       int(bool(noneg)), _ptr(ws), ws.numel(), _stream_handle()),
                      "Atb:")
        return x

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.tigre_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sart_residual(proj, w, ax):
    """ax <- w * (proj - ax) in place (float32, same shapes)."""
    lib = _native.lib()
    _native.check(lib.tigre_sart_residual(_ptr(proj), _ptr(w), _ptr(ax), ax.numel(), _stream_handle()),
                  "TIGRE: SART residual failed")
    return ax


def _dtype_code(dtype, torch):
    codes = {torch.float32: _native.DTYPE_F32, torch.bfloat16: _native.DTYPE_BF16, torch.float16: _native.DTYPE_F16}
    if dtype not in codes:
        raise TypeError("volumes and projections are stored as float32, bfloat16 or float16, not " + str(dtype))
    return codes[dtype]


class Autotuner:
    """Picks plan parameters by timing candidates on the device.

    Forward (views per group x slab planes) and backward (slices per lane)
    choices are independent kernels and are timed separately on random data
    of the real shapes: ``reps`` launches per candidate after one warm-up,
    HIP events on the current stream.  Winners are remembered per shape key
    (volume shape, detector shape, views in the subset) and, when ``path`` is
    given, kept in a JSON file between runs."""

    def __init__(self, path=None, reps=3, projection_type=_native.PROJ_SIDDON,
                 backprojection_type=_native.BACKPROJ_FDK):
        self.path = path
        self.reps = int(reps)
        self.ptype = projection_type
        self.btype = backprojection_type
        self.table = {}
        self.timings = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.table = {k: tuple(v) for k, v in json.load(f).items()}

    @staticmethod
    def key(geo, nviews):
        nv = tuple(int(v) for v in np.asarray(geo.nVoxel).reshape(3))
        nd = tuple(int(v) for v in np.asarray(geo.nDetector).reshape(2))
        return "vol%dx%dx%d_det%dx%d_views%d" % (nv + nd + (int(nviews),))

    def _time(self, fn):
        torch = _torch()
        fn()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(self.reps):
            fn()
        stop.record()
        stop.synchronize()
        return start.elapsed_time(stop) / self.reps

    def choose(self, geo, angles):
        """(views_per_group, max_slab_planes, slices_per_thread) for this geometry's shape."""
        torch = _torch()
        k = self.key(geo, len(angles))
        if k in self.table:
            return self.table[k]
        gen = torch.Generator(device="cuda").manual_seed(0)
        nv = tuple(int(v) for v in np.asarray(geo.nVoxel).reshape(3))
        nd = (len(angles),) + tuple(int(v) for v in np.asarray(geo.nDetector).reshape(2))
        img = torch.rand(nv, generator=gen, device="cuda")
        proj = torch.rand(nd, generator=gen, device="cuda")
        vol = torch.empty_like(img)
        out = torch.empty_like(proj)
        best_f, best_b = None, None
        times = {}
        for vpg in VIEWS_PER_GROUP:
            for planes in SLAB_PLANES:
                p = Plan(geo, angles, vpg, planes, 16)
                ws = p.workspace()
                t = self._time(lambda: p.ax(img, out, ws, self.ptype))
                times["ax", vpg, planes] = t
                if best_f is None or t < best_f[0]:
                    best_f = (t, vpg, planes)
                p.close()
        for z in SLICES_PER_THREAD:
            p = Plan(geo, angles, 4, 8, z)
            ws = p.workspace()
            t = self._time(lambda: p.atb(proj, vol, ws, self.btype))
            times["atb", z] = t
            if best_b is None or t < best_b[0]:
                best_b = (t, z)
            p.close()
        choice = (best_f[1], best_f[2], best_b[1])
        self.table[k] = choice
        self.timings[k] = times
        if self.path:
            os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
            with open(self.path, "w") as f:
                json.dump({kk: list(v) for kk, v in self.table.items()}, f, indent=1)
        return choice
