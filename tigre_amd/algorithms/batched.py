"""Many independent OS-SART reconstructions on one device, batched per subset step.

Each case is an OS-SART problem (art_family_algorithms.py:47-62 with the
subset update of iterative_recon_alg.py:326-345 / :389-399): for every subset
j, r = W_j (p_j - Ax_j(x)), then x = max(x + (lambda / V_j) Atb_j(r), 0).

Cases advance in waves: wave w applies subset w mod n_c of every case c that
still has iterations left (n_c subsets per iteration).  A wave is one
``tigre_batch_sart`` call, five launches that each cover every case of the
wave (staging copies, beam projections, residuals, quad copies, FDK
backprojections with the update epilogue), so the small problems run beside
the large ones instead of after them.  The job tables of every wave are built
on the device once; ``run`` then only launches kernels, and ``capture``
records all waves into one CUDA/HIP graph replayed with one host call.

``run_sequential`` launches the same plan kernels case after case from the
host.  Both give the bits of :func:`tigre_amd.algorithms.ossart` when the
plans use the default parameters, since the kernels and their rounding steps
are the same.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .. import _native
from ..operators import _stream_handle
from ..plan import Autotuner, Plan, sart_residual
from . import _SartState, _sub_geometry, _torch


class _Case:
    def __init__(self, proj, geo, angles, blocksize, lmbda, lmbda_red, niter, noneg, order, tuner, params):
        torch = _torch()
        st = _SartState(proj, geo, angles, blocksize, order=order, lmbda=lmbda, noneg=noneg, use_plans=False)
        self.blocks = st.blocks
        self.proj_blocks = st.proj_blocks
        self.W_blocks = st.W_blocks
        self.noneg = noneg
        if params is None:
            params = tuner.choose(_sub_geometry(st.geo, self.blocks[0]), st.angles[self.blocks[0]]) \
                if tuner is not None else (4, 8, 16)
        self.params = tuple(params)
        self.plans = [Plan(_sub_geometry(st.geo, idx), st.angles[idx], *self.params) for idx in self.blocks]
        # lambda / V_j, the expression the eager loop evaluates; lmbda_red is not
        # applied, as in the reference's SART family (iterative_recon_alg.py:143, :389-399)
        self.inv_v = [(lmbda / st.V[j]).contiguous() for j in range(len(self.blocks))]
        self.ws = torch.empty(max(max(p.workspace_bytes() for p in self.plans), 256), dtype=torch.uint8,
                              device="cuda")
        nmax = max(len(b) for b in self.blocks)
        nV, nU = self.plans[0].proj_shape[1:]
        self.axbuf = torch.empty((nmax, nV, nU), dtype=torch.float32, device="cuda")
        self.res = torch.zeros(self.plans[0].vol_shape, dtype=torch.float32, device="cuda")
        self.work = float(np.prod(self.res.shape)) * len(angles)

    def iteration(self, it):
        for j, p in enumerate(self.plans):
            ax = self.axbuf[: p.nviews]
            p.ax(self.res, ax, self.ws, _native.PROJ_SIDDON)
            sart_residual(self.proj_blocks[j], self.W_blocks[j], ax)
            p.sart_update(ax, self.res, self.inv_v[j], self.ws, self.noneg, _native.BACKPROJ_FDK)


class BatchedOSSART:
    """OS-SART on a list of ``(proj, geo, angles)`` cases, ``niter`` iterations each.

    ``tuner``: an :class:`~tigre_amd.plan.Autotuner` picking the plan
    parameters per case shape; ``params``: fixed ``(views_per_group,
    max_slab_planes, slices_per_thread)`` for every case instead; neither:
    the defaults (4, 8, 16)."""

    def __init__(self, cases, niter, blocksize=20, lmbda=1.0, lmbda_red=1.0, noneg=True, OrderStrategy=None,
                 tuner: Autotuner | None = None, params=None):
        self.niter = int(niter)
        self.cases = [_Case(p, g, a, blocksize, lmbda, lmbda_red, self.niter, noneg, OrderStrategy, tuner, params)
                      for (p, g, a) in cases]
        self.graph = None
        self.waves = None
        self._lib = _native.lib()

    def reset(self):
        for c in self.cases:
            c.res.zero_()

    def results(self):
        return [c.res for c in self.cases]

    def _build_waves(self):
        """One batch per wave: subset w mod n_c of every case with iterations left."""
        steps_per_case = [self.niter * len(c.plans) for c in self.cases]
        self.waves = []
        for w in range(max(steps_per_case)):
            items = []
            for c, n in zip(self.cases, steps_per_case):
                if w >= n:
                    continue
                j = w % len(c.plans)
                p = c.plans[j]
                items.append(_native.TigreBatchStep(
                    p._h.value, c.res.data_ptr(), c.proj_blocks[j].data_ptr(), c.W_blocks[j].data_ptr(),
                    c.axbuf.data_ptr(), c.inv_v[j].data_ptr(), c.ws.data_ptr(), c.ws.numel(), int(c.noneg)))
            arr = (_native.TigreBatchStep * len(items))(*items)
            h = ctypes.c_void_p()
            _native.check(self._lib.tigre_batch_create(arr, len(items), ctypes.byref(h)), "TIGRE: batch creation")
            self.waves.append(h)

    def run(self):
        """Every case's ``niter`` iterations from a zero image, one batch launch per wave, on the current stream."""
        if self.waves is None:
            self._build_waves()
        self.reset()
        stream = _stream_handle()
        for h in self.waves:
            _native.check(self._lib.tigre_batch_sart(h, stream), "TIGRE: batched SART step")
        return self.results()

    def run_sequential(self):
        """All iterations of case 0, then case 1, ... from the host, on the current stream (from the
        current images)."""
        for c in self.cases:
            for it in range(self.niter):
                c.iteration(it)
        return self.results()

    def capture(self):
        """Record ``run`` (zeroing included) into one graph."""
        torch = _torch()
        self.run()  # builds the tables and loads every kernel before capture
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.run()
        torch.cuda.synchronize()
        return self.graph

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()
        return self.results()

    def close(self):
        self.graph = None
        for h in self.waves or []:
            self._lib.tigre_batch_destroy(h)
        self.waves = None
        for c in self.cases:
            for p in c.plans:
                p.close()
