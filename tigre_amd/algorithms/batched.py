"""Many independent OS-SART reconstructions on one device, captured in one graph.

Each case is an OS-SART problem (art_family_algorithms.py:47-62 with the
subset update of iterative_recon_alg.py:326-345 / :389-399): for every subset
j, r = W_j (p_j - Ax_j(x)), then x = max(x + (lambda / V_j) Atb_j(r), 0).
Per subset the plan launches stage / beam-projection kernels, the fused
residual and the backprojector whose epilogue applies the update, so one
subset costs a handful of kernel launches and no host work.

``run_sequential`` launches case after case from the host (the same kernels,
one Python call per launch); ``capture`` records every case's iterations into
one CUDA/HIP graph with the cases spread over a few streams (independent
branches of the graph), and ``replay`` runs it with one host call;
``capture_lanes`` / ``replay_lanes`` record one graph per stream instead and
replay them on concurrent streams, which lets small cases overlap (the
branches of one graph execute one after another on this runtime).  All give
the same bits as :func:`tigre_amd.algorithms.ossart` when the plans use the
default parameters, since the kernels and their rounding steps are the same.
"""
from __future__ import annotations

import numpy as np

from .. import _native
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
                 tuner: Autotuner | None = None, params=None, streams=8):
        self.niter = int(niter)
        self.cases = [_Case(p, g, a, blocksize, lmbda, lmbda_red, self.niter, noneg, OrderStrategy, tuner, params)
                      for (p, g, a) in cases]
        self.nstreams = max(1, int(streams))
        self.graph = None

    def reset(self):
        for c in self.cases:
            c.res.zero_()

    def results(self):
        return [c.res for c in self.cases]

    def run(self):
        """Every case's ``niter`` iterations from a zero image, on the current stream."""
        self.reset()
        return self.run_sequential()

    def run_sequential(self):
        """All iterations of case 0, then case 1, ... from the host, on the current stream."""
        for c in self.cases:
            for it in range(self.niter):
                c.iteration(it)
        return self.results()

    def _assign(self):
        # longest cases first, each to the least-loaded stream
        load = [0.0] * self.nstreams
        lanes = [[] for _ in range(self.nstreams)]
        for i in sorted(range(len(self.cases)), key=lambda k: -self.cases[k].work):
            s = int(np.argmin(load))
            lanes[s].append(i)
            load[s] += self.cases[i].work
        return lanes

    def capture(self):
        """Record every case's ``niter`` iterations into one graph (results start from zero at replay)."""
        torch = _torch()
        # one eager pass loads every kernel before capture
        self.run_sequential()
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        side = [torch.cuda.Stream() for _ in range(self.nstreams)]
        cap = torch.cuda.Stream()
        lanes = self._assign()
        with torch.cuda.graph(self.graph, stream=cap):
            main = torch.cuda.current_stream()
            for c in self.cases:
                c.res.zero_()
            for s, idx in zip(side, lanes):
                if not idx:
                    continue
                s.wait_stream(main)
                with torch.cuda.stream(s):
                    for i in idx:
                        for it in range(self.niter):
                            self.cases[i].iteration(it)
            for s, idx in zip(side, lanes):
                if idx:
                    main.wait_stream(s)
        torch.cuda.synchronize()
        return self.graph

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()
        return self.results()

    def capture_lanes(self):
        """One graph per stream (its share of the cases), so that replays on
        different streams can run at the same time (the branches of a single
        graph run one after another on this runtime)."""
        torch = _torch()
        self.run_sequential()
        torch.cuda.synchronize()
        self.lanes = []
        for idx in self._assign():
            if not idx:
                continue
            s = torch.cuda.Stream()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=s):
                for i in idx:
                    self.cases[i].res.zero_()
                    for it in range(self.niter):
                        self.cases[i].iteration(it)
            self.lanes.append((s, g))
        torch.cuda.synchronize()
        return self.lanes

    def replay_lanes(self):
        """Replay every lane's graph on its own stream, joined to the current stream."""
        torch = _torch()
        if not getattr(self, "lanes", None):
            self.capture_lanes()
        main = torch.cuda.current_stream()
        for s, g in self.lanes:
            s.wait_stream(main)
            with torch.cuda.stream(s):
                g.replay()
        for s, _ in self.lanes:
            main.wait_stream(s)
        return self.results()

    def close(self):
        self.graph = None
        self.lanes = []
        for c in self.cases:
            for p in c.plans:
                p.close()
