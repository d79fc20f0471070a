"""Learned proximal step for ASD-POCS: a residual 3D CNN on BF16 MFMA kernels.

The TV step of ASD-POCS (minimizeTV, pocs_algorithms.py:145-149 calling
GD_TV.hpp:58) moves the image by at most ``tviter`` steps of size dtvg.  The
learned step replaces it with

    P(x, s) = x + s * F(x),    F = conv6 o relu o conv5 ... relu o conv1,

six 3x3x3 convolutions with 32 channels between them, BF16 weights and
activations, FP32 accumulation, trained so that x + F(x) matches
x + (minimizeTV(x, s) - x) / s on pairs that the plain TV step generates.
Every convolution, forward and backward, runs in the library's hand-written
MFMA kernels (tigre_cnn_conv / tigre_cnn_wgrad); torch only holds device
memory, the small FP32 master weights and the Adam update.

``LearnedProx`` is the ``regularizer`` callable of
:func:`tigre_amd.algorithms.asd_pocs`; plain TV is its fallback, taken when the
learned step does not lower the TV objective or raises the data residual.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .operators import _stream_handle

C = 32
LAYERS = 6


def _torch():
    import torch

    return torch


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class ResidualCNN:
    """The network, its FP32 master weights and its kernels.

    Master weights W[l] have shape (27, Cout, Cin) (taps (dz, dy, dx) index
    9(dz+1) + 3(dy+1) + dx+1), Cin = 1 for the first layer, Cout = 1 for the
    last; biases b[l] have Cout entries."""

    def __init__(self, seed=0, device="cuda", data_scale=1.0):
        torch = _torch()
        # images are multiplied by data_scale on the way in and divided by it on
        # the way out (``apply``), so the network sees values of order one
        self.data_scale = float(data_scale)
        g = torch.Generator().manual_seed(int(seed))
        shapes = [(27, C, 1)] + [(27, C, C)] * 4 + [(27, 1, C)]
        self.W, self.b = [], []
        for l, s in enumerate(shapes):
            fan_in = 27 * s[2]
            std = (2.0 / fan_in) ** 0.5 if l < LAYERS - 1 else 0.1 * (1.0 / fan_in) ** 0.5
            self.W.append((torch.randn(s, generator=g) * std).to(device))
            self.b.append(torch.zeros(s[1], device=device))
        self.device = device
        self.lib = _native.lib()
        self._adam = None
        self.refresh()

    # ---- kernel-side copies ------------------------------------------------
    def refresh(self):
        """Rebuild the bf16 kernel weights (padded to 32 x 32 per tap), the
        flipped / transposed data-gradient weights and padded biases."""
        torch = _torch()
        self.Wk, self.Wd, self.bk = [], [], []
        for l in range(LAYERS):
            w = torch.zeros((27, C, C), device=self.device)
            co, ci = self.W[l].shape[1], self.W[l].shape[2]
            w[:, :co, :ci] = self.W[l]
            wk = w.to(torch.bfloat16).contiguous()
            self.Wk.append(wk)
            # data gradient: conv of dy with taps mirrored and channels swapped
            self.Wd.append(wk.flip(0).transpose(1, 2).contiguous())
            bb = torch.zeros(C, device=self.device)
            bb[:co] = self.b[l]
            self.bk.append(bb)

    def parameters(self):
        return self.W + self.b

    def state(self):
        return {"W": [w.cpu().numpy() for w in self.W], "b": [v.cpu().numpy() for v in self.b],
                "data_scale": self.data_scale}

    def load(self, state):
        torch = _torch()
        self.data_scale = float(state.get("data_scale", 1.0))
        self.W = [torch.as_tensor(np.asarray(w, dtype=np.float32), device=self.device) for w in state["W"]]
        self.b = [torch.as_tensor(np.asarray(v, dtype=np.float32), device=self.device) for v in state["b"]]
        self.refresh()

    # ---- launches ----------------------------------------------------------
    def _conv(self, x, x_mode, w, bias, aux, y, epi, shape):
        n, d, h, wd = shape
        st = self.lib.tigre_cnn_conv(_p(x), x_mode, _p(w), _p(bias), _p(aux), _p(y), epi, n, d, h, wd,
                                     _stream_handle())
        _native.check(st, "TIGRE: CNN convolution failed")

    def _wgrad(self, x, x_mode, dy, dy_mode, shape, ws, out):
        n, d, h, wd = shape
        st = self.lib.tigre_cnn_wgrad(_p(x), x_mode, _p(dy), dy_mode, _p(out), n, d, h, wd, _p(ws), ws.numel(),
                                      _stream_handle())
        _native.check(st, "TIGRE: CNN weight gradient failed")

    def _shape(self, img):
        s = tuple(int(v) for v in img.shape)
        if len(s) == 3:
            s = (1,) + s
        if len(s) != 4:
            raise ValueError("expected a volume (D, H, W) or a batch (N, D, H, W)")
        return s

    def forward(self, img, scale=1.0, keep=False):
        """img + scale * F(img) for an FP32 device volume or batch; with keep,
        the activations are kept for ``backward``."""
        torch = _torch()
        if img.dtype != torch.float32 or not img.is_contiguous():
            img = img.float().contiguous()
        shape = self._shape(img)
        acts = []
        bufs = [torch.empty(shape + (C,), dtype=torch.bfloat16, device=img.device) for _ in range(5 if keep else 2)]
        x, x_mode = img, _native.CNN_IN_F32X1
        for l in range(LAYERS - 1):
            y = bufs[l] if keep else bufs[l % 2]
            self._conv(x, x_mode, self.Wk[l], self.bk[l], None, y, _native.CNN_EPI_BIAS_RELU, shape)
            acts.append(y)
            x, x_mode = y, _native.CNN_IN_BF16X32
        out = torch.empty_like(img)
        wl, bl = self.Wk[5], self.bk[5]
        if scale != 1.0:
            wl = (self.W[5] * scale)
            w = torch.zeros((27, C, C), device=img.device)
            w[:, :1, :] = wl
            wl = w.to(torch.bfloat16).contiguous()
            bl = self.bk[5] * scale
        self._conv(x, x_mode, wl, bl, img, out, _native.CNN_EPI_RESIDUAL, shape)
        if keep:
            self._saved = (img, acts, shape)
        return out

    __call__ = forward

    def apply(self, img):
        """The trained map on an image in its own units: forward(img * s) / s, s = data_scale."""
        s = self.data_scale
        if s == 1.0:
            return self.forward(img)
        out = self.forward(img * s)
        return out.div_(s)

    def backward(self, dout):
        """Gradients (dW, db) of sum(out * dout) after ``forward(..., keep=True)`` with scale 1."""
        torch = _torch()
        img, acts, shape = self._saved
        nbytes = ctypes.c_size_t()
        _native.check(self.lib.tigre_cnn_wgrad_workspace_size(*shape, ctypes.byref(nbytes)), "wgrad workspace")
        ws = torch.empty(max(int(nbytes.value), 256), dtype=torch.uint8, device=img.device)
        dout = dout.float().contiguous()
        dW, db = [None] * LAYERS, [None] * LAYERS
        raw = torch.empty((28, C, C), device=img.device)
        # last layer: input h5, gradient dout (one FP32 channel)
        self._wgrad(acts[4], _native.CNN_IN_BF16X32, dout, _native.CNN_IN_F32X1, shape, ws, raw)
        dW[5] = raw[:27, :1, :].clone()
        db[5] = raw[27, :1, 0].clone()
        g = torch.empty(shape + (C,), dtype=torch.bfloat16, device=img.device)
        self._conv(dout, _native.CNN_IN_F32X1, self.Wd[5], None, acts[4], g, _native.CNN_EPI_RELU_MASK, shape)
        g2 = torch.empty_like(g)
        for l in range(LAYERS - 2, 0, -1):
            self._wgrad(acts[l - 1], _native.CNN_IN_BF16X32, g, _native.CNN_IN_BF16X32, shape, ws, raw)
            dW[l] = raw[:27].clone()
            db[l] = raw[27, :, 0].clone()
            self._conv(g, _native.CNN_IN_BF16X32, self.Wd[l], None, acts[l - 1], g2, _native.CNN_EPI_RELU_MASK, shape)
            g, g2 = g2, g
        self._wgrad(img, _native.CNN_IN_F32X1, g, _native.CNN_IN_BF16X32, shape, ws, raw)
        dW[0] = raw[:27, :, :1].clone()
        db[0] = raw[27, :, 0].clone()
        self._saved = None
        return dW, db

    # ---- training ----------------------------------------------------------
    def train_step(self, x, target, lr=1e-3, betas=(0.9, 0.999), eps=1e-8):
        """One Adam step on the mean squared error of forward(x) against target; returns the loss."""
        torch = _torch()
        out = self.forward(x, keep=True)
        diff = out - target
        loss = float(torch.mean(diff.double() ** 2))
        dW, db = self.backward(diff * (2.0 / diff.numel()))
        grads = dW + db
        params = self.W + self.b
        if self._adam is None:
            self._adam = {"t": 0, "m": [torch.zeros_like(p) for p in params], "v": [torch.zeros_like(p) for p in params]}
        a = self._adam
        a["t"] += 1
        b1, b2 = betas
        c1, c2 = 1.0 - b1 ** a["t"], 1.0 - b2 ** a["t"]
        for p, g, m, v in zip(params, grads, a["m"], a["v"]):
            m.mul_(b1).add_(g, alpha=1.0 - b1)
            v.mul_(b2).addcmul_(g, g, value=1.0 - b2)
            p.sub_(lr * (m / c1) / ((v / c2).sqrt() + eps))
        self.refresh()
        return loss


def tv_value(x, eps=1e-8):
    """Isotropic TV objective sum sqrt(|forward differences|^2 + eps) (diagnostic of the fallback rule)."""
    torch = _torch()
    d0 = torch.zeros_like(x)
    d1 = torch.zeros_like(x)
    d2 = torch.zeros_like(x)
    d0[:-1] = x[1:] - x[:-1]
    d1[:, :-1] = x[:, 1:] - x[:, :-1]
    d2[:, :, :-1] = x[:, :, 1:] - x[:, :, :-1]
    return float(torch.sqrt(d0.double() ** 2 + d1.double() ** 2 + d2.double() ** 2 + eps).sum())


class LearnedProx:
    """ASD-POCS ``regularizer``: the learned step, with plain TV as fallback.

    ``judge="input"``: the learned result y = net(x) (``step_scaled``: x +
    dtvg F(x), the form trained to imitate one TV call) is kept when

    * its TV objective is below that of the step's input x, and
    * if ``residual`` (a callable image -> data residual norm |A y - b|) is
      given, its data residual does not exceed the data residual of the
      previous step's output by more than the fraction ``tol`` (the first step
      has no previous output and passes);

    otherwise the step is minTV(x, dtvg, tviter), plain TV.

    ``judge="tv"`` (for a network trained to jump ahead of the iteration):
    plain TV's own step t = minTV(x, dtvg, tviter) is computed every time and
    judges y by the problem ASD-POCS solves, the smallest TV objective among
    images whose data residual is at most ``epsilon``.  y is kept when it is
    closer to that data constraint than t (max(|A y - b| - epsilon, 0) below
    t's), or when both meet it and y's TV objective does not exceed t's;
    otherwise t is returned.  It needs ``residual`` and ``epsilon``.

    ``attempts`` bounds the number of calls that try the network (None: all);
    with ``latch`` the first fallback ends the attempts.  The calls after the
    attempts are plain minTV, with no network or judge work.

    ``taken`` / ``fallbacks`` count the two outcomes of the attempts;
    ``history`` holds one (residual of y, residual of t, TV of y, TV of t,
    kept) tuple per attempt under ``judge="tv"``."""

    def __init__(self, net, tviter=20, residual=None, tol=1e-3, step_scaled=False, judge="input", epsilon=None,
                 attempts=None, latch=False):
        if judge not in ("input", "tv"):
            raise ValueError("judge must be 'input' or 'tv'")
        if judge == "tv" and (residual is None or epsilon is None):
            raise ValueError("judge='tv' needs residual and epsilon")
        self.net = net
        self.tviter = tviter
        self.residual = residual
        self.tol = tol
        self.step_scaled = step_scaled
        self.judge = judge
        self.epsilon = epsilon
        self.taken = 0
        self.fallbacks = 0
        self.last_residual = None
        self.history = []
        self.attempts = attempts
        self.latch = latch
        self._closed = False

    def __call__(self, img, dtvg):
        from .tv import minTV

        if self._closed or (self.attempts is not None and self.taken + self.fallbacks >= self.attempts):
            return minTV(img, dtvg, self.tviter)
        if self.latch and self.fallbacks > 0:
            self._closed = True
            return minTV(img, dtvg, self.tviter)
        y = self.net.forward(img, scale=float(dtvg)) if self.step_scaled else self.net.apply(img)
        if self.judge == "tv":
            t = minTV(img, dtvg, self.tviter)
            ry, rt = self.residual(y), self.residual(t)
            iy, it = max(ry - self.epsilon, 0.0), max(rt - self.epsilon, 0.0)
            if iy > 0.0 or it > 0.0:
                ok, ty, tt = iy < it, None, None
            else:
                ty, tt = tv_value(y), tv_value(t)
                ok = ty <= tt
            self.history.append((ry, rt, ty, tt, bool(ok)))
            if ok:
                self.taken += 1
                return y
            self.fallbacks += 1
            return t
        ok = tv_value(y) < tv_value(img)
        r = None
        if ok and self.residual is not None:
            r = self.residual(y)
            ok = self.last_residual is None or r <= self.last_residual * (1.0 + self.tol)
        if ok:
            self.taken += 1
            self.last_residual = r
            return y
        self.fallbacks += 1
        out = minTV(img, dtvg, self.tviter)
        if self.residual is not None:
            self.last_residual = self.residual(out)
        return out


def data_epsilon(proj, geo, angles, residual=None):
    """The data-residual tolerance asd_pocs uses when maxl2err is not given:
    0.2 |A FDK(proj) - proj| (pocs_algorithms.py:115-121 restated in
    tigre_amd.algorithms.asd_pocs)."""
    from .algorithms import fdk

    r = residual if residual is not None else data_residual(proj, geo, angles)
    return 0.2 * r(fdk(proj, geo, angles, return_tensor=True))


def data_residual(proj, geo, angles, projection_type="Siddon"):
    """Callable image -> |Ax(image) - proj| (FP64 norm) over all views, for the
    fallback test of :class:`LearnedProx`."""
    import numpy as np

    from .algorithms import ViewSets, _norm, _to_device

    vs = ViewSets(geo, angles, [np.arange(len(angles))])
    b = _to_device(proj)
    kind = _native.PROJ_INTERPOLATED if projection_type == "interpolated" else _native.PROJ_SIDDON
    return lambda img: _norm(vs.ax(img.contiguous(), 0, kind) - b)


# ---- data for training and evaluation ----------------------------------------
def poisson_sinogram(proj, I0=1e5, seed=0):
    """Noisy line integrals -ln(Poisson(I0 e^{-p}) / I0), counts clamped at 1."""
    torch = _torch()
    g = torch.Generator(device=proj.device).manual_seed(int(seed))
    counts = torch.poisson(I0 * torch.exp(-proj.double()), generator=g).clamp_(min=1.0)
    return (-torch.log(counts / I0)).float().contiguous()


def collect_pairs(proj, geo, angles, niter, tviter=20, alpha=0.002, **kwargs):
    """(pre-TV image, minimizeTV output, dtvg) of every iteration of plain ASD-POCS."""
    from .algorithms import asd_pocs
    from .tv import minTV

    pairs = []

    def recording(img, step):
        out = minTV(img, step, tviter)
        pairs.append((img.clone(), out.clone(), float(step)))
        return out

    asd_pocs(proj, geo, angles, niter, alpha=alpha, tviter=tviter, regularizer=recording, return_tensor=True,
             **kwargs)
    return pairs


def collect_trajectory(proj, geo, angles, niter, tviter=20, alpha=0.002, **kwargs):
    """The images of a plain ASD-POCS run: (pre-TV images, post-TV images), one
    of each per iteration (the run may stop before ``niter`` on the reference's
    criteria)."""
    from .algorithms import asd_pocs
    from .tv import minTV

    pre, post = [], []

    def recording(img, step):
        out = minTV(img, step, tviter)
        pre.append(img.clone())
        post.append(out.clone())
        return out

    asd_pocs(proj, geo, angles, niter, alpha=alpha, tviter=tviter, regularizer=recording, return_tensor=True,
             **kwargs)
    return pre, post


def jump_pairs(trajectory, ahead, first=None):
    """Pairs a plain ASD-POCS run generates further ahead: (pre-TV image of
    iteration k, post-TV image of iteration k + ``ahead`` (the run's last if it
    stopped earlier), 1.0) for the first ``first`` iterations k (all by
    default).  With make_patches the target of a crop of x_k is the same crop
    ``ahead`` TV steps and data sweeps later, so the network learns to take the
    iterate where plain ASD-POCS would arrive only after them."""
    pre, post = trajectory
    last = len(post) - 1
    ks = range(len(pre) if first is None else min(first, len(pre)))
    return [(pre[k], post[min(k + ahead, last)], 1.0) for k in ks]


def collect_far_pairs(proj, geo, angles, niter, first, tviter=20, alpha=0.002, **kwargs):
    """(pre-TV image of iteration k, the run's final post-TV image, 1.0) for
    k = 1 .. ``first`` of a plain ASD-POCS run of ``niter`` iterations."""
    pre, post = collect_trajectory(proj, geo, angles, niter, tviter, alpha, **kwargs)
    return jump_pairs((pre, post), niter, first)


def make_patches(pairs, count, size, seed=0, object_level=None, air_keep=0.2):
    """``count`` random size^3 crops: inputs x and training targets x + (post - x) / step
    (post = minimizeTV(x, dtvg) with step = dtvg, or a later iterate with step = 1).
    With ``object_level``, a crop whose target mean is below it (mostly air
    around the object) is kept only with probability ``air_keep``."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    while len(xs) < count:
        pre, post, step = pairs[int(rng.integers(len(pairs)))]
        D, H, W = pre.shape
        z, y, x = (int(rng.integers(0, max(1, n - size + 1))) for n in (D, H, W))
        a = pre[z:z + size, y:y + size, x:x + size]
        b = post[z:z + size, y:y + size, x:x + size]
        if object_level is not None and float(b.mean()) < object_level and rng.random() >= air_keep:
            continue
        xs.append(a)
        ts.append(a + (b - a) / step)
    return torch.stack(xs).contiguous(), torch.stack(ts).contiguous()


def train(net, x, target, steps, batch=8, lr=1e-3, seed=0, log=None, lr_final=None):
    """Adam on random minibatches of (x, target), both multiplied by the net's
    data_scale; the learning rate falls from lr to lr_final (default: lr) along
    a cosine.  Returns the loss history (in scaled units)."""
    import math

    torch = _torch()
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    losses = []
    n = x.shape[0]
    ds = net.data_scale
    lr_end = lr if lr_final is None else lr_final
    for s in range(steps):
        idx = torch.randint(0, n, (min(batch, n),), generator=g).to(x.device)
        xb, tb = x[idx], target[idx]
        if ds != 1.0:
            xb, tb = xb * ds, tb * ds
        rate = lr_end + 0.5 * (lr - lr_end) * (1.0 + math.cos(math.pi * s / max(1, steps - 1)))
        losses.append(net.train_step(xb.contiguous(), tb.contiguous(), lr=rate))
        if log is not None and (s % 50 == 0 or s == steps - 1):
            log("step %d loss %.4e" % (s, losses[-1]))
    return losses
