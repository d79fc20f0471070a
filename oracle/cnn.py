"""FP64 restatement of the learned proximal network (test infrastructure only).

The network has no counterpart in the reference: it stands in for the TV
step (minimizeTV, GD_TV.hpp:58) inside ASD-POCS (pocs_algorithms.py:15-185).
This module states, in numpy FP64, what the device kernels compute:

* ``conv3d(x, w)``: y[v][o] = sum_tap sum_i x[v + tap][i] w[tap][o][i] over the
  27 taps (dz, dy, dx) in {-1, 0, 1}^3 (index 9(dz+1) + 3(dy+1) + (dx+1)),
  zero outside the volume; x is (N, D, H, W, C) channel-last.
* ``conv3d_data_grad`` / ``conv3d_weight_grad``: the adjoints of conv3d in x
  and in w.
* ``forward`` / ``backward``: the 6-layer residual network
  out = img + s (conv6(h5) + b6), h_l = relu(conv_l(h_{l-1}) + b_l), h0 = img,
  with the device's storage rounding (bf16 round-to-nearest-even of the
  network input, of every activation and of every activation gradient) when
  ``bf16=True``, and no rounding at all when False.

Pinned (tests/test_oracle_cnn.py) by: scipy.ndimage.correlate for one
channel, the impulse response (a delta input reproduces the kernel, mirrored),
the identity kernel, finite differences of the unrounded network for every
parameter gradient, and the adjoint identity <conv(x), y> = <x, conv^T(y)>.
"""
from __future__ import annotations

import numpy as np

TAPS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def bf16_round(a):
    """Round to the nearest bfloat16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = (u & 0x7F800000) == 0x7F800000
    r = np.where(nan, u & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _shifted(x, dz, dy, dx):
    """x[v + (dz, dy, dx)] with zeros outside, x: (N, D, H, W, C)."""
    p = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    D, H, W = x.shape[1:4]
    return p[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :]


def conv3d(x, w):
    """w: (27, Cout, Cin) -> (N, D, H, W, Cout)."""
    y = 0.0
    for t, (dz, dy, dx) in enumerate(TAPS):
        y = y + _shifted(x, dz, dy, dx) @ w[t].T
    return y


def conv3d_data_grad(g, w):
    """Adjoint of conv3d in x: dx[v][i] = sum_tap sum_o g[v - tap][o] w[tap][o][i]."""
    out = 0.0
    for t, (dz, dy, dx) in enumerate(TAPS):
        out = out + _shifted(g, -dz, -dy, -dx) @ w[t]
    return out


def conv3d_weight_grad(x, g):
    """Adjoint of conv3d in w: dw[tap][o][i] = sum_v x[v + tap][i] g[v][o]."""
    C_out, C_in = g.shape[-1], x.shape[-1]
    dw = np.zeros((27, C_out, C_in))
    gf = g.reshape(-1, C_out)
    for t, (dz, dy, dx) in enumerate(TAPS):
        dw[t] = gf.T @ _shifted(x, dz, dy, dx).reshape(-1, C_in)
    return dw


def forward(img, W, b, scale=1.0, bf16=True):
    """img: (N, D, H, W) float; W: 6 arrays (27, Cout, Cin) (Cin 1 for the first,
    Cout 1 for the last); b: 6 bias vectors.  Returns (out, activations)."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    img = np.asarray(img, dtype=np.float64)
    h = [rnd(img)[..., None]]
    for l in range(5):
        pre = conv3d(h[-1], W[l]) + b[l]
        h.append(rnd(np.maximum(pre, 0.0)))
    out = img + scale * (conv3d(h[-1], W[5])[..., 0] + b[5][0])
    return out, h


def backward(dout, h, W, scale=1.0, bf16=True):
    """Gradients of sum(out * dout) w.r.t. W and b (lists as in forward)."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    g = rnd(np.asarray(dout, dtype=np.float64) * scale)[..., None]
    dW, db = [None] * 6, [None] * 6
    dW[5] = conv3d_weight_grad(h[5], g)
    db[5] = g.reshape(-1, 1).sum(axis=0)
    for l in range(5, 0, -1):
        # gradient w.r.t. h_l (the input of conv l+1), masked by relu'(pre_l)
        g = rnd(conv3d_data_grad(g, W[l]) * (h[l] > 0))
        dW[l - 1] = conv3d_weight_grad(h[l - 1], g)
        db[l - 1] = g.reshape(-1, g.shape[-1]).sum(axis=0)
    return dW, db
