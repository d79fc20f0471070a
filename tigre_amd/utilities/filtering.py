"""FDK pre-processing filter, Python/tigre/utilities/filtering.py:15-87.

The filter's frequency response is built exactly as the reference builds it
(spatial ramp kernel h[0] = 1/4, h[odd n] = -1/(pi n)^2, |FFT| * 2, optional
window, zero above the Nyquist band, mirrored), then applied to zero-padded
detector rows on the device with a real FFT (rocFFT through torch.fft),
scaled by (DSD/DSO) (2 pi / nangles) / (4 dU).
"""
from __future__ import annotations

import numpy as np


def _nextpow2(n):
    i = 1
    while 2 ** i <= n:
        i += 1
    return i


def ramp_filter(n_u, filter_name=None):
    """(filter, padded length) for detectors n_u pixels wide."""
    filt_len = max(64, 2 ** _nextpow2(2 * int(n_u)))
    nn = np.arange(-filt_len / 2, filt_len / 2)
    h = np.zeros(nn.shape, dtype=np.float32)
    h[int(filt_len / 2)] = 1 / 4
    odd = nn % 2 == 1
    h[odd] = -1 / (np.pi * nn[odd]) ** 2
    resp = np.abs(np.fft.fft(h)) * 2
    filt = resp[: int(filt_len / 2 + 1)]
    w = 2 * np.pi * np.arange(len(filt)) / filt_len
    d = 1
    if filter_name in ("ram_lak", None):
        pass
    elif filter_name == "shepp_logan":
        filt[1:] *= np.sin(w[1:] / (2 * d)) / (w[1:] / (2 * d))
    elif filter_name == "cosine":
        filt[1:] *= np.cos(w[1:] / (2 * d))
    elif filter_name == "hamming":
        filt[1:] *= 0.54 + 0.46 * np.cos(w[1:] / d)
    elif filter_name == "hann":
        filt[1:] *= (1 + np.cos(w[1:] / d)) / 2
    else:
        raise ValueError("filter not recognized: " + str(filter_name))
    filt[w > np.pi * d] = 0
    return np.hstack((filt, filt[1:-1][::-1])).astype(np.float32), filt_len


def filtering(proj, geo, angles, parker=False, verbose=False):
    """Filter cosine-weighted projections (torch tensor on the GPU, (n, nV, nU))."""
    import torch

    if parker:
        # as the reference: weights applied to the (views, u, v) transpose
        from .parkerweight import parkerweight

        proj = parkerweight(proj.transpose(1, 2).contiguous(), geo, angles, parker).transpose(1, 2).contiguous()
    n_u = int(geo.nDetector[1])
    filt, filt_len = ramp_filter(n_u, getattr(geo, "filter", None))
    pad = int((filt_len - n_u) // 2)
    DSD0 = float(np.asarray(geo.DSD).reshape(-1)[0])
    DSO0 = float(np.asarray(geo.DSO).reshape(-1)[0])
    scale = (DSD0 / DSO0) * (2 * np.pi / len(angles)) / (4 * float(geo.dDetector[1]))
    # the mirrored response is real and even: the half spectrum drives an rfft/irfft pair
    half = torch.as_tensor(filt[: filt_len // 2 + 1], device=proj.device)
    buf = torch.zeros(proj.shape[0], proj.shape[1], filt_len, dtype=torch.float32, device=proj.device)
    buf[:, :, pad:pad + n_u] = proj
    spec = torch.fft.rfft(buf, dim=2)
    spec *= half
    out = torch.fft.irfft(spec, n=filt_len, dim=2)[:, :, pad:pad + n_u]
    return (out * scale).contiguous()
