"""Parker short-scan redundancy weights (Wesarg, Ebert, Bortfeld, "Parker
weights revisited", Med. Phys. 29(3) 2002), as Python/tigre/utilities/
parkerweight.py:23-78 computes them.

For view ii at gantry angle beta = |angle_ii - angle_0| and fan angle alpha of
each detector position along the first detector axis (alpha = -atan(position /
mean DSD)), with delta = half the fan, epsilon = max(scan - (pi + 2 delta), 0):

    w = 1/2 [ S(beta / b(alpha) - 1/2) + S((beta - 2 delta + 2 alpha - epsilon) / b(alpha) + 1/2)
            - S((beta - pi + 2 alpha) / b(-alpha) - 1/2) - S((beta - pi - 2 delta - epsilon) / b(-alpha) + 1/2) ]

with b(alpha) = q (2 delta - 2 alpha + epsilon) and S the smooth step
S(x) = 0 (x <= -1/2), (1 + sin(pi x)) / 2 (|x| < 1/2), 1 (x >= 1/2).
The weight row of length nDetector[0] multiplies proj[ii] tiled along its
first axis (the reference's layout convention, kept as is).
"""
from __future__ import annotations

import warnings

import numpy as np


def s_function(x):
    x = np.asarray(x, dtype=np.float64)
    return np.where(x <= -0.5, 0.0, np.where(x >= 0.5, 1.0, 0.5 * (1.0 + np.sin(np.pi * x))))


def b_function(alpha, delta, epsilon):
    return 2 * delta - 2 * alpha + epsilon


def b_subf(alpha, delta, epsilon, q):
    return q * b_function(alpha, delta, epsilon)


def parker_weights(geo, angles, q):
    """(nviews, nDetector[0]) array of weights."""
    n = int(round(float(geo.sDetector[0]) / float(geo.dDetector[0])))
    pos = -float(geo.sDetector[0]) / 2 + float(geo.dDetector[0]) / 2 + float(geo.dDetector[0]) * np.arange(n)
    alpha = -np.arctan(pos / np.mean(geo.DSD))
    delta = abs(alpha[0] - alpha[-1]) / 2
    angles = np.asarray(angles, dtype=np.float64).reshape(len(angles), -1)[:, 0]
    total = abs(angles[0] - angles[-1])
    if total > 2 * np.pi:
        warnings.warn("Computing Parker weights for scanning angle equal or bigger than 2*pi "
                      "Consider disabling Parker weights.")
    if total < np.pi + 2 * delta:
        warnings.warn("Scanning angles smaller than pi+cone_angle. This is limited angle tomography, \n"
                      "there is no sufficient data, thus weighting for data redundancy is not required.")
    epsilon = max(total - (np.pi + 2 * delta), 0)
    bp, bm = b_subf(alpha, delta, epsilon, q), b_subf(-alpha, delta, epsilon, q)
    out = np.empty((len(angles), n))
    for ii in range(len(angles)):
        beta = abs(angles[ii] - angles[0])
        out[ii] = 0.5 * (s_function(beta / bp - 0.5) + s_function((beta - 2 * delta + 2 * alpha - epsilon) / bp + 0.5)
                         - s_function((beta - np.pi + 2 * alpha) / bm - 0.5)
                         - s_function((beta - np.pi - 2 * delta - epsilon) / bm + 0.5))
    return out


def parkerweight(proj, geo, angles, q):
    """Weights proj in place (NumPy array or torch tensor, views first) and returns it."""
    w = parker_weights(geo, angles, q)
    try:
        import torch

        if isinstance(proj, torch.Tensor):
            wt = torch.as_tensor(w, dtype=proj.dtype, device=proj.device)
            for ii in range(proj.shape[0]):
                proj[ii] *= wt[ii].repeat(proj.shape[-1], 1)
            return proj
    except ImportError:  # pragma: no cover - torch is part of the image
        pass
    for ii in range(proj.shape[0]):
        proj[ii] *= np.tile(w[ii], [proj.shape[-1], 1]).astype(proj.dtype)
    return proj
