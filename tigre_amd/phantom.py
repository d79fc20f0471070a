"""Seeded ellipsoid phantoms (synthetic inputs for tests and benchmarks).

A phantom is the sum of N ellipsoids: centres uniform in the inner 60 % of
the volume box, semi-axes U[5, 40] mm, orientation uniform on SO(3) (random
unit quaternion), density U[0.1, 1].  A voxel gets an ellipsoid's density when
its centre lies inside.  Line integrals of the continuous phantom have a
closed form (``ellipsoid_line_integral``).
"""
from __future__ import annotations

import numpy as np


def ellipsoid_params(n, s_voxel, seed, axes_mm=(5.0, 40.0)):
    """Return a list of (centre(z,y,x) mm, semi-axes(3) mm, rotation 3x3, density)."""
    rng = np.random.default_rng(seed)
    s = np.asarray(s_voxel, dtype=np.float64)
    out = []
    for _ in range(n):
        c = (rng.random(3) - 0.5) * 0.6 * s
        ax = rng.uniform(axes_mm[0], axes_mm[1], size=3)
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        d = rng.uniform(0.1, 1.0)
        out.append((c, ax, R, d))
    return out


def ellipsoid_phantom(n_voxel, s_voxel, n_ellipsoids, seed, params=None):
    """Voxelise the phantom; (nZ, nY, nX) float32, voxel centres in mm about 0."""
    n_voxel = [int(v) for v in n_voxel]
    s = np.asarray(s_voxel, dtype=np.float64)
    d = s / np.asarray(n_voxel)
    axes = [(np.arange(n_voxel[i]) + 0.5) * d[i] - s[i] / 2 for i in range(3)]
    img = np.zeros(n_voxel, dtype=np.float32)
    if params is None:
        params = ellipsoid_params(n_ellipsoids, s_voxel, seed)
    for c, ax, R, dens in params:
        r = float(np.max(ax))
        sl = []
        for i in range(3):
            lo = np.searchsorted(axes[i], c[i] - r - d[i])
            hi = np.searchsorted(axes[i], c[i] + r + d[i])
            sl.append(slice(lo, hi))
        zz, yy, xx = np.meshgrid(axes[0][sl[0]] - c[0], axes[1][sl[1]] - c[1], axes[2][sl[2]] - c[2],
                                 indexing="ij")
        p = np.stack([zz, yy, xx], axis=-1) @ R          # body coordinates
        inside = ((p / ax) ** 2).sum(axis=-1) <= 1.0
        img[sl[0], sl[1], sl[2]] += (inside * dens).astype(np.float32)
    return img


def ellipsoid_phantom_torch(n_voxel, s_voxel, n_ellipsoids, seed, device, params=None):
    """Same phantom voxelised with torch on ``device`` (fast at 512^3 and above)."""
    import torch

    n_voxel = [int(v) for v in n_voxel]
    s = np.asarray(s_voxel, dtype=np.float64)
    d = s / np.asarray(n_voxel)
    axes = [(np.arange(n_voxel[i]) + 0.5) * d[i] - s[i] / 2 for i in range(3)]
    img = torch.zeros(n_voxel, dtype=torch.float32, device=device)
    if params is None:
        params = ellipsoid_params(n_ellipsoids, s_voxel, seed)
    for c, ax, R, dens in params:
        r = float(np.max(ax))
        sl = []
        for i in range(3):
            lo = int(np.searchsorted(axes[i], c[i] - r - d[i]))
            hi = int(np.searchsorted(axes[i], c[i] + r + d[i]))
            sl.append(slice(lo, hi))
        if any(q.stop <= q.start for q in sl):
            continue
        grids = [torch.as_tensor(axes[i][sl[i]] - c[i], dtype=torch.float64, device=device) for i in range(3)]
        zz, yy, xx = torch.meshgrid(*grids, indexing="ij")
        p = torch.stack([zz, yy, xx], dim=-1) @ torch.as_tensor(R, dtype=torch.float64, device=device)
        inside = ((p / torch.as_tensor(ax, dtype=torch.float64, device=device)) ** 2).sum(dim=-1) <= 1.0
        img[sl[0], sl[1], sl[2]] += inside.to(torch.float32) * float(dens)
    return img


def ellipsoid_line_integral(params, p0, direction):
    """Closed-form integral of the continuous phantom along p0 + t*direction (mm, z,y,x order)."""
    direction = np.asarray(direction, dtype=np.float64)
    u = direction / np.linalg.norm(direction)
    total = 0.0
    for c, ax, R, dens in params:
        o = (np.asarray(p0, dtype=np.float64) - c) @ R / ax
        v = u @ R / ax
        a = v @ v
        b = 2 * o @ v
        cc = o @ o - 1
        disc = b * b - 4 * a * cc
        if disc > 0:
            total += dens * np.sqrt(disc) / a
    return total
