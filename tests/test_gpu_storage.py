"""Narrow storage on the GPU: volumes and projections in bfloat16 / float16,
float32 arithmetic.  Reference: the FP64 oracle applied to the same rounded
inputs, so the measured error is the operator's plus the output rounding.

Tolerance (north-star precision rule): relative L2 <= 3e-3 for BF16 storage;
FP16 (11-bit significand) is held to the same bound.
"""
import numpy as np
import pytest

import oracle
from oracle import algorithms as oalg
import tigre_amd
from tigre_amd import algorithms as algs
from tigre_amd.geometry import cone_geometry, prepared
from tigre_amd.phantom import ellipsoid_phantom

pytestmark = pytest.mark.gpu
TOL = 3e-3


def _case():
    geo = cone_geometry((64, 64, 64), (64.0, 64.0, 64.0), (64, 80), (1.6, 1.6), 1000.0, 1536.0)
    angles = np.linspace(0, 2 * np.pi, 48, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 8, seed=31)
    return geo, angles, img, prepared(geo, angles)


@pytest.mark.parametrize("dtype", ["bfloat16", "float16"])
@pytest.mark.parametrize("ptype", ["Siddon", "interpolated"])
def test_ax_storage(cuda, dtype, ptype):
    import torch

    geo, angles, img, gp = _case()
    t = torch.from_numpy(img).to(cuda).to(getattr(torch, dtype))
    got = tigre_amd.Ax(t, geo, angles, ptype)
    assert got.dtype == t.dtype
    ref = oracle.ax(t.float().cpu().numpy(), gp, ptype)
    err = oracle.rel_l2(got.float().cpu().numpy(), ref)
    print("Ax %s %s rel-L2 %.3e" % (ptype, dtype, err))
    assert err <= TOL


@pytest.mark.parametrize("dtype", ["bfloat16", "float16"])
@pytest.mark.parametrize("btype", ["FDK", "matched"])
def test_atb_storage(cuda, dtype, btype):
    import torch

    geo, angles, img, gp = _case()
    proj = oracle.ax(img, gp).astype(np.float32)
    t = torch.from_numpy(proj).to(cuda).to(getattr(torch, dtype))
    got = tigre_amd.Atb(t, geo, angles, btype)
    assert got.dtype == t.dtype
    ref = oracle.atb(t.float().cpu().numpy(), gp, btype)
    err = oracle.rel_l2(got.float().cpu().numpy(), ref)
    print("Atb %s %s rel-L2 %.3e" % (btype, dtype, err))
    assert err <= TOL


# Bounds of the final CGLS iterate (8 iterations).  The data themselves are
# stored narrow, so the reference is the FP64 oracle's CGLS on the same rounded
# sinogram (as for the operators: the oracle of the same rounded inputs); the
# rounding of the data alone moves the FP64 iterate by ~4e-3 (BF16) / ~5e-4
# (FP16) from the FP32-data iterate, which no arithmetic can recover.
# Float32 iterate with narrow working vectors: the north-star bound for BF16
# (3e-3), and 1e-3 for FP16.  Narrow iterate: rounding x after each of the 8
# updates adds up to 2^-8 (BF16) / 2^-11 (FP16) relative error each time,
# bounded at 1e-2 / 2e-3.
CGLS_TOL = {"bfloat16": (3e-3, 1e-2), "float16": (1e-3, 2e-3)}


@pytest.mark.parametrize("dtype", ["bfloat16", "float16"])
def test_cgls_storage(cuda, dtype):
    """CGLS with narrow storage against the FP64 oracle CGLS on the same
    rounded data, beside the float32-storage run: the default keeps the
    iterate in float32 (working vectors narrow); narrow_iterate=True stores
    the iterate narrow too."""
    import torch

    geo, angles, img, gp = _case()
    proj = oracle.ax(img, gp).astype(np.float32)
    dt = getattr(torch, dtype)
    proj_q = torch.from_numpy(proj).to(dt).float().numpy()  # the sinogram as stored
    ref, _ = oalg.cgls(proj, geo, angles, 8)
    ref_q, _ = oalg.cgls(proj_q, geo, angles, 8)
    x32 = algs.cgls(proj, geo, angles, 8)
    xm = algs.cgls(proj, geo, angles, 8, storage=dt)
    xn = algs.cgls(proj, geo, angles, 8, storage=dt, narrow_iterate=True)
    e32 = oracle.rel_l2(x32, ref)
    em, en = (oracle.rel_l2(x, ref_q) for x in (xm, xn))
    print("CGLS 8 it: float32 %.3e; %s data vs FP32 data (FP64 oracle) %.3e; %s working vectors + float32 "
          "iterate %.3e, all %s %.3e (vs the oracle on the rounded data)"
          % (e32, dtype, oracle.rel_l2(ref_q, ref), dtype, em, dtype, en))
    assert e32 <= 1e-4
    assert em <= CGLS_TOL[dtype][0]
    assert en <= CGLS_TOL[dtype][1]
