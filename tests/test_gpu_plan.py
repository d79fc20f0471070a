"""GPU: projection plans, the fused SART steps, graph-batched OS-SART and the autotuner.

Bit-exact where the same kernels and rounding steps run (plan vs per-call
entry points at the default parameters; any views-per-group / slab depth,
since each ray's sum runs in the same order; graph replay vs host launches);
rel-L2 1e-6 where only the backprojector's slices per lane change rounding;
the batched reconstructions against the FP64 oracle's OS-SART at 1e-4, the
tolerance of the single-case OS-SART test.
"""
import numpy as np
import pytest

import oracle
import oracle.algorithms as oalg
import tigre_amd
from tigre_amd import _native
from tigre_amd import algorithms as algs
from tigre_amd.algorithms.batched import BatchedOSSART
from tigre_amd.geometry import cone_geometry, prepared
from tigre_amd.phantom import ellipsoid_phantom
from tigre_amd.plan import Autotuner, Plan, sart_residual

pytestmark = pytest.mark.gpu


def _t(a):
    import torch

    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")


@pytest.fixture(scope="module")
def case():
    geo = cone_geometry((40, 36, 44), (44.0, 36.0, 40.0), (48, 56), (1.1, 1.0), 300.0, 450.0)
    angles = np.linspace(0.1, 2 * np.pi + 0.1, 37, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 5, seed=7)
    return geo, angles, img


@pytest.mark.parametrize("kind", ["Siddon", "interpolated"])
def test_plan_ax_matches_entry_point(cuda, case, kind):
    import torch

    geo, angles, img = case
    x = _t(img)
    ref = tigre_amd.Ax(x, geo, angles, kind)
    ptype = _native.PROJ_SIDDON if kind == "Siddon" else _native.PROJ_INTERPOLATED
    for vpg in (1, 2, 4):
        for planes in (1, 2, 4, 8):
            p = Plan(geo, angles, vpg, planes, 16)
            ws = p.workspace()
            out = torch.full(p.proj_shape, float("nan"), device="cuda")
            p.ax(x, out, ws, ptype)
            torch.cuda.synchronize()
            print("ax %s vpg %d planes %d max |diff| %.3e" % (kind, vpg, planes, float((out - ref).abs().max())))
            assert torch.equal(out, ref)
            p.close()


@pytest.mark.parametrize("kind", ["FDK", "matched"])
def test_plan_atb_slices(cuda, case, kind):
    import torch

    geo, angles, img = case
    proj = tigre_amd.Ax(_t(img), geo, angles, "Siddon")
    ref = tigre_amd.Atb(proj, geo, angles, kind)
    btype = _native.BACKPROJ_FDK if kind == "FDK" else _native.BACKPROJ_MATCHED
    for z in (16, 8, 4):
        p = Plan(geo, angles, 4, 8, z)
        out = torch.full(p.vol_shape, float("nan"), device="cuda")
        p.atb(proj, out, p.workspace(), btype)
        torch.cuda.synchronize()
        err = oracle.rel_l2(out.cpu().numpy(), ref.cpu().numpy())
        print("atb %s slices %d rel-L2 vs entry point %.3e" % (kind, z, err))
        if z == 16:
            assert torch.equal(out, ref)
        else:
            assert err <= 1e-6


def test_plan_narrow_storage(cuda, case):
    import torch

    geo, angles, img = case
    x = _t(img).to(torch.bfloat16)
    p = Plan(geo, angles)
    ws = p.workspace(torch.bfloat16)
    out = torch.empty(p.proj_shape, dtype=torch.bfloat16, device="cuda")
    p.ax(x, out, ws)
    ref = tigre_amd.Ax(x, geo, angles, "Siddon")
    vol = torch.empty(p.vol_shape, dtype=torch.bfloat16, device="cuda")
    p.atb(ref, vol, ws)
    vref = tigre_amd.Atb(ref, geo, angles, "FDK")
    torch.cuda.synchronize()
    assert torch.equal(out, ref) and torch.equal(vol, vref)


def test_fused_sart_steps(cuda, case):
    """Residual and update kernels against the tensor expressions of the eager loop, bit for bit."""
    import torch

    geo, angles, img = case
    g = torch.Generator(device="cuda").manual_seed(3)
    p = Plan(geo, angles)
    ws = p.workspace()
    x = _t(img) * 0.5
    measured = torch.rand(p.proj_shape, generator=g, device="cuda")
    w = torch.rand(p.proj_shape, generator=g, device="cuda") + 0.5
    V = torch.rand(p.vol_shape[1:], generator=g, device="cuda") + 0.5
    inv_v = (0.7 / V).contiguous()
    ax = torch.empty(p.proj_shape, device="cuda")
    p.ax(x, ax, ws)
    ref_r = w * (measured - ax)
    sart_residual(measured, w, ax)
    assert torch.equal(ax, ref_r)
    ref_x = torch.clamp(x + inv_v * tigre_amd.Atb(ref_r, geo, angles, "FDK"), min=0)
    p.sart_update(ax, x, inv_v, ws, noneg=True)
    torch.cuda.synchronize()
    assert torch.equal(x, ref_x)


def _cases(n, seed, tilted=False):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        edge = int(rng.choice([16, 24, 32]))
        nang = int(rng.integers(20, 50))
        dso = float(rng.uniform(200.0, 300.0))
        geo = cone_geometry((edge,) * 3, (float(edge),) * 3, (edge, edge), (1.5, 1.5), dso, 1.5 * dso)
        angles = np.linspace(0, 2 * np.pi, nang, endpoint=False)
        if tilted and k == 0:  # out-of-plane views: the backprojection runs outside the staged multi-job kernel
            angles = np.stack([angles, np.full(nang, 0.1), np.zeros(nang)], axis=1)
        img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 4, seed=100 + k)
        proj = oracle.ax(img, prepared(geo, angles)).astype(np.float32)
        out.append((proj, geo, angles))
    return out


@pytest.fixture(scope="module")
def batch():
    return _cases(5, seed=11, tilted=True)


def test_batched_matches_ossart(cuda, batch):
    """Wave batches (every case's subset step in one launch per stage; cases
    with fewer subsets drop out of later waves) give ossart()'s bits, eagerly
    and from three graph replays with eager work between them."""
    import torch

    bo = BatchedOSSART([(_t(p), g, a) for p, g, a in batch], niter=2, blocksize=8)
    print("waves: %d" % max(2 * len(c.plans) for c in bo.cases), flush=True)
    got = [r.clone() for r in bo.run()]
    torch.cuda.synchronize()
    refs = []
    for k, ((p, g, a), r) in enumerate(zip(batch, got)):
        ref = algs.ossart(p, g, a, 2, blocksize=8, return_tensor=True)
        refs.append(ref)
        assert torch.equal(r, ref), k
    bo.reset()
    seq = [r.clone() for r in bo.run_sequential()]
    for r, ref in zip(seq, refs):
        assert torch.equal(r, ref)
    bo.capture()
    for rep in range(3):
        bo.reset()
        torch.cuda.synchronize()
        scratch = torch.randn(1000, device="cuda").sum()  # eager work between replays
        got = bo.replay()
        torch.cuda.synchronize()
        assert float(scratch) == float(scratch)
        for r, ref in zip(got, refs):
            assert torch.equal(r, ref), "replay %d differs" % rep
    for k in (0, 1):
        ora = oalg.os_sart(batch[k][0], batch[k][1], batch[k][2], 2, blocksize=8)
        err = oracle.rel_l2(got[k].cpu().numpy(), ora)
        print("batched OS-SART case %d rel-L2 vs oracle %.3e" % (k, err))
        assert err <= 1e-4
    bo.close()


def test_autotuned_batch(cuda, batch):
    tuner = Autotuner(reps=2)
    bo = BatchedOSSART([(_t(p), g, a) for p, g, a in batch[1:4]], niter=2, blocksize=8, tuner=tuner)
    for c in bo.cases:
        vpg, planes, z = c.params
        assert vpg in (1, 2, 4) and planes in (2, 4, 8) and z in (4, 8, 16)
    print("tuned", {k: v for k, v in tuner.table.items()})
    got = [r.cpu().numpy() for r in bo.replay()]
    for (p, g, a), r in zip(batch[1:4], got):
        ref = oalg.os_sart(p, g, a, 2, blocksize=8)
        assert oracle.rel_l2(r, ref) <= 1e-4
    bo.close()
