"""GPU parity at the full sizes of the five benchmark workloads.

The FP64 oracle cannot run the large workloads whole in test time, so each
case is checked the way its size allows:

* 128^3 / 128 views (DSO 1000, DSD 1536 mm, 1 mm voxels, 128^2 detector at
  1.6 mm, seeded 10-ellipsoid phantom): FDK with the ram-lak filter in full
  against the oracle's FDK (rel-L2 <= 1e-5).
* 512^3 / 360 views (0.5 mm voxels, 512^2 detector at 0.8 mm, 30 ellipsoids):
  the four operators of the full 360-view calls on a fixed 8-view subset
  against the oracle (<= 1e-5); OS-SART (blocksize 20, lambda 1) on the same
  geometry at 128^3 for 2 iterations against the oracle (<= 1e-4); at 512^3,
  two OS-SART iterations on projection plans equal the plain per-call loop
  bit for bit and lower the error against the phantom.
* 1024^3 / 720 views (0.25 mm voxels, 1024 x 768 detector at 0.4 mm, per-view
  detector offsets jittered +-2 pixels): Ax Siddon and Atb matched stored as
  FP32, BF16 and FP16 on 2 views against the oracle of the same rounded inputs
  (FP32 <= 1e-5, BF16 <= 3e-3, FP16 <= 1e-3); CGLS with narrow working vectors
  and a float32 iterate against the float32-storage run after 4 iterations
  (<= 3e-3, the BF16 bound).
* 64 uneven OS-SART cases (edge 64-256, 60-360 views, DSO 800-1200 mm, 10
  iterations, 20-view subsets): the batched runner equals ossart() on every
  case bit for bit; the two smallest cases against the oracle's OS-SART (<= 1e-4).
* the learned ASD-POCS step at 256^3 / 180 views is covered in
  tests/test_gpu_learned.py.
"""
import numpy as np
import pytest

import oracle
from oracle import algorithms as oalg
import tigre_amd
from tigre_amd import algorithms as algs
from tigre_amd.algorithms import _sub_geometry
from tigre_amd.geometry import cone_geometry, prepared
from tigre_amd.phantom import ellipsoid_phantom, ellipsoid_phantom_torch

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _say(*a):
    print(*a, flush=True)


# ---------------------------------------------------------------------------
# 128^3 / 128 views: FDK with ram-lak
# ---------------------------------------------------------------------------
def test_fdk_ramlak_128(cuda):
    import torch

    geo = cone_geometry((128, 128, 128), (128.0, 128.0, 128.0), (128, 128), (1.6, 1.6), 1000.0, 1536.0)
    angles = np.linspace(0, 2 * np.pi, 128, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 10, seed=20261202 + 1)
    proj = oracle.ax(img, prepared(geo, angles)).astype(np.float32)
    _say("FDK ram-lak 128^3 / 128 views")
    got = algs.fdk(torch.from_numpy(proj).to(cuda), geo, angles, filter="ram_lak", return_tensor=True)
    torch.cuda.synchronize()
    ref, _ = oalg.fdk(proj, geo, angles, filter_name="ram_lak")
    err = oracle.rel_l2(got.cpu().numpy(), ref)
    _say("FDK ram-lak 128^3 rel-L2 %.3e" % err)
    assert err <= TOL


# ---------------------------------------------------------------------------
# 512^3 / 360 views
# ---------------------------------------------------------------------------
def _geo512(n=512):
    return cone_geometry((n, n, n), (256.0,) * 3, (n, n), (0.8 * 512 / n,) * 2, 1000.0, 1536.0)


def _views(geo, angles, idx):
    """The geometry with its per-view arrays expanded over all views, restricted to views idx."""
    import copy

    g = copy.deepcopy(geo)
    g.check_geo(angles)
    return _sub_geometry(g, idx)


SUB512 = np.array([k * 360 // 8 for k in range(8)])


@pytest.fixture(scope="module")
def case512(cuda):
    import torch

    geo = _geo512()
    angles = np.linspace(0, 2 * np.pi, 360, endpoint=False)
    img = ellipsoid_phantom_torch(geo.nVoxel, geo.sVoxel, 30, seed=20261202 + 2, device=cuda)
    torch.cuda.synchronize()
    return geo, angles, img


@pytest.mark.parametrize("ptype", ["Siddon", "interpolated"])
def test_ax_512_full_call_on_view_subset(cuda, case512, ptype):
    """All 360 views in one call (beam groups, slab planning at full size); the
    fixed 8-view subset against the oracle."""
    import torch

    geo, angles, img = case512
    _say("Ax %s 512^3 / 360 views" % ptype)
    got = tigre_amd.Ax(img, geo, angles, ptype)[torch.as_tensor(SUB512, device=cuda)].cpu().numpy()
    ref = oracle.ax(img.cpu().numpy(), prepared(_views(geo, angles, SUB512), angles[SUB512]), ptype)
    err = oracle.rel_l2(got, ref)
    _say("Ax %s 512^3 8-view subset rel-L2 %.3e" % (ptype, err))
    assert err <= TOL


@pytest.mark.parametrize("btype", ["FDK", "matched"])
def test_atb_512_view_subset(cuda, case512, btype):
    import torch

    geo, angles, img = case512
    sub_geo = _views(geo, angles, SUB512)
    gp = prepared(sub_geo, angles[SUB512])
    img_h = img.cpu().numpy()
    proj = oracle.ax(img_h, gp).astype(np.float32)
    _say("Atb %s 512^3 on 8 views" % btype)
    got = tigre_amd.Atb(torch.from_numpy(proj).to(cuda), sub_geo, angles[SUB512], btype).cpu().numpy()
    ref = oracle.atb(proj, gp, btype)
    err = oracle.rel_l2(got, ref)
    _say("Atb %s 512^3 8 views rel-L2 %.3e" % (btype, err))
    assert err <= TOL


def test_ossart_512_geometry_at_128(cuda):
    """OS-SART of the 512^3 / 360-view workload (blocksize 20, lambda 1) with
    the volume and detector at 128: 2 iterations against the oracle."""
    geo = _geo512(128)
    angles = np.linspace(0, 2 * np.pi, 360, endpoint=False)
    img = ellipsoid_phantom(geo.nVoxel, geo.sVoxel, 30, seed=20261202 + 2)
    proj = oracle.ax(img, prepared(geo, angles)).astype(np.float32)
    _say("OS-SART 360 views at 128^3, 2 iterations")
    got = algs.ossart(proj, geo, angles, 2, blocksize=20)
    ref = oalg.os_sart(proj, geo, angles, 2, blocksize=20)
    err = oracle.rel_l2(got, ref)
    _say("OS-SART 128^3 / 360 views rel-L2 %.3e" % err)
    assert err <= 1e-4


def test_ossart_512_plans_equal_plain_loop(cuda, case512):
    """At full size the plan-based sweep (records uploaded once, fused residual
    and update) gives the bits of the per-call loop, and the error against the
    phantom falls with the iterations."""
    import torch

    from tigre_amd.algorithms import _SartState

    geo, angles, img = case512
    proj = tigre_amd.Ax(img, geo, angles, "Siddon")
    errs = []
    states = [_SartState(proj, geo, angles, 20, use_plans=u) for u in (True, False)]
    for it in range(2):
        _say("OS-SART 512^3 iteration %d (plans and per-call loop)" % it)
        for st in states:
            st.sweep()
        torch.cuda.synchronize()
        assert torch.equal(states[0].res, states[1].res)
        errs.append(float(torch.linalg.vector_norm((states[0].res - img).double())
                          / torch.linalg.vector_norm(img.double())))
    _say("OS-SART 512^3 error vs phantom %s" % errs)
    assert errs[1] < errs[0] < 1.0


# ---------------------------------------------------------------------------
# 1024^3 / 720 views, narrow storage
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def case1024(cuda):
    import torch

    nviews = 720
    geo = cone_geometry((1024,) * 3, (256.0,) * 3, (768, 1024), (0.4, 0.4), 1000.0, 1536.0)
    rng = np.random.default_rng(20261202 + 3)
    geo.offDetector = rng.uniform(-2.0, 2.0, (nviews, 2)) * np.asarray(geo.dDetector)
    angles = np.linspace(0, 2 * np.pi, nviews, endpoint=False)
    img = ellipsoid_phantom_torch(geo.nVoxel, geo.sVoxel, 30, seed=20261202 + 3, device=cuda)
    torch.cuda.synchronize()
    return geo, angles, img


STORAGE_TOL = {"float32": 1e-5, "bfloat16": 3e-3, "float16": 1e-3}


@pytest.mark.parametrize("dtype", ["float32", "bfloat16", "float16"])
def test_storage_1024_two_views(cuda, case1024, dtype):
    import torch

    geo, angles, img = case1024
    sub = np.array([0, 359])
    sub_geo = _views(geo, angles, sub)
    gp = prepared(sub_geo, angles[sub])
    t = img.to(getattr(torch, dtype))
    _say("Ax/Atb 1024^3 %s on 2 views" % dtype)
    ax = tigre_amd.Ax(t, sub_geo, angles[sub], "Siddon")
    assert ax.dtype == t.dtype
    ref_ax = oracle.ax(t.float().cpu().numpy(), gp, "Siddon")
    e_ax = oracle.rel_l2(ax.float().cpu().numpy(), ref_ax)
    p = torch.from_numpy(ref_ax.astype(np.float32)).to(cuda).to(t.dtype)
    atb = tigre_amd.Atb(p, sub_geo, angles[sub], "matched")
    ref_atb = oracle.atb(p.float().cpu().numpy(), gp, "matched")
    e_atb = oracle.rel_l2(atb.float().cpu().numpy(), ref_atb)
    _say("1024^3 %s: Ax rel-L2 %.3e, Atb matched rel-L2 %.3e" % (dtype, e_ax, e_atb))
    assert e_ax <= STORAGE_TOL[dtype] and e_atb <= STORAGE_TOL[dtype]


def test_cgls_1024_narrow_vectors(cuda, case1024):
    """CGLS at full size with BF16 working vectors and a float32 iterate stays
    within the BF16 bound of the float32-storage run."""
    import torch

    geo, angles, img = case1024
    proj = tigre_amd.Ax(img, geo, angles, "Siddon")
    _say("CGLS 1024^3 / 720 views, 4 iterations, float32 storage")
    x32 = algs.cgls(proj, geo, angles, 4, return_tensor=True)
    _say("CGLS 1024^3 / 720 views, 4 iterations, bfloat16 working vectors")
    xb = algs.cgls(proj, geo, angles, 4, storage=torch.bfloat16, return_tensor=True)
    err = float(torch.linalg.vector_norm((xb - x32).double()) / torch.linalg.vector_norm(x32.double()))
    _say("CGLS 1024^3 bf16 working vectors vs float32 run rel-L2 %.3e" % err)
    assert err <= 3e-3


# ---------------------------------------------------------------------------
# 64 uneven OS-SART cases
# ---------------------------------------------------------------------------
def batched_cases(device, count=64):
    """The 64 seeded cases of the batched workload: (proj, geo, angles, phantom)."""
    import tigre_amd as T

    rng = np.random.default_rng(20261202 + 4)
    out = []
    for k in range(count):
        edge = int(rng.choice([64, 96, 128, 192, 256]))
        nviews = int(rng.integers(60, 361))
        dso = float(rng.uniform(800.0, 1200.0))
        geo = cone_geometry((edge,) * 3, (float(edge),) * 3, (edge, edge), (1.6, 1.6), dso, 1.536 * dso)
        angles = np.linspace(0, 2 * np.pi, nviews, endpoint=False)
        img = ellipsoid_phantom_torch(geo.nVoxel, geo.sVoxel, 10, seed=20261202 + 4 + 1000 * k, device=device)
        out.append((T.Ax(img, geo, angles, "Siddon"), geo, angles, img))
    return out


def test_batched_64_cases(cuda):
    import torch

    from tigre_amd.algorithms.batched import BatchedOSSART

    cases = batched_cases(cuda)
    niter = 10
    _say("batched OS-SART: 64 cases, %d iterations" % niter)
    bo = BatchedOSSART([(p, g, a) for p, g, a, _ in cases], niter=niter, blocksize=20)
    bo.run()
    torch.cuda.synchronize()
    got = [r.clone() for r in bo.results()]
    bo.close()
    for k, (p, g, a, _) in enumerate(cases):
        ref = algs.ossart(p, g, a, niter, blocksize=20, return_tensor=True)
        assert torch.equal(got[k], ref), k
    _say("batched OS-SART: 64 cases bit-identical to ossart()")
    # the two smallest problems against the oracle
    sizes = [int(np.prod(g.nVoxel)) * len(a) for _, g, a, _ in cases]
    for k in np.argsort(sizes, kind="stable")[:2]:
        p, g, a, _ = cases[k]
        _say("batched case %d (%d^3, %d views) vs oracle" % (k, int(g.nVoxel[0]), len(a)))
        ref = oalg.os_sart(p.cpu().numpy(), g, a, niter, blocksize=20)
        err = oracle.rel_l2(got[k].cpu().numpy(), ref)
        _say("batched case %d rel-L2 vs oracle %.3e" % (k, err))
        assert err <= 1e-4
