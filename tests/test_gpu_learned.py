"""GPU: the learned proximal step inside ASD-POCS (small scale).

A network trained for a few hundred steps on (pre-TV, minimizeTV) pairs of
plain ASD-POCS runs must (1) do the TV step's job better than no step on
held-out pairs, (2) give ASD-POCS a lower image RMSE than the same loop with
no regularization step, within 10 % of plain ASD-POCS, and (3) hand over to
plain TV, exactly, when its step would raise the TV objective.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, VIEWS, ITERS = 48, 60, 8


@pytest.fixture(scope="module")
def setup():
    import torch

    import tigre_amd
    from tigre_amd import learned as L
    from tigre_amd.geometry import cone_geometry
    from tigre_amd.phantom import ellipsoid_phantom_torch

    geo = cone_geometry((N, N, N), (float(N),) * 3, (N, N), (1.6, 1.6), 1000.0, 1536.0)
    angles = np.linspace(0, 2 * np.pi, VIEWS, endpoint=False)

    def case(seed):
        img = 0.02 * ellipsoid_phantom_torch(geo.nVoxel, geo.sVoxel, 40, seed=seed, device="cuda")
        proj = tigre_amd.Ax(img, geo, angles, "Siddon")
        return img, L.poisson_sinogram(proj, 1e5, seed=seed + 1)

    pairs = []
    for s in (31, 32, 33):
        pairs += L.collect_pairs(case(s)[1], geo, angles, ITERS)
    x, t = L.make_patches(pairs, 400, 20, seed=2)
    net = L.ResidualCNN(seed=1)
    losses = L.train(net, x, t, 300, batch=8, lr=1e-3)
    torch.cuda.synchronize()
    return geo, angles, case, net, losses


def _rmse(a, b):
    import torch

    return float(torch.sqrt(torch.mean((a - b).double() ** 2)))


def test_network_approximates_tv_step(cuda, setup):
    from tigre_amd import learned as L

    geo, angles, case, net, losses = setup
    assert losses[-1] < losses[0]
    val = L.collect_pairs(case(41)[1], geo, angles, 4)
    for pre, post, step in val:
        y = net.forward(pre, scale=step)
        e_net, e_id = float((y - post).norm()), float((pre - post).norm())
        print("held-out: |net - TV| %.3e vs |x - TV| %.3e" % (e_net, e_id))
        assert e_net < 0.7 * e_id


def test_learned_asd_pocs(cuda, setup):
    from tigre_amd import algorithms as algs
    from tigre_amd import learned as L

    geo, angles, case, net, _ = setup
    img, b = case(42)
    out = {}
    learned = L.LearnedProx(net, 20, step_scaled=True)  # trained on one TV call: its step scales F
    for name, reg in (("plain", None), ("learned", learned), ("none", lambda x, s: x)):
        res = algs.asd_pocs(b, geo, angles, ITERS, regularizer=reg, return_tensor=True)
        out[name] = _rmse(res, img)
        if name == "learned":
            print("learned step taken %d, fallbacks %d" % (reg.taken, reg.fallbacks))
            assert reg.taken > 0
    print("RMSE plain %.5f learned %.5f no step %.5f" % (out["plain"], out["learned"], out["none"]))
    assert out["learned"] < out["none"]
    assert out["learned"] <= 1.1 * out["plain"]


def test_fallback_is_plain_tv(cuda, setup):
    import torch

    import tigre_amd
    from tigre_amd import learned as L

    geo, angles, case, net, _ = setup
    bad = L.ResidualCNN(seed=5)
    bad.W[5] = bad.W[5] * 0 + 1.0  # a large positive last layer: adds structure, raises TV
    bad.refresh()
    prox = L.LearnedProx(bad, 20)
    x = case(43)[0] + 0.001
    y = prox(x, 1e-3)
    assert prox.fallbacks == 1 and prox.taken == 0
    assert torch.equal(y, tigre_amd.minTV(x, 1e-3, 20))
