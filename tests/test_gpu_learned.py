"""GPU: the learned proximal step inside ASD-POCS (small scale).

A network trained for a few hundred steps on (pre-TV, minimizeTV) pairs of
plain ASD-POCS runs must (1) do the TV step's job better than no step on
held-out pairs, (2) give ASD-POCS a lower image RMSE than the same loop with
no regularization step, within 10 % of plain ASD-POCS, and (3) hand over to
plain TV, exactly, when its step would raise the TV objective.  A network
trained to jump ahead (pre-TV images of the first iterations -> the final
image of a longer plain run), tried in the first iterations and judged by
plain TV's own step, must (4) at the learned-step workload's size (256^3,
180 views) leave ASD-POCS closer to the phantom than plain ASD-POCS after the
iterations in which it was tried and after all 20, reaching plain's final
RMSE in fewer iterations, and (5) lose to plain TV, which is then returned
exactly, when it moves away from the data.
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


# ---------------------------------------------------------------------------
# The network trained to jump ahead of the iteration, judged by plain TV, at
# the size of the learned-step workload: 256^3, 180 views, 256^2 detector,
# Poisson noise at I0 = 1e5, ASD-POCS with tviter 20 and alpha 0.002
# ---------------------------------------------------------------------------
def test_jump_ahead_step_at_256(cuda):
    """Trained on the pre-TV images of the first 3 iterations of plain ASD-POCS
    runs against their final images (4 phantoms, 40 iterations each), tried in
    the first 3 iterations until plain TV wins once: on a held-out phantom the
    learned run is closer to the phantom than plain ASD-POCS after the
    iterations in which it was tried and after all 20, and reaches plain's
    final RMSE in fewer iterations."""
    import torch

    import tigre_amd
    from tigre_amd import algorithms as algs
    from tigre_amd import learned as L
    from tigre_amd.geometry import cone_geometry
    from tigre_amd.phantom import ellipsoid_phantom_torch

    n, niter, first = 256, 20, 3
    geo = cone_geometry((n, n, n), (float(n),) * 3, (n, n), (1.6, 1.6), 1000.0, 1536.0)
    angles = np.linspace(0, 2 * np.pi, 180, endpoint=False)
    seed0 = 20261202 + 5

    def case(k):
        img = 0.02 * ellipsoid_phantom_torch(geo.nVoxel, geo.sVoxel, 40, seed=seed0 + 100 * k, device="cuda")
        return img, L.poisson_sinogram(tigre_amd.Ax(img, geo, angles, "Siddon"), 1e5, seed=seed0 + 100 * k + 1)

    pairs = []
    for k in range(1, 5):
        pre, post = L.collect_trajectory(case(k)[1], geo, angles, 40)
        pairs += [(x, post[-1], 1.0) for x in pre[:first]]
        del pre, post
    x, t = L.make_patches(pairs, 4000, 32, seed=seed0, object_level=2e-3)
    del pairs
    net = L.ResidualCNN(seed=seed0, data_scale=50.0)
    losses = L.train(net, x, t, 6000, batch=8, lr=3e-4, seed=seed0, lr_final=3e-5)
    del x, t
    assert np.mean(losses[-200:]) < 0.5 * np.mean(losses[:200])
    img, b = case(0)
    resid = L.data_residual(b, geo, angles)
    prox = L.LearnedProx(net, 20, residual=resid, judge="tv", epsilon=L.data_epsilon(b, geo, angles, resid),
                         attempts=first, latch=True)
    curves = {}
    for name, reg in (("plain", None), ("learned", prox)):
        curves[name] = []
        algs.asd_pocs(b, geo, angles, niter, regularizer=reg, return_tensor=True,
                      callback=lambda k, r, c=curves[name]: c.append(_rmse(r, img)))
    torch.cuda.synchronize()
    print("RMSE per iteration: plain   %s" % ["%.3g" % v for v in curves["plain"]])
    print("                    learned %s (taken %d, fallbacks %d)"
          % (["%.3g" % v for v in curves["learned"]], prox.taken, prox.fallbacks))
    assert prox.taken > 0
    tried = prox.taken + prox.fallbacks
    assert curves["learned"][tried - 1] < curves["plain"][tried - 1]
    assert curves["learned"][-1] <= curves["plain"][-1]
    reach = [i for i, v in enumerate(curves["learned"]) if v <= curves["plain"][-1]]
    assert reach and reach[0] + 1 < niter


def test_tv_judge_returns_plain_tv_exactly(cuda, setup):
    """A step that moves away from the data loses to plain TV, whose own
    result is returned bit for bit."""
    import torch

    import tigre_amd
    from tigre_amd import learned as L

    geo, angles, case, _, _ = setup
    bad = L.ResidualCNN(seed=5)
    bad.W[5] = bad.W[5] * 0 + 1.0
    bad.refresh()
    img, b = case(45)
    resid = L.data_residual(b, geo, angles)
    prox = L.LearnedProx(bad, 20, residual=resid, judge="tv", epsilon=0.0)
    x = img + 0.001
    y = prox(x, 1e-3)
    assert prox.fallbacks == 1 and prox.taken == 0
    assert torch.equal(y, tigre_amd.minTV(x, 1e-3, 20))
