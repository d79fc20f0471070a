"""GPU: the learned proximal network's MFMA convolutions against the FP64 oracle.

Inputs are bf16-representable, so a single convolution differs from the FP64
oracle only by FP32 accumulation plus the final bf16 rounding of stored
activations: rel-L2 <= 2e-3 for bf16 outputs (half a bf16 ulp is 2^-9), 1e-5
for FP32 outputs (weight gradients, the residual output).  Through the whole
network, rounding flips of single activations propagate: outputs 5e-3,
gradients 2e-2.
"""
import numpy as np
import pytest

import oracle
from oracle import cnn as ocnn
from tigre_amd import _native

pytestmark = pytest.mark.gpu

SHAPE = (2, 5, 6, 37)  # N, D, H, W: W neither a multiple of 16 nor of 64


def _dev(a, dtype=None):
    import torch

    t = torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")
    return t.to(dtype) if dtype is not None else t


def _lib():
    return _native.lib()


def _conv(x, x_mode, w, bias, aux, y, epi, shape):
    import ctypes

    from tigre_amd.operators import _stream_handle

    p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    st = _lib().tigre_cnn_conv(p(x), x_mode, p(w), p(bias), p(aux), p(y), epi, *shape, _stream_handle())
    _native.check(st, "conv")


def _wgrad(x, x_mode, dy, dy_mode, shape):
    import ctypes

    import torch

    from tigre_amd.operators import _stream_handle

    n = ctypes.c_size_t()
    _native.check(_lib().tigre_cnn_wgrad_workspace_size(*shape, ctypes.byref(n)), "ws")
    ws = torch.empty(n.value, dtype=torch.uint8, device="cuda")
    out = torch.full((28, 32, 32), float("nan"), device="cuda")
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    st = _lib().tigre_cnn_wgrad(p(x), x_mode, p(dy), dy_mode, p(out), *shape, p(ws), n.value, _stream_handle())
    _native.check(st, "wgrad")
    return out


def _rand_bf16(rng, shape, scale=1.0):
    return ocnn.bf16_round(rng.standard_normal(shape) * scale)


def _rel(a, b):
    return oracle.rel_l2(np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64))


@pytest.mark.parametrize("x_mode", ["bf16x32", "f32x1"])
def test_conv_bias_relu(cuda, x_mode):
    import torch

    rng = np.random.default_rng(1)
    cin = 32 if x_mode == "bf16x32" else 1
    x = _rand_bf16(rng, SHAPE + (cin,))
    w = _rand_bf16(rng, (27, 32, cin), 0.2)
    b = rng.standard_normal(32).astype(np.float32) * 0.1
    wk = np.zeros((27, 32, 32))
    wk[:, :, :cin] = w
    y = torch.empty(SHAPE + (32,), dtype=torch.bfloat16, device="cuda")
    if x_mode == "bf16x32":
        xd, mode = _dev(x, torch.bfloat16), _native.CNN_IN_BF16X32
    else:
        xd, mode = _dev(x[..., 0]), _native.CNN_IN_F32X1
    _conv(xd, mode, _dev(wk, torch.bfloat16), _dev(b), None, y, _native.CNN_EPI_BIAS_RELU, SHAPE)
    ref = ocnn.bf16_round(np.maximum(ocnn.conv3d(x, w) + b, 0.0))
    err = _rel(y.float().cpu().numpy(), ref)
    print("conv %s rel-L2 %.3e" % (x_mode, err))
    assert err <= 2e-3


def test_conv_relu_mask_data_gradient(cuda):
    import torch

    rng = np.random.default_rng(2)
    g = _rand_bf16(rng, SHAPE + (32,))
    w = _rand_bf16(rng, (27, 32, 32), 0.2)
    act = _rand_bf16(rng, SHAPE + (32,))
    act = np.maximum(act, 0.0)  # activations after relu: the mask is act > 0
    wd = torch.as_tensor(w, dtype=torch.float32).flip(0).transpose(1, 2).contiguous()
    y = torch.empty(SHAPE + (32,), dtype=torch.bfloat16, device="cuda")
    _conv(_dev(g, torch.bfloat16), _native.CNN_IN_BF16X32, wd.cuda().to(torch.bfloat16), None,
          _dev(act, torch.bfloat16), y, _native.CNN_EPI_RELU_MASK, SHAPE)
    ref = ocnn.bf16_round(ocnn.conv3d_data_grad(g, w) * (act > 0))
    err = _rel(y.float().cpu().numpy(), ref)
    print("data gradient rel-L2 %.3e" % err)
    assert err <= 2e-3


@pytest.mark.parametrize("epi", ["bias_relu", "relu_mask"])
def test_conv_32ch_plane_chunks(cuda, epi):
    """32-channel convolution over more planes than one workgroup's chunk (70 = 32 + 32 + 6), with a
    ragged last x tile (40 = 32 + 8) and a ragged last row block (9 = 4 + 4 + 1)."""
    import torch

    shape = (1, 70, 9, 40)
    rng = np.random.default_rng(11)
    x = _rand_bf16(rng, shape + (32,))
    w = _rand_bf16(rng, (27, 32, 32), 0.2)
    y = torch.empty(shape + (32,), dtype=torch.bfloat16, device="cuda")
    if epi == "bias_relu":
        b = rng.standard_normal(32).astype(np.float32) * 0.1
        _conv(_dev(x, torch.bfloat16), _native.CNN_IN_BF16X32, _dev(w, torch.bfloat16), _dev(b), None, y,
              _native.CNN_EPI_BIAS_RELU, shape)
        ref = ocnn.bf16_round(np.maximum(ocnn.conv3d(x, w) + b, 0.0))
    else:
        act = np.maximum(_rand_bf16(rng, shape + (32,)), 0.0)
        _conv(_dev(x, torch.bfloat16), _native.CNN_IN_BF16X32, _dev(w, torch.bfloat16), None,
              _dev(act, torch.bfloat16), y, _native.CNN_EPI_RELU_MASK, shape)
        ref = ocnn.bf16_round(ocnn.conv3d(x, w) * (act > 0))
    got = y.float().cpu().numpy()
    err = _rel(got, ref)
    print("32-channel %s over 70 planes rel-L2 %.3e" % (epi, err))
    assert err <= 2e-3
    # every plane of every chunk is written and matches on its own
    per_plane = [_rel(got[0, z], ref[0, z]) for z in range(shape[1])]
    assert max(per_plane) <= 5e-3, int(np.argmax(per_plane))


def test_conv_residual_output(cuda):
    import torch

    rng = np.random.default_rng(3)
    h = np.maximum(_rand_bf16(rng, SHAPE + (32,)), 0.0)
    w = _rand_bf16(rng, (27, 1, 32), 0.2)
    b = np.array([0.25], dtype=np.float32)
    img = rng.standard_normal(SHAPE).astype(np.float32)
    wk = np.zeros((27, 32, 32))
    wk[:, :1, :] = w
    bk = np.zeros(32, dtype=np.float32)
    bk[0] = b[0]
    out = torch.empty(SHAPE, device="cuda")
    _conv(_dev(h, torch.bfloat16), _native.CNN_IN_BF16X32, _dev(wk, torch.bfloat16), _dev(bk), _dev(img), out,
          _native.CNN_EPI_RESIDUAL, SHAPE)
    ref = img + ocnn.conv3d(h, w)[..., 0] + b[0]
    err = _rel(out.cpu().numpy(), ref)
    print("residual output rel-L2 %.3e" % err)
    assert err <= 1e-5


@pytest.mark.parametrize("modes", [("bf16x32", "bf16x32"), ("f32x1", "bf16x32"), ("bf16x32", "f32x1")])
def test_weight_gradient(cuda, modes):
    import torch

    rng = np.random.default_rng(4)
    xm, gm = modes
    cin = 32 if xm == "bf16x32" else 1
    cout = 32 if gm == "bf16x32" else 1
    x = _rand_bf16(rng, SHAPE + (cin,))
    g = _rand_bf16(rng, SHAPE + (cout,))
    xd = _dev(x, torch.bfloat16) if cin == 32 else _dev(x[..., 0])
    gd = _dev(g, torch.bfloat16) if cout == 32 else _dev(g[..., 0])
    mode = lambda s: _native.CNN_IN_BF16X32 if s == "bf16x32" else _native.CNN_IN_F32X1  # noqa: E731
    out = _wgrad(xd, mode(xm), gd, mode(gm), SHAPE).cpu().numpy()
    ref = ocnn.conv3d_weight_grad(x, g)
    err = _rel(out[:27, :cout, :cin], ref)
    berr = _rel(out[27, :cout, 0], g.reshape(-1, cout).sum(axis=0))
    print("weight gradient %s rel-L2 %.3e, bias %.3e" % (modes, err, berr))
    assert err <= 1e-5 and berr <= 1e-5
    if cin == 1:
        assert not np.any(out[:27, :cout, 1:]) and not np.any(np.isnan(out))


def _net_and_oracle(seed=5):
    from tigre_amd.learned import ResidualCNN

    net = ResidualCNN(seed=seed)
    # non-trivial last layer and biases, bf16-representable weights in the oracle
    import torch

    g = torch.Generator().manual_seed(seed + 1)
    net.W[5] = (torch.randn(net.W[5].shape, generator=g) * 0.05).cuda()
    for l in range(6):
        net.b[l] = (torch.randn(net.b[l].shape, generator=g) * 0.02).cuda()
    net.refresh()
    W = [ocnn.bf16_round(w.cpu().numpy()) for w in net.W]
    b = [v.cpu().numpy().astype(np.float64) for v in net.b]
    return net, W, b


def test_network_forward_backward(cuda):
    import torch

    net, W, b = _net_and_oracle()
    rng = np.random.default_rng(6)
    img = rng.random((2, 6, 7, 19)).astype(np.float32)
    out = net.forward(_dev(img), keep=True)
    ref, h = ocnn.forward(img, W, b, 1.0, bf16=True)
    err = _rel(out.cpu().numpy(), ref)
    dout = rng.standard_normal(img.shape).astype(np.float32)
    dW, db = net.backward(_dev(dout))
    rW, rb = ocnn.backward(dout, h, W, 1.0, bf16=True)
    gerr = max(_rel(dW[l].cpu().numpy(), rW[l]) for l in range(6))
    berr = max(_rel(db[l].cpu().numpy(), rb[l]) for l in range(6))
    print("network output rel-L2 %.3e, weight grads max %.3e, bias grads max %.3e" % (err, gerr, berr))
    assert err <= 5e-3 and gerr <= 2e-2 and berr <= 2e-2
    # inference with a scale: img + s F(img)
    s = 0.37
    out_s = net.forward(_dev(img), scale=s).cpu().numpy()
    ref_s, _ = ocnn.forward(img, W, b, s, bf16=True)
    assert _rel(out_s - img, ref_s - img) <= 5e-3
    torch.cuda.synchronize()


def test_training_lowers_loss(cuda):
    """Fit the network to a fixed smoothing target on random volumes."""
    import torch

    from tigre_amd.learned import ResidualCNN

    net = ResidualCNN(seed=7)
    g = torch.Generator(device="cuda").manual_seed(8)
    x = torch.rand((4, 12, 12, 12), generator=g, device="cuda")
    # target: x plus a fixed local-average correction (a smoothing step)
    pad = torch.nn.functional.pad(x[:, None], (1, 1, 1, 1, 1, 1), mode="replicate")
    avg = torch.nn.functional.avg_pool3d(pad, 3, stride=1)[:, 0]
    target = x + 0.5 * (avg - x)
    losses = [net.train_step(x, target, lr=3e-3) for _ in range(250)]
    print("loss %.3e -> %.3e" % (losses[0], losses[-1]))
    assert losses[-1] < 0.3 * losses[0]
