"""CPU: pins of the FP64 network oracle (oracle/cnn.py).

Each check is independent of the oracle's own formulas: scipy's correlate,
impulse and identity responses, the adjoint identity, and central finite
differences of the unrounded network for every parameter gradient.
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle import cnn


def _rng(seed):
    return np.random.default_rng(seed)


def test_single_channel_matches_scipy_correlate():
    r = _rng(0)
    x = r.standard_normal((1, 5, 6, 7, 1))
    w = r.standard_normal((27, 1, 1))
    k = w[:, 0, 0].reshape(3, 3, 3)  # k[dz+1, dy+1, dx+1]
    ref = ndimage.correlate(x[0, ..., 0], k, mode="constant", cval=0.0)
    np.testing.assert_allclose(cnn.conv3d(x, w)[0, ..., 0], ref, rtol=1e-12, atol=1e-12)


def test_impulse_response_is_mirrored_kernel():
    x = np.zeros((1, 5, 5, 5, 1))
    x[0, 2, 2, 2, 0] = 1.0
    w = np.arange(27, dtype=np.float64).reshape(27, 1, 1) + 1.0
    y = cnn.conv3d(x, w)[0, ..., 0]
    for t, (dz, dy, dx) in enumerate(cnn.TAPS):
        # y[v] picks x[v + tap]: the delta at c is seen from v = c - tap
        assert y[2 - dz, 2 - dy, 2 - dx] == w[t, 0, 0]


def test_identity_kernel_and_zero_padding():
    r = _rng(1)
    x = r.standard_normal((2, 3, 4, 5, 3))
    w = np.zeros((27, 3, 3))
    w[13] = np.eye(3)
    np.testing.assert_array_equal(cnn.conv3d(x, w), x)
    # a pure +x shift drops the last column
    w = np.zeros((27, 3, 3))
    w[14] = np.eye(3)  # (dz, dy, dx) = (0, 0, 1)
    y = cnn.conv3d(x, w)
    np.testing.assert_array_equal(y[..., :-1, :], x[..., 1:, :])
    assert not y[..., -1, :].any()


def test_adjoints():
    r = _rng(2)
    x = r.standard_normal((2, 4, 3, 5, 4))
    g = r.standard_normal((2, 4, 3, 5, 6))
    w = r.standard_normal((27, 6, 4))
    lhs = np.sum(cnn.conv3d(x, w) * g)
    assert np.isclose(lhs, np.sum(x * cnn.conv3d_data_grad(g, w)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(w * cnn.conv3d_weight_grad(x, g)), rtol=1e-12)


def _net(seed, c=4):
    r = _rng(seed)
    shapes = [(27, c, 1)] + [(27, c, c)] * 4 + [(27, 1, c)]
    W = [r.standard_normal(s) * 0.3 for s in shapes]
    b = [r.standard_normal(s[1]) * 0.1 + 0.05 for s in shapes]
    return W, b


def test_network_gradients_finite_differences():
    W, b = _net(3)
    r = _rng(4)
    img = r.standard_normal((1, 4, 4, 5))
    dout = r.standard_normal(img.shape)
    scale = 0.7

    def loss(W_, b_):
        return np.sum(cnn.forward(img, W_, b_, scale, bf16=False)[0] * dout)

    out, h = cnn.forward(img, W, b, scale, bf16=False)
    dW, db = cnn.backward(dout, h, W, scale, bf16=False)
    eps = 1e-6
    for l in range(6):
        for idx in [(0, 0, 0), (13, 0, 0), (26, W[l].shape[1] - 1, W[l].shape[2] - 1)]:
            Wp = [w.copy() for w in W]
            Wm = [w.copy() for w in W]
            Wp[l][idx] += eps
            Wm[l][idx] -= eps
            num = (loss(Wp, b) - loss(Wm, b)) / (2 * eps)
            assert np.isclose(num, dW[l][idx], rtol=1e-5, atol=1e-7), (l, idx, num, dW[l][idx])
        bp = [v.copy() for v in b]
        bm = [v.copy() for v in b]
        bp[l][0] += eps
        bm[l][0] -= eps
        num = (loss(W, bp) - loss(W, bm)) / (2 * eps)
        assert np.isclose(num, db[l][0], rtol=1e-5, atol=1e-7), (l, num, db[l][0])


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], dtype=np.float32)
    got = cnn.bf16_round(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1); 1 + 3 2^-8 ties up to 1 + 2^-6
    np.testing.assert_array_equal(got[:4], [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6])
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8
