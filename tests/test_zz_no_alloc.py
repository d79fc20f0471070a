"""GPU: the allocation-free weight gradient (tigre_cnn_wgrad_noalloc) against
tigre_cnn_wgrad with a caller workspace, through the C ABI.

Two calls in a row, the second on a smaller volume, so the second reads a
static scratch array that still holds the first call's partial sums.  Both must
equal the workspace entry point bit for bit (same chunking, same reduction
order) and, for the first, the FP64 oracle within the FP32-accumulation bound
of the other weight-gradient tests (1e-5).
"""
import ctypes

import numpy as np
import pytest

from oracle import cnn as ocnn
from tigre_amd import _native

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _with_workspace(x, dy, shape, x_mode, dy_mode):
    import torch

    from tigre_amd.operators import _stream_handle

    lib = _native.lib()
    n = ctypes.c_size_t()
    _native.check(lib.tigre_cnn_wgrad_workspace_size(*shape, ctypes.byref(n)), "workspace size")
    ws = torch.empty(n.value, dtype=torch.uint8, device="cuda")
    out = torch.full((28, 32, 32), float("nan"), device="cuda")
    _native.check(lib.tigre_cnn_wgrad(_p(x), x_mode, _p(dy), dy_mode, _p(out), *shape, _p(ws), n.value,
                                      _stream_handle()), "wgrad")
    return out


def _no_alloc(x, dy, shape, x_mode, dy_mode):
    import torch

    from tigre_amd.operators import _stream_handle

    out = torch.full((28, 32, 32), float("nan"), device="cuda")
    _native.check(_native.lib().tigre_cnn_wgrad_noalloc(_p(x), x_mode, _p(dy), dy_mode, _p(out), *shape,
                                                        _stream_handle()), "wgrad_noalloc")
    return out


def _bf16_data(rng, shape):
    import torch

    a = rng.standard_normal(shape).astype(np.float32)
    t = torch.as_tensor(a, device="cuda").to(torch.bfloat16)
    return t, t.float().cpu().numpy().astype(np.float64)


def test_limit_is_the_largest_workspace():
    lib = _native.lib()
    lim, big = ctypes.c_size_t(), ctypes.c_size_t()
    _native.check(lib.tigre_cnn_wgrad_noalloc_limit(ctypes.byref(lim)), "limit")
    _native.check(lib.tigre_cnn_wgrad_workspace_size(4, 256, 256, 256, ctypes.byref(big)), "workspace size")
    assert lim.value == big.value == 256 * 28 * 32 * 32 * 4


def test_no_alloc_matches_workspace_entry_point(cuda):
    import torch

    rng = np.random.default_rng(11)
    B = _native.CNN_IN_BF16X32
    # a volume that fills every one of the 256 chunks, then a smaller one (fewer chunks: the
    # scratch still holds the first call's partials beyond them)
    for shape in ((2, 24, 32, 70), (1, 3, 5, 37)):
        print("no-alloc weight gradient, shape %s" % (shape,), flush=True)
        x, xh = _bf16_data(rng, shape + (32,))
        g, gh = _bf16_data(rng, shape + (32,))
        torch.cuda.synchronize()
        ref = _with_workspace(x, g, shape, B, B)
        got = _no_alloc(x, g, shape, B, B)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), shape
        want = ocnn.conv3d_weight_grad(xh, gh)
        err = float(np.linalg.norm(got[:27].cpu().numpy() - want) / np.linalg.norm(want))
        print("shape %s: bit-identical to the workspace entry point, rel-L2 vs oracle %.3e" % (shape, err),
              flush=True)
        assert err <= 1e-5


def test_single_channel_modes(cuda):
    import torch

    rng = np.random.default_rng(12)
    F = _native.CNN_IN_F32X1
    shape = (1, 6, 7, 40)
    x = torch.as_tensor(rng.standard_normal(shape).astype(np.float32), device="cuda")
    g, _ = _bf16_data(rng, shape + (32,))
    ref = _with_workspace(x, g, shape, F, _native.CNN_IN_BF16X32)
    got = _no_alloc(x, g, shape, F, _native.CNN_IN_BF16X32)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)


def test_rejects_empty_volume():
    import torch

    lib = _native.lib()
    st = lib.tigre_cnn_wgrad_noalloc(ctypes.c_void_p(16), 0, ctypes.c_void_p(16), 0, ctypes.c_void_p(16), 0, 1, 1, 1,
                                     None)
    assert st != 0
    assert b"empty" in lib.tigre_last_error()
