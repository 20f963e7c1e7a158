"""Warm-start U-Net kernels (GPU) vs torch on the CPU.

The forward output and the gradient of the mean-squared error w.r.t. every
parameter come from this library's BF16 MFMA kernels; the checker is
torch.nn.functional (conv3d, conv_transpose3d, avg_pool3d) with autograd in
FP64 on the same BF16-rounded weights and activations.  Tolerances are
normwise relative errors: 2e-2 for outputs (BF16 activations, FP32
accumulation in a different order), 5e-2 per layer for gradients (the device
also rounds every back-propagated activation gradient to BF16).
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _net(shape=(16, 16, 16), batch=2, levels=3, base=16, seed=3):
    from lbm3d_bb_amd.warmstart import UNet

    return UNet(shape, batch=batch, levels=levels, base=base, seed=seed)


def _data(shape, batch, seed=0):
    rng = np.random.default_rng(seed)
    x = np.zeros((batch, 2) + shape, dtype=np.float32)
    x[:, 0] = (rng.random((batch,) + shape) < 0.3)
    x[:, 1] = rng.random((batch, 1, 1, 1))
    t = rng.standard_normal((batch, 4) + shape).astype(np.float32) * 0.1
    return x, t


@pytest.mark.parametrize("levels,shape,base", [(2, (8, 8, 16), 16), (3, (16, 16, 16), 16), (4, (16, 16, 24), 16),
                                               (2, (8, 8, 16), 48), (2, (8, 16, 16), 80)])
def test_forward_matches_torch(levels, shape, base):
    """Bases 48 and 80 give channel counts (48, 96, 80, 160) that are not
    16 times a power of two: every output channel tile must still be computed."""
    import torch

    from unet_reference import forward

    net = _net(shape, batch=2, levels=levels, base=base)
    x, _ = _data(shape, 2)
    y = net.forward(x)
    p = torch.tensor(net.get_params(), dtype=torch.float64)
    yr = forward(p, net.layers(), torch.tensor(x, dtype=torch.float64), levels).numpy()
    assert np.abs(yr).max() > 1e-3
    assert _rel(y, yr) < 2e-2, _rel(y, yr)


@pytest.mark.parametrize("levels,shape,base", [(3, (16, 16, 16), 16), (2, (8, 8, 16), 48)])
def test_gradient_matches_torch_autograd(levels, shape, base):
    from unet_reference import loss_and_grad

    net = _net(shape, batch=2, levels=levels, base=base)
    x, t = _data(shape, 2, seed=5)
    loss, g = net.gradient(x, t)
    lr, gr, _ = loss_and_grad(net.get_params(), net.layers(), x, t, levels)
    assert abs(loss - lr) / lr < 1e-2
    for i, L in enumerate(net.layers()):
        nw = (8 if L.kind == 1 else 1) * L.cout * (27 if L.kind == 0 else 1) * L.cin
        gw, gwr = g[L.w_off:L.w_off + nw], gr[L.w_off:L.w_off + nw]
        gb, gbr = g[L.b_off:L.b_off + L.cout], gr[L.b_off:L.b_off + L.cout]
        assert _rel(gw, gwr) < 5e-2, (i, L, _rel(gw, gwr))
        assert _rel(gb, gbr) < 5e-2, (i, L, _rel(gb, gbr))


def test_training_reduces_loss():
    shape = (16, 16, 16)
    net = _net(shape, batch=2, levels=3)
    x, t = _data(shape, 2, seed=9)
    first = net.train_step(x, t)
    for _ in range(60):
        last = net.train_step(x, t)
    assert last < 0.5 * first, (first, last)


@pytest.mark.parametrize("base", [32, 64])
def test_wide_channel_tiles_match_torch(base):
    """At 32^3 and batch 4 the finest level has 131072 rows (512 blocks of
    256), so its convolutions keep the 32-channel (base 32) and 64-channel
    (base 64) output tiles instead of narrowing to 16 channels; the smaller
    shapes above only reach the 16-channel tile."""
    import torch

    from unet_reference import forward, loss_and_grad

    shape = (32, 32, 32)
    net = _net(shape, batch=4, levels=2, base=base, seed=11)
    x, t = _data(shape, 4, seed=13)
    y = net.forward(x)
    p = torch.tensor(net.get_params(), dtype=torch.float32)
    yr = forward(p, net.layers(), torch.tensor(x), 2).double().numpy()
    assert np.abs(yr).max() > 1e-3
    assert _rel(y, yr) < 2e-2, _rel(y, yr)
    if base == 32:
        loss, g = net.gradient(x, t)
        lr, gr, _ = loss_and_grad(net.get_params(), net.layers(), x, t, 2)
        assert abs(loss - lr) / lr < 1e-2
        for i, L in enumerate(net.layers()):
            nw = (8 if L.kind == 1 else 1) * L.cout * (27 if L.kind == 0 else 1) * L.cin
            assert _rel(g[L.w_off:L.w_off + nw], gr[L.w_off:L.w_off + nw]) < 5e-2, (i, L)
            assert _rel(g[L.b_off:L.b_off + L.cout], gr[L.b_off:L.b_off + L.cout]) < 5e-2, (i, L)


@pytest.mark.parametrize("periodic", [6, 7])
def test_periodic_padding_matches_torch(periodic):
    """Convolutions that wrap around along y and z (and x) instead of padding
    with zeros, forward and gradient, against torch with circular padding."""
    import torch

    from unet_reference import forward, loss_and_grad
    from lbm3d_bb_amd.warmstart import UNet

    shape = (8, 8, 32)
    net = UNet(shape, batch=2, levels=2, base=16, seed=17, periodic=periodic)
    x, t = _data(shape, 2, seed=19)
    y = net.forward(x)
    p = torch.tensor(net.get_params(), dtype=torch.float64)
    yr = forward(p, net.layers(), torch.tensor(x, dtype=torch.float64), 2, periodic=periodic).numpy()
    y0 = forward(p, net.layers(), torch.tensor(x, dtype=torch.float64), 2, periodic=0).numpy()
    assert _rel(y0, yr) > 0.05  # the wrap changes the answer
    assert _rel(y, yr) < 2e-2, _rel(y, yr)
    loss, g = net.gradient(x, t)
    lr, gr, _ = loss_and_grad(net.get_params(), net.layers(), x, t, 2, periodic=periodic)
    assert abs(loss - lr) / lr < 1e-2
    for i, L in enumerate(net.layers()):
        nw = (8 if L.kind == 1 else 1) * L.cout * (27 if L.kind == 0 else 1) * L.cin
        assert _rel(g[L.w_off:L.w_off + nw], gr[L.w_off:L.w_off + nw]) < 5e-2, (i, L)
        assert _rel(g[L.b_off:L.b_off + L.cout], gr[L.b_off:L.b_off + L.cout]) < 5e-2, (i, L)


def test_periodic_padding_on_a_grid_too_small_fails():
    from lbm3d_bb_amd import _ffi
    from lbm3d_bb_amd.warmstart import UNet

    with pytest.raises(_ffi.LbmError):
        UNet((8, 8, 16), batch=1, levels=2, base=16, seed=1, periodic=6)  # level 1 is 8 voxels wide



def test_gradient_bitwise_reproducible():
    """The weight gradient is a fixed-order split-K sum (block sums per chunk of
    rows, added in chunk order), so two gradient calls on the same inputs
    return the same bits; 32^3 batch 4 runs hundreds of chunks per layer."""
    shape = (32, 32, 32)
    net = _net(shape, batch=4, levels=3, base=16, seed=31)
    x, t = _data(shape, 4, seed=37)
    l1, g1 = net.gradient(x, t)
    l2, g2 = net.gradient(x, t)
    net.close()
    assert np.abs(g1).max() > 0
    assert l1 == l2
    assert np.array_equal(g1, g2)


@pytest.mark.parametrize("base", [16, 32])
def test_weight_gradient_block_order_bitwise(base, monkeypatch):
    """The weight-gradient blocks numbered per XCD compute the same chunk sums
    as the plain (chunk, tap, tile) grid, so the gradients are bitwise equal;
    batch 3 at 32^3 gives chunk counts that are not multiples of 8."""
    shape = (32, 32, 32)
    x, t = _data(shape, 3, seed=41)
    res = []
    for order in ("0", "1"):
        monkeypatch.setenv("LBM_UNET_WGRAD_XCD", order)
        net = _net(shape, batch=3, levels=3, base=base, seed=43)
        res.append(net.gradient(x, t))
        net.close()
    assert np.abs(res[0][1]).max() > 0
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1])


@pytest.mark.parametrize("shape,batch,base,periodic", [((32, 16, 48), 3, 16, 0), ((32, 16, 48), 3, 16, 6),
                                                       ((64, 64, 64), 1, 32, 6)])
def test_weight_gradient_brick_kernel_matches_row_chunks(shape, batch, base, periodic, monkeypatch):
    """The all-taps brick kernel of the 3x3x3 layers against the row-chunk
    kernel: the same sums in another order, so the gradients agree to FP32
    rounding.  Batch 3 at 32 x 16 x 48 (base 16: 8, 16 and 32 input channels
    into 16) gives bricks that overhang the grid in x and a block count that
    does not divide the bricks evenly; 64^3 with base 32 runs the 32-output-
    channel layers of level 0 (1024 bricks) in 16 x 16 channel slices."""
    from lbm3d_bb_amd.warmstart import UNet

    x, t = _data(shape, batch, seed=47)
    res = []
    for brick in ("0", "1"):
        monkeypatch.setenv("LBM_UNET_WGRAD_BRICK", brick)
        net = UNet(shape, batch=batch, levels=2, base=base, seed=53, periodic=periodic)
        res.append(net.gradient(x, t))
        net.close()
    g0, g1 = res[0][1], res[1][1]
    assert np.abs(g0).max() > 0
    assert res[0][0] == res[1][0]
    assert _rel(g1, g0) < 1e-5, _rel(g1, g0)
