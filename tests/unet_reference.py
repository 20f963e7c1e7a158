"""TEST INFRASTRUCTURE: the warm-start U-Net restated with torch on the CPU.

Builds the same graph as unet_capi.hip from the library's flat parameter
vector (layer table from lbm_unet_param_layout) with torch.nn.functional
conv3d / conv_transpose3d / avg_pool3d, so forward outputs and autograd
gradients pin the HIP kernels.  Activations can be rounded to BF16 at the same
points as the device graph.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F


def _bf(x: torch.Tensor, on: bool) -> torch.Tensor:
    return x.to(torch.bfloat16).to(x.dtype) if on else x


def layer_weights(params: torch.Tensor, L) -> tuple[torch.Tensor, torch.Tensor]:
    """Library layout [npar][cout][taps][cin] -> torch layouts."""
    b = params[L.b_off:L.b_off + L.cout]
    if L.kind == 0:
        w = params[L.w_off:L.w_off + L.cout * 27 * L.cin].reshape(L.cout, 3, 3, 3, L.cin).permute(0, 4, 1, 2, 3)
    elif L.kind == 1:
        w = params[L.w_off:L.w_off + 8 * L.cout * L.cin].reshape(2, 2, 2, L.cout, L.cin).permute(4, 3, 0, 1, 2)
    else:
        w = params[L.w_off:L.w_off + L.cout * L.cin].reshape(L.cout, L.cin, 1, 1, 1)
    return w, b


def pad3(h: torch.Tensor, periodic: int) -> torch.Tensor:
    """One voxel of padding on every side of (B, C, D, H, W): wrapped around
    along the axes periodic marks (bit 0 x = W, bit 1 y = H, bit 2 z = D),
    zeros along the others."""
    hp = F.pad(h, (1, 1, 1, 1, 1, 1), mode="circular")
    keep = torch.ones(hp.shape[2:], dtype=hp.dtype)
    for bit, ax in ((1, 2), (2, 1), (4, 0)):
        if not periodic & bit:
            idx = [slice(None)] * 3
            idx[ax] = 0
            keep[tuple(idx)] = 0
            idx[ax] = -1
            keep[tuple(idx)] = 0
    return hp * keep


def forward(params: torch.Tensor, layers, x: torch.Tensor, levels: int, bf16: bool = True,
            periodic: int = 0) -> torch.Tensor:
    """x: (B, cin, D, H, W); the layer order matches the library's table."""
    it = iter(layers)

    def conv3(h, L, relu=True):
        w, b = layer_weights(params, L)
        y = F.conv3d(pad3(h, periodic), _bf(w, bf16), b)
        return _bf(F.relu(y) if relu else y, bf16)

    enc = []
    h = _bf(x, bf16)
    Ls = list(layers)
    # table order: enc1_0, enc2_0, enc1_1, enc2_1, ..., then per decoder level up, dec1, dec2, then head
    k = 0
    skips = []
    for l in range(levels):
        h = conv3(h, Ls[k]); k += 1
        h = conv3(h, Ls[k]); k += 1
        if l < levels - 1:
            skips.append(h)
            h = _bf(F.avg_pool3d(h, 2), bf16)
    for l in range(levels - 2, -1, -1):
        w, b = layer_weights(params, Ls[k]); k += 1
        up = _bf(F.relu(F.conv_transpose3d(h, _bf(w, bf16), b, stride=2)), bf16)
        h = torch.cat([up, skips[l]], dim=1)
        h = conv3(h, Ls[k]); k += 1
        h = conv3(h, Ls[k]); k += 1
    w, b = layer_weights(params, Ls[k])
    del it, enc
    return F.conv3d(h, _bf(w, bf16), b)


def loss_and_grad(params_np: np.ndarray, layers, x_np, t_np, levels, bf16=True, periodic=0):
    p = torch.tensor(params_np, dtype=torch.float64, requires_grad=True)
    x = torch.tensor(x_np, dtype=torch.float64)
    t = torch.tensor(t_np, dtype=torch.float64)
    y = forward(p, layers, x, levels, bf16=bf16, periodic=periodic)
    loss = ((y - t) ** 2).mean()
    loss.backward()
    return float(loss.detach()), p.grad.numpy(), y.detach().numpy()
