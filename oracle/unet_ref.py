"""CPU oracle of the learned preconditioner's U-Net (test infrastructure
only; the product network runs on the GPU kernels of libotfwi_unet.so).

PyTorch-CPU FP32 restatement of the same network (otfwi_amd/unet.py) on the
same flat parameter vector.  With ``bf16=True`` it rounds to BF16 exactly
where the GPU kernels store BF16: the packed input, the convolution weights,
every post-ReLU activation and pooled tensor, and in the backward pass the
gradient at every convolution output, at every decoder concatenation and at
every pooled tensor.  With ``bf16=False`` it is plain FP32, which measures
the network's BF16 error.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

from otfwi_amd.unet import CONVS, IN_CH, PAD_CH, param_layout


def _bf(x):
    return x.to(torch.bfloat16).to(torch.float32)


class _RoundAct(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return _bf(x)

    @staticmethod
    def backward(ctx, g):
        return g


class _RoundGrad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return x.clone()

    @staticmethod
    def backward(ctx, g):
        return _bf(g)


def _weights(p, bf16):
    """{name: (torch weight (Cout, Cin, 3, 3), bias)} as leaf tensors."""
    lay, _ = param_layout()
    out = {}
    for name, ca, cb, co in CONVS:
        off, nw, boff, nb = lay[name]
        w = torch.tensor(p[off:off + nw].reshape(co, 3, 3, ca + cb))
        w = w.permute(0, 3, 1, 2).contiguous()          # [co][tap=3dy+dx][ci] -> [co][ci][dy][dx]
        if bf16:
            w = _bf(w)
        out[name] = (w.requires_grad_(), torch.tensor(p[boff:boff + nb]).requires_grad_())
    off, nw, boff, _ = lay["head"]
    out["head"] = (torch.tensor(p[off:off + nw]).requires_grad_(),
                   torch.tensor(p[boff:boff + 1]).requires_grad_())
    return out


def forward_backward(p, inp, target, bf16=True):
    """p: flat FP32 parameters; inp (N, 2, H, W); target (N, H, W).
    Returns (out (N,H,W), loss = mean squared error, flat gradient of the loss)."""
    ra = _RoundAct.apply if bf16 else (lambda x: x)
    rg = _RoundGrad.apply if bf16 else (lambda x: x)
    Wt = _weights(np.asarray(p, np.float32), bf16)
    x = torch.tensor(np.asarray(inp, np.float32))
    N, _, H, W = x.shape
    x0 = torch.zeros(N, PAD_CH, H, W)
    x0[:, :IN_CH] = x
    x0 = ra(x0)

    def conv(name, t):
        w, b = Wt[name]
        y = F.conv2d(t, w, b, padding=1)
        y = rg(y)
        return ra(F.relu(y))

    def pool(t):
        return rg(ra(F.avg_pool2d(t, 2)))

    e0 = conv("enc0b", conv("enc0a", x0))
    e1 = conv("enc1b", conv("enc1a", pool(e0)))
    e2 = conv("enc2b", conv("enc2a", pool(e1)))
    bt = conv("botb", conv("bota", pool(e2)))

    def up_cat(skip, coarse):
        return rg(torch.cat([skip, F.interpolate(coarse, scale_factor=2, mode="nearest")], 1))

    d2 = conv("dec2b", conv("dec2a", up_cat(e2, bt)))
    d1 = conv("dec1b", conv("dec1a", up_cat(e1, d2)))
    d0 = conv("dec0b", conv("dec0a", up_cat(e0, d1)))
    hw, hb = Wt["head"]
    out = torch.einsum("nchw,c->nhw", d0, hw) + hb
    t = torch.tensor(np.asarray(target, np.float32))
    loss = ((out - t) ** 2).mean()
    loss.backward()
    lay, n = param_layout()
    g = np.zeros(n, np.float32)
    for name, ca, cb, co in CONVS:
        off, nw, boff, nb = lay[name]
        w, b = Wt[name]
        g[off:off + nw] = w.grad.permute(0, 2, 3, 1).reshape(-1).numpy()
        g[boff:boff + nb] = b.grad.numpy()
    off, nw, boff, _ = lay["head"]
    g[off:off + nw] = hw.grad.numpy()
    g[boff] = hb.grad.numpy()[0]
    return out.detach().numpy(), loss.item(), g


def conv3x3(A, B, w, bias, relu, b_half=True):
    """Single layer: NHWC FP32 inputs (B nearest-upsampled when b_half),
    weights [Cout][9][Cin] -> NHWC output (no rounding)."""
    a = torch.tensor(np.asarray(A, np.float32)).permute(0, 3, 1, 2)
    if B is not None:
        bb = torch.tensor(np.asarray(B, np.float32)).permute(0, 3, 1, 2)
        if b_half:
            bb = F.interpolate(bb, scale_factor=2, mode="nearest")
        a = torch.cat([a, bb], 1)
    co, _, cin = w.shape
    wt = torch.tensor(np.asarray(w, np.float32).reshape(co, 3, 3, cin)).permute(0, 3, 1, 2)
    y = F.conv2d(a, wt, None if bias is None else torch.tensor(np.asarray(bias, np.float32)),
                 padding=1)
    if relu:
        y = F.relu(y)
    return y.permute(0, 2, 3, 1).numpy()
