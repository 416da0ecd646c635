"""CPU checks of the U-Net oracle (oracle/unet_ref.py) and of the parameter
layout the GPU kernels share with it: a naive loop pins the weight layout
[Cout][9][Cin] with tap 3*dy + dx and the nearest upsampling of the second
input; a central finite difference pins the oracle's backward pass."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import unet_ref as R  # noqa: E402
from otfwi_amd import unet as U  # noqa: E402


def test_conv3x3_layout_naive_loop():
    rng = np.random.default_rng(0)
    N, H, W, CA, CB, Co = 1, 6, 8, 3, 2, 4
    A = rng.standard_normal((N, H, W, CA))
    B = rng.standard_normal((N, H // 2, W // 2, CB))
    w = rng.standard_normal((Co, 9, CA + CB))
    b = rng.standard_normal(Co)
    got = R.conv3x3(A, B, w, b, relu=False)
    ref = np.zeros((N, H, W, Co))
    for y in range(H):
        for x in range(W):
            for dy in range(3):
                for dx in range(3):
                    yy, xx = y + dy - 1, x + dx - 1
                    if not (0 <= yy < H and 0 <= xx < W):
                        continue
                    v = np.concatenate([A[0, yy, xx], B[0, yy // 2, xx // 2]])
                    ref[0, y, x] += w[:, 3 * dy + dx, :] @ v
            ref[0, y, x] += b
    assert np.abs(got - ref).max() < 1e-5


def test_param_layout_counts():
    lay, n = U.param_layout()
    total = sum(co * 9 * (ca + cb) + co for _, ca, cb, co in U.CONVS) + 33
    assert n == total
    assert lay["head"][2] == n - 1


def test_oracle_backward_finite_difference():
    rng = np.random.default_rng(1)
    p = U.init_params(seed=3)
    lay, n = U.param_layout()
    off, nw, _, _ = lay["head"]
    p[off:off + nw] = rng.standard_normal(nw).astype(np.float32) * 0.3
    H = W = 256
    zz, xx = np.mgrid[0:H, 0:W] / H
    inp = np.stack([np.sin(6 * xx + 2 * zz), np.cos(3 * zz)])[None].astype(np.float32)
    tgt = (0.2 * np.sin(3 * xx + zz))[None].astype(np.float32)
    _, L0, g = R.forward_backward(p, inp, tgt, bf16=False)
    d = np.zeros(n, np.float32)
    # a direction touching an encoder, a decoder concatenation, biases and the head
    for name in ("enc1b", "dec1a", "head"):
        o, nw_, bo, nb = lay[name]
        d[o:o + nw_] = rng.standard_normal(nw_)
        d[bo:bo + nb] = rng.standard_normal(nb)
    d /= np.linalg.norm(d)
    h = 1e-2
    Lp = R.forward_backward(p + h * d, inp, tgt, bf16=False)[1]
    Lm = R.forward_backward(p - h * d, inp, tgt, bf16=False)[1]
    fd = (Lp - Lm) / (2 * h)
    an = float(np.dot(g.astype(np.float64), d))
    assert abs(fd - an) < 2e-3 * max(abs(an), 1e-12) + 1e-7, (fd, an)


def test_train_roofline_counts():
    """The per-kernel roofline of a training step counts the same FLOPs as
    train_flops, and classifies the 32-channel full-resolution convolutions
    as HBM-bound and the 256-channel bottleneck as MFMA-bound."""
    from otfwi_amd import unet as U

    r = U.train_roofline(8, 1024, 1024)
    assert r["flops"] == U.train_flops(8, 1024, 1024)
    assert r["hbm_bound_convs"] + r["mfma_bound_convs"] == 3 * len(U.CONVS) - 1
    px = 8 * 1024 * 1024
    b = 2.0 * px * 32 * 2                          # enc0b forward: 32 channels in and out
    f = 2.0 * px * 32 * 9 * 32
    assert b / 8e12 > f / 2.5e15                   # HBM-bound
    px3 = px / 64
    assert 2.0 * px3 * 512 / 8e12 < 2.0 * px3 * 256 * 9 * 256 / 2.5e15   # botb: MFMA-bound
    assert 0 < r["conv_ms"] < r["ms"]
