"""Learned preconditioner: a 4-level U-Net (32/64/128/256 channels, BF16
weights and activations, FP32 accumulation) run by the hand-written gfx950
kernels of libotfwi_unet.so, forward, backward and Adam.

The network maps (masked raw gradient, illumination) -> model update:

    enc0: conv 16->32, conv 32->32         (input: 2 channels padded to 16)
    enc1: pool, conv 32->64, conv 64->64
    enc2: pool, conv 64->128, conv 128->128
    bot:  pool, conv 128->256, conv 256->256
    dec2: conv [enc2 | up(bot)] 384->128, conv 128->128
    dec1: conv [enc1 | up(dec2)] 192->64, conv 64->64
    dec0: conv [enc0 | up(dec1)] 96->32, conv 32->32
    head: 1x1 conv 32->1 (FP32 output)

every 3x3 convolution followed by bias and ReLU; pooling is 2x2 average,
upsampling nearest.  PyTorch only allocates device memory and supplies the
stream; every operation is a call into the library.

Parameters live in one flat FP32 vector (per convolution: weights
[Cout][9][Cin], tap = 3*dy + dx, then bias [Cout]; then the head's w[32],
b[1]) with Adam moments beside it and a BF16 copy that the kernels read.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib

CONVS = (
    ("enc0a", 16, 0, 32), ("enc0b", 32, 0, 32),
    ("enc1a", 32, 0, 64), ("enc1b", 64, 0, 64),
    ("enc2a", 64, 0, 128), ("enc2b", 128, 0, 128),
    ("bota", 128, 0, 256), ("botb", 256, 0, 256),
    ("dec2a", 128, 256, 128), ("dec2b", 128, 0, 128),
    ("dec1a", 64, 128, 64), ("dec1b", 64, 0, 64),
    ("dec0a", 32, 64, 32), ("dec0b", 32, 0, 32),
)
CONV = {c[0]: c for c in CONVS}
IN_CH = 2          # real input channels (gradient, illumination)
PAD_CH = 16        # input channels as stored
# work-groups per weight-gradient launch (each sums several 256-pixel tiles);
# in the fixed-order form every work-group's sums are one workspace row that
# wgrad_reduce_kernel reads back, so fewer groups mean fewer bytes reduced.
# Training step at 8 x 1024^2 (MI355X): 16.28-16.30 ms with 1024, 16.38-16.42
# with 2048, 16.96-17.01 with 512.
WGRAD_GROUPS = 1024


def param_layout():
    """{name: (w_offset, w_count, b_offset, b_count)} in the flat vector, and its length."""
    off = 0
    lay = {}
    for name, ca, cb, co in CONVS:
        nw = co * 9 * (ca + cb)
        lay[name] = (off, nw, off + nw, co)
        off += nw + co
    lay["head"] = (off, 32, off + 32, 1)
    return lay, off + 33


def init_params(seed=0):
    """He-normal convolution weights (fan-in = 9 x real input channels), zero
    biases, a small head so that the initial update is near zero."""
    rng = np.random.default_rng(seed)
    lay, n = param_layout()
    p = np.zeros(n, np.float32)
    for name, ca, cb, co in CONVS:
        off, nw, _, _ = lay[name]
        cin = ca + cb
        fan = 9 * (IN_CH if name == "enc0a" else cin)
        w = rng.standard_normal((co, 9, cin)) * math.sqrt(2.0 / fan)
        if name == "enc0a":
            w[:, :, IN_CH:] = 0.0
        p[off:off + nw] = w.ravel()
    off, nw, _, _ = lay["head"]
    p[off:off + nw] = rng.standard_normal(nw) * 0.01
    return p


def conv_flops(N, H, W):
    """Multiply-add FLOPs (2 per MAC) of the 3x3 convolutions for one
    forward pass of N images of H x W, per layer name; the input layer counts
    its 16 stored channels.  The data gradient of a layer costs the same as
    its forward pass (not needed for the input layer), the weight gradient
    too."""
    lvl = {"enc0": 0, "enc1": 1, "enc2": 2, "bot": 3, "dec2": 2, "dec1": 1, "dec0": 0}
    out = {}
    for name, ca, cb, co in CONVS:
        l = lvl[name[:-1]]
        px = N * (H >> l) * (W >> l)
        out[name] = 2.0 * px * co * 9 * (ca + cb)
    return out


def train_flops(N, H, W):
    f = conv_flops(N, H, W)
    return sum(f.values()) * 3 - f["enc0a"]


def train_roofline(N, H, W, hbm=8.0e12, mfma=2.5e15):
    """Per-kernel roofline of one training step (forward, backward, Adam) of
    N images of H x W: each 3x3 convolution kernel (forward, data gradient,
    weight gradient) is bound by max(HBM bytes / hbm, FLOPs / mfma) with its
    operands read and its result written once in BF16; the elementwise
    kernels (input packing, pooling, upsample and skip gradients, head,
    Adam) by their bytes.  Returns {"ms", "conv_ms", "other_ms", "bytes",
    "flops", "hbm_bound_convs", "mfma_bound_convs"}."""
    lvl = {"enc0": 0, "enc1": 1, "enc2": 2, "bot": 3, "dec2": 2, "dec1": 1, "dec0": 0}
    t_conv = t_other = 0.0
    tot_b = tot_f = 0.0
    nh = nm = 0

    def kern(b, f):
        nonlocal t_conv, tot_b, tot_f, nh, nm
        t_conv += max(b / hbm, f / mfma)
        tot_b += b
        tot_f += f
        if b / hbm >= f / mfma:
            nh += 1
        else:
            nm += 1

    def elem(b):
        nonlocal t_other, tot_b
        t_other += b / hbm
        tot_b += b

    for name, ca, cb, co in CONVS:
        l = lvl[name[:-1]]
        px = N * (H >> l) * (W >> l)
        cin = ca + cb
        x_bytes = 2.0 * px * ca + 2.0 * (px / 4) * cb    # [skip | upsampled coarse]
        f = 2.0 * px * co * 9 * cin
        kern(x_bytes + 2.0 * px * co, f)                 # forward
        kern(2.0 * px * co + x_bytes, f)                 # weight gradient
        if name != "enc0a":                              # data gradient (+ ReLU mask)
            masked = name.endswith("b") or name == "botb"
            kern(2.0 * px * co + 2.0 * px * cin * (2 if masked else 1), f)
        if name.endswith("a") and l > 0 and not name.startswith("dec"):
            elem(2.0 * (4 * px) * ca + 2.0 * px * ca)     # 2x2 pooling of the finer level
        if cb:
            # upsample gradient: fine gradient and coarse mask read, coarse written
            elem(2.0 * px * cb + 2.0 * 2 * (px / 4) * cb)
            # skip combine of the encoder output: skip gradient, mask and result at
            # this level, the pooled path's gradient at the next coarser one
            elem(2.0 * px * ca * 3 + 2.0 * (px / 4) * ca)
    px0 = N * H * W
    elem(4.0 * 2 * px0 + 2.0 * PAD_CH * px0)             # input packing
    elem(2.0 * 32 * px0 + 4.0 * px0)                     # head forward
    elem(2.0 * 32 * px0 + 8.0 * px0 + 2.0 * 32 * px0)    # head backward
    n_par = param_layout()[1]
    elem(n_par * (4.0 * 7 + 2.0))                        # Adam: p, g, m, v read/written, BF16 copy
    return {"ms": 1e3 * (t_conv + t_other), "conv_ms": 1e3 * t_conv,
            "other_ms": 1e3 * t_other, "bytes": tot_b, "flops": tot_f,
            "hbm_bound_convs": nh, "mfma_bound_convs": nm}


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class UNet:
    """Device-resident network.  forward() keeps the activations that
    backward() needs; train_step() = forward + backward + Adam.  With
    deterministic=True (default) the weight gradients are summed in a fixed
    order from per-work-group partial sums (otfwi_unet_wgrad_work), so a
    training run is bitwise reproducible; False adds them with FP32 atomics."""

    def __init__(self, params=None, seed=0, device="cuda", deterministic=True):
        import torch

        self.torch = torch
        self.lib = _lib.load_unet()
        self.dev = torch.device(device)
        self.lay, self.n = param_layout()
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        if p.shape != (self.n,):
            raise ValueError(f"parameter vector must have {self.n} entries")
        self.p = torch.tensor(p, device=self.dev)
        self.g = torch.zeros_like(self.p)
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.pb = torch.empty(self.n, dtype=torch.bfloat16, device=self.dev)
        self.step = 0
        self.flip = {}
        for name, ca, cb, co in CONVS:
            if name != "enc0a":
                self.flip[name] = torch.empty(co * 9 * (ca + cb), dtype=torch.bfloat16,
                                              device=self.dev)
        self._bufs = {}
        self.deterministic = deterministic
        self._ws = None        # weight-gradient workspace (FP32), grown on demand
        self.loss_acc = torch.zeros(1, dtype=torch.float64, device=self.dev)
        self._sync_bf16()

    # ---- plumbing -----------------------------------------------------------
    def _st(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.dev).cuda_stream)

    def _buf(self, key, shape, dtype=None):
        dtype = dtype or self.torch.bfloat16
        t = self._bufs.get(key)
        if t is None or tuple(t.shape) != tuple(shape) or t.dtype != dtype:
            t = self.torch.empty(shape, dtype=dtype, device=self.dev)
            self._bufs[key] = t
        return t

    def _wb(self, name):
        """BF16 weights of a convolution (device pointer)."""
        return ctypes.c_void_p(self.pb.data_ptr() + 2 * self.lay[name][0])

    def _bias(self, name):
        return ctypes.c_void_p(self.p.data_ptr() + 4 * self.lay[name][2])

    def _gw(self, name):
        return ctypes.c_void_p(self.g.data_ptr() + 4 * self.lay[name][0])

    def _gb(self, name):
        return ctypes.c_void_p(self.g.data_ptr() + 4 * self.lay[name][2])

    def _sync_bf16(self):
        _lib.ucheck(self.lib.otfwi_unet_to_bf16(_vp(self.p), self.n, _vp(self.pb), self._st()))
        for name, ca, cb, co in CONVS:
            if name in self.flip:
                _lib.ucheck(self.lib.otfwi_unet_flip_weights(self._wb(name), co, ca + cb,
                                                             _vp(self.flip[name]), self._st()))

    def params(self):
        """FP32 parameter vector (host copy)."""
        return self.p.detach().cpu().numpy()

    def load(self, p):
        self.p.copy_(self.torch.as_tensor(np.asarray(p, np.float32)))
        self._sync_bf16()

    # ---- layers -------------------------------------------------------------
    def _conv(self, name, A, B, Y, N, H, W, relu=1, mask=None, flipped=False):
        _, ca, cb, co = CONV[name]
        if flipped:
            # data gradient: input = output gradient (co channels), output = ca+cb channels
            rc = self.lib.otfwi_unet_conv3x3(_vp(A), co, None, 0, 0, _vp(self.flip[name]), None,
                                             _vp(mask), _vp(Y), N, H, W, ca + cb, 0, self._st())
        else:
            rc = self.lib.otfwi_unet_conv3x3(_vp(A), ca, _vp(B), cb, 1 if cb else 0,
                                             self._wb(name), self._bias(name), _vp(mask), _vp(Y),
                                             N, H, W, co, relu, self._st())
        _lib.ucheck(rc)

    def _wgrad(self, name, A, B, G, N, H, W):
        _, ca, cb, co = CONV[name]
        tiles = N * H * W // 256
        other = max(1, co // 64) * ((ca + cb + 31) // 32)
        tpw = max(1, -(-tiles * other // WGRAD_GROUPS))
        if not self.deterministic:
            _lib.ucheck(self.lib.otfwi_unet_wgrad(_vp(A), ca, _vp(B), cb, 1 if cb else 0, _vp(G),
                                                  self._gw(name), self._gb(name), N, H, W, co,
                                                  tpw, self._st()))
            return
        need = int(self.lib.otfwi_unet_wgrad_work_elems(N, H, W, ca + cb, co, tpw))
        if need <= 0:
            raise ValueError(f"bad weight-gradient shape for {name}")
        if self._ws is None or self._ws.numel() < need:
            self._ws = self.torch.empty(need, dtype=self.torch.float32, device=self.dev)
        _lib.ucheck(self.lib.otfwi_unet_wgrad_work(_vp(A), ca, _vp(B), cb, 1 if cb else 0, _vp(G),
                                                   self._gw(name), self._gb(name), N, H, W, co,
                                                   tpw, _vp(self._ws), self._ws.numel(),
                                                   self._st()))

    def _pool(self, X, Y, N, H, W, C):
        _lib.ucheck(self.lib.otfwi_unet_pool2(_vp(X), N, H, W, C, _vp(Y), self._st()))

    # ---- forward / backward -------------------------------------------------
    def forward(self, inp):
        """inp: (N, 2, H, W) FP32 device tensor; H, W multiples of 64 (or
        32 at the coarsest level).  Returns the (N, H, W) FP32 output."""
        t = self.torch
        N, C, H, W = inp.shape
        if C != IN_CH:
            raise ValueError("input must have 2 channels")
        if H % 32 or W % 32 or (H // 8) % 8 or (W // 8) % 32:
            raise ValueError("H and W must be multiples of 256 / 8 tiles (e.g. 256, 512, 1024)")
        inp = inp.contiguous()
        b = self._buf
        a = {}
        a["x0"] = b("x0", (N, H, W, PAD_CH))
        _lib.ucheck(self.lib.otfwi_unet_pack_input(_vp(inp), N, H, W, _vp(a["x0"]), self._st()))
        h, w = H, W
        prev = a["x0"]
        for lvl, (ca_name, cb_name) in enumerate((("enc0a", "enc0b"), ("enc1a", "enc1b"),
                                                  ("enc2a", "enc2b"), ("bota", "botb"))):
            co = CONV[ca_name][3]
            if lvl:
                pin = CONV[ca_name][1]
                a[f"p{lvl - 1}"] = b(f"p{lvl - 1}", (N, h // 2, w // 2, pin))
                self._pool(prev, a[f"p{lvl - 1}"], N, h, w, pin)
                h, w = h // 2, w // 2
                prev = a[f"p{lvl - 1}"]
            a[ca_name] = b(ca_name, (N, h, w, co))
            self._conv(ca_name, prev, None, a[ca_name], N, h, w)
            a[cb_name] = b(cb_name, (N, h, w, co))
            self._conv(cb_name, a[ca_name], None, a[cb_name], N, h, w)
            prev = a[cb_name]
        for skip, (da, db) in zip(("enc2b", "enc1b", "enc0b"),
                                  (("dec2a", "dec2b"), ("dec1a", "dec1b"), ("dec0a", "dec0b"))):
            h, w = h * 2, w * 2
            co = CONV[da][3]
            a[da] = b(da, (N, h, w, co))
            self._conv(da, a[skip], prev, a[da], N, h, w)
            a[db] = b(db, (N, h, w, co))
            self._conv(db, a[da], None, a[db], N, h, w)
            prev = a[db]
        out = b("out", (N, H, W), t.float32)
        hw, hb = self.lay["head"][0], self.lay["head"][2]
        _lib.ucheck(self.lib.otfwi_unet_head_fwd(
            _vp(prev), ctypes.c_void_p(self.p.data_ptr() + 4 * hw),
            ctypes.c_void_p(self.p.data_ptr() + 4 * hb), N * H * W, _vp(out), self._st()))
        self._acts = a
        self._shape = (N, H, W)
        return out

    def backward(self, out, target):
        """Gradients of mean((out - target)^2) into self.g (zeroed first).
        Returns the device FP64 sum of squared errors."""
        t = self.torch
        a = self._acts
        N, H, W = self._shape
        b = self._buf
        self.g.zero_()
        self.loss_acc.zero_()
        target = target.contiguous()
        npx = N * H * W
        hw, hb = self.lay["head"][0], self.lay["head"][2]
        g = b("g_dec0b", (N, H, W, 32))
        hwork = b("head_work", (_lib.UNET_HEAD_WORK,), t.float32) if self.deterministic else None
        _lib.ucheck(self.lib.otfwi_unet_head_bwd_work(
            _vp(a["dec0b"]), ctypes.c_void_p(self.p.data_ptr() + 4 * hw), _vp(out), _vp(target),
            npx, 2.0 / npx, _vp(g), ctypes.c_void_p(self.g.data_ptr() + 4 * hw),
            ctypes.c_void_p(self.g.data_ptr() + 4 * hb), _vp(self.loss_acc), _vp(hwork),
            0 if hwork is None else hwork.numel(), self._st()))
        # decoder, finest level first
        dxc = {}
        h, w = H, W
        for lvl, (da, db, skip, coarse) in ((0, ("dec0a", "dec0b", "enc0b", "dec1b")),
                                            (1, ("dec1a", "dec1b", "enc1b", "dec2b")),
                                            (2, ("dec2a", "dec2b", "enc2b", "botb"))):
            co = CONV[db][3]
            self._wgrad(db, a[da], None, g, N, h, w)
            ga = b(f"g_{da}", (N, h, w, co))
            self._conv(db, g, None, ga, N, h, w, mask=a[da], flipped=True)
            _, ca, cb, _ = CONV[da]
            self._wgrad(da, a[skip], a[coarse], ga, N, h, w)
            dxc[lvl] = b(f"dxc{lvl}", (N, h, w, ca + cb))
            self._conv(da, ga, None, dxc[lvl], N, h, w, flipped=True)
            gc = b(f"g_{coarse}", (N, h // 2, w // 2, cb))
            _lib.ucheck(self.lib.otfwi_unet_upsum_mask(_vp(dxc[lvl]), ca + cb, ca, _vp(a[coarse]),
                                                       N, h // 2, w // 2, cb, _vp(gc), self._st()))
            g = gc
            h, w = h // 2, w // 2
        # bottleneck (h, w = H/8)
        self._wgrad("botb", a["bota"], None, g, N, h, w)
        ga = b("g_bota", (N, h, w, 256))
        self._conv("botb", g, None, ga, N, h, w, mask=a["bota"], flipped=True)
        self._wgrad("bota", a["p2"], None, ga, N, h, w)
        gp = b("g_p2", (N, h, w, 128))
        self._conv("bota", ga, None, gp, N, h, w, flipped=True)
        # encoder, coarsest first
        for lvl, (ea, eb, pin) in ((2, ("enc2a", "enc2b", "p1")), (1, ("enc1a", "enc1b", "p0")),
                                   (0, ("enc0a", "enc0b", "x0"))):
            h, w = h * 2, w * 2
            _, ca, cb, co = CONV[eb]
            _, dca, dcb, _ = CONV[("dec0a", "dec1a", "dec2a")[lvl]]
            ge = b(f"g_{eb}", (N, h, w, co))
            _lib.ucheck(self.lib.otfwi_unet_skip_combine(_vp(dxc[lvl]), dca + dcb, 0, _vp(gp),
                                                         _vp(a[eb]), N, h, w, co, _vp(ge),
                                                         self._st()))
            self._wgrad(eb, a[ea], None, ge, N, h, w)
            ga = b(f"g_{ea}", (N, h, w, co))
            self._conv(eb, ge, None, ga, N, h, w, mask=a[ea], flipped=True)
            self._wgrad(ea, a[pin], None, ga, N, h, w)
            if lvl:
                cin = CONV[ea][1]
                gp = b(f"g_{pin}", (N, h, w, cin))
                self._conv(ea, ga, None, gp, N, h, w, flipped=True)
        return self.loss_acc

    def adam(self, lr=1e-4, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        _lib.ucheck(self.lib.otfwi_unet_adam(_vp(self.p), _vp(self.g), _vp(self.m), _vp(self.v),
                                             _vp(self.pb), self.n, lr, b1, b2, eps, self.step,
                                             self._st()))
        for name, ca, cb, co in CONVS:
            if name in self.flip:
                _lib.ucheck(self.lib.otfwi_unet_flip_weights(self._wb(name), co, ca + cb,
                                                             _vp(self.flip[name]), self._st()))

    def train_step(self, inp, target, lr=1e-4):
        """One Adam step on mean((net(inp) - target)^2); returns the device
        FP64 sum of squared errors of the batch (before the step)."""
        out = self.forward(inp)
        loss = self.backward(out, target)
        self.adam(lr)
        return loss

    def predict(self, inp):
        return self.forward(inp)
