"""Batched propagator handle: the native model behind Cuda_interface_PML.

One :class:`PMLModel` holds a velocity model, its damping profiles, sources,
receivers and boundary ring on the GPU, and runs every shot of a batch in one
launch per time step.  All arrays use the reference's layouts
(regression_PML_single.cu:9, cuda_interface_pml.py:40-53): models (Z, X),
traces (n_shots, n_rx, T), point lists in caller order.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, f32, i32, ptr


class PMLModel:
    """Handle on one native model (precision 4 = FP32, 8 = FP64 validation)."""

    def __init__(self, X, Z, T, precision=4, constvec=(1.0, 1.0, 1.0)):
        self._lib = _lib.load()
        self.X, self.Z, self.T = int(X), int(Z), int(T)
        self.precision = precision
        h = ctypes.c_void_p()
        cv = f32(constvec)
        check(self._lib.otfwi_create(ctypes.byref(h), self.X, self.Z, self.T, precision, ptr(cv)))
        self._h = h
        self.n_src = self.n_rx = self.n_ring = 0

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.otfwi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up ---------------------------------------------------------
    def set_damping(self, d_x, d_z):
        dx, dz = f32(d_x), f32(d_z)
        if dx.shape != (self.Z, self.X) or dz.shape != (self.Z, self.X):
            raise ValueError("damping arrays must be (Z, X)")
        check(self._lib.otfwi_set_damping(self._h, ptr(dx), ptr(dz)))

    def set_cquad(self, cquad):
        cq = f32(cquad).reshape(-1)
        if cq.size != self.X * self.Z:
            raise ValueError(f"cquad has {cq.size} values, expected {self.X * self.Z}")
        check(self._lib.otfwi_set_cquad(self._h, ptr(cq)))

    def set_cquad_device(self, dev_ptr):
        """c^2 from a device address (FP32, Z*X contiguous), e.g. a torch tensor."""
        check(self._lib.otfwi_set_cquad_device(self._h, ctypes.c_void_p(int(dev_ptr))))

    def set_sources(self, xs, zs, wavelets):
        xs, zs = i32(xs), i32(zs)
        w = f32(wavelets).reshape(len(xs), self.T)
        check(self._lib.otfwi_set_sources(self._h, len(xs), ptr(xs), ptr(zs), ptr(w)))
        self.n_src = len(xs)

    def set_receivers(self, xs, zs):
        xs, zs = i32(xs), i32(zs)
        check(self._lib.otfwi_set_receivers(self._h, len(xs), ptr(xs), ptr(zs)))
        self.n_rx = len(xs)

    def set_ring(self, xs, zs):
        xs, zs = i32(xs), i32(zs)
        check(self._lib.otfwi_set_ring(self._h, len(xs), ptr(xs), ptr(zs)))
        self.n_ring = len(xs)

    def set_observed(self, obs):
        o = f32(obs)
        if o.ndim == 2:
            o = o[None]
        if o.shape[1:] != (self.n_rx, self.T):
            raise ValueError(f"observed must be (n_shots, {self.n_rx}, {self.T})")
        check(self._lib.otfwi_set_observed(self._h, o.shape[0], ptr(o)))

    def set_tuning(self, th_fwd=0, th_adj=0, max_batch=0):
        check(self._lib.otfwi_set_tuning(self._h, th_fwd, th_adj, max_batch))

    def set_launch(self, persistent_fwd=False, persistent_adj=False):
        """Persistent (prefetching) work-groups for the forward / adjoint steps."""
        check(self._lib.otfwi_set_launch(self._h, int(bool(persistent_fwd)),
                                         int(bool(persistent_adj))))

    def set_graph(self, use_graph=True):
        """Replay the step loops as captured hipGraphs (default on)."""
        check(self._lib.otfwi_set_graph(self._h, int(bool(use_graph))))

    def autotune(self, n_shots, steps=20):
        """Time every tile height / launch style on a batch of n_shots and keep
        the fastest, then choose the steps per launch of both loops.  Returns
        the choice and the best one-step per-launch times (ms)."""
        ch = np.zeros(4, np.int32)
        ms = np.zeros(2, np.float32)
        check(self._lib.otfwi_autotune(self._h, int(n_shots), int(steps), ptr(ch), ptr(ms)))
        return dict(th_fwd=int(ch[0]), persistent_fwd=bool(ch[1]), th_adj=int(ch[2]),
                    persistent_adj=bool(ch[3]), fwd_ms=float(ms[0]), adj_ms=float(ms[1]),
                    fwd_steps=self.get_fwd_steps(), adj_steps=self.get_adj_steps())

    def set_blocking(self, steps_fwd=1, shots_fwd=0, steps_adj=1, shots_adj=0):
        """Cache blocking of the step loops (see otfwi_set_blocking)."""
        check(self._lib.otfwi_set_blocking(self._h, int(steps_fwd), int(shots_fwd),
                                           int(steps_adj), int(shots_adj)))

    def get_blocking(self):
        out = np.zeros(4, np.int32)
        check(self._lib.otfwi_get_blocking(self._h, ptr(out)))
        return tuple(int(v) for v in out)

    def set_fwd_steps(self, steps):
        """Forward time steps per launch: 1, 2 or 0 = by launch size (see
        otfwi_set_fwd_steps)."""
        check(self._lib.otfwi_set_fwd_steps(self._h, int(steps)))

    def get_fwd_steps(self):
        v = ctypes.c_int()
        check(self._lib.otfwi_get_fwd_steps(self._h, ctypes.byref(v)))
        return int(v.value)

    def kernel_times(self, n_shots, reps=10):
        """Device ms per launch of the step kernels on a batch of n_shots
        (see otfwi_kernel_times): dict fwd, adj, fwd_pair, rec_pair,
        lam_pair (None where a kernel does not apply)."""
        ms = np.zeros(5, np.float32)
        check(self._lib.otfwi_kernel_times(self._h, int(n_shots), int(reps), ptr(ms)))
        keys = ("fwd", "adj", "fwd_pair", "rec_pair", "lam_pair")
        return {k: (float(v) if v >= 0 else None) for k, v in zip(keys, ms)}

    def set_adj_steps(self, steps):
        """Adjoint time steps per launch pair: 1, 2 or 0 = by launch size (see
        otfwi_set_adj_steps)."""
        check(self._lib.otfwi_set_adj_steps(self._h, int(steps)))

    def get_adj_steps(self):
        v = ctypes.c_int()
        check(self._lib.otfwi_get_adj_steps(self._h, ctypes.byref(v)))
        return int(v.value)

    def steps_per_launch(self, n_shots):
        """(forward, adjoint) steps per launch used for a batch of n_shots."""
        f, a = ctypes.c_int(), ctypes.c_int()
        check(self._lib.otfwi_steps_per_launch(self._h, int(n_shots), ctypes.byref(f),
                                               ctypes.byref(a)))
        return int(f.value), int(a.value)

    def graph_modes(self, n_shots):
        """How the forward and adjoint loops of n_shots shots run:
        {"fwd": "graph" | "direct", "fwd_streams": k, "adj": ..., "adj_streams": k}."""
        out = (ctypes.c_int * 4)()
        check(self._lib.otfwi_loop_modes(self._h, int(n_shots), out))
        return {"fwd": "graph" if out[1] else "direct", "fwd_streams": int(out[0]),
                "adj": "graph" if out[3] else "direct", "adj_streams": int(out[2])}

    def set_pair_streams(self, k=0):
        """Streams of the two-step loops: 1, 2, 4 or 0 (by size); with graphs
        on (set_graph) each stream group's two-step launches are captured and
        replayed on that group's stream."""
        check(self._lib.otfwi_set_pair_streams(self._h, int(k)))

    def set_pair_width(self, nw=0):
        """Work-group width of the two-step kernels: 128, 192, 256, 384 or 0 (256)."""
        check(self._lib.otfwi_set_pair_width(self._h, int(nw)))

    def set_steps_per_launch(self, fwd=2, adj=2):
        """Both loops at once (see set_fwd_steps, set_adj_steps)."""
        self.set_fwd_steps(fwd)
        self.set_adj_steps(adj)

    def get_launch(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        check(self._lib.otfwi_get_launch(self._h, ctypes.byref(a), ctypes.byref(b)))
        return bool(a.value), bool(b.value)

    def get_tuning(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(self._lib.otfwi_get_tuning(self._h, ctypes.byref(a), ctypes.byref(b),
                                          ctypes.byref(c)))
        return a.value, b.value, c.value

    def timing(self):
        """Device time of the last gradient call from HIP events on the
        model's stream: dict(fwd_ms, misfit_ms, adj_ms, n_fwd, n_adj)."""
        out = np.zeros(5, np.float64)
        check(self._lib.otfwi_get_timing(self._h, ptr(out)))
        return dict(fwd_ms=out[0], misfit_ms=out[1], adj_ms=out[2], n_fwd=int(out[3]),
                    n_adj=int(out[4]))

    # ---- runs -------------------------------------------------------------
    @staticmethod
    def _shots(shot_src):
        s = i32(np.atleast_1d(shot_src))
        return s, len(s)

    def forward(self, shot_src=(-1,)):
        """Traces (n_shots, n_rx, T) FP32; shot k fires source shot_src[k]
        (-1: all sources at once)."""
        s, n = self._shots(shot_src)
        rec = np.zeros((n, self.n_rx, self.T), np.float32)
        check(self._lib.otfwi_forward(self._h, n, ptr(s), ptr(rec)))
        return rec

    def gradient(self, shot_src=(-1,), misfit="w2", want_rec=False):
        """Device misfit + adjoint gradient.  Returns (grad (Z,X) FP32 mean
        over shots, per-shot misfit FP64 [, traces])."""
        kind = _lib.MISFIT_KINDS[misfit]
        s, n = self._shots(shot_src)
        mis = np.zeros(n, np.float64)
        grad = np.zeros((self.Z, self.X), np.float32)
        rec = np.zeros((n, self.n_rx, self.T), np.float32) if want_rec else None
        check(self._lib.otfwi_gradient(self._h, n, ptr(s), kind, ptr(mis), ptr(grad), ptr(rec)))
        return (grad, mis, rec) if want_rec else (grad, mis)

    def gradient_device(self, shot_src, misfit, grad_dev_ptr):
        """Same as gradient(), writing the (Z,X) FP32 mean gradient to device memory."""
        kind = _lib.MISFIT_KINDS[misfit]
        s, n = self._shots(shot_src)
        mis = np.zeros(n, np.float64)
        check(self._lib.otfwi_gradient_device(self._h, n, ptr(s), kind, ptr(mis),
                                              ctypes.c_void_p(int(grad_dev_ptr))))
        return mis

    def set_cquad64(self, cquad):
        """c^2 (Z,X) in FP64 (see otfwi_set_cquad64)."""
        c = np.ascontiguousarray(cquad, dtype=np.float64)
        if c.size != self.X * self.Z:
            raise ValueError("cquad must have Z*X entries")
        check(self._lib.otfwi_set_cquad64(self._h, ptr(c)))

    def gradient_exact(self, shot_src=(-1,), misfit="l2", snap_every=0, fp64=False):
        """Exact discrete-adjoint gradient dJ/d(c^2) of J = mean over shots of
        the misfit value, with snapshot checkpointing every snap_every steps
        (0: sqrt(T-1)).  Returns (grad (Z,X) FP32 or FP64, per-shot misfit)."""
        kind = _lib.MISFIT_KINDS[misfit]
        s, n = self._shots(shot_src)
        mis = np.zeros(n, np.float64)
        if fp64:
            g = np.zeros((self.Z, self.X), np.float64)
            check(self._lib.otfwi_gradient_exact(self._h, n, ptr(s), kind, int(snap_every),
                                                 ptr(mis), None, ptr(g)))
        else:
            g = np.zeros((self.Z, self.X), np.float32)
            check(self._lib.otfwi_gradient_exact(self._h, n, ptr(s), kind, int(snap_every),
                                                 ptr(mis), ptr(g), None))
        return g, mis

    def gradient_exact_device(self, shot_src, misfit, snap_every, grad_dev_ptr):
        """Same as gradient_exact(), writing the (Z,X) FP32 gradient to device
        memory; returns the per-shot misfits."""
        kind = _lib.MISFIT_KINDS[misfit]
        s, n = self._shots(shot_src)
        mis = np.zeros(n, np.float64)
        check(self._lib.otfwi_gradient_exact_device(self._h, n, ptr(s), kind, int(snap_every),
                                                    ptr(mis), ctypes.c_void_p(int(grad_dev_ptr))))
        return mis

    def objective(self, shot_src=(-1,), misfit="w2"):
        """Per-shot misfit (FP64) of the current model: forward runs and the
        device misfit, no adjoint."""
        kind = _lib.MISFIT_KINDS[misfit]
        s, n = self._shots(shot_src)
        mis = np.zeros(n, np.float64)
        check(self._lib.otfwi_objective(self._h, n, ptr(s), kind, ptr(mis)))
        return mis

    def set_illum(self, stride):
        """Collect the source-side illumination (sum of P^2 over shots and
        every stride-th step, times stride) during the next gradient calls;
        0 switches it off."""
        check(self._lib.otfwi_set_illum(self._h, int(stride)))

    def illum(self):
        """(Z,X) FP32 illumination map of the last gradient call."""
        out = np.zeros((self.Z, self.X), np.float32)
        check(self._lib.otfwi_get_illum(self._h, ptr(out), 0))
        return out

    def illum_device(self, dev_ptr):
        check(self._lib.otfwi_get_illum(self._h, ctypes.c_void_p(int(dev_ptr)), 1))

    def forward_for_grad(self, shot_src=(-1,)):
        """Stage 1 of a host-misfit gradient: traces of a forward run with the
        boundary ring saved on the device."""
        s, n = self._shots(shot_src)
        rec = np.zeros((n, self.n_rx, self.T), np.float32)
        check(self._lib.otfwi_forward_for_grad(self._h, n, ptr(s), ptr(rec)))
        return rec

    def adjoint_from_host(self, adj):
        """Stage 2: mean gradient from host adjoint sources (n_shots, n_rx, T)."""
        a = f32(adj)
        grad = np.zeros((self.Z, self.X), np.float32)
        check(self._lib.otfwi_adjoint_from_host(self._h, ptr(a), ptr(grad)))
        return grad


def misfit_exact(kind, sim, obs):
    """Per-trace misfit value and its exact derivative w.r.t. sim on the GPU (FP64)."""
    return misfit(kind, sim, obs, _exact=True)


def misfit(kind, sim, obs, _exact=False):
    """Per-trace misfit and adjoint source of (n, T) traces on the GPU (FP64)."""
    k = _lib.MISFIT_KINDS[kind]
    s = np.ascontiguousarray(sim, dtype=np.float64)
    o = np.ascontiguousarray(obs, dtype=np.float64)
    if s.ndim == 1:
        s, o = s[None], o[None]
    if s.shape != o.shape:
        raise ValueError("sim and obs shapes differ")
    n, T = s.shape
    mis = np.zeros(n, np.float64)
    adj = np.zeros((n, T), np.float64)
    fn = _lib.load().otfwi_misfit_exact if _exact else _lib.load().otfwi_misfit
    check(fn(k, n, T, ptr(s), ptr(o), ptr(mis), ptr(adj)))
    return mis, adj


def gradient_multi(models, shots, misfit="w2", workers=4, fused=False):
    """Gradients of several independent PMLModels (uneven sizes) at once.
    fused=False: scheduled longest-first over `workers` host threads, each
    model on its own stream (otfwi_gradient_multi).  fused=True: every
    two-step launch covers all models, the two-step phases captured into one
    graph each (otfwi_gradient_fused; FP32 models running two steps per
    launch, shots in one batch).  shots: one shot_src list per model;
    misfit: one kind or a list.  Returns [(grad (Z,X) FP32, per-shot misfit
    FP64)] in model order."""
    n = len(models)
    kinds = [misfit] * n if isinstance(misfit, str) else list(misfit)
    code = _lib.MISFIT_KINDS
    ss = [i32(np.atleast_1d(s)) for s in shots]
    grads = [np.zeros((m.Z, m.X), np.float32) for m in models]
    mis = [np.zeros(len(s), np.float64) for s in ss]
    P = ctypes.c_void_p
    arr_m = (P * n)(*[m._h.value for m in models])
    arr_n = (ctypes.c_int * n)(*[len(s) for s in ss])
    arr_s = (P * n)(*[s.ctypes.data for s in ss])
    arr_k = (ctypes.c_int * n)(*[code[k] for k in kinds])
    arr_mis = (P * n)(*[x.ctypes.data for x in mis])
    arr_g = (P * n)(*[g.ctypes.data for g in grads])
    if fused:
        check(_lib.load().otfwi_gradient_fused(n, arr_m, arr_n, arr_s, arr_k, arr_mis, arr_g))
    else:
        check(_lib.load().otfwi_gradient_multi(n, arr_m, arr_n, arr_s, arr_k, arr_mis, arr_g,
                                               int(workers)))
    return list(zip(grads, mis))
