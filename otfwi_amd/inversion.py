"""Inversion drivers over the propagator.

* :class:`FWI` - the plain method: an illumination-preconditioned gradient
  step on the velocity, accepted only if the misfit decreases (otherwise the
  step is halved and tried again), and a hook for a learned step that is
  judged by the same acceptance test and falls back to the plain step.
* :func:`lbfgs` - the reference's own driver flow (regression.py:34-99):
  scipy L-BFGS-B on ``1e3*mse`` with gradient ``grad_scale*gmask*grad`` and
  bounds ``[vmin, vmax]`` over :class:`Cuda_interface_PML`.

The reference has no plain step, no illumination and no acceptance test: it
hands its gradient to L-BFGS-B.  Its gradient (regression_PML_single.cu:
390-445) is the imaging condition with respect to c^2; the chain rule to the
velocity multiplies it by 2c, which keeps its sign.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .cuda_interface_pml import Cuda_interface_PML


def array_mask(Z, X, pml, rows_blocked, taper=8, band=8):
    """Update mask for an array on the top row: zero in the PML, on rows
    < rows_blocked (the array and its near field) and in a band of ``band``
    cells along the bottom and side PML (where the boundary ring is restored
    and the imaging condition carries edge artefacts), a cosine taper over
    ``taper`` rows below rows_blocked, one elsewhere."""
    m = np.zeros((Z, X), np.float32)
    m[pml:Z - pml - band, pml + band:X - pml - band] = 1.0
    m[:rows_blocked, :] = 0.0
    for k in range(taper):
        z = rows_blocked + k
        if z < Z - pml - band:
            m[z, :] *= 0.5 - 0.5 * np.cos(np.pi * (k + 1) / (taper + 1))
    return m


def plain_direction(grad, illum, c, mask, eps=1e-2):
    """Preconditioned descent direction on the velocity, scaled to max |d| = 1:
    d = -mask * 2c*grad / (illum + eps*max(illum)).  FP64."""
    g = 2.0 * np.asarray(c, np.float64) * np.asarray(grad, np.float64)
    ill = np.asarray(illum, np.float64)
    d = -np.asarray(mask, np.float64) * g / (ill + eps * ill.max())
    n = np.abs(d).max()
    return d / n if n > 0 else d


@dataclass
class Record:
    t: float            # wall seconds since the start of the run
    it: int
    misfit: float
    err: float          # relative L2 velocity error (nan without a true model)
    kind: str           # "init", "plain", "learned"
    alpha: float = 0.0


@dataclass
class History:
    records: list = field(default_factory=list)
    n_grad: int = 0     # gradient evaluations (forward + adjoint)
    n_obj: int = 0      # misfit-only evaluations (forward) of trial steps

    def time_to(self, tol):
        """Wall time at which the velocity error first fell to tol (None if never)."""
        for r in self.records:
            if r.err <= tol:
                return r.t
        return None

    @property
    def best_err(self):
        return min(r.err for r in self.records)


class FWI:
    """Plain illumination-preconditioned inversion on a Cuda_interface_PML
    whose init_regression has been called (observed data and ring set).

    A trial step is judged by the misfit alone (forward runs of all shots and
    the device misfit, no adjoint); the gradient and illumination are
    computed only at an accepted iterate, for the next iteration.  The
    forward-only misfit equals the gradient call's misfit bit for bit, so the
    iterates are those of judging every trial by a full gradient.
    """

    def __init__(self, cp: Cuda_interface_PML, mask, vmin, vmax, misfit="w2", illum_stride=4,
                 eps=1e-2, alpha0=0.02, alpha_max=0.1, max_halvings=3, c_true=None):
        self.cp = cp
        self.m = cp.model
        self.mask = np.asarray(mask, np.float32)
        self.vmin, self.vmax = float(vmin), float(vmax)
        self.misfit = misfit
        self.eps = eps
        self.alpha = alpha0
        self.alpha_max = alpha_max
        self.max_halvings = max_halvings
        self.c_true = None if c_true is None else np.asarray(c_true, np.float64)
        self.shots = np.arange(cp.n_source, dtype=np.int32)
        self.m.set_illum(illum_stride)

    def err(self, c):
        if self.c_true is None:
            return float("nan")
        return float(np.linalg.norm(c - self.c_true) / np.linalg.norm(self.c_true))

    def evaluate(self, c):
        """(mean misfit over shots, grad (Z,X) FP32, illumination (Z,X) FP32) at c."""
        self.m.set_cquad((np.asarray(c, np.float64) ** 2).astype(np.float32))
        g, mis = self.m.gradient(self.shots, self.misfit)
        return float(np.mean(mis)), g, self.m.illum()

    def objective(self, c):
        """Mean misfit over shots at c (forward runs only)."""
        self.m.set_cquad((np.asarray(c, np.float64) ** 2).astype(np.float32))
        return float(np.mean(self.m.objective(self.shots, self.misfit)))

    def clip(self, c):
        return np.clip(c, self.vmin, self.vmax)

    def run(self, c0, n_iter, learned=None, tol=None, time_limit=None, patience=1,
            learned_scales=(1.0, 0.5), progress=None):
        """Iterate from c0.  learned(grad, illum, c) -> proposed velocity update
        (or None); the proposal, scaled by each of learned_scales in turn, is
        kept at the first scale that decreases the misfit, otherwise the plain
        step is taken; after `patience` rejected proposals no more are asked
        for.  Stops after n_iter iterations, at the velocity-error
        tolerance tol, at time_limit seconds, or when no step size decreases
        the misfit.  progress(record), when given, is called with every
        iteration's record."""
        hist = History()
        t0 = time.perf_counter()
        c = np.asarray(c0, np.float64).copy()
        J, g, ill = self.evaluate(c)
        hist.n_grad += 1
        hist.records.append(Record(time.perf_counter() - t0, 0, J, self.err(c), "init"))
        rejected = 0
        for it in range(1, n_iter + 1):
            if tol is not None and hist.records[-1].err <= tol:
                break
            if time_limit is not None and time.perf_counter() - t0 > time_limit:
                break
            accepted = None
            if learned is not None and rejected < patience:
                dc = learned(g, ill, c)
                for scale in (learned_scales if dc is not None else ()):
                    ct = self.clip(c + scale * dc)
                    Jt = self.objective(ct)
                    hist.n_obj += 1
                    if Jt < J:
                        accepted = ("learned", ct, Jt, scale)
                        break
                if dc is not None and accepted is None:
                    rejected += 1
            if accepted is None:
                d = plain_direction(g, ill, c, self.mask, self.eps)
                a = self.alpha
                for _ in range(self.max_halvings + 1):
                    ct = self.clip(c + a * d)
                    Jt = self.objective(ct)
                    hist.n_obj += 1
                    if Jt < J:
                        accepted = ("plain", ct, Jt, a)
                        self.alpha = min(1.5 * a, self.alpha_max)
                        break
                    a *= 0.5
                else:
                    self.alpha = a
            if accepted is None:
                break
            kind, c, J, a = accepted
            hist.records.append(Record(time.perf_counter() - t0, it, J, self.err(c), kind, a))
            if progress is not None:
                progress(hist.records[-1])
            # the gradient at the accepted iterate, for the next iteration
            if it < n_iter and not (tol is not None and hist.records[-1].err <= tol):
                J, g, ill = self.evaluate(c)
                hist.n_grad += 1
        self.c = c
        return c, hist


def regression_objective(cp: Cuda_interface_PML, gmask, grad_scale):
    """The objective regression.py:45-49 hands to L-BFGS-B: calc_grad at the
    velocity ``params`` (flattened (Z, X)), value ``1e3*mse``, gradient
    ``grad_scale*(gmask*grad)`` flattened."""
    Z, X = cp.Z, cp.X

    def objfun(params):
        cp.calc_grad(c=np.asarray(params).reshape(Z, X))
        grad = grad_scale * (gmask * cp.grad).flatten()
        return 1e3 * float(cp.mse[0]), grad.astype(np.float64)

    return objfun


def model_misfit(params, c_true):
    """Model misfit the reference's callback logs per iteration
    (regression.py:76-77): ||params - c_true||^2 / (Nz*Nx)."""
    c_true = np.asarray(c_true, np.float64)
    return float(np.linalg.norm(np.asarray(params).reshape(c_true.shape) - c_true) ** 2
                 / c_true.size)


def lbfgs(cp: Cuda_interface_PML, c0, vmin, vmax, gmask, grad_scale, maxiter=200, callback=None,
          options=None, c_true=None):
    """regression.py:45-87 on this library: minimize 1e3*mse over the
    velocity with L-BFGS-B, jac = grad_scale*(gmask*grad), bounds [vmin, vmax]
    (options gtol = ftol = 1e-8, maxcor 20, maxls 10).  With c_true, the
    result carries ``model_misfit``: the model misfit after each iteration,
    as the reference's callback records it.  Returns scipy's OptimizeResult."""
    from scipy.optimize import Bounds, minimize

    objfun = regression_objective(cp, gmask, grad_scale)
    log = []

    def cb(x):
        if c_true is not None:
            log.append(model_misfit(x, c_true))
        if callback is not None:
            callback(x)

    opt = {"gtol": 1e-8, "ftol": 1e-8, "maxiter": maxiter, "maxcor": 20, "maxls": 10}
    if options:
        opt.update(options)
    r = minimize(objfun, np.asarray(c0, np.float64).flatten(), jac=True, method="L-BFGS-B",
                 bounds=Bounds(vmin, vmax), options=opt, callback=cb)
    r.model_misfit = np.array(log)
    return r
