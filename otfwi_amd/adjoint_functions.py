"""Misfits with adjoint sources, computed on the GPU.

Same names, signatures and return values as the reference's
adjoint_functions.py: ``f(sim, obs) -> (misfit, adjoint_source (n_rx, T)
float64)``.

* :func:`L2` -- adjoint_functions.py:8-11: (sum (sim-obs)^2, sim-obs).
* :func:`wasserstein` -- adjoint_functions.py:25-67: per-trace 1D W2 of the
  affinely shifted, normalized traces, with the unit-spacing transport-map
  integral as adjoint source; one wavefront per trace with FP64 CDF scans and
  inverse-CDF searches.
* :func:`quad_envelope` -- adjoint_functions.py:14-22: squared-envelope
  difference E = |hilbert(sim)|^2 - |hilbert(obs)|^2, misfit sum E^2, adjoint
  source sim E - Im(hilbert(E Im(hilbert(sim)))); the Hilbert transforms are
  FP64 circular convolutions with scipy.signal.hilbert's kernel.
* :func:`wasserstein_quad`, :func:`wasserstein_experimental`,
  :func:`wasserstein_pm`, :func:`wasserstein_yang` --
  adjoint_functions.py:70-266: the further Wasserstein-type misfits
  (squared-trace densities; the t-weighted transport term; positive parts;
  the suffix-sum adjoint of Yang's formulation), one wavefront per trace in
  FP64, computed from FP64 copies of the traces.

:class:`otfwi_amd.cuda_interface_pml.Cuda_interface_PML` recognizes these
functions and keeps the whole gradient on the device; any other callable is
called on the host once per shot, as the reference's callback does.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import check, ptr

__all__ = ["L2", "wasserstein", "quad_envelope", "wasserstein_quad", "wasserstein_experimental",
           "wasserstein_pm", "wasserstein_yang"]


def _run(kind, sim, obs):
    s = np.ascontiguousarray(sim, dtype=np.float64)
    o = np.ascontiguousarray(obs, dtype=np.float64)
    squeeze = s.ndim == 1
    if squeeze:
        s, o = s[None], o[None]
    if s.shape != o.shape or s.ndim != 2:
        raise ValueError(f"sim {s.shape} and obs {o.shape} must be equal (n_rx, T) arrays")
    n, T = s.shape
    mis = np.zeros(n, np.float64)
    adj = np.zeros((n, T), np.float64)
    check(_lib.load().otfwi_misfit(kind, n, T, ptr(s), ptr(o), ptr(mis), ptr(adj)))
    # traces are accumulated in trace order (adjoint_functions.py:33, 64)
    total = 0.0
    for v in mis:
        total += v
    return total, (adj[0] if squeeze else adj)


def L2(sim, obs):
    """(sum((sim-obs)^2), sim-obs) -- adjoint_functions.py:8-11."""
    return _run(_lib.MISFIT_L2, sim, obs)


def wasserstein(sim, obs):
    """1D Wasserstein-2 misfit and adjoint source -- adjoint_functions.py:25-67."""
    return _run(_lib.MISFIT_W2, sim, obs)


def quad_envelope(sim, obs):
    """Squared-envelope misfit and adjoint source -- adjoint_functions.py:14-22."""
    return _run(_lib.MISFIT_ENV, sim, obs)


def wasserstein_quad(sim, obs):
    """W2 of the squared traces shifted by 5 % of max(obs^2), rectangle-rule
    CDFs -- adjoint_functions.py:70-104."""
    return _run(_lib.MISFIT_W2_QUAD, sim, obs)


def wasserstein_experimental(sim, obs):
    """W2 weighted by F^-1(t), adjoint t((t - q)^2 - 2tq) -- adjoint_functions.py:107-141."""
    return _run(_lib.MISFIT_W2_EXPERIMENTAL, sim, obs)


def wasserstein_pm(sim, obs):
    """W2 of the positive parts, counted twice as the reference's negative
    half reuses them -- adjoint_functions.py:144-223."""
    return _run(_lib.MISFIT_W2_PM, sim, obs)


def wasserstein_yang(sim, obs):
    """W2 with the suffix-sum adjoint of Yang's formulation -- adjoint_functions.py:226-266."""
    return _run(_lib.MISFIT_W2_YANG, sim, obs)


# markers the gradient driver uses to keep these misfits on the device
L2._otfwi_misfit = "l2"
wasserstein._otfwi_misfit = "w2"
quad_envelope._otfwi_misfit = "env"
wasserstein_quad._otfwi_misfit = "w2_quad"
wasserstein_experimental._otfwi_misfit = "w2_experimental"
wasserstein_pm._otfwi_misfit = "w2_pm"
wasserstein_yang._otfwi_misfit = "w2_yang"
