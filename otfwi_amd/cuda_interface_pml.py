"""Drop-in replacement for the reference's ``cuda_interface_pml.Cuda_interface_PML``.

Same constructor, methods, attributes and array layouts as
cuda_interface_pml.py:33-249, running on the MI355X library through the
batched handle API instead of the per-shot ctypes loop:

* ``simulate`` runs every shot of a multishot setup in one batched forward
  (cuda_interface_pml.py:185-206 loops over ``cuda_simulate``);
* ``calc_grad`` with any of the seven ``adjoint_functions`` misfits (either
  the reference's module or :mod:`otfwi_amd.adjoint_functions`, recognized by
  name) computes the misfit and adjoint source on the device and reduces the
  per-shot gradients in FP64 (cuda_interface_pml.py:208-236 reduces in FP32);
  any other ``adj_fun`` is called on the host once per shot with
  ``(simulated float64, observed)`` exactly as ``adj_wrapper`` calls it.

Extra keyword arguments (not in the reference): ``R`` (PML reflection
coefficient, default 1e-5 as hard-coded at cuda_interface_pml.py:117),
``vmax`` (default 0.5), ``precision`` (4 = FP32, 8 = FP64 validation build)
and ``max_batch`` (shots resident at once, 0 = all).
"""
from __future__ import annotations

import time

import numpy as np

from .model import PMLModel

__all__ = ["Cuda_interface_PML"]


def _device_kind(adj_fun):
    kind = getattr(adj_fun, "_otfwi_misfit", None)
    if kind is not None:
        return kind
    mod = getattr(adj_fun, "__module__", "") or ""
    name = getattr(adj_fun, "__name__", "")
    known = {"L2": "l2", "wasserstein": "w2", "quad_envelope": "env",
             "wasserstein_quad": "w2_quad", "wasserstein_experimental": "w2_experimental",
             "wasserstein_pm": "w2_pm", "wasserstein_yang": "w2_yang"}
    if mod.split(".")[-1] == "adjoint_functions" and name in known:
        return known[name]
    return None


class Cuda_interface_PML:
    def __init__(self, X, Z, T, PML_size, c, n_source, pos_source, n_sensor, pos_sensor, source,
                 initial=None, multishot=False, *, R=1e-5, vmax=0.5, precision=4, max_batch=0,
                 verbose=False):
        self.pos_sensor = np.where(pos_sensor)
        self.pos_source = np.where(pos_source)
        self.X = X
        self.Z = Z
        self.T = T
        self.PML_size = PML_size
        self.R = R
        self.Vmax = vmax
        self.cquad = (np.asarray(c) ** 2).astype(np.single)
        self.n_source = n_source
        self.pos_source_x = self.pos_source[1]
        self.pos_source_z = self.pos_source[0]
        self.n_sensor = n_sensor
        self.pos_sensor_x = self.pos_sensor[1]
        self.pos_sensor_z = self.pos_sensor[0]
        self.source = np.asarray(source).astype(np.single)
        self.d_x, self.d_z = self._calcPML()
        self.constvec = np.asarray([1, 1, 1]).astype(np.single)
        self._pos_revert = None
        self._pos_revert_x = None
        self._pos_revert_z = None
        self._tammaskrev = 0
        self.grad = np.zeros((Z, X)).astype(np.single)
        self.mse = np.zeros(1)
        self._adj_fun = None
        self._multishot = multishot
        if multishot:
            self.recording = np.zeros((n_source, n_sensor, T)).astype(np.single)
        else:
            self.recording = np.zeros((n_sensor, T)).astype(np.single)
        self.idx_source = -1
        # the reference accepts an initial state it never uses (regression_PML_single.cu:146)
        self.initial = np.zeros((Z, X, 2)) if initial is None else initial
        self.observed = None
        self.verbose = verbose
        if len(self.pos_source_x) != n_source or len(self.pos_sensor_x) != n_sensor:
            raise ValueError("n_source / n_sensor do not match the masks")
        self._model = PMLModel(X, Z, T, precision=precision, constvec=self.constvec)
        self._model.set_damping(self.d_x, self.d_z)
        self._model.set_cquad(self.cquad)
        self._model.set_sources(self.pos_source_x, self.pos_source_z,
                                self.source.reshape(n_source, T))
        self._model.set_receivers(self.pos_sensor_x, self.pos_sensor_z)
        if max_batch:
            self._model.set_tuning(0, 0, max_batch)

    @property
    def model(self):
        """The native :class:`otfwi_amd.model.PMLModel` behind this wrapper."""
        return self._model

    def _calcPML(self):
        """Quadratic damping profile (cuda_interface_pml.py:116-132)."""
        L = self.PML_size
        d0 = 3 * self.Vmax * np.log(1 / self.R) / (2 * L ** 3)
        prof = d0 * np.linspace(0, L, L) ** 2
        d_x = np.zeros((self.Z, self.X))
        d_z = np.zeros((self.Z, self.X))
        d_z[:L, :] = prof[::-1, None]
        d_z[self.Z - L:, :] = prof[:, None]
        d_x[:, self.X - L:] = prof[None, :]
        d_x[:, :L] = prof[None, ::-1]
        return d_x, d_z

    def init_regression(self, observed, adj_fun, prec_deriv=8):
        """Boundary ring + observed data + misfit (cuda_interface_pml.py:134-159)."""
        margem = prec_deriv + 1
        L = self.PML_size
        mask = np.full((self.Z, self.X), False)
        mask[L:L + margem, L:-L] = True
        mask[-L - margem:-L, L:-L] = True
        mask[L:-L, L:L + margem] = True
        mask[L:-L, -L - margem:-L] = True
        mask[self.pos_source] = True
        self._pos_revert = np.where(mask)
        self._pos_revert_x = self._pos_revert[1]
        self._pos_revert_z = self._pos_revert[0]
        self._tammaskrev = int(np.sum(mask))
        self.observed = observed
        self._adj_fun = adj_fun
        self._model.set_ring(self._pos_revert_x, self._pos_revert_z)
        obs = np.asarray(observed, dtype=np.float32)
        self._model.set_observed(obs if self._multishot else obs.reshape(1, self.n_sensor, self.T))

    def _set_c_quad(self, cquad):
        self.cquad = np.asarray(cquad).astype(np.single)
        self._model.set_cquad(self.cquad.reshape(self.Z, self.X))

    def _set_source(self, n_source, pos_source, source):
        """New sources from a boolean mask (the reference routes this to a
        library symbol it never exports, cuda_interface_pml.py:166-183)."""
        self.pos_source = np.where(pos_source)
        self.source = np.asarray(source).astype(np.single)
        self.n_source = n_source
        self.pos_source_x = self.pos_source[1]
        self.pos_source_z = self.pos_source[0]
        self._model.set_sources(self.pos_source_x, self.pos_source_z,
                                self.source.reshape(n_source, self.T))

    def _shots(self, idx_source):
        self.idx_source = -1 if idx_source is None else idx_source
        if not self._multishot:
            return [int(self.idx_source)]
        return list(range(self.n_source))

    def simulate(self, output=False, c=None, idx_source=None):
        """Forward run(s) into ``self.recording`` (cuda_interface_pml.py:185-206).
        ``output`` (video frames) is accepted and ignored."""
        if c is not None:
            self._set_c_quad(np.asarray(c) ** 2)
        shots = self._shots(idx_source)
        rec = self._model.forward(shots)
        if self._multishot:
            self.recording[:, :, :] = rec
        else:
            self.recording[:, :] = rec[0]

    def calc_grad(self, c=None, idx_source=None):
        """Misfit and gradient into ``self.mse`` / ``self.grad``
        (cuda_interface_pml.py:208-236): one shot, or the mean over every
        source of a multishot setup."""
        if self._adj_fun is None:
            raise RuntimeError("init_regression must be called before calc_grad")
        if c is not None:
            self._set_c_quad(np.asarray(c) ** 2)
        shots = self._shots(idx_source)
        start = time.time()
        kind = _device_kind(self._adj_fun)
        if kind is not None:
            grad, mis = self._model.gradient(shots, misfit=kind)
        else:
            sim = self._model.forward_for_grad(shots)
            adj = np.zeros_like(sim)
            mis = np.zeros(len(shots))
            for k in range(len(shots)):
                self.idx_source = shots[k] if self._multishot else self.idx_source
                mis[k] = self._host_adjoint(sim[k], adj[k])
            grad = self._model.adjoint_from_host(adj)
        if self.verbose:
            print(time.time() - start)
        self.grad = grad
        self.mse[0] = float(np.mean(mis)) if self._multishot else float(mis[0])

    def _host_adjoint(self, sim32, adj_out):
        sim = sim32.astype(np.float64)
        if self._multishot:
            mse, src = self._adj_fun(sim, self.observed[self.idx_source, :, :])
        else:
            mse, src = self._adj_fun(sim, self.observed)
        adj_out[:, :] = np.asarray(src).reshape(self.n_sensor, self.T)
        return mse

    def adj_wrapper(self, sim_c, adjoint_source_c):
        """Callback with the reference's ctypes signature (cuda_interface_pml.py:238-249):
        reads n_sensor*T floats from sim_c, writes the adjoint source into
        adjoint_source_c and sets ``self.mse[0]``."""
        n = self.T * self.n_sensor
        sim = np.ctypeslib.as_array(sim_c, shape=(n,)).reshape(self.n_sensor, self.T)
        out = np.ctypeslib.as_array(adjoint_source_c, shape=(n,))
        tmp = np.zeros((self.n_sensor, self.T), np.float32)
        self.mse[0] = self._host_adjoint(sim, tmp)
        out[:] = tmp.reshape(-1)
