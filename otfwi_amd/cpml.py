"""Second-order C-PML pressure scheme (simulator/src/regression_CPML.cu) on
the GPU: ctypes binding of libotfwi_cpml.so (include/otfwi_cpml.h).

:class:`CPMLModel` runs batches of shots through the handle API; the
library also exports the reference's own C ABI (init_memory_sim,
cuda_simulate2, setCquad, set_source), so code that loads
cuda_regression_CPML.so can load it instead.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import OtfwiError, _runtime_first

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libotfwi_cpml.so")

_vp, _i = ctypes.c_void_p, ctypes.c_int
SIGS = [
    ("otfwi_cpml_last_error", ctypes.c_char_p, []),
    ("otfwi_cpml_create", _i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    ("otfwi_cpml_destroy", _i, [_vp]),
    ("otfwi_cpml_set_cquad", _i, [_vp, _vp]),
    ("otfwi_cpml_set_points", _i, [_vp, _i, _vp, _vp, _vp, _i, _vp, _vp]),
    ("otfwi_cpml_forward", _i, [_vp, _i, _vp, _vp]),
    ("init_memory_sim", None, None),
    ("setCquad", None, [_vp]),
    ("set_source", None, [_i, _vp, _vp]),
    ("cuda_simulate2", None, [_i, _i]),
    ("def_python_adj", None, [_vp]),
]
EXPORTED = [s[0] for s in SIGS]
_LIB = None

# profile keys in the order otfwi_cpml_create takes them
PROFILES = ("K_x", "a_x", "b_x", "K_xh", "a_xh", "b_xh", "K_z", "a_z", "b_z", "K_zh", "a_zh", "b_zh")


def load():
    global _LIB
    if _LIB is None:
        _runtime_first()
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "(make -C otfwi_amd/csrc); there is no CPU fallback")
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGS:
            f = getattr(lib, name)
            f.restype = res
            if args is not None:
                f.argtypes = args
        _LIB = lib
    return _LIB


def _check(rc):
    if rc != 0:
        msg = load().otfwi_cpml_last_error()
        raise OtfwiError(rc, msg.decode() if msg else "")


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ptrs(arrs):
    return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


class CPMLModel:
    """One medium, sources and receivers; forward(shots) runs the shots
    together.  Arguments mirror init_memory_sim (regression_CPML.cu:349-361):
    cquad and kappa ([z][x]), rho at the half points in x and z, the twelve
    K / a / b profiles (full and half points along x, then z), dx, dz, dt^2,
    the source scale, pixels per element and the initial state (P at t = 0, 1)."""

    def __init__(self, X, Z, T, *, cquad, kappa, rho_hx, rho_hz, profiles, sx, sz, src, rx, rz,
                 dx=1.0, dz=1.0, dt=1.0, dt2=1.0, qt_pi=4 * np.pi, ppe=1, initial=None):
        self._lib = load()
        self.X, self.Z, self.T = int(X), int(Z), int(T)
        self._keep = [_f32(a).reshape(Z, X) for a in (cquad, kappa, rho_hx, rho_hz)]
        prof = [_f32(profiles[k]) for k in PROFILES]
        for k, a in zip(PROFILES, prof):
            n = X if "_x" in k else Z
            if a.shape != (n,):
                raise ValueError(f"profile {k} must have {n} values, has {a.shape}")
        consts = np.array([dx, dz, dt, dt2, qt_pi], np.float32)
        init = None if initial is None else _f32(initial).reshape(2, Z, X)
        h = ctypes.c_void_p()
        _check(self._lib.otfwi_cpml_create(ctypes.byref(h), self.X, self.Z, self.T, int(ppe),
                                           _ptr(consts), _ptrs(self._keep), _ptrs(prof),
                                           None if init is None else _ptr(init)))
        self._h = h
        self.sx, self.sz, self.rx, self.rz = _i32(sx), _i32(sz), _i32(rx), _i32(rz)
        s = _f32(src).reshape(len(self.sx), self.T)
        _check(self._lib.otfwi_cpml_set_points(self._h, len(self.sx), _ptr(self.sx), _ptr(self.sz),
                                               _ptr(s), len(self.rx), _ptr(self.rx),
                                               _ptr(self.rz)))

    @classmethod
    def from_problem(cls, p):
        """From a dict with the keys of tests/problems.cpml_problem."""
        return cls(p["X"], p["Z"], p["T"], cquad=p["cquad"], kappa=p["kappa"],
                   rho_hx=p["rho_hx"], rho_hz=p["rho_hz"], profiles=p, sx=p["sx"], sz=p["sz"],
                   src=p["src"], rx=p["rx"], rz=p["rz"], dx=p["dx"], dz=p["dz"],
                   dt=p.get("dt", 1.0), dt2=p["dt2"], qt_pi=p["qt_pi"], ppe=p["ppe"],
                   initial=p.get("initial"))

    def set_cquad(self, cquad):
        c = _f32(cquad).reshape(self.Z, self.X)
        _check(self._lib.otfwi_cpml_set_cquad(self._h, _ptr(c)))

    def forward(self, shots):
        """Traces (n_shots, n_rx, T) float32; shot k fires source shots[k]
        (-1: every source; every source regardless when ppe != 1)."""
        idx = _i32(np.atleast_1d(shots))
        out = np.zeros((len(idx), len(self.rx), self.T), np.float32)
        _check(self._lib.otfwi_cpml_forward(self._h, len(idx), _ptr(idx), _ptr(out)))
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.otfwi_cpml_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
