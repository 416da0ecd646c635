"""The reference's own ctypes binding on libotfwi_pml.so.

The reference loads its library with CDLL, sets restype = None on the five
entry points and declares no argtypes (cuda_interface_pml.py:9-30); it hands
arrays over as ctypes arrays built from Python lists, receives the library's
pinned result buffers through float** out-parameters stored in a numpy scalar
(:86-114, 151-159), and passes the adjoint callback as a CFUNCTYPE(None,
float*, float*) cast to c_void_p (:67-68), which converts the simulated
traces to a Python list and writes the adjoint source back through a cast
pointer (:238-249).  These tests drive the library exactly that way:
init_memory_sim -> init_memory_reg -> setCquad -> cuda_simulate ->
cuda_grad_ext, and compare traces and gradients with the oracle.
"""
import ctypes
from ctypes import CDLL, CFUNCTYPE, POINTER, c_float, c_int

import numpy as np
import pytest

from oracle import oracle as O
from tests.problems import rel, small_problem

pytestmark = pytest.mark.gpu

ADJ_FN = CFUNCTYPE(None, POINTER(c_float), POINTER(c_float))


def _bind():
    from otfwi_amd import _lib

    lib = CDLL(_lib.LIB_PATH)          # the same shared object, no argtypes declared
    for name in ("init_memory_sim", "init_memory_reg", "setCquad", "cuda_simulate",
                 "cuda_grad_ext"):
        getattr(lib, name).restype = None
    return lib


def _farr(a):
    a = np.asarray(a, np.float32).ravel()
    return (c_float * len(a))(*a)


def _iarr(a):
    a = np.asarray(a).ravel()
    return (c_int * len(a))(*[int(v) for v in a])


def _outptr():
    # a float** out-parameter in an 8-byte numpy scalar, as the reference keeps it
    return np.asarray(0).ctypes.data_as(POINTER(POINTER(c_float)))


class RefStyle:
    """The call sequence of the reference wrapper on one process-global model."""

    def __init__(self, lib, p, cq):
        self.lib, self.p = lib, p
        self.n_rx, self.T = len(p["rx_xz"][0]), p["T"]
        self.recptr, self.gradptr = _outptr(), _outptr()
        sx, sz = p["src_xz"]
        rx, rz = p["rx_xz"]
        init = np.zeros((p["Z"], p["X"], 2))
        lib.init_memory_sim(c_int(p["X"]), c_int(p["Z"]), c_int(p["T"]), _farr(cq),
                            _farr([1, 1, 1]), _farr(p["d_x"]), _farr(p["d_z"]), c_int(len(sx)),
                            _iarr(sx), _iarr(sz), c_int(len(rx)), _iarr(rx), _iarr(rz),
                            _farr(p["src"]), _farr(init), self.recptr)

    def init_reg(self, obs):
        vx, vz = self.p["ring_xz"]
        self.lib.init_memory_reg(_farr(obs), self.gradptr, _iarr(vx), _iarr(vz), c_int(len(vx)))

    def set_cquad(self, cq):
        cq = np.ascontiguousarray(cq, np.float32)
        self.lib.setCquad(cq.ctypes.data_as(POINTER(c_float)))

    def simulate(self, s):
        self.lib.cuda_simulate(0, c_int(s))
        out = np.zeros((self.n_rx, self.T), np.float32)
        ctypes.memmove(out.ctypes.data_as(POINTER(c_float)), self.recptr[0],
                       self.n_rx * self.T * ctypes.sizeof(c_float))
        return out

    def grad(self, s, callback):
        mse_f = (c_float * 1)(0.0)
        self.lib.cuda_grad_ext(mse_f, callback, c_int(s))
        g = np.zeros((self.p["Z"], self.p["X"]), np.float32)
        ctypes.memmove(g.ctypes.data_as(POINTER(c_float)), self.gradptr[0],
                       self.p["Z"] * self.p["X"] * ctypes.sizeof(c_float))
        return g, mse_f[0]


def _list_callback(n_rx, T, obs, misfit, log):
    """The reference's adj_wrapper conversions: traces through a Python list
    (float64), adjoint source written through a cast c_float array."""
    def cb(sim_c, adj_c):
        sim = np.asarray(sim_c[:T * n_rx]).reshape((n_rx, T))
        mse, src = misfit(sim, obs)
        dst = ctypes.cast(adj_c, POINTER(c_float * (T * n_rx)))[0]
        dst[:T * n_rx] = src.flatten()
        log.append((mse, sim))
    fn = ADJ_FN(cb)
    return fn, ctypes.cast(fn, ctypes.c_void_p)


def test_reference_binding_traces_and_gradients(gpu):
    p = small_problem(seed=21, n_src=2, n_rx=6)
    lib = _bind()
    ref = RefStyle(lib, p, p["cq_true"])
    obs = []
    for s in (0, 1):
        rec = ref.simulate(s)
        o = O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"],
                      p["T"], idx=s)["rec"]
        e = rel(rec, o)
        print(f"reference binding, shot {s}: traces vs oracle {e:.3e}", flush=True)
        assert e <= 1e-5 and np.all(rec[:, 0] == 0)
        obs.append(rec)
    obs = np.stack(obs)
    ref.init_reg(obs)
    ref.set_cquad(p["cq_bg"])
    for s, misfit, bar in ((0, O.l2, 1e-4), (1, O.l2, 1e-4), (1, O.wasserstein, 2e-3)):
        log = []
        fn, cb = _list_callback(len(p["rx_xz"][0]), p["T"], obs[s], misfit, log)
        g, mse_f = ref.grad(s, cb)
        assert len(log) == 1 and mse_f == 0.0      # the reference leaves *mse at 0
        g_ref, mse_ref, _ = O.gradient(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                                       p["rx_xz"], p["ring_xz"], obs[s], p["T"], misfit,
                                       shots=(s,))
        eg = rel(g, g_ref)
        em = abs(log[0][0] - mse_ref) / abs(mse_ref)
        print(f"reference binding, shot {s}, {misfit.__name__}: gradient vs oracle {eg:.3e}, "
              f"misfit {em:.3e}", flush=True)
        assert eg <= bar and em <= (1e-6 if misfit is O.l2 else 1e-5)
        del fn


def test_wrapper_adj_wrapper_as_the_library_callback(gpu):
    """Cuda_interface_PML.adj_wrapper (cuda_interface_pml.py:238-249) handed to
    cuda_grad_ext as the reference hands its own: same gradient as the
    wrapper's calc_grad and as the oracle."""
    from otfwi_amd.cuda_interface_pml import Cuda_interface_PML

    p = small_problem(seed=22, n_src=2, n_rx=6)
    obs = np.stack([O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                              p["rx_xz"], p["T"], idx=s)["rec"] for s in (0, 1)])
    c_bg = np.sqrt(p["cq_bg"].astype(np.float64))
    iface = Cuda_interface_PML(p["X"], p["Z"], p["T"], p["pml"], c_bg, 2, p["src_mask"], 6,
                               p["rx_mask"], p["src"], multishot=True, R=1e-3)
    iface.init_regression(obs, O.l2, prec_deriv=4)
    lib = _bind()
    ref = RefStyle(lib, p, p["cq_bg"])
    ref.init_reg(obs)
    fn = ADJ_FN(iface.adj_wrapper)
    cb = ctypes.cast(fn, ctypes.c_void_p)
    iface.idx_source = 1
    g, _ = ref.grad(1, cb)
    g_ref, mse_ref, _ = O.gradient(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                                   p["rx_xz"], p["ring_xz"], obs[1], p["T"], O.l2, shots=(1,))
    print(f"adj_wrapper callback: gradient vs oracle {rel(g, g_ref):.3e}", flush=True)
    assert rel(g, g_ref) <= 1e-4
    assert abs(iface.mse[0] - mse_ref) <= 1e-6 * abs(mse_ref)
    iface.calc_grad(idx_source=1)                 # multishot: mean over both sources
    g2, _, _ = O.gradient(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"],
                          p["ring_xz"], obs, p["T"], O.l2, shots=(0, 1))
    assert rel(iface.grad, g2) <= 1e-4
    del fn
