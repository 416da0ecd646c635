"""ctypes binding of libotfwi_pml.so (include/otfwi_pml.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make -C
otfwi_amd/csrc``).  There is no fallback: importing the package without the
built library, or calling it without a GPU, raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libotfwi_pml.so")

OK, E_ARG, E_HIP, E_MEM, E_STATE, E_NUMERIC = 0, -1, -2, -3, -4, -5
MISFIT_L2, MISFIT_W2, MISFIT_ENV = 0, 1, 2
UNET_HEAD_WORK = 1024 * 34   # floats of the head-backward workspace (include/otfwi_unet.h)
MISFIT_W2_QUAD, MISFIT_W2_EXPERIMENTAL, MISFIT_W2_PM, MISFIT_W2_YANG = 3, 4, 5, 6
# misfit names of the Python layers -> kind codes of the C ABI
MISFIT_KINDS = {"l2": MISFIT_L2, "w2": MISFIT_W2, "env": MISFIT_ENV,
                "w2_quad": MISFIT_W2_QUAD, "w2_experimental": MISFIT_W2_EXPERIMENTAL,
                "w2_pm": MISFIT_W2_PM, "w2_yang": MISFIT_W2_YANG}

# (name, restype, argtypes) of every exported symbol declared in include/otfwi_pml.h
_vp = ctypes.c_void_p
_i = ctypes.c_int
_SIGS = [
    ("init_memory_sim", None, [_i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _i, _vp, _vp, _vp,
                               _vp, _vp]),
    ("init_memory_reg", None, [_vp, _vp, _vp, _vp, _i]),
    ("setCquad", None, [_vp]),
    ("cuda_simulate", None, [_i, _i]),
    ("cuda_grad_ext", None, [_vp, _vp, _i]),
    ("def_python_adj", None, [_vp]),
    ("otfwi_last_error", ctypes.c_char_p, []),
    ("otfwi_device_info", _i, [ctypes.c_char_p, _i]),
    ("otfwi_create", _i, [_vp, _i, _i, _i, _i, _vp]),
    ("otfwi_destroy", _i, [_vp]),
    ("otfwi_set_damping", _i, [_vp, _vp, _vp]),
    ("otfwi_set_cquad", _i, [_vp, _vp]),
    ("otfwi_set_cquad_device", _i, [_vp, _vp]),
    ("otfwi_set_sources", _i, [_vp, _i, _vp, _vp, _vp]),
    ("otfwi_set_receivers", _i, [_vp, _i, _vp, _vp]),
    ("otfwi_set_ring", _i, [_vp, _i, _vp, _vp]),
    ("otfwi_set_observed", _i, [_vp, _i, _vp]),
    ("otfwi_set_tuning", _i, [_vp, _i, _i, _i]),
    ("otfwi_get_tuning", _i, [_vp, _vp, _vp, _vp]),
    ("otfwi_set_launch", _i, [_vp, _i, _i]),
    ("otfwi_get_launch", _i, [_vp, _vp, _vp]),
    ("otfwi_set_graph", _i, [_vp, _i]),
    ("otfwi_set_blocking", _i, [_vp, _i, _i, _i, _i]),
    ("otfwi_get_blocking", _i, [_vp, _vp]),
    ("otfwi_set_fwd_steps", _i, [_vp, _i]),
    ("otfwi_get_fwd_steps", _i, [_vp, _vp]),
    ("otfwi_set_adj_steps", _i, [_vp, _i]),
    ("otfwi_kernel_times", _i, [_vp, _i, _i, _vp]),
    ("otfwi_get_adj_steps", _i, [_vp, _vp]),
    ("otfwi_set_pair_width", _i, [_vp, _i]),
    ("otfwi_set_pair_streams", _i, [_vp, _i]),
    ("otfwi_steps_per_launch", _i, [_vp, _i, _vp, _vp]),
    ("otfwi_loop_modes", _i, [_vp, _i, _vp]),
    ("otfwi_autotune", _i, [_vp, _i, _i, _vp, _vp]),
    ("otfwi_get_timing", _i, [_vp, _vp]),
    ("otfwi_forward", _i, [_vp, _i, _vp, _vp]),
    ("otfwi_gradient", _i, [_vp, _i, _vp, _i, _vp, _vp, _vp]),
    ("otfwi_gradient_device", _i, [_vp, _i, _vp, _i, _vp, _vp]),
    ("otfwi_forward_for_grad", _i, [_vp, _i, _vp, _vp]),
    ("otfwi_adjoint_from_host", _i, [_vp, _vp, _vp]),
    ("otfwi_objective", _i, [_vp, _i, _vp, _i, _vp]),
    ("otfwi_gradient_multi", _i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _i]),
    ("otfwi_gradient_fused", _i, [_i, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("otfwi_gradient_exact", _i, [_vp, _i, _vp, _i, _i, _vp, _vp, _vp]),
    ("otfwi_gradient_exact_device", _i, [_vp, _i, _vp, _i, _i, _vp, _vp]),
    ("otfwi_set_cquad64", _i, [_vp, _vp]),
    ("otfwi_set_illum", _i, [_vp, _i]),
    ("otfwi_get_illum", _i, [_vp, _vp, _i]),
    ("otfwi_misfit", _i, [_i, _i, _i, _vp, _vp, _vp, _vp]),
    ("otfwi_misfit_exact", _i, [_i, _i, _i, _vp, _vp, _vp, _vp]),
    ("otfwi_misfit_noalloc", _i, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
]

EXPORTED = [s[0] for s in _SIGS]

_LIB = None


class OtfwiError(RuntimeError):
    """Error code + message from libotfwi_pml."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _runtime_first():
    """Import PyTorch (when installed) before the first library load.  Its
    wheel ships its own HIP and HSA runtimes; loaded first, they are the ones
    this library binds to as well, so one process has one runtime.  Loaded
    after this library has initialised /opt/rocm's runtime, PyTorch finds no
    GPU."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load():
    """Load the library and declare every signature.  Raises ImportError when
    it has not been built."""
    global _LIB
    if _LIB is None:
        _runtime_first()
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(make -C otfwi_amd/csrc); there is no CPU fallback")
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _LIB = lib
    return _LIB


def check(rc):
    if rc != OK:
        msg = load().otfwi_last_error()
        raise OtfwiError(rc, msg.decode() if msg else "")
    return rc


def ptr(a):
    """Data pointer of a C-contiguous numpy array (None stays None)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(ctypes.c_void_p)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


UNET_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libotfwi_unet.so")
_ll = ctypes.c_longlong
_f = ctypes.c_float
UNET_SIGS = [
    ("otfwi_unet_last_error", ctypes.c_char_p, []),
    ("otfwi_unet_conv3x3", _i, [_vp, _i, _vp, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    ("otfwi_unet_set_conv_tiles", _i, [_i]),
    ("otfwi_unet_wgrad", _i, [_vp, _i, _vp, _i, _i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    ("otfwi_unet_wgrad_work", _i, [_vp, _i, _vp, _i, _i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp,
                                   ctypes.c_size_t, _vp]),
    ("otfwi_unet_wgrad_work_elems", ctypes.c_longlong, [_i, _i, _i, _i, _i, _i]),
    ("otfwi_unet_flip_weights", _i, [_vp, _i, _i, _vp, _vp]),
    ("otfwi_unet_pool2", _i, [_vp, _i, _i, _i, _i, _vp, _vp]),
    ("otfwi_unet_upsum_mask", _i, [_vp, _i, _i, _vp, _i, _i, _i, _i, _vp, _vp]),
    ("otfwi_unet_skip_combine", _i, [_vp, _i, _i, _vp, _vp, _i, _i, _i, _i, _vp, _vp]),
    ("otfwi_unet_pack_input", _i, [_vp, _i, _i, _i, _vp, _vp]),
    ("otfwi_unet_head_fwd", _i, [_vp, _vp, _vp, _ll, _vp, _vp]),
    ("otfwi_unet_head_bwd", _i, [_vp, _vp, _vp, _vp, _ll, _f, _vp, _vp, _vp, _vp, _vp]),
    ("otfwi_unet_head_bwd_work", _i, [_vp, _vp, _vp, _vp, _ll, _f, _vp, _vp, _vp, _vp, _vp,
                                      ctypes.c_size_t, _vp]),
    ("otfwi_unet_bias_grad", _i, [_vp, _ll, _i, _vp, _vp]),
    ("otfwi_unet_adam", _i, [_vp, _vp, _vp, _vp, _vp, _ll, _f, _f, _f, _f, _i, _vp]),
    ("otfwi_unet_to_bf16", _i, [_vp, _ll, _vp, _vp]),
]
UNET_EXPORTED = [s[0] for s in UNET_SIGS]
_ULIB = None


def load_unet():
    """Load libotfwi_unet.so (the learned preconditioner's kernels).  Raises
    ImportError when it has not been built; there is no fallback."""
    global _ULIB
    if _ULIB is None:
        _runtime_first()
        if not os.path.exists(UNET_LIB_PATH):
            raise ImportError(
                f"{UNET_LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(make -C otfwi_amd/csrc); there is no CPU fallback")
        lib = ctypes.CDLL(UNET_LIB_PATH)
        for name, res, args in UNET_SIGS:
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _ULIB = lib
    return _ULIB


def ucheck(rc):
    if rc != 0:
        msg = load_unet().otfwi_unet_last_error()
        raise OtfwiError(rc, msg.decode() if msg else "")
    return rc


def device_info():
    buf = ctypes.create_string_buffer(256)
    check(load().otfwi_device_info(buf, 256))
    return buf.value.decode()
