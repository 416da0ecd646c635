"""MI355X-native 2D acoustic FWI hot path (split-field PML propagator,
boundary-saving adjoint gradient, L2 / Wasserstein-2 misfits) behind the call
surface of the reference's ``cuda_interface_pml`` and ``adjoint_functions``.

The compute lives in ``libotfwi_pml.so`` (hand-written HIP for gfx950, built
in-tree by ``__graft_entry__.build()``); importing the package loads it and
fails loudly when it is missing.
"""
from . import _lib

_lib.load()

from .model import PMLModel, misfit  # noqa: E402
from .cuda_interface_pml import Cuda_interface_PML  # noqa: E402
from . import adjoint_functions, configs  # noqa: E402

__all__ = ["PMLModel", "misfit", "Cuda_interface_PML", "adjoint_functions", "configs"]
