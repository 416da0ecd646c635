"""W2 golden fixtures (tests/golden/w2_cases.npz, made by
tests/golden/make_w2_golden.py through the scipy calls the reference makes)
against the oracle (CPU) and the device kernel (GPU)."""
import os

import numpy as np
import pytest

from oracle import oracle as O

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "w2_cases.npz"))
CASES = sorted({int(k[3:]) for k in G.files if k.startswith("sim")})


@pytest.mark.parametrize("k", CASES)
def test_oracle_matches_golden(k):
    w, a = O.wasserstein(G[f"sim{k}"], G[f"obs{k}"])
    assert w == G[f"w2_{k}"]
    assert np.array_equal(a, G[f"adj{k}"])


@pytest.mark.gpu
@pytest.mark.parametrize("k", CASES)
def test_device_w2_matches_golden(gpu, k):
    from otfwi_amd import adjoint_functions as AF

    w, a = AF.wasserstein(G[f"sim{k}"], G[f"obs{k}"])
    w0, a0 = float(G[f"w2_{k}"]), G[f"adj{k}"]
    assert abs(w - w0) <= 1e-10 * abs(w0)
    assert np.abs(a - a0).max() <= 1e-9 * np.abs(a0).max()
