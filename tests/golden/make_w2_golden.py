"""Writes tests/golden/w2_cases.npz: inputs and outputs of the reference's W2
misfit computed through the library calls adjoint_functions.wasserstein makes
(numpy + scipy.interpolate.interp1d + scipy.integrate cumulative trapezoid,
adjoint_functions.py:25-67; SciPy >= 1.14 names cumtrapz
cumulative_trapezoid).  The reference module itself is not imported.

Run from the repository root:  python tests/golden/make_w2_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests.test_oracle_w2 import direct_scipy_w2, traces  # noqa: E402


def main():
    out = {}
    cases = [(1, 64, False, 11), (4, 300, False, 12), (2, 400, True, 13)]
    for k, (n, T, neg, seed) in enumerate(cases):
        sim, obs = traces(n, T, seed, neg=neg)
        sim = sim.astype(np.float32).astype(np.float64)   # FP32 traces, as the propagator gives
        obs = obs.astype(np.float32).astype(np.float64)
        w, a = direct_scipy_w2(sim, obs)
        out[f"sim{k}"], out[f"obs{k}"], out[f"w2_{k}"], out[f"adj{k}"] = sim, obs, np.float64(w), a
    np.savez_compressed(os.path.join(HERE, "w2_cases.npz"), **out)
    print("wrote", len(cases), "cases")


if __name__ == "__main__":
    main()
