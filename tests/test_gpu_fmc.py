"""The 64-shot 1024^2 full-matrix-capture launch shape against the oracle.

All 64 shots are resident, so the library runs the two-step kernels split
over two streams (32 shots each), each half's two-step launches replayed
from its own captured hipGraph: the configuration the 64-shot W2 gradient
benchmark measures.  T is truncated so the oracle
finishes in seconds; the true model has one shallow blob so that the
residuals are nonzero within T.

Bars (fixed):
* traces: relative L2 <= 1e-5 against the reference-precision oracle;
* L2 misfit <= 1e-6 (relative); L2 gradient <= 1e-4 against the FP64 oracle
  over the whole grid, and <= 1e-4 against the reference-precision oracle
  below the array row.  On the array row and above, every cell is within a
  few cells of a colocated source and receiver, where the imaging condition
  multiplies two near-singular fields: there the reference's own FP32
  arithmetic is 1.6e-4 from FP64 over the whole grid (2e-5 below the array),
  so two FP32 evaluations with different rounding sequences differ by about
  2e-4; the whole-grid distance to the reference-precision oracle is held to
  3e-4;
* W2 adjoint propagation, given the oracle's own FP32 adjoint source:
  gradient <= 1e-4 below the array row, 3e-4 over the whole grid;
* W2 end to end (device traces, device W2, device adjoint): gradient
  <= 2e-3.  The W2 adjoint source (adjoint_functions.py:25-67) amplifies the
  FP32 rounding of the traces about 300x, so no FP32 forward run gives a W2
  gradient within 1e-4 of FP64 end to end (test_gpu_gradient.py states the
  figures).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.problems import fmc_problem, make_model, rel

pytestmark = pytest.mark.gpu


def _split_shape(m, n):
    modes = m.graph_modes(n)
    print(f"loop modes at {n} shots: {modes}, steps per launch {m.steps_per_launch(n)}",
          flush=True)
    assert m.steps_per_launch(n) == (2, 2)
    assert modes["fwd_streams"] == 2 and modes["adj_streams"] == 2
    assert modes["fwd"] == "graph" and modes["adj"] == "graph"


ARRAY_ROW = 20


def test_fmc64_traces_and_l2_gradient(gpu):
    p = fmc_problem(T=400)
    S = 64
    chosen = [25, 40]                      # one shot in each half of the batch, 90 cells
    #                                        either side of the blob
    shots = list(range(S))
    m = make_model(p, cq="cq_bg")
    _split_shape(m, S)
    sim_bg = m.forward(shots)
    m.set_cquad(p["cq_true"])
    rec_true = m.forward(shots)
    for s in chosen:
        ref = O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"],
                        p["T"], idx=s)["rec"]
        e = rel(rec_true[s], ref)
        print(f"shot {s}: traces vs oracle {e:.3e}", flush=True)
        assert e <= 1e-5
    # the other shots observe exactly what the background model predicts:
    # their residual, misfit and gradient contribution are zero
    obs = sim_bg.copy()
    obs[chosen] = rec_true[chosen]
    m.set_cquad(p["cq_bg"])
    m.set_observed(obs)
    grad, mis = m.gradient(shots, misfit="l2")
    others = [s for s in shots if s not in chosen]
    assert np.all(mis[others] == 0)
    g_ref, mse_ref, _ = O.gradient(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                                   p["rx_xz"], p["ring_xz"], obs[chosen], p["T"], O.l2,
                                   shots=chosen)
    g64, mse64, _ = O.gradient(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                               p["rx_xz"], p["ring_xz"], obs[chosen].astype(np.float64), p["T"],
                               O.l2, shots=chosen, precision="f64")
    g = grad.astype(np.float64) * S / len(chosen)
    below = slice(ARRAY_ROW + 1, None)
    e64, eref, ebelow = rel(g, g64), rel(g, g_ref), rel(g[below], g_ref[below])
    em = abs(mis[chosen].mean() - mse_ref) / abs(mse_ref)
    print(f"fmc64 L2: gradient vs FP64 {e64:.3e}, vs reference precision {eref:.3e} (below the "
          f"array {ebelow:.3e}; oracle FP32 vs FP64 {rel(g_ref, g64):.3e}, below the array "
          f"{rel(g_ref[below], g64[below]):.3e}); misfit {em:.3e}, vs FP64 "
          f"{abs(mis[chosen].mean() - mse64) / abs(mse64):.3e}", flush=True)
    assert e64 <= 1e-4 and ebelow <= 1e-4 and eref <= 3e-4 and em <= 1e-6


def _oracle_w2(p, obs, s, precision):
    dt = np.float32 if precision == "ref" else np.float64
    fw = O.forward(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"], p["T"],
                   idx=s, ring_xz=p["ring_xz"], want_seed=True, precision=precision)
    mse, adj = O.wasserstein(fw["rec"].astype(np.float64), obs)
    adj = adj.astype(dt)
    g = O.adjoint(p["cq_bg"], p["d_x"], p["d_z"], p["rx_xz"], adj, p["ring_xz"], fw["ring"],
                  fw["seed"], precision=precision)
    return g.astype(np.float64), mse, adj


def test_fmc64_w2_gradient(gpu):
    # every shot fires one of two central elements, so every receiver sees
    # the wave within T (W2 needs nonzero observed traces): shots 0-31 fire
    # element 31, shots 32-63 element 33
    p = fmc_problem(T=900)
    S = 64
    shots = [31] * 32 + [33] * 32
    m = make_model(p, cq="cq_true")
    _split_shape(m, S)
    obs = m.forward(shots)
    assert np.abs(obs).max(axis=2).min() > 0
    m.set_cquad(p["cq_bg"])
    m.set_observed(obs)
    grad, mis = m.gradient(shots, misfit="w2")

    ref = {s: _oracle_w2(p, obs[k], s, "ref") for k, s in ((0, 31), (32, 33))}
    g_ref = 0.5 * (ref[31][0] + ref[33][0])
    mse_ref = 0.5 * (ref[31][1] + ref[33][1])
    e_ref = rel(grad, g_ref)
    em = abs(mis.mean() - mse_ref) / abs(mse_ref)
    print(f"fmc64 W2 end to end: gradient vs reference-precision oracle {e_ref:.3e}; misfit "
          f"{em:.3e}", flush=True)
    assert e_ref <= 2e-3 and em <= 1e-6

    # adjoint propagation alone: the oracle's own FP32 adjoint sources
    m.forward_for_grad(shots)
    adj = np.stack([ref[31][2]] * 32 + [ref[33][2]] * 32)
    g_prop = m.adjoint_from_host(adj).astype(np.float64)
    below = slice(ARRAY_ROW + 1, None)
    e_prop, e_below = rel(g_prop, g_ref), rel(g_prop[below], g_ref[below])
    print(f"fmc64 W2 adjoint propagation vs oracle: {e_prop:.3e} (below the array {e_below:.3e})",
          flush=True)
    assert e_below <= 1e-4 and e_prop <= 3e-4
