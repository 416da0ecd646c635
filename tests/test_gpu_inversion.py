"""Plain inversion step on the GPU: illumination map against the oracle's
forward fields, the objective-only entry point against the gradient's
misfit, the preconditioned direction against the same direction built from
oracle gradient + oracle illumination, and the misfit-decrease acceptance
over a few iterations."""
import numpy as np
import pytest

from oracle import oracle as O
from otfwi_amd import inversion
from tests.problems import make_model, rel, small_problem

pytestmark = pytest.mark.gpu


def observed(p, shots):
    return np.stack([O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                               p["rx_xz"], p["T"], idx=i)["rec"] for i in shots])


@pytest.mark.parametrize("stride", [1, 4])
def test_illumination_vs_oracle(gpu, stride):
    p = small_problem(seed=21)
    shots = [0, 1]
    m = make_model(p, cq="cq_bg")
    m.set_observed(observed(p, shots))
    m.set_illum(stride)
    m.gradient(shots, misfit="l2")
    ill = m.illum()
    ref = O.illumination(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["T"], shots,
                         stride)
    e = rel(ill, ref)
    print(f"illumination stride {stride}: rel {e:.3e}")
    assert e < 1e-5
    # a later gradient call without collection leaves no stale map behind
    m.set_illum(0)
    m.gradient(shots, misfit="l2")
    with pytest.raises(Exception):
        m.illum()


@pytest.mark.parametrize("T,stride", [(400, 4), (401, 4), (402, 2), (401, 3)])
def test_illumination_with_two_step_launches(gpu, T, stride):
    """Illumination collected inside two-step forward loops (pairs start at an
    odd step, sampled steps read from the pair's output, one stream) equals
    the one-step-per-launch illumination bit for bit, and so do the gradient
    and misfits; an odd stride runs the forward one step per launch."""
    p = small_problem(X=420, Z=300, T=T, seed=27, n_src=5, n_rx=20)
    shots = [0, 1, 2, 3, 4]
    obs = observed(p, shots)
    out = []
    for steps in (1, 2):
        m = make_model(p, cq="cq_bg")
        m.set_observed(obs)
        m.set_steps_per_launch(steps, steps)
        m.set_pair_streams(2)
        m.set_illum(stride)
        g, mis = m.gradient(shots, "l2")
        out.append((g, mis, m.illum()))
    for u, v in zip(*out):
        assert np.array_equal(u, v)
    assert np.abs(out[0][2]).max() > 0


@pytest.mark.parametrize("kind", ["l2", "w2"])
def test_objective_equals_gradient_misfit(gpu, kind):
    p = small_problem(X=330, Z=160, T=760, seed=22)
    shots = [1, 0]
    m = make_model(p, cq="cq_bg", max_batch=1)
    m.set_observed(observed(p, shots))
    _, mis_g = m.gradient(shots, misfit=kind)
    mis_o = m.objective(shots, misfit=kind)
    print(kind, mis_g, mis_o)
    assert np.array_equal(mis_g, mis_o)


def test_plain_direction_vs_oracle(gpu):
    p = small_problem(seed=23)
    shots = [0, 1]
    obs = observed(p, shots)
    m = make_model(p, cq="cq_bg")
    m.set_observed(obs)
    m.set_illum(2)
    grad, _ = m.gradient(shots, misfit="l2")
    ill = m.illum()
    g_ref, _, _ = O.gradient(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"],
                             p["ring_xz"], obs, p["T"], O.l2, shots=shots)
    ill_ref = O.illumination(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["T"],
                             shots, 2)
    c = np.sqrt(p["cq_bg"].astype(np.float64))
    mask = inversion.array_mask(p["Z"], p["X"], p["pml"], p["pml"] + 6, taper=4)
    d = inversion.plain_direction(grad, ill, c, mask)
    d_ref = inversion.plain_direction(g_ref, ill_ref, c, mask)
    e = rel(d, d_ref)
    print(f"plain direction rel {e:.3e}")
    assert e < 1e-4


@pytest.mark.parametrize("T,stride", [(400, 4), (401, 4), (402, 2), (401, 3)])
def test_illumination_with_two_step_launches(gpu, T, stride):
    """Illumination collected inside two-step forward loops (pairs start at an
    odd step, sampled steps read from the pair's output, one stream) equals
    the one-step-per-launch illumination bit for bit, and so do the gradient
    and misfits; an odd stride runs the forward one step per launch."""
    p = small_problem(X=420, Z=300, T=T, seed=27, n_src=5, n_rx=20)
    shots = [0, 1, 2, 3, 4]
    obs = observed(p, shots)
    out = []
    for steps in (1, 2):
        m = make_model(p, cq="cq_bg")
        m.set_observed(obs)
        m.set_steps_per_launch(steps, steps)
        m.set_pair_streams(2)
        m.set_illum(stride)
        g, mis = m.gradient(shots, "l2")
        out.append((g, mis, m.illum()))
    for u, v in zip(*out):
        assert np.array_equal(u, v)
    assert np.abs(out[0][2]).max() > 0


@pytest.mark.parametrize("kind", ["l2", "w2"])
def test_objective_equals_gradient_misfit(gpu, kind):
    """The forward-only misfit the inversion judges trial steps by equals the
    gradient call's misfit bit for bit: with two steps per launch asked for,
    the objective's forward runs two-step launches and the gradient's forward,
    collecting illumination, one step per launch."""
    p = small_problem(X=160, Z=128, T=400, seed=26, n_src=3, n_rx=12)
    shots = [0, 1, 2]
    m = make_model(p, cq="cq_bg")
    m.set_observed(observed(p, shots))
    m.set_steps_per_launch(2, 2)
    m.set_illum(4)
    _, mis_g = m.gradient(shots, kind)
    mis_o = m.objective(shots, kind)
    print(kind, mis_g, mis_o)
    assert np.all(mis_g > 0) and np.array_equal(mis_g, mis_o)


def test_plain_inversion_accepts_only_decreases(gpu):
    from otfwi_amd import Cuda_interface_PML, adjoint_functions

    p = small_problem(X=160, Z=128, T=400, seed=24, n_src=3, n_rx=12)
    shots = list(range(3))
    obs = observed(p, shots)
    c_true = np.sqrt(p["cq_true"].astype(np.float64))
    c0 = np.sqrt(p["cq_bg"].astype(np.float64))
    cp = Cuda_interface_PML(p["X"], p["Z"], p["T"], p["pml"], c0, 3, p["src_mask"],
                            int(p["rx_mask"].sum()), p["rx_mask"], p["src"], multishot=True,
                            R=1e-3)
    cp.init_regression(obs, adjoint_functions.L2, prec_deriv=4)
    mask = inversion.array_mask(p["Z"], p["X"], p["pml"], p["pml"] + 6, taper=4)
    fwi = inversion.FWI(cp, mask, vmin=0.25, vmax=0.5, misfit="l2", c_true=c_true,
                        alpha0=0.02)
    _, hist = fwi.run(c0, n_iter=4)
    J = [r.misfit for r in hist.records]
    E = [r.err for r in hist.records]
    print("misfit", J, "err", E)
    assert len(J) >= 3
    assert all(b < a for a, b in zip(J, J[1:]))
    assert E[-1] < E[0]


def test_lbfgs_regression_flow(gpu):
    """regression.py:45-87 on the drop-in: L-BFGS-B on 1e3*mse with
    grad_scale*gmask*grad and velocity bounds lowers the misfit."""
    from otfwi_amd import Cuda_interface_PML, adjoint_functions

    p = small_problem(X=128, Z=112, T=360, seed=25, n_src=2, n_rx=10)
    c_true = np.sqrt(p["cq_true"].astype(np.float64))
    c0 = np.sqrt(p["cq_bg"].astype(np.float64))
    cp = Cuda_interface_PML(p["X"], p["Z"], p["T"], p["pml"], c_true, 2, p["src_mask"],
                            int(p["rx_mask"].sum()), p["rx_mask"], p["src"], multishot=True,
                            R=1e-3)
    cp.simulate()
    cp.init_regression(cp.recording.copy(), adjoint_functions.L2, prec_deriv=4)
    gmask = inversion.array_mask(p["Z"], p["X"], p["pml"], p["pml"] + 6, taper=4)
    seen = []
    cp.calc_grad(c=c0.flatten())
    j0 = 1e3 * float(cp.mse[0])
    # d(1e3 J)/dc = 1e3 * 2c * dJ/d(c^2) and the reference-mode L2 gradient is
    # about half of dJ/d(c^2): a scalar grad_scale of 4e3 * c_background matches
    r = inversion.lbfgs(cp, c0, 0.25, 0.5, gmask, grad_scale=4e3 * float(c0.mean()), maxiter=5,
                        callback=lambda x: seen.append(1))
    print("L-BFGS-B:", j0, "->", r.fun, r.message, len(seen))
    assert r.fun < j0
    assert np.all(r.x >= 0.25 - 1e-12) and np.all(r.x <= 0.5 + 1e-12)


@pytest.mark.parametrize("kind", ["l2", "w2"])
def test_regression_objective_matches_oracle(gpu, kind):
    """The value and gradient regression.py:45-49 hands to L-BFGS-B
    (1e3*mse, grad_scale*gmask*grad over a multishot drop-in) against the
    oracle's mean misfit and mean gradient over the shots, with the same
    mask and scale.  Bars: L2 value 1e-6, gradient 1e-4; W2 value 1e-6,
    gradient 2e-3 (the W2 bar of test_gpu_gradient.py)."""
    from otfwi_amd import Cuda_interface_PML, adjoint_functions

    p = small_problem(X=128, Z=112, T=360, seed=26, n_src=2, n_rx=10)
    c_true = np.sqrt(p["cq_true"].astype(np.float64))
    c0 = np.sqrt(p["cq_bg"].astype(np.float64))
    cp = Cuda_interface_PML(p["X"], p["Z"], p["T"], p["pml"], c_true, 2, p["src_mask"],
                            int(p["rx_mask"].sum()), p["rx_mask"], p["src"], multishot=True,
                            R=1e-3)
    cp.simulate()
    obs = cp.recording.copy()
    fn = adjoint_functions.L2 if kind == "l2" else adjoint_functions.wasserstein
    cp.init_regression(obs, fn, prec_deriv=4)
    gmask = inversion.array_mask(p["Z"], p["X"], p["pml"], p["pml"] + 6, taper=4)
    scale = 4e3 * float(c0.mean())
    val, grad = inversion.regression_objective(cp, gmask, scale)(c0.flatten())
    # the oracle at the same c^2 (FP32, as _set_c_quad stores it) and observed traces
    cq = (c0 ** 2).astype(np.float32)
    ofn = O.l2 if kind == "l2" else O.wasserstein
    g_ref, mse_ref, _ = O.gradient(cq, p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"],
                                   p["ring_xz"], obs, p["T"], ofn, shots=(0, 1))
    want_val = 1e3 * mse_ref
    want_grad = (scale * (gmask * g_ref)).flatten()
    ev = abs(val - want_val) / abs(want_val)
    eg = rel(grad, want_grad)
    print(f"regression objective {kind}: value {ev:.3e}, gradient {eg:.3e}")
    assert ev <= 1e-6 and eg <= (1e-4 if kind == "l2" else 2e-3)
    assert inversion.model_misfit(c_true.flatten(), c_true) == 0.0
    want_mm = np.sum((c0 - c_true) ** 2) / c0.size
    assert abs(inversion.model_misfit(c0.flatten(), c_true) - want_mm) <= 1e-12 * want_mm
