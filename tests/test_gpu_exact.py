"""Exact discrete adjoint on the GPU (checkpointed, FP64 and FP32 builds):
against the NumPy restatement of the transposed recursion, against central
finite differences of the L2 and W2 misfit values in FP64, invariance under
the checkpoint interval, the autograd wrapper, and its relation to the
reference's imaging-condition gradient."""
import numpy as np
import pytest

from oracle import oracle as O
from otfwi_amd import model as M
from tests import np_scheme as NP
from tests.problems import make_model, rel, small_problem

pytestmark = pytest.mark.gpu


def observed(p, shots):
    return np.stack([O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                               p["rx_xz"], p["T"], idx=i)["rec"] for i in shots])


def _traces(n, T, seed):
    rng = np.random.default_rng(seed)
    tt = np.arange(T)
    sim, obs = np.zeros((n, T)), np.zeros((n, T))
    for i in range(n):
        t0, f = rng.uniform(0.3 * T, 0.6 * T), rng.uniform(0.02, 0.05)
        a = (np.pi * f * (tt - t0)) ** 2
        obs[i] = (1 - 2 * a) * np.exp(-a)
        a = (np.pi * f * (tt - t0 - rng.uniform(-8, 8))) ** 2
        sim[i] = rng.uniform(0.9, 1.1) * (1 - 2 * a) * np.exp(-a)
    return sim, obs


@pytest.mark.parametrize("T", [300, 1500])
def test_misfit_exact_vs_oracle(gpu, T):
    sim, obs = _traces(4, T, T)
    mis, rho = M.misfit_exact("w2", sim, obs)
    w_ref, rho_ref = O.wasserstein_exact(sim, obs)
    print(f"T={T}: W2 rel {abs(mis.sum() - w_ref) / w_ref:.2e}, rho rel {rel(rho, rho_ref):.2e}")
    assert abs(mis.sum() - w_ref) < 1e-12 * w_ref
    assert rel(rho, rho_ref) < 1e-9
    mis2, rho2 = M.misfit_exact("l2", sim, obs)
    assert np.allclose(rho2, 2 * (sim - obs), rtol=0, atol=1e-15)


def _tiny():
    p = small_problem(X=40, Z=36, T=90, pml=6, n_src=1, n_rx=5, seed=2, f0=0.07)
    cq = 0.5 * (p["cq_bg"].astype(np.float64) + p["cq_true"].astype(np.float64))
    return p, cq


@pytest.mark.parametrize("kind", ["l2", "w2"])
def test_exact_gradient_vs_numpy(gpu, kind):
    p, cq = _tiny()
    obs = NP.forward(p["cq_true"].astype(np.float64), p["d_x"], p["d_z"], p["src_xz"], p["src"],
                     p["rx_xz"], p["T"], idx=0)[0]
    fn = None if kind == "l2" else (lambda rec: O.wasserstein_exact(rec, obs))
    J_ref, g_ref, _ = NP.exact_gradient(cq, p["d_x"], p["d_z"], p["src_xz"], p["src"],
                                        p["rx_xz"], obs, p["T"], idx=0, rho_fn=fn)
    m = make_model(p, precision=8)
    m.set_cquad64(cq)
    m.set_observed(obs[None].astype(np.float32))
    # the observed traces go through FP32 on upload: compare on those values
    obs32 = obs.astype(np.float32).astype(np.float64)
    fn32 = None if kind == "l2" else (lambda rec: O.wasserstein_exact(rec, obs32))
    J_ref, g_ref, _ = NP.exact_gradient(cq, p["d_x"], p["d_z"], p["src_xz"], p["src"],
                                        p["rx_xz"], obs32, p["T"], idx=0, rho_fn=fn32)
    g, mis = m.gradient_exact([0], kind, snap_every=9, fp64=True)
    e, ej = rel(g, g_ref), abs(mis[0] - J_ref) / J_ref
    print(f"{kind}: grad rel {e:.2e}, J rel {ej:.2e}")
    assert e < 1e-9 and ej < 1e-11


@pytest.mark.parametrize("kind", ["l2", "w2"])
def test_exact_gradient_central_fd_fp64(gpu, kind):
    p = small_problem(seed=31)
    shots = [0, 1]
    obs = observed(p, shots)
    m = make_model(p, precision=8)
    m.set_observed(obs)
    cq = 0.5 * (p["cq_bg"].astype(np.float64) + p["cq_true"].astype(np.float64))
    m.set_cquad64(cq)
    g, mis = m.gradient_exact(shots, kind, fp64=True)
    J = mis.mean()
    zz, xx = np.mgrid[0:p["Z"], 0:p["X"]]
    rng = np.random.default_rng(4)
    worst = 0.0
    for _ in range(3):
        cz, cx = rng.uniform(0.35 * p["Z"], 0.8 * p["Z"]), rng.uniform(0.25 * p["X"], 0.75 * p["X"])
        d = np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * 4.0 ** 2))
        h = 1e-5
        m.set_cquad64(cq + h * d)
        jp = m.objective(shots, kind).mean()
        m.set_cquad64(cq - h * d)
        jm = m.objective(shots, kind).mean()
        fd = (jp - jm) / (2 * h)
        an = float((g * d).sum())
        e = abs(fd - an) / abs(fd)
        print(f"{kind}: J {J:.6e} fd {fd:.10e} adjoint {an:.10e} rel {e:.2e}")
        worst = max(worst, e)
    assert worst < 1e-6


@pytest.mark.parametrize("precision", [4, 8])
def test_checkpoint_interval_invariance(gpu, precision):
    p = small_problem(seed=32)
    shots = [1, 0]
    m = make_model(p, precision=precision, cq="cq_bg")
    m.set_observed(observed(p, shots))
    gs = [m.gradient_exact(shots, "l2", snap_every=k, fp64=True)[0]
          for k in (0, 1, 7, p["T"] - 1)]
    for g in gs[1:]:
        assert np.array_equal(g, gs[0])


def test_fp32_exact_close_to_fp64(gpu):
    p = small_problem(seed=33)
    shots = [0, 1]
    obs = observed(p, shots)
    g = {}
    for prec in (4, 8):
        m = make_model(p, precision=prec, cq="cq_bg")
        m.set_observed(obs)
        g[prec] = m.gradient_exact(shots, "w2", fp64=True)[0]
    e = rel(g[4], g[8])
    print(f"FP32 exact gradient vs FP64: {e:.2e}")
    assert e < 1e-3


def test_autograd_wrapper(gpu):
    import torch

    from otfwi_amd.autograd import propagator_misfit

    p = small_problem(seed=34)
    shots = [0, 1]
    m = make_model(p, precision=8)
    m.set_observed(observed(p, shots))
    cq = torch.tensor(p["cq_bg"].astype(np.float64), requires_grad=True)
    J = propagator_misfit(cq, m, shots, "l2", "exact")
    (3.0 * J).backward()
    m.set_cquad64(p["cq_bg"].astype(np.float64))
    g, mis = m.gradient_exact(shots, "l2", fp64=True)
    assert abs(J.item() - mis.mean()) <= 1e-14 * mis.mean()
    assert np.array_equal(cq.grad.numpy(), 3.0 * g)
    # directional derivative through autograd against finite differences
    zz, xx = np.mgrid[0:p["Z"], 0:p["X"]]
    d = torch.tensor(np.exp(-((zz - 55) ** 2 + (xx - 48) ** 2) / 32.0))
    h = 1e-5
    with torch.no_grad():
        jp = propagator_misfit(cq + h * d, m, shots, "l2", "exact")
        jm = propagator_misfit(cq - h * d, m, shots, "l2", "exact")
    fd = (jp - jm).item() / (2 * h)
    an = float((cq.grad / 3.0 * d).sum())
    assert abs(fd - an) < 1e-6 * abs(fd)


@pytest.mark.parametrize("mode", ["exact", "reference"])
def test_autograd_device_tensor(gpu, mode):
    """A float32 CUDA tensor on an FP32 model: c^2 is read from the tensor's
    memory and the gradient written into a device tensor (the model array
    formed on the device), bitwise equal to the same values through host
    arrays; a negative c^2 on the device is refused."""
    import torch

    from otfwi_amd._lib import OtfwiError
    from otfwi_amd.autograd import propagator_misfit

    p = small_problem(seed=35)
    shots = [0, 1]
    m = make_model(p)
    m.set_observed(observed(p, shots))
    cq_h = torch.tensor(p["cq_bg"], dtype=torch.float32, requires_grad=True)
    J_h = propagator_misfit(cq_h, m, shots, "l2", mode)
    J_h.backward()
    cq_d = torch.tensor(p["cq_bg"], dtype=torch.float32, device="cuda", requires_grad=True)
    J_d = propagator_misfit(cq_d * 1.0, m, shots, "l2", mode)
    (2.0 * J_d).backward()
    assert J_d.is_cuda and cq_d.grad.is_cuda
    assert J_d.item() == J_h.item()
    assert np.array_equal(cq_d.grad.cpu().numpy(), 2.0 * cq_h.grad.numpy())
    bad = torch.tensor(p["cq_bg"], dtype=torch.float32, device="cuda")
    bad[5, 7] = -1.0
    with pytest.raises(OtfwiError):
        m.set_cquad_device(bad.data_ptr())


def test_exact_vs_reference_gradient(gpu):
    p = small_problem(seed=21)
    shots = [0, 1]
    m = make_model(p, cq="cq_bg")
    m.set_observed(observed(p, shots))
    g_ref, _ = m.gradient(shots, "l2")
    g_ex, _ = m.gradient_exact(shots, "l2", fp64=True)
    pml, Z, X = p["pml"], p["Z"], p["X"]
    inner = np.zeros((Z, X), bool)
    inner[pml + 6:Z - pml - 6, pml + 6:X - pml - 6] = True
    a, b = g_ref[inner].astype(np.float64), g_ex[inner]
    cos = a @ b / np.linalg.norm(a) / np.linalg.norm(b)
    scale = (a @ b) / (a @ a)
    print(f"interior cosine {cos:.4f}, exact / reference scale {scale:.3f}")
    # the reference L2 adjoint source is (sim - obs), half of dJ/dsim
    assert cos > 0.9 and 1.5 < scale < 2.5
