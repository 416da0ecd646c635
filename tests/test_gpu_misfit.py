"""Device W2 / L2 against the numpy oracle on identical traces
(adjoint_functions.py:8-11, 25-67)."""
import numpy as np
import pytest

from oracle import oracle as O
from otfwi_amd import adjoint_functions as AF

pytestmark = pytest.mark.gpu


def traces(n, T, seed, neg=False):
    rng = np.random.default_rng(seed)
    t = np.arange(T)
    obs = np.zeros((n, T))
    sim = np.zeros((n, T))
    for i in range(n):
        for arr in (obs, sim):
            k0 = rng.uniform(0.1, 0.7) * T
            w = rng.uniform(0.01, 0.05) * T
            arr[i] = np.exp(-((t - k0) / w) ** 2) * np.cos((t - k0) / (0.3 * w)) * rng.uniform(0.5, 2)
    if neg:   # make sim + shift negative somewhere: non-monotone CDF
        sim[:, T // 3] = -5.0 * np.abs(obs).max(axis=1)
    return sim.astype(np.float32).astype(np.float64), obs.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("n,T", [(1, 64), (4, 1500), (3, 4000)])
def test_w2_matches_oracle(gpu, n, T):
    sim, obs = traces(n, T, seed=T + n)
    w, a = AF.wasserstein(sim, obs)
    w0, a0 = O.wasserstein(sim, obs)
    print(f"n={n} T={T}: W2 rel {abs(w - w0) / abs(w0):.2e}, adj rel "
          f"{np.linalg.norm(a - a0) / np.linalg.norm(a0):.2e}")
    assert abs(w - w0) <= 1e-10 * abs(w0)
    assert np.linalg.norm(a - a0) <= 1e-9 * np.linalg.norm(a0)


def test_w2_non_monotone_cdf(gpu):
    sim, obs = traces(2, 700, seed=9, neg=True)
    w, a = AF.wasserstein(sim, obs)
    w0, a0 = O.wasserstein(sim, obs)
    assert abs(w - w0) <= 1e-9 * abs(w0)
    assert np.linalg.norm(a - a0) <= 1e-8 * np.linalg.norm(a0)


def test_w2_identical_traces(gpu):
    sim, obs = traces(3, 500, seed=3)
    w, a = AF.wasserstein(obs, obs)
    assert abs(w) < 1e-20 and np.abs(a).max() < 1e-12


def test_l2(gpu):
    sim, obs = traces(5, 333, seed=4)
    w, a = AF.L2(sim, obs)
    w0, a0 = O.l2(sim, obs)
    assert abs(w - w0) <= 1e-13 * abs(w0)
    assert np.array_equal(a, a0)


def test_w2_all_zero_observed_is_nan_like_reference(gpu):
    # adjoint_functions.py:38-43: obs == 0 gives c = 0 and g /= 0 -> NaN
    sim, obs = traces(2, 300, seed=5)
    obs[1] = 0.0
    w, a = AF.wasserstein(sim, obs)
    assert np.isnan(w) and np.isnan(a[1]).all() and np.isfinite(a[0]).all()


@pytest.mark.parametrize("n,T", [(1, 64), (4, 1499), (3, 4000)])
def test_envelope_matches_oracle(gpu, n, T):
    """Squared-envelope misfit (adjoint_functions.py:14-22) on the device
    against the oracle in FP64 (obs given as float64: 1e-10), and against the
    reference's own precision mix (obs float32, so its analytic signal is
    complex64: 1e-5)."""
    sim, obs = traces(n, T, seed=T + 7 * n)
    m, a = AF.quad_envelope(sim, obs)
    m0, a0 = O.quad_envelope(sim, obs)
    m1, a1 = O.quad_envelope(sim, obs.astype(np.float32))
    em, ea = abs(m - m0) / abs(m0), np.linalg.norm(a - a0) / np.linalg.norm(a0)
    em1, ea1 = abs(m - m1) / abs(m1), np.linalg.norm(a - a1) / np.linalg.norm(a1)
    print(f"n={n} T={T}: vs FP64 misfit {em:.2e} adj {ea:.2e}; vs reference mix "
          f"misfit {em1:.2e} adj {ea1:.2e}")
    assert em <= 1e-10 and ea <= 1e-10
    assert em1 <= 1e-5 and ea1 <= 1e-5


def test_envelope_identical_traces(gpu):
    sim, _ = traces(2, 500, seed=5)
    m, a = AF.quad_envelope(sim, sim.copy())
    assert m == 0.0 and not np.any(a)


VARIANTS = {"quad": (AF.wasserstein_quad, O.wasserstein_quad),
            "experimental": (AF.wasserstein_experimental, O.wasserstein_experimental),
            "pm": (AF.wasserstein_pm, O.wasserstein_pm),
            "yang": (AF.wasserstein_yang, O.wasserstein_yang)}


@pytest.mark.parametrize("form", sorted(VARIANTS))
@pytest.mark.parametrize("n,T,neg", [(1, 64, False), (4, 1500, False), (2, 700, True),
                                     (2, 4000, False)])
def test_w2_variants_match_oracle(gpu, form, n, T, neg):
    """wasserstein_quad / _experimental / _pm / _yang (adjoint_functions.py:
    70-266) on the device against the oracle on identical FP64 traces (1e-9;
    neg: a sample of sim far below zero, so the experimental form's F and the
    others' shifted densities leave their usual range).  With the observed
    traces as the reference stores them (float32), the reference squares,
    clips and shifts them in float32 while the device works in FP64: 1e-5
    (yang's adjoint 1e-4: it divides by the float32 observed density at
    G^-1(F(t)))."""
    dev, orc = VARIANTS[form]
    sim, obs = traces(n, T, seed=T + n + len(form), neg=neg)
    if form == "yang":
        obs = obs + 0.01 * np.abs(obs).max()    # positive observed traces: c < 0, G may decrease
    w, a = dev(sim, obs)
    w0, a0 = orc(sim, obs)
    w1, a1 = orc(sim, obs.astype(np.float32))
    ew, ea = abs(w - w0) / abs(w0), np.linalg.norm(a - a0) / np.linalg.norm(a0)
    ew1, ea1 = abs(w - w1) / abs(w1), np.linalg.norm(a - a1) / np.linalg.norm(a1)
    print(f"{form} n={n} T={T} neg={neg}: vs FP64 misfit {ew:.2e} adj {ea:.2e}; vs float32 "
          f"observed misfit {ew1:.2e} adj {ea1:.2e}")
    assert ew <= 1e-9 and ea <= 1e-9
    assert ew1 <= 1e-5 and ea1 <= (1e-4 if form == "yang" else 1e-5)


@pytest.mark.parametrize("form", ["quad", "experimental", "pm", "yang"])
def test_w2_variant_gradient_on_device_equals_host_callback(gpu, form):
    """Cuda_interface_PML keeps the variant on the device; the oracle's
    restatement as a host callback (FP64 observed traces) gives the same
    gradient and misfit."""
    from otfwi_amd.cuda_interface_pml import Cuda_interface_PML
    from tests.problems import small_problem

    dev, orc = VARIANTS[form]
    p = small_problem(seed=17)
    obs = np.stack([O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                              p["rx_xz"], p["T"], idx=s)["rec"] for s in (0, 1)])
    c = np.sqrt(p["cq_bg"].astype(np.float64))
    runs = []
    for fn in (dev, lambda s, o: orc(s, np.asarray(o, np.float64))):
        cp = Cuda_interface_PML(p["X"], p["Z"], p["T"], p["pml"], c, 2, p["src_mask"],
                                len(p["rx_xz"][0]), p["rx_mask"], p["src"], multishot=True, R=1e-3)
        cp.init_regression(obs, fn, prec_deriv=4)
        cp.calc_grad()
        runs.append((cp.grad.astype(np.float64), float(cp.mse[0])))
    (gd, md), (gh, mh) = runs
    eg = np.linalg.norm(gd - gh) / np.linalg.norm(gh)
    print(f"{form}: device vs host callback gradient {eg:.2e}, misfit {abs(md - mh) / abs(mh):.2e}")
    assert eg <= 1e-5 and abs(md - mh) <= 1e-9 * abs(mh)
