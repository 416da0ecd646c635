"""Pins the C oracle (oracle/oracle_pml.c) of the propagator and gradient.

* an independent vectorized NumPy restatement (tests/np_scheme.py) on tiny
  grids, FP64: forward traces, boundary ring, reverse seed and gradient;
* invariants: zero source -> zero field; linearity in the source amplitude;
  superposition of two sources; the stability limit of the 8th-order
  staggered pair (c*dt/dx < 1/(sqrt(2) sum |c_k|) = 0.5497); PML attenuation;
* the coefficient literals against the rationals they truncate.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import np_scheme as NP
from tests.problems import small_problem


def tiny(seed=0, X=40, Z=36, T=60, pml=6):
    return small_problem(X=X, Z=Z, T=T, pml=pml, n_src=2, n_rx=4, seed=seed, f0=0.08)


def test_coefficients_truncate_the_exact_rationals():
    exact = (1225 / 1024, -245 / 3072, 49 / 5120, -5 / 7168)
    for c, e in zip(NP.C, exact):
        assert abs(c - e) < 6e-7 * abs(e)
    # first-derivative consistency of the staggered pair: sum_k c_k (2k-1) = 1
    assert abs(sum(c * (2 * k - 1) for k, c in enumerate(NP.C, 1)) - 1) < 2e-6


@pytest.mark.parametrize("idx", [-1, 0, 1])
def test_forward_matches_numpy_restatement(idx):
    p = tiny(seed=1)
    o = O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"], p["T"],
                  idx=idx, ring_xz=p["ring_xz"], precision="f64", want_seed=True)
    rec, ring, seed = NP.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                                 p["rx_xz"], p["T"], idx=idx, ring_xz=p["ring_xz"])
    scale = np.abs(rec).max()
    assert scale > 0
    assert np.abs(o["rec"] - rec).max() <= 1e-12 * scale
    assert np.abs(o["ring"] - ring).max() <= 1e-12 * scale
    assert np.abs(o["seed"] - seed).max() <= 1e-12 * np.abs(seed).max()
    assert np.all(o["rec"][:, 0] == 0)


def test_gradient_matches_numpy_restatement():
    p = tiny(seed=2, T=50)
    fw = O.forward(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"], p["T"],
                   ring_xz=p["ring_xz"], precision="f64", want_seed=True)
    rng = np.random.default_rng(0)
    adj = rng.standard_normal(fw["rec"].shape)
    g = O.adjoint(p["cq_bg"], p["d_x"], p["d_z"], p["rx_xz"], adj, p["ring_xz"], fw["ring"],
                  fw["seed"], precision="f64")
    g_np = NP.adjoint(p["cq_bg"], p["d_x"], p["d_z"], p["rx_xz"], adj, p["ring_xz"], fw["ring"],
                      fw["seed"])
    assert np.abs(g - g_np).max() <= 1e-11 * np.abs(g_np).max()


def test_reference_precision_close_to_fp64():
    p = tiny(seed=3)
    a = O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"], p["T"])
    b = O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"], p["T"],
                  precision="f64")
    assert a["rec"].dtype == np.float32
    e = np.linalg.norm(a["rec"] - b["rec"]) / np.linalg.norm(b["rec"])
    assert 0 < e < 1e-5


def test_zero_source_and_linearity():
    p = tiny(seed=4)
    args = (p["cq_true"], p["d_x"], p["d_z"], p["src_xz"])
    zero = O.forward(*args, np.zeros_like(p["src"]), p["rx_xz"], p["T"], precision="f64")
    assert np.all(zero["rec"] == 0)
    src = p["src"].astype(np.float64)
    r1 = O.forward(*args, src, p["rx_xz"], p["T"], precision="f64")["rec"]
    r3 = O.forward(*args, 3.0 * src, p["rx_xz"], p["T"], precision="f64")["rec"]
    assert np.abs(r3 - 3 * r1).max() <= 1e-12 * np.abs(r1).max()
    # superposition: both sources at once = each alone, summed
    a = O.forward(*args, p["src"], p["rx_xz"], p["T"], idx=0, precision="f64")["rec"]
    b = O.forward(*args, p["src"], p["rx_xz"], p["T"], idx=1, precision="f64")["rec"]
    assert np.abs(r1 - (a + b)).max() <= 1e-12 * np.abs(r1).max()


def test_stability_limit():
    # 1/(sqrt 2 * sum|c_k|) = 0.5497: c = 0.54 stays bounded, c = 0.57 diverges
    p = tiny(seed=5, T=400, pml=6)
    cl = 1 / (np.sqrt(2) * sum(abs(c) for c in NP.C))
    assert abs(cl - 0.5497) < 1e-3
    res = {}
    for c in (0.54, 0.57):
        cq = np.full_like(p["cq_true"], np.float32(c * c))
        r = O.forward(cq, np.zeros_like(p["d_x"]), np.zeros_like(p["d_z"]), p["src_xz"], p["src"],
                      p["rx_xz"], p["T"], precision="f64")["rec"]
        res[c] = np.abs(r).max()
    assert np.isfinite(res[0.54]) and res[0.54] < 1e3
    assert not np.isfinite(res[0.57]) or res[0.57] > 1e6


def test_pml_attenuates():
    # late-time trace energy with the PML is far below the rigid-box case
    p = tiny(seed=6, X=60, Z=60, T=500, pml=12)
    args = (p["cq_bg"],)
    with_pml = O.forward(*args, p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"], p["T"],
                         precision="f64")["rec"]
    no_pml = O.forward(*args, np.zeros_like(p["d_x"]), np.zeros_like(p["d_z"]), p["src_xz"],
                       p["src"], p["rx_xz"], p["T"], precision="f64")["rec"]
    late = slice(350, None)
    assert np.abs(with_pml[:, late]).max() < 0.1 * np.abs(no_pml[:, late]).max()


@pytest.mark.parametrize("stride", [1, 3])
def test_illumination_matches_numpy_fields(stride):
    p = small_problem(X=34, Z=30, T=70, pml=6, n_src=2, n_rx=3, seed=5, f0=0.08)
    Z, X = p["Z"], p["X"]
    zz, xx = np.mgrid[0:Z, 0:X]
    every = (xx.ravel(), zz.ravel())
    ill = np.zeros(Z * X)
    for idx in (0, 1):
        rec, _, _ = NP.forward(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"], every,
                               p["T"], idx=idx)
        for t in range(1, p["T"]):
            if t % stride == 0:
                ill += stride * rec[:, t] ** 2
    got = O.illumination(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["T"], [0, 1],
                         stride, precision="f64")
    assert np.abs(got.ravel() - ill).max() <= 1e-12 * np.abs(ill).max()
    assert ill.max() > 0


def _exact_fd_case():
    p = small_problem(X=40, Z=36, T=90, pml=6, n_src=1, n_rx=5, seed=2, f0=0.07)
    cq0 = p["cq_bg"].astype(np.float64)
    cqt = p["cq_true"].astype(np.float64)
    obs = NP.forward(cqt, p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"], p["T"])[0]
    return p, 0.5 * (cq0 + cqt), obs


@pytest.mark.parametrize("kind", ["l2", "w2"])
def test_exact_adjoint_central_fd(kind):
    """The exact discrete adjoint restated in NumPy (the transpose of the
    forward recursion) against central finite differences of J in FP64."""
    p, cq, obs = _exact_fd_case()
    fn = None if kind == "l2" else (lambda rec: O.wasserstein_exact(rec, obs))
    args = (p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"], obs, p["T"])
    _, g, _ = NP.exact_gradient(cq, *args, rho_fn=fn)
    zz, xx = np.mgrid[0:p["Z"], 0:p["X"]]
    for cz, cx in ((20, 18), (14, 25)):
        d = np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * 3.0 ** 2))
        h = 1e-5
        jp = NP.exact_gradient(cq + h * d, *args, rho_fn=fn)[0]
        jm = NP.exact_gradient(cq - h * d, *args, rho_fn=fn)[0]
        fd = (jp - jm) / (2 * h)
        assert abs(fd - (g * d).sum()) < 1e-6 * abs(fd)


def test_w2_exact_trace_derivative():
    rng = np.random.default_rng(0)
    T = 300
    tt = np.arange(T)

    def ricker(t0, f=0.03):
        a = (np.pi * f * (tt - t0)) ** 2
        return (1 - 2 * a) * np.exp(-a)

    obs, sim = ricker(120), 1.05 * ricker(126)
    w, rho = O.wasserstein_exact_trace(sim, obs)
    assert w == O.wasserstein_trace(sim, obs)[0]
    for _ in range(3):
        d = rng.standard_normal(T)
        h = 1e-6
        fd = (O.wasserstein_exact_trace(sim + h * d, obs)[0]
              - O.wasserstein_exact_trace(sim - h * d, obs)[0]) / (2 * h)
        assert abs(fd - rho @ d) < 1e-7 * abs(fd)
