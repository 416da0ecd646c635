"""Pins the W2 / L2 oracle (oracle/oracle.py) against the library the
reference calls (scipy.interpolate.interp1d, scipy.integrate's cumulative
trapezoid; adjoint_functions.py:49-54, 61) and against closed forms."""
import numpy as np
import pytest
import scipy.integrate as integ
import scipy.interpolate as ip

from oracle import oracle as O


def traces(n, T, seed, neg=False):
    rng = np.random.default_rng(seed)
    t = np.arange(T)
    out = []
    for _ in range(2):
        a = np.zeros((n, T))
        for i in range(n):
            k0, w = rng.uniform(0.1, 0.7) * T, rng.uniform(0.01, 0.05) * T
            a[i] = np.exp(-((t - k0) / w) ** 2) * np.cos((t - k0) / (0.3 * w)) * rng.uniform(0.5, 2)
        out.append(a)
    sim, obs = out
    if neg:
        sim[:, T // 3] = -5 * np.abs(obs).max(axis=1)
    return sim, obs


@pytest.mark.parametrize("seed", range(4))
def test_interp_matches_scipy(seed):
    rng = np.random.default_rng(seed)
    xs = np.sort(rng.uniform(0, 1, 50))
    xs[10] = xs[11]                      # a tie
    if seed % 2:
        xs = rng.permutation(xs)         # non-monotone: scipy sorts (mergesort) first
    ys = rng.standard_normal(50)
    q = np.concatenate([rng.uniform(-0.2, 1.2, 200), xs[:5]])
    ref = ip.interp1d(xs, ys, fill_value="extrapolate")(q)
    got = O._interp_linear(xs, ys, q)
    assert np.array_equal(got, ref)


def test_cumtrapz_matches_scipy():
    rng = np.random.default_rng(0)
    y = rng.standard_normal(77)
    t = np.linspace(0, 1, 77)
    assert np.array_equal(O._cumtrapz0(y, np.diff(t)), integ.cumulative_trapezoid(y, t, initial=0))
    assert np.array_equal(O._cumtrapz0(y, 1.0), integ.cumulative_trapezoid(y, initial=0))


def direct_scipy_w2(sim, obs):
    """adjoint_functions.wasserstein's steps written with the scipy calls the
    reference makes (cumtrapz is cumulative_trapezoid in SciPy >= 1.14)."""
    n, T = sim.shape
    t = np.linspace(0, 1, T)
    dt = t[1] - t[0]
    W2, out = 0, np.zeros((n, T))
    for s in range(n):
        c = np.maximum(-1.1 * np.min(obs[s]), 0.1 * np.max(obs[s]))
        g, f = obs[s] + c, sim[s] + c
        g = g / np.sum(g * dt)
        f = f / np.sum(f * dt)
        G = integ.cumulative_trapezoid(g, t, initial=0)
        F = integ.cumulative_trapezoid(f, t, initial=0)
        Fi = ip.interp1d(t, F, fill_value="extrapolate")
        F1 = ip.interp1d(F, t, fill_value="extrapolate")
        G1 = ip.interp1d(G, t, fill_value="extrapolate")
        dist = ip.interp1d(t, 2 * integ.cumulative_trapezoid(t - G1(Fi(t)), initial=0))
        out[s] = dist(t) - dist(1)
        W2 += sum(np.abs(F1(t) - G1(t)) ** 2)
    return W2, out


@pytest.mark.parametrize("n,T,neg", [(1, 64, False), (3, 500, False), (2, 400, True)])
def test_wasserstein_matches_scipy_path(n, T, neg):
    sim, obs = traces(n, T, seed=T, neg=neg)
    w, a = O.wasserstein(sim, obs)
    w0, a0 = direct_scipy_w2(sim, obs)
    assert w == w0
    assert np.array_equal(a, a0)


def test_wasserstein_closed_forms():
    sim, obs = traces(3, 300, seed=1)
    w, a = O.wasserstein(obs, obs)
    assert w == 0 and np.abs(a).max() < 1e-13
    # positive scaling of both traces leaves misfit and adjoint source unchanged
    w1, a1 = O.wasserstein(sim, obs)
    w2, a2 = O.wasserstein(7.5 * sim, 7.5 * obs)
    assert abs(w2 - w1) <= 1e-12 * abs(w1)
    assert np.abs(a2 - a1).max() <= 1e-10 * np.abs(a1).max()
    assert w1 > 0


def test_wasserstein_time_shift_direction():
    # sim arriving later than obs: the transport map moves mass back in time,
    # t - G^-1(F(t)) >= 0 where the CDFs rise, so the adjoint integral is
    # non-decreasing there
    T = 400
    t = np.arange(T)
    obs = np.exp(-((t - 150) / 12.0) ** 2)[None]
    sim = np.exp(-((t - 170) / 12.0) ** 2)[None]
    w, a = O.wasserstein(sim, obs)
    assert w > 0
    assert np.all(np.diff(a[0, 160:180]) >= -1e-12)


def test_l2_closed_form():
    sim, obs = traces(2, 100, seed=2)
    w, r = O.l2(sim, obs)
    assert np.array_equal(r, sim - obs)
    assert w == np.sum((sim - obs) ** 2)
