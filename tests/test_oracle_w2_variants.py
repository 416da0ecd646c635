"""Pins the oracle of the four further Wasserstein-type misfits
(adjoint_functions.wasserstein_quad / _experimental / _pm / _yang,
adjoint_functions.py:70-266) against the library the reference calls
(scipy.interpolate.interp1d with linear extrapolation; numpy cumsum) and
against closed forms:

* identical traces give a zero misfit; the adjoint source is zero for quad,
  pm and yang (G^-1(F(t)) = t), and 2 - 2 t^3 for the experimental form
  (Y = t((t - q)^2 - 2tq) = -2t^3 at q = t, minus its value at t = 1);
* every form is invariant under a common positive scaling of sim and obs
  (each density is normalized; the shifts scale with the data);
* float32 observed traces keep the reference's float32 density arithmetic
  (quad, pm, yang square, clip or shift them in their own dtype).

The scipy transcriptions below follow the reference's statement order; they
are test fixtures, not product code.
"""
import numpy as np
import pytest
import scipy.interpolate as ip

from oracle import oracle as O
from tests.test_oracle_w2 import traces


def _ip(x, y):
    return ip.interp1d(x, y, fill_value="extrapolate")


def scipy_quad(sim, obs):
    n, T = sim.shape
    t = np.linspace(0, 1, T)
    dt = t[1] - t[0]
    W2, out = 0, np.zeros((n, T))
    for s in range(n):
        g, f = obs[s, :] ** 2, sim[s, :] ** 2
        c = np.max(g) * 0.05
        g += c
        f += c
        g /= np.sum(g * dt)
        f /= np.sum(f * dt)
        G, F = np.cumsum(g * dt), np.cumsum(f * dt)
        Fi, F1, G1 = _ip(t, F), _ip(F, t), _ip(G, t)
        dist = ip.interp1d(t, 2 * np.cumsum(t - G1(Fi(t))) * dt)
        out[s, :] = dist(t) - dist(1)
        W2 += sum(np.abs(F1(t) - G1(t)) ** 2)
    return W2, out


def scipy_experimental(sim, obs):
    n, T = sim.shape
    t = np.linspace(0, 1, T)
    dt = t[1] - t[0]
    W2, out = 0, np.zeros((n, T))
    for s in range(n):
        fg = np.concatenate([np.zeros(0), sim[s, :]])
        gg = np.concatenate([np.zeros(0), obs[s, :]])
        c = np.maximum(-1.1 * np.min(gg), 0.1 * np.max(gg))
        g, f = gg + c, fg + c
        g /= np.sum(g * dt)
        f /= np.sum(f * dt)
        G, F = np.cumsum(g * dt), np.cumsum(f * dt)
        Fi, F1, G1 = _ip(t, F), _ip(F, t), _ip(G, t)
        dist = ip.interp1d(t, t * ((t - G1(Fi(t))) ** 2 - 2 * t * (G1(Fi(t)))))
        out[s, :] = dist(t) - dist(1)
        W2 += sum(np.abs(F1(t) - G1(t)) ** 2 * F1(t))
    return W2, out


def scipy_pm(sim, obs):
    n, T = sim.shape
    t = np.linspace(0, 1, T)
    dt = t[1] - t[0]
    W2, out = 0, np.zeros((n, T))
    for s in range(n):
        g_p, f_p = np.maximum(0, obs[s, :]), np.maximum(0, sim[s, :])
        c_p = np.max(g_p) * 0.01
        g_p += c_p
        f_p += c_p
        g_p /= np.sum(g_p * dt)
        f_p /= np.sum(f_p * dt)
        G_p, F_p = np.cumsum(g_p * dt), np.cumsum(f_p * dt)
        G_n, F_n = np.cumsum(g_p * dt), np.cumsum(f_p * dt)   # as the reference builds them
        F_ip, G_1p, F_1p = _ip(t, F_p), _ip(G_p, t), _ip(F_p, t)
        F_in, G_1n, F_1n = _ip(t, F_n), _ip(G_n, t), _ip(F_n, t)
        dist = ip.interp1d(t, np.cumsum(t - G_1p(F_ip(t))) * dt + np.cumsum(t - G_1n(F_in(t))) * dt)
        out[s, :] = dist(t) - dist(t[-1])
        W2 += sum(np.abs(F_1p(t) - G_1p(t)) ** 2) + sum(np.abs(F_1n(t) - G_1n(t)) ** 2)
    return W2, out


def scipy_yang(ff, gg):
    n, T = ff.shape
    t = np.linspace(0, 1, T)
    dt = t[1] - t[0]
    W2, out = 0, np.zeros((n, T))
    for s in range(n):
        c = -1.1 * np.min(gg[s, :])
        g, f = gg[s, :] + c, ff[s, :] + c
        g /= np.sum(g * dt)
        f /= np.sum(f * dt)
        G, F = np.cumsum(g * dt), np.cumsum(f * dt)
        f_i, g_i, F_i = _ip(t, f), _ip(t, g), _ip(t, F)
        G_1 = _ip(G, t)
        U = np.triu(np.ones((T, T)))
        out[s, :] = (U @ np.diag(-2 * f_i(t) / g_i(G_1(F_i(t)))) + np.diag(t - G_1(F_i(t)))) \
            @ (t - G_1(F_i(t)))
        W2 += sum(np.abs(t - G_1(F)) ** 2 * f)
    return W2, out


FORMS = {"quad": (O.wasserstein_quad, scipy_quad),
         "experimental": (O.wasserstein_experimental, scipy_experimental),
         "pm": (O.wasserstein_pm, scipy_pm),
         "yang": (O.wasserstein_yang, scipy_yang)}


@pytest.mark.parametrize("form", sorted(FORMS))
@pytest.mark.parametrize("n,T,neg,f32", [(1, 64, False, False), (3, 300, False, True),
                                         (2, 200, True, False)])
def test_matches_scipy_path(form, n, T, neg, f32):
    sim, obs = traces(n, T, seed=T + 7, neg=neg)
    if f32:
        obs = obs.astype(np.float32)   # the stored observed traces are float32
    got = FORMS[form][0](sim, obs)
    want = FORMS[form][1](sim, obs)
    assert abs(got[0] - want[0]) <= 1e-12 * abs(want[0])
    assert np.abs(got[1] - want[1]).max() <= 1e-12 * np.abs(want[1]).max()


@pytest.mark.parametrize("form", sorted(FORMS))
def test_identical_traces(form):
    _, obs = traces(3, 257, seed=5)
    w, a = FORMS[form][0](obs.copy(), obs)
    assert abs(w) <= 1e-20
    if form == "experimental":
        t = np.linspace(0, 1, obs.shape[1])
        assert np.abs(a - (2 - 2 * t ** 3)[None]).max() <= 1e-12
    else:
        assert np.abs(a).max() <= 1e-12


@pytest.mark.parametrize("form", sorted(FORMS))
def test_scale_invariance(form):
    sim, obs = traces(2, 200, seed=9)
    w0, a0 = FORMS[form][0](sim, obs)
    w1, a1 = FORMS[form][0](3.0 * sim, 3.0 * obs)
    assert abs(w1 - w0) <= 1e-10 * abs(w0)
    assert np.abs(a1 - a0).max() <= 1e-9 * np.abs(a0).max()
