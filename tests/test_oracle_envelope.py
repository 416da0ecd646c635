"""The oracle's squared-envelope misfit (adjoint_functions.py:14-22) and its
Hilbert transform, pinned by scipy.signal.hilbert (the routine the
reference calls), closed forms and finite differences."""
import numpy as np
import pytest
import scipy.signal as sig

from oracle import oracle as O


@pytest.mark.parametrize("N", [64, 65, 1500, 1499])
def test_hilbert_matches_scipy(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((3, N))
    a, b = O.hilbert(x), sig.hilbert(x)
    assert a.dtype == b.dtype == np.complex128
    assert np.abs(a - b).max() < 1e-12 * np.abs(b).max()
    x32 = x.astype(np.float32)
    a, b = O.hilbert(x32), sig.hilbert(x32)
    assert a.dtype == b.dtype == np.complex64
    assert np.abs(a - b).max() < 1e-5 * np.abs(b).max()


@pytest.mark.parametrize("N", [64, 1500, 1499])
def test_hilbert_closed_form(N):
    """hilbert(cos(w n)) = cos(w n) + i sin(w n) for w = 2 pi k / N, 0 < k < N/2."""
    n = np.arange(N)
    for k in (1, 7, (N - 1) // 2 - 1):
        w = 2 * np.pi * k / N
        h = O.hilbert(np.cos(w * n))
        assert np.abs(h.real - np.cos(w * n)).max() < 1e-12
        assert np.abs(h.imag - np.sin(w * n)).max() < 1e-12


def test_envelope_zero_and_scaling():
    rng = np.random.default_rng(3)
    sim = rng.standard_normal((4, 200))
    obs = rng.standard_normal((4, 200))
    m0, a0 = O.quad_envelope(sim, sim.copy())
    assert m0 == 0.0 and not np.any(a0)
    m, a = O.quad_envelope(sim, obs)
    m2, a2 = O.quad_envelope(2 * sim, 2 * obs)      # misfit ~ a^4, adjoint ~ a^3
    assert m2 == pytest.approx(16 * m, rel=1e-12)
    assert np.allclose(a2, 8 * a, rtol=1e-12, atol=1e-12 * np.abs(a).max())


@pytest.mark.parametrize("N", [32, 33])
def test_envelope_adjoint_is_quarter_gradient(N):
    """The Hilbert circulant is antisymmetric, so d(sum E^2)/d sim = 4 adj:
    central differences of the misfit pin every term of the adjoint source."""
    rng = np.random.default_rng(N)
    sim = rng.standard_normal((2, N))
    obs = rng.standard_normal((2, N))
    _, adj = O.quad_envelope(sim, obs)
    h = 1e-6
    fd = np.zeros_like(sim)
    for r in range(2):
        for i in range(N):
            p, q = sim.copy(), sim.copy()
            p[r, i] += h
            q[r, i] -= h
            fd[r, i] = (O.quad_envelope(p, obs)[0] - O.quad_envelope(q, obs)[0]) / (2 * h)
    assert np.abs(fd - 4 * adj).max() < 1e-6 * np.abs(fd).max()
