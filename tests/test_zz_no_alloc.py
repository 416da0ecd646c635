"""otfwi_misfit_noalloc (device arrays, no allocation) against otfwi_misfit
(host arrays) through the C ABI: identical kernels, so identical results,
on two calls in a row (the second smaller), for L2 and W2, including a W2
call that runs its traces in several chunks of the fixed device array."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _traces(rng, n, T):
    t = np.arange(T)
    out = np.zeros((n, T))
    for i in range(n):
        t0, f = rng.uniform(0.3 * T, 0.6 * T), rng.uniform(0.01, 0.05)
        a = (np.pi * f * (t - t0)) ** 2
        out[i] = (1 - 2 * a) * np.exp(-a) * rng.uniform(0.5, 2.0)
    return out


def _host(L, kind, sim, obs):
    n, T = sim.shape
    mis = np.zeros(n)
    adj = np.zeros((n, T))
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert L.otfwi_misfit(kind, n, T, P(sim), P(obs), P(mis), P(adj)) == 0
    return mis, adj


def _noalloc(L, kind, sim, obs):
    import torch

    n, T = sim.shape
    ds = torch.from_numpy(sim).cuda()
    do = torch.from_numpy(obs).cuda()
    dm = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    da = torch.full((n, T), float("nan"), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rc = L.otfwi_misfit_noalloc(kind, n, T, ds.data_ptr(), do.data_ptr(), dm.data_ptr(),
                                da.data_ptr(), stream)
    assert rc == 0, L.otfwi_last_error()
    return dm.cpu().numpy(), da.cpu().numpy()


@pytest.mark.parametrize("kind", [0, 1])
def test_noalloc_matches_misfit(gpu, kind):
    from otfwi_amd import _lib

    L = _lib.load()
    rng = np.random.default_rng(kind)
    for n, T in ((700, 4000), (37, 1500)):      # 700 W2 traces at T=4000: two chunks and a part
        print(f"kind {kind}: {n} traces of {T} samples", flush=True)
        obs = _traces(rng, n, T)
        sim = obs + 0.05 * _traces(rng, n, T)
        m0, a0 = _host(L, kind, sim, obs)
        m1, a1 = _noalloc(L, kind, sim, obs)
        assert np.isfinite(m0).all() and np.isfinite(a0).all()
        assert np.array_equal(m0, m1) and np.array_equal(a0, a1)


def test_noalloc_rejects_what_it_cannot_do(gpu):
    from otfwi_amd import _lib

    L = _lib.load()
    assert L.otfwi_misfit_noalloc(2, 4, 100, 1, 1, 1, 1, None) == _lib.E_ARG
    assert L.otfwi_misfit_noalloc(1, 4, 20000, 1, 1, 1, 1, None) == _lib.E_ARG
    assert b"T <=" in L.otfwi_last_error()
