"""Launch machinery: hipGraph replay of the step loops, autotuner, launch
styles -- all must give results bitwise equal to plain per-step launches."""
import time

import numpy as np
import pytest

from oracle import oracle as O
from tests.problems import config1_problem, make_model, small_problem

pytestmark = pytest.mark.gpu


def test_graph_replay_bitwise(gpu):
    p = small_problem(X=200, Z=150, T=400, pml=10, seed=21, n_src=2, n_rx=5)
    obs = np.stack([O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                              p["rx_xz"], p["T"], idx=i)["rec"] for i in (0, 1)])
    res = []
    for graph in (False, True):
        m = make_model(p, cq="cq_bg")
        m.set_graph(graph)
        m.set_observed(obs)
        r1 = m.forward([0, 1])
        g1, mis1 = m.gradient([0, 1], misfit="w2")
        g2, mis2 = m.gradient([0, 1], misfit="w2")      # replay of the cached graphs
        m.set_tuning(16, 32, 0)                           # new tiles -> re-capture
        g3, mis3 = m.gradient([0, 1], misfit="w2")
        res.append((r1, g1, mis1, g2, g3, mis3))
    for a, b in zip(*res):
        assert np.array_equal(a, b)
    assert np.array_equal(res[1][1], res[1][3]) and np.array_equal(res[1][1], res[1][4])


def test_autotune_keeps_results(gpu):
    p = small_problem(X=330, Z=160, T=300, pml=10, seed=22, n_src=2, n_rx=5)
    obs = np.stack([O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                              p["rx_xz"], p["T"], idx=i)["rec"] for i in (0, 1)])
    m = make_model(p, cq="cq_bg")
    m.set_observed(obs)
    g0, mis0 = m.gradient([0, 1], misfit="l2")
    choice = m.autotune(2, steps=10)
    print("autotune:", choice)
    assert choice["th_fwd"] in (16, 32, 64) and choice["th_adj"] in (16, 32, 64)
    assert m.get_tuning()[:2] == (choice["th_fwd"], choice["th_adj"])
    g1, mis1 = m.gradient([0, 1], misfit="l2")
    assert np.array_equal(g0, g1) and np.array_equal(mis0, mis1)


def test_config1_graph_speedup(gpu):
    """Config 1 (256^2, one shot, 1499 steps) is launch-bound; report both."""
    p = config1_problem(seed=0)
    m = make_model(p, cq="cq_bg")
    obs = m.forward([-1])
    m.set_observed(obs)
    out = {}
    for graph in (False, True):
        m.set_graph(graph)
        m.gradient([-1], misfit="l2")
        t0 = time.perf_counter()
        for _ in range(3):
            g, mis = m.gradient([-1], misfit="l2")
        out[graph] = (time.perf_counter() - t0) / 3
    print(f"config1 L2 gradient: {out[False] * 1e3:.1f} ms direct, {out[True] * 1e3:.1f} ms graph")
    assert out[True] <= 1.2 * out[False]


def test_cache_blocking_bitwise(gpu):
    p = small_problem(X=200, Z=160, T=300, seed=51, n_src=3, n_rx=8)
    shots = [0, 1, 2, 1]
    obs = np.stack([O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                              p["rx_xz"], p["T"], idx=i)["rec"] for i in shots])
    m = make_model(p, cq="cq_bg")
    m.set_observed(obs)
    g0, mis0 = m.gradient(shots, "l2")
    for kf, gf, ka, ga in ((3, 2, 5, 1), (7, 3, 2, 2)):
        m.set_blocking(kf, gf, ka, ga)
        assert m.get_blocking() == (kf, gf, ka, ga)
        g, mis = m.gradient(shots, "l2")
        assert np.array_equal(g, g0) and np.array_equal(mis, mis0)
