"""Two time steps per launch (otfwi_set_fwd_steps / otfwi_set_adj_steps):
traces, ring, reference-mode and exact gradients must be bitwise equal to
one step per launch, for grids spanning several column blocks (224 owned
columns) and row chunks (256 rows), with event cells in the blocks' halos,
odd and even step counts, and both precisions."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.problems import make_model, small_problem

pytestmark = pytest.mark.gpu


def _obs(p, shots):
    return np.stack([O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                               p["rx_xz"], p["T"], idx=i)["rec"] for i in shots])


def _run(p, shots, obs, steps, precision, w2, adj_steps=1):
    m = make_model(p, precision=precision, cq="cq_bg")
    m.set_fwd_steps(steps)
    m.set_adj_steps(adj_steps)
    assert m.get_fwd_steps() == steps and m.get_adj_steps() == adj_steps
    m.set_observed(obs)
    out = {"rec": m.forward(shots)}
    out["g_l2"], out["mis_l2"] = m.gradient(shots, "l2")
    if w2:
        out["g_w2"], out["mis_w2"] = m.gradient(shots, "w2")
    out["g_exact"], out["mis_exact"] = m.gradient_exact(shots, "l2", snap_every=7)
    return out


@pytest.mark.parametrize("X,Z,T,precision", [
    (200, 160, 301, 4),     # one block, one chunk, odd step count
    (500, 600, 240, 4),     # 3 column blocks x 3 row chunks
    (470, 300, 6, 4),       # pairs only (5 steps after the first)
    (96, 80, 5, 4),         # too few steps for an even number of pairs
    (300, 280, 120, 8),     # FP64
])
def test_fwd2_bitwise(gpu, X, Z, T, precision):
    p = small_problem(X=X, Z=Z, T=T, pml=10, seed=X + Z, n_src=5, n_rx=min(40, X // 8))
    shots = [0, 3, 4, 1]
    long = T > 20
    obs = _obs(p, shots) if long else np.zeros((len(shots), len(p["rx_xz"][0]), T), np.float32)
    w2 = long and bool((np.abs(obs).max(axis=-1) > 0).all())   # W2 needs nonzero traces
    a = _run(p, shots, obs, 1, precision, w2=w2)
    for fs, ads in ((2, 1), (1, 2), (2, 2)):
        b = _run(p, shots, obs, fs, precision, w2=w2, adj_steps=ads)
        assert a.keys() == b.keys()
        for name in a:
            assert np.array_equal(a[name], b[name]), (fs, ads, name)
    if long:
        assert np.abs(a["rec"]).max() > 0 and np.abs(a["g_l2"]).max() > 0


def test_fwd2_rejects_bad_value(gpu):
    p = small_problem()
    m = make_model(p)
    assert m.get_fwd_steps() == 0 and m.get_adj_steps() == 0     # default: by launch size
    with pytest.raises(Exception):
        m.set_fwd_steps(3)
    with pytest.raises(Exception):
        m.set_adj_steps(-1)
    assert m.get_fwd_steps() == 0 and m.get_adj_steps() == 0


@pytest.mark.parametrize("X,Z,T", [(700, 300, 121), (1100, 270, 60)])
def test_pair_width_bitwise(gpu, X, Z, T):
    """Every two-step block width gives the one-step results."""
    p = small_problem(X=X, Z=Z, T=T, pml=10, seed=X + T, n_src=5, n_rx=40)
    shots = [4, 0, 2]
    obs = _obs(p, shots)
    a = _run(p, shots, obs, 1, 4, w2=False)
    for nw in (128, 192, 256, 384, 0):
        m = make_model(p, cq="cq_bg")
        m.set_pair_width(nw)
        m.set_steps_per_launch(2, 2)
        m.set_observed(obs)
        assert np.array_equal(m.forward(shots), a["rec"]), nw
        g, mis = m.gradient(shots, "l2")
        assert np.array_equal(g, a["g_l2"]) and np.array_equal(mis, a["mis_l2"]), nw
    with pytest.raises(Exception):
        m.set_pair_width(320)


def test_pairs_batches_and_all_sources(gpu):
    """Two-step launches with the shots split over batches (the last one
    smaller than the buffers' capacity) and with every source firing at
    once: bitwise equal to one step per launch."""
    p = small_problem(X=480, Z=300, T=130, pml=10, seed=7, n_src=5, n_rx=30)
    shots = [0, 1, 2, 3, 4]
    obs = _obs(p, shots)
    out = []
    for steps in (1, 2):
        m = make_model(p, cq="cq_bg", max_batch=2)
        m.set_steps_per_launch(steps, steps)
        m.set_observed(obs)
        r = m.forward(shots)
        g, mis = m.gradient(shots, "l2")
        m2 = make_model(p, cq="cq_bg")
        m2.set_steps_per_launch(steps, steps)
        r_all = m2.forward([-1, 2, -1])
        out.append((r, g, mis, r_all))
    for u, v in zip(*out):
        assert np.array_equal(u, v)
    assert np.abs(out[0][3]).max() > 0


@pytest.mark.parametrize("graph", [
    pytest.param(True, marks=pytest.mark.skip(
        reason="a run ended with a host segmentation fault in the first forward call of this "
               "case when its loops were captured as one graph across 2 or 4 streams; no cause "
               "found in the library, so that capture is no longer built and this case stays "
               "off (per-stream-group capture: test_split_loops_captured_per_group)")),
    False])
def test_pair_streams_bitwise(gpu, graph):
    """The two-step phase split over two or four streams (groups of
    consecutive shots) gives the one-step results when launched directly,
    for an odd shot count."""
    p = small_problem(X=420, Z=300, T=140, pml=10, seed=11, n_src=5, n_rx=30)
    shots = [3, 0, 4, 1, 2]
    obs = _obs(p, shots)
    ref = None
    for steps, k in ((1, 1), (2, 1), (2, 2), (2, 4)):
        m = make_model(p, cq="cq_bg")
        m.set_graph(graph)
        m.set_steps_per_launch(steps, steps)
        m.set_pair_streams(k)
        m.set_observed(obs)
        out = [m.forward(shots)]
        for _ in range(2):                      # second call replays the graphs
            g, mis = m.gradient(shots, "l2")
            out += [g, mis]
        if ref is None:
            ref = out
        for u, v in zip(ref, out):
            assert np.array_equal(u, v), (steps, k)
    with pytest.raises(Exception):
        m.set_pair_streams(3)


@pytest.mark.parametrize("k", [2, 4])
def test_split_loops_captured_per_group(gpu, k):
    """Graphs on (the default) with the two-step phase split over k stream
    groups: each group's pairs are captured by themselves on the group's own
    stream and replayed there.  Results stay bitwise those of one step per
    launch and of direct split launches over repeated calls (graph replays),
    for an odd shot count, after a re-capture for another batch, and for a
    batch of 3 shots over 4 groups (one group empty)."""
    p = small_problem(X=420, Z=300, T=140, pml=10, seed=12, n_src=5, n_rx=30)
    shots = [1, 4, 0, 3, 2]
    obs = _obs(p, shots)
    out = {}
    for steps, kk, graph in ((1, 1, True), (2, k, False), (2, k, True)):
        m = make_model(p, cq="cq_bg")
        m.set_steps_per_launch(steps, steps)
        m.set_pair_streams(kk)
        m.set_graph(graph)
        if steps == 2:
            modes = m.graph_modes(len(shots))
            assert modes["fwd_streams"] == kk and modes["adj_streams"] == kk, modes
            want = "graph" if graph else "direct"
            assert modes["fwd"] == want and modes["adj"] == want, modes
        m.set_observed(obs)
        res = [m.forward(shots)]
        for _ in range(3):
            g, mis = m.gradient(shots, "l2")
            res += [g, mis, m.forward(shots)]
        m.set_observed(obs[:3])
        res.append(m.forward(shots[:3]))
        for _ in range(2):
            g, mis = m.gradient(shots[:3], "l2")
            res += [g, mis]
        m.set_observed(obs)
        g, mis = m.gradient(shots, "l2")
        res += [g, mis]
        out[(steps, graph)] = res
    ref = out[(1, True)]
    for key in ((2, False), (2, True)):
        for i, (u, v) in enumerate(zip(ref, out[key])):
            assert np.array_equal(u, v), (key, i)


def test_pairs_batch_bound(gpu):
    """Two-step launches address both slots of a field from one shot's base
    with 32-bit buffer offsets below 0x7FFFFFF0 bytes; a larger batch runs one
    step per launch (2048^2: padded plane 2176 x 2064, at most 118 shots)."""
    p = small_problem(X=2048, Z=2048, T=10, pml=20, seed=3, n_src=4, n_rx=8)
    m = make_model(p, cq="cq_bg")
    plane = 2176 * 2064
    cap = 0x7FFFFFF0 // (4 * plane) - 1
    assert cap == 118
    assert m.steps_per_launch(cap) == (2, 2)
    assert m.steps_per_launch(cap + 1) == (1, 1)
