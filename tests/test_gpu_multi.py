"""Uneven multi-model batching: several models of different grid size, T and
shot count run concurrently (longest first, one stream each), or with every
two-step launch covering all of them (fused), and give the same results, bit
for bit, as separate gradient calls."""
import numpy as np
import pytest

from oracle import oracle as O
from otfwi_amd.model import gradient_multi
from tests.problems import make_model, small_problem

pytestmark = pytest.mark.gpu


def observed(p, shots):
    return np.stack([O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                               p["rx_xz"], p["T"], idx=i)["rec"] for i in shots])


@pytest.mark.parametrize("workers", [1, 3])
def test_multi_model_equals_separate_runs(gpu, workers):
    probs = [small_problem(X=96, Z=80, T=240, seed=41, n_src=2),
             small_problem(X=160, Z=128, T=400, seed=42, n_src=3, n_rx=10),
             small_problem(X=128, Z=96, T=320, seed=43, n_src=2, n_rx=8)]
    shots = [[0, 1], [2, 0, 1], [1]]
    kinds = ["w2", "l2", "w2"]
    models, ref = [], []
    for p, s, k in zip(probs, shots, kinds):
        m = make_model(p, cq="cq_bg")
        m.set_observed(observed(p, s))
        ref.append(m.gradient(s, k))
        models.append(m)
    out = gradient_multi(models, shots, kinds, workers=workers)
    for (g, mis), (g0, mis0) in zip(out, ref):
        assert np.array_equal(g, g0)
        assert np.array_equal(mis, mis0)


def test_fused_equals_separate_runs(gpu):
    """One two-step launch per pair index over all models (uneven grids, step
    counts, so models drop out of the later launches, shot counts and
    misfits), in one stream group and in two (a model of one shot sits in
    one of them), repeated: bitwise the separate gradient calls, which run
    these small batches one step per launch.  A model held to one step per
    launch is refused."""
    from otfwi_amd._lib import OtfwiError

    probs = [small_problem(X=96, Z=80, T=240, seed=41, n_src=2),
             small_problem(X=500, Z=300, T=401, seed=42, n_src=3, n_rx=10),
             small_problem(X=128, Z=96, T=320, seed=43, n_src=2, n_rx=8),
             small_problem(X=260, Z=530, T=90, seed=44, n_src=4, n_rx=12)]
    shots = [[0, 1], [2, 0, 1], [1], [3, 1, 0, 2]]
    kinds = ["w2", "l2", "w2", "l2"]
    models, ref = [], []
    for p, s, k in zip(probs, shots, kinds):
        m = make_model(p, cq="cq_bg")
        m.set_observed(observed(p, s))
        ref.append(m.gradient(s, k))
        models.append(m)
    for groups in (1, 2, 2, 0):
        models[0].set_pair_streams(groups)      # stream groups of the fused launches
        out = gradient_multi(models, shots, kinds, fused=True)
        for i, ((g, mis), (g0, mis0)) in enumerate(zip(out, ref)):
            assert np.array_equal(g, g0), (groups, i)
            assert np.array_equal(mis, mis0), (groups, i)
    assert any(np.abs(g).max() > 0 for g, _ in out)
    models[2].set_steps_per_launch(1, 1)
    with pytest.raises(OtfwiError):
        gradient_multi(models, shots, kinds, fused=True)


def test_fused_one_shot_models_two_groups(gpu):
    """Two stream groups forced on models of one shot each: every model sits
    in the second group, the first has no launches and is skipped; results
    bitwise those of separate calls."""
    probs = [small_problem(X=96, Z=80, T=240, seed=45, n_src=2),
             small_problem(X=160, Z=128, T=301, seed=46, n_src=3, n_rx=10)]
    shots = [[1], [2]]
    models, ref = [], []
    for p, s in zip(probs, shots):
        m = make_model(p, cq="cq_bg")
        m.set_observed(observed(p, s))
        ref.append(m.gradient(s, "l2"))
        models.append(m)
    models[0].set_pair_streams(2)
    out = gradient_multi(models, shots, "l2", fused=True)
    for (g, mis), (g0, mis0) in zip(out, ref):
        assert np.array_equal(g, g0) and np.array_equal(mis, mis0)
    assert all(np.abs(g).max() > 0 for g, _ in out)
