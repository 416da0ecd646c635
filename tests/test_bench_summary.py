"""The bench's time-to-tolerance summary over per-model inversion results
(learned.summarize + bench.inversion_summary), on synthetic records: the
plain method stalls above the tolerance on every model, the learned method
reaches it on two of three."""
import pytest

import bench
from otfwi_amd import learned as L


def _run(err0, best, t_tol, wall, iters):
    return dict(err0=err0, best_err=best, time_to_tol=t_tol, wall=wall, iters=iters, n_grad=0,
                n_obj=0, kinds=[], errs=[], times=[], time_to_plain_final=None)


def test_summary_of_stalled_plain_runs():
    res = [
        {"seed": 1, "plain": _run(0.0239, 0.0229, None, 30.8, 6),
         "learned": _run(0.0239, 0.0185, 6.9, 6.9, 1)},
        {"seed": 2, "plain": _run(0.0217, 0.0216, None, 9.6, 2),
         "learned": _run(0.0217, 0.0197, 6.9, 6.9, 1)},
        {"seed": 3, "plain": _run(0.0225, 0.0224, None, 30.0, 30),
         "learned": _run(0.0225, 0.0224, None, 19.1, 2)},
    ]
    s = bench.inversion_summary(L.summarize(res), res, n_iter=30)
    assert s["models"] == 3 and s["plain_reached_tol"] == 0 and s["learned_reached_tol"] == 2
    assert s["plain_time_to_tol"] is None and s["time_to_tol_ratio"] is None
    assert s["learned_time_to_tol"] == pytest.approx(6.9)
    assert s["time_to_tol_ratio_lower_bound_per_model"] == pytest.approx([30.8 / 6.9, 9.6 / 6.9])
    assert s["time_to_tol_ratio_lower_bound"] == pytest.approx(9.6 / 6.9)
    # the third plain run used its whole budget: not counted as stalled
    assert s["plain_stalled_above_tol"] == 2
    assert [m["seed"] for m in s["per_model"]] == [1, 2, 3]


def test_summary_when_both_reach():
    res = [{"seed": 1, "plain": _run(0.03, 0.019, 40.0, 50.0, 12),
            "learned": _run(0.03, 0.015, 8.0, 8.0, 1)}]
    s = bench.inversion_summary(L.summarize(res), res, n_iter=30)
    assert s["time_to_tol_ratio"] == pytest.approx(5.0)
    assert s["time_to_tol_ratio_lower_bound"] is None and s["plain_stalled_above_tol"] == 0
