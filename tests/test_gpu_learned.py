"""What the trained preconditioner gains over the plain method: on held-out
random-defect models (seeds never used in training), inversions with the
learned step (judged by the same W2 misfit-decrease test, falling back to the
plain step) end with a lower velocity error than the plain method after the
same number of iterations, and reach the plain method's final error sooner."""
import os

import numpy as np
import pytest

from otfwi_amd import learned as L

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(L.WEIGHTS), reason="no trained weights")
def test_learned_beats_plain_on_held_out_models(gpu):
    from otfwi_amd.unet import UNet

    prob = L.Problem()
    net = UNet(params=L.load_weights())
    res = L.compare(prob, net, [100000 + k for k in range(4)], n_iter=10, log=print)
    s = L.summarize(res)
    print(s)
    assert s["learned_best_err"] < 0.9 * s["plain_best_err"]
    assert s["learned_reached_plain_final"] >= 3
    assert s["matched_error_speedup"] > 1.0
