"""Adjoint gradient on the GPU against the CPU oracle (reference phase 2,
regression_PML_single.cu:704-846, with adjoint_functions.L2 / wasserstein).

Bars (fixed):
* L2: misfit within 1e-6 and gradient within 1e-4 of the reference-precision
  oracle; the FP64 build within 1e-9 / 1e-6 of the FP64 oracle;
* W2 adjoint propagation given the oracle's own FP32 adjoint source (the
  library's host-adjoint entry point): the FP32 build within 1e-4 of the
  FP64 oracle (the FP64 build within 1e-6; check_w2_propagation);
* W2 end to end: gradient within 2e-3 of both the reference-precision and
  the FP64 oracle, misfit within 1e-6 of the reference-precision oracle.
  The W2 adjoint source amplifies the FP32 rounding of the traces about
  300x (relative noise 1e-7 on the traces moves it by 6e-4), so no FP32
  forward run gives a W2 gradient within 1e-4 of FP64 end to end: this
  library's is 8.4e-4 from FP64 at the 256^2 configuration, the reference's
  own arithmetic 1.2e-3.  The W2 stage itself is checked on identical traces
  in test_gpu_misfit.py.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.problems import config1_problem, make_model, rel, small_problem

pytestmark = pytest.mark.gpu


def observed(p, shots):
    return np.stack([O.forward(p["cq_true"], p["d_x"], p["d_z"], p["src_xz"], p["src"],
                               p["rx_xz"], p["T"], idx=i)["rec"] for i in shots])


def oracle_grad(p, obs, shots, misfit, precision="ref"):
    return O.gradient(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"],
                      p["ring_xz"], obs, p["T"], misfit, shots=shots, precision=precision)


def oracle_adjoint_sources(p, obs, shots, misfit):
    """The reference-precision oracle's own FP32 adjoint source per shot."""
    out = []
    for k, idx in enumerate(shots):
        rec = O.forward(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"],
                        p["T"], idx=idx)["rec"]
        out.append(misfit(rec.astype(np.float64), obs[k])[1].astype(np.float32))
    return np.stack(out)


def oracle_adjoint(p, shots, adj, precision):
    """Mean over shots of the oracle's phase 2 for given adjoint sources."""
    gs = []
    for k, idx in enumerate(shots):
        fw = O.forward(p["cq_bg"], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"], p["T"],
                       idx=idx, ring_xz=p["ring_xz"], want_seed=True, precision=precision)
        gs.append(O.adjoint(p["cq_bg"], p["d_x"], p["d_z"], p["rx_xz"], adj[k], p["ring_xz"],
                            fw["ring"], fw["seed"], precision=precision).astype(np.float64))
    return np.mean(gs, axis=0)


def check_w2_propagation(p, obs, shots, label):
    """Adjoint propagation of a W2 adjoint source alone, through the
    library's two-stage host-adjoint entry points, with the oracle's own
    FP32 W2 adjoint sources: the FP64 build within 1e-6 of the FP64 oracle,
    the FP32 build within 1e-4 of it.  A W2 adjoint source has a large mean,
    so its running sum builds a slowly varying field around the receivers
    that the imaging sum over t nearly cancels; the reference's own FP32
    arithmetic (derivatives rounded to FP32 from FP64) is 8.4e-4 from FP64 on
    the 256^2 configuration, so the distance to the reference-precision
    oracle is printed, not held to 1e-4."""
    adj = oracle_adjoint_sources(p, obs, shots, O.wasserstein)
    g64 = oracle_adjoint(p, shots, adj.astype(np.float64), "f64")
    gref = oracle_adjoint(p, shots, adj, "ref")
    out = {}
    for prec in (8, 4):
        m = make_model(p, precision=prec, cq="cq_bg")
        m.forward_for_grad(shots)
        out[prec] = m.adjoint_from_host(adj).astype(np.float64)
    e8, e4, er = rel(out[8], g64), rel(out[4], g64), rel(out[4], gref)
    print(f"{label} W2 adjoint propagation: FP64 build vs FP64 oracle {e8:.3e}; FP32 build vs "
          f"FP64 oracle {e4:.3e}, vs reference precision {er:.3e} (reference precision vs FP64 "
          f"{rel(gref, g64):.3e})")
    assert e8 <= 1e-6 and e4 <= 1e-4


@pytest.mark.parametrize("kind", ["l2", "w2"])
def test_small_gradient(gpu, kind):
    p = small_problem(seed=11)
    shots = [0, 1]
    obs = observed(p, shots)
    m = make_model(p, cq="cq_bg")
    m.set_observed(obs)
    grad, mis = m.gradient(shots, misfit=kind)
    fn = O.l2 if kind == "l2" else O.wasserstein
    g_ref, mse_ref, _ = oracle_grad(p, obs, shots, fn)
    g64, mse64, _ = oracle_grad(p, obs.astype(np.float64), shots, fn, precision="f64")
    eg, em = rel(grad, g_ref), abs(mis.mean() - mse_ref) / abs(mse_ref)
    print(f"{kind}: grad {eg:.3e} (oracle ref vs f64 {rel(g_ref, g64):.3e}), "
          f"misfit {em:.3e} (ref vs f64 {abs(mse_ref - mse64) / abs(mse64):.3e})")
    if kind == "l2":
        assert eg < 1e-4 and em < 1e-6
    else:
        print(f"w2: grad vs f64 {rel(grad, g64):.3e}")
        assert eg <= 2e-3 and rel(grad, g64) <= 2e-3 and em <= 1e-6
        check_w2_propagation(p, obs, shots, "small")


@pytest.mark.parametrize("kind", ["l2", "w2"])
def test_small_gradient_fp64(gpu, kind):
    p = small_problem(seed=12)
    shots = [1, 0]
    obs = observed(p, shots)
    m = make_model(p, precision=8, cq="cq_bg")
    m.set_observed(obs)
    grad, mis = m.gradient(shots, misfit=kind)
    fn = O.l2 if kind == "l2" else O.wasserstein
    g64, mse64, _ = oracle_grad(p, obs.astype(np.float64), shots, fn, precision="f64")
    eg, em = rel(grad, g64), abs(mis.mean() - mse64) / abs(mse64)
    print(f"fp64 {kind}: grad {eg:.3e} misfit {em:.3e}")
    assert eg < 1e-6 and em < 1e-9


def test_config1_l2_and_w2(gpu):
    p = config1_problem(seed=0)
    obs = observed(p, [-1])
    m = make_model(p, cq="cq_bg")
    m.set_observed(obs)
    for kind, fn in (("l2", O.l2), ("w2", O.wasserstein)):
        grad, mis = m.gradient([-1], misfit=kind)
        g_ref, mse_ref, _ = oracle_grad(p, obs, [-1], fn)
        g64, mse64, _ = oracle_grad(p, obs.astype(np.float64), [-1], fn, precision="f64")
        eg, em = rel(grad, g_ref), abs(mis[0] - mse_ref) / abs(mse_ref)
        print(f"config1 {kind}: grad vs ref-oracle {eg:.3e}, vs f64 {rel(grad, g64):.3e} "
              f"(ref vs f64 {rel(g_ref, g64):.3e}); misfit vs ref {em:.3e}, vs f64 "
              f"{abs(mis[0] - mse64) / abs(mse64):.3e}")
        if kind == "l2":
            assert eg < 1e-4 and em < 1e-6
        else:
            assert eg <= 2e-3 and rel(grad, g64) <= 2e-3 and em <= 1e-6
            check_w2_propagation(p, obs, [-1], "config1")


def test_host_adjoint_matches_device(gpu):
    p = small_problem(seed=13)
    shots = [0, 1]
    obs = observed(p, shots)
    m = make_model(p, cq="cq_bg")
    m.set_observed(obs)
    g_dev, mis = m.gradient(shots, misfit="l2")
    sim = m.forward_for_grad(shots)
    adj = sim.astype(np.float64) - obs
    g_host = m.adjoint_from_host(adj)
    assert rel(g_host, g_dev) < 1e-6
    # another run in between reuses the buffers: the second stage refuses
    m.forward_for_grad(shots)
    m.forward(shots)
    with pytest.raises(Exception):
        m.adjoint_from_host(adj)


def test_batched_gradient_equals_chunked(gpu):
    p = small_problem(seed=14, n_src=3)
    shots = [0, 1, 2]
    obs = observed(p, shots)
    a = make_model(p, cq="cq_bg")
    a.set_observed(obs)
    b = make_model(p, cq="cq_bg", max_batch=1)
    b.set_observed(obs)
    ga, ma = a.gradient(shots, misfit="w2")
    gb, mb = b.gradient(shots, misfit="w2")
    assert np.array_equal(ma, mb)
    assert rel(ga, gb) < 1e-6   # FP64 cross-shot sums in another grouping


def test_tile_invariance_with_interior_tiles(gpu):
    """Interior tiles (no PML) and strip tiles must round identically, so the
    forward traces and the whole gradient are bitwise independent of the
    tile heights, and repeated calls are deterministic."""
    # T long enough for every receiver to see the wave: an all-zero observed
    # trace makes the reference's W2 shift 0 and its normalization 0/0 (NaN)
    p = small_problem(X=330, Z=160, T=760, pml=10, seed=15, n_src=2, n_rx=5)
    shots = [0, 1]
    obs = observed(p, shots)
    base = None
    for thf, tha in [(32, 16), (16, 16), (64, 32), (32, 64)]:
        m = make_model(p, cq="cq_bg", th_fwd=thf, th_adj=tha)
        m.set_observed(obs)
        g1, mis1, rec = m.gradient(shots, misfit="w2", want_rec=True)
        g2, mis2 = m.gradient(shots, misfit="w2")
        assert np.array_equal(g1, g2) and np.array_equal(mis1, mis2)
        if base is None:
            base = (g1, mis1, rec)
        else:
            assert np.array_equal(rec, base[2]), (thf, tha)
            assert np.array_equal(mis1, base[1]), (thf, tha)
            assert np.array_equal(g1, base[0]), (thf, tha)
    g_ref, mse_ref, _ = oracle_grad(p, obs, shots, O.wasserstein)
    g64, mse64, _ = oracle_grad(p, obs.astype(np.float64), shots, O.wasserstein, precision="f64")
    e, spread = rel(base[0], g_ref), rel(g_ref, g64)
    print(f"330x160 W2: gradient vs reference-precision oracle {e:.3e} (oracle ref vs f64 "
          f"{spread:.3e}); misfit {base[1].mean():.6e} vs {mse_ref:.6e} (f64 {mse64:.6e})")
    assert e <= 2e-3 and rel(base[0], g64) <= 2e-3


def test_all_zero_observed_trace_is_reported(gpu):
    from otfwi_amd._lib import OtfwiError

    p = small_problem(X=200, Z=80, T=60, seed=16, n_rx=6)   # far receivers see nothing in 60 steps
    obs = observed(p, [0])
    assert (np.abs(obs[0]).max(axis=1) == 0).any()
    m = make_model(p, cq="cq_bg")
    m.set_observed(obs)
    with pytest.raises(OtfwiError, match="all zero"):
        m.gradient([0], misfit="w2")
    g, mis = m.gradient([0], misfit="l2")     # L2 is well defined there
    assert np.isfinite(g).all()


@pytest.mark.parametrize("thf,tha", [(16, 16), (32, 32)])
def test_persistent_launch_bitwise(gpu, thf, tha):
    """Persistent work-groups compute the same cells with the same arithmetic."""
    p = small_problem(X=330, Z=160, T=300, pml=10, seed=17, n_src=3, n_rx=5)
    shots = [0, 1, 2]
    obs = observed(p, shots)
    out = []
    for pers in (False, True):
        m = make_model(p, cq="cq_bg", th_fwd=thf, th_adj=tha)
        m.set_launch(pers, pers)
        m.set_observed(obs)
        out.append(m.gradient(shots, misfit="l2", want_rec=True))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_envelope_gradient(gpu):
    """Device gradient with the squared-envelope misfit against the oracle's
    gradient with the reference's quad_envelope steps (host adjoint source
    rounded to FP32, as adj_wrapper does).  Fixed bars: gradient within
    5e-3, misfit within 5e-5 (the envelope difference amplifies FP32 trace
    rounding like W2: the oracle's own FP32 misfit is 1.5e-5 from FP64 here;
    the spreads are printed)."""
    p = small_problem(seed=14)
    shots = [0, 1]
    obs = observed(p, shots)
    m = make_model(p, cq="cq_bg")
    m.set_observed(obs)
    grad, mis = m.gradient(shots, misfit="env")
    g_ref, mse_ref, _ = oracle_grad(p, obs, shots, O.quad_envelope)
    g64, mse64, _ = oracle_grad(p, obs.astype(np.float64), shots, O.quad_envelope,
                                precision="f64")
    eg, em = rel(grad, g_ref), abs(mis.mean() - mse_ref) / abs(mse_ref)
    print(f"env: grad {eg:.3e} (oracle ref vs f64 {rel(g_ref, g64):.3e}), "
          f"misfit {em:.3e} (ref vs f64 {abs(mse_ref - mse64) / abs(mse64):.3e})")
    assert eg <= 5e-3 and rel(grad, g64) <= 5e-3
    assert em <= 5e-5


def test_envelope_drop_in_device_path(gpu):
    """Cuda_interface_PML keeps adjoint_functions.quad_envelope on the
    device.  A host callback with the same steps in FP64 (the observed traces
    promoted before the Hilbert transform) gives the same gradient and misfit;
    the reference's own precision mix (a complex64 analytic signal of the
    float32 observed traces) is printed beside it: where sim and obs nearly
    agree the envelope difference cancels and that rounding shows in the
    gradient at the 1e-3 level."""
    from otfwi_amd import adjoint_functions as AF
    from otfwi_amd.cuda_interface_pml import Cuda_interface_PML

    p = small_problem(seed=15)
    obs = observed(p, [0, 1])
    c = np.sqrt(p["cq_bg"]).astype(np.float64)
    runs = []
    for fn in (AF.quad_envelope,
               lambda s, o: O.quad_envelope(s, np.asarray(o, np.float64)),
               lambda s, o: O.quad_envelope(s, o)):
        cp = Cuda_interface_PML(p["X"], p["Z"], p["T"], p["pml"], c, len(p["src_xz"][0]),
                                p["src_mask"], len(p["rx_xz"][0]), p["rx_mask"], p["src"],
                                multishot=True, R=1e-3)
        cp.init_regression(obs, fn, prec_deriv=4)
        cp.calc_grad()
        runs.append((cp.grad.copy(), float(cp.mse[0])))
    (gd, md), (gh, mh), (gr, mr) = runs
    print(f"drop-in env: vs FP64 host grad {rel(gd, gh):.3e} misfit {abs(md - mh) / abs(mh):.3e}; "
          f"vs reference mix grad {rel(gd, gr):.3e} misfit {abs(md - mr) / abs(mr):.3e}")
    assert rel(gd, gh) < 1e-5 and abs(md - mh) <= 1e-9 * abs(mh)
