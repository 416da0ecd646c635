"""The oracle of the second-order C-PML scheme (regression_CPML.cu,
oracle/oracle_cpml.c) pinned by properties the scheme must have: linearity,
superposition, mirror symmetry, arrival times at the medium's velocity,
invariance of kappa / rho, and absorption by the layer."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.problems import cpml_problem


def homogeneous(p, c=0.4):
    p = dict(p)
    Z, X = p["cquad"].shape
    for k in ("cquad",):
        p[k] = np.full((Z, X), c * c, np.float32)
    p["kappa"] = np.full((Z, X), c * c, np.float32)
    for k in ("rho_hx", "rho_hz"):
        p[k] = np.ones((Z, X), np.float32)
    return p


def test_zero_source_zero_traces():
    p = cpml_problem(T=120)
    p["src"] = np.zeros_like(p["src"])
    assert not np.any(O.cpml_forward(p))


def test_linearity_and_superposition():
    p = cpml_problem(T=200, seed=1)
    r = O.cpml_forward(p)
    q = dict(p, src=2 * p["src"])
    assert np.array_equal(O.cpml_forward(q), 2 * r)       # power-of-2 scaling is exact
    parts = sum(O.cpml_forward(p, idx=s).astype(np.float64) for s in range(len(p["sx"])))
    assert np.abs(parts - r).max() <= 1e-5 * np.abs(r).max()
    # pixel_per_element != 1: every source fires whatever idx says
    assert np.array_equal(O.cpml_forward(dict(p, ppe=4), idx=1), r)


def test_mirror_symmetry():
    """A centred source in a homogeneous medium: mirrored receivers record
    the same trace until waves come back from the grid's edges.  (The scheme
    is not mirror-symmetric at the edges: the half-point derivative fields are
    kept on indices 4 .. X-5, i.e. half points 4.5 .. X-4.5, whose mirror is
    3.5 .. X-5.5.)"""
    X = 97
    p = homogeneous(cpml_problem(X=X, Z=80, T=260, n_src=1, n_rx=8, seed=2))
    p["sx"] = np.array([X // 2], np.int32)
    p["rx"] = np.array([20, 30, X - 1 - 30, X - 1 - 20], np.int32)
    p["rz"] = np.full(4, p["rz"][0], np.int32)
    r = O.cpml_forward(p)[:, :130]
    assert np.abs(r).max() > 0.1
    assert np.abs(r[0] - r[3]).max() <= 1e-6 * np.abs(r).max()
    assert np.abs(r[1] - r[2]).max() <= 1e-6 * np.abs(r).max()


def test_arrival_time_matches_velocity():
    """Peak of the trace at horizontal offset D arrives about D / c after
    the wavelet's own peak (t0 = 1.2 / f0 = 24 samples), for c = 0.4 and 0.3."""
    for c in (0.4, 0.3):
        p = homogeneous(cpml_problem(X=160, Z=80, T=300, n_src=1, n_rx=2, seed=3), c=c)
        p["sx"] = np.array([40], np.int32)
        p["rx"] = np.array([40 + 30, 40 + 60], np.int32)
        r = O.cpml_forward(p)
        for k, D in enumerate((30, 60)):
            t_peak = int(np.argmax(np.abs(r[k])))
            assert abs(t_peak - (24 + D / c)) <= 8, (c, D, t_peak)


def test_kappa_over_rho_invariance():
    """P_tt = kappa div(grad P / rho): doubling kappa and rho together leaves
    the traces unchanged (the source term uses cquad, not kappa)."""
    p = cpml_problem(T=220, seed=4)
    r = O.cpml_forward(p)
    q = dict(p, kappa=2 * p["kappa"], rho_hx=2 * p["rho_hx"], rho_hz=2 * p["rho_hz"])
    assert np.abs(O.cpml_forward(q) - r).max() <= 1e-5 * np.abs(r).max()
    q = dict(p, kappa=np.float32(1.2) * p["kappa"])   # a real change must show
    assert np.abs(O.cpml_forward(q) - r).max() > 1e-2 * np.abs(r).max()


def test_layer_absorbs():
    """Late-time energy with the C-PML layer is far below that of the same
    grid with the layer switched off (a = 0, b = 1, K = 1: a rigid box)."""
    p = cpml_problem(T=900, seed=5)
    r = O.cpml_forward(p)
    q = dict(p)
    for ax in ("x", "z"):
        for h in ("", "h"):
            q[f"a_{ax}{h}"] = np.zeros_like(p[f"a_{ax}{h}"])
            q[f"b_{ax}{h}"] = np.ones_like(p[f"b_{ax}{h}"])
            q[f"K_{ax}{h}"] = np.ones_like(p[f"K_{ax}{h}"])
    r0 = O.cpml_forward(q)
    late = slice(600, 900)
    assert np.abs(r[:, late]).max() < 0.1 * np.abs(r0[:, late]).max()
