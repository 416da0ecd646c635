"""Pins the C oracle of the split-field scheme (oracle/oracle_pml_impl.h) by
what the mathematics of regression_PML_single.cu fixes, independently of any
second restatement:

* closed-form first samples (first_samples): at a source cell rec[1] = s[1]
  and rec[2] = s1 (1 - (dx~ + dz~) dt/2) - 4 m (dt/dx)^2 s1 sum a_k^2 + s1 + s2,
  with the damping read one cell further along its axis (dx~ = d_x[z][x+1],
  dz~ = d_z[z+1][x], .cu:505-506); at the four neighbours
  rec[2] = m (dt/dx)^2 s1 (a1^2 - 2 (a1 a2 + a2 a3 + a3 a4)).  This fixes the
  1/2 + 1/2 running-sum injection that never uses s[0] (.cu:286-291, 661),
  the damping shift of the P step and the D- o D+ pairing with its literals
  (deriv_macros.h:85-108);
* damping placement of the A step (damping_placement): a damping column at
  d_x[:, x0] acts, through the shifted lookup, on the A cell x0-1 only, which
  moves rec[3] at (x0-2, z0) by exactly -m a1 a2 s1 (dt/dx)^2 delta dt
  (a_k with their signs)
  (.cu:592-593); likewise along z;
* exact reconstruction without damping (reconstruction): the reverse run from
  the seeds P(T-1), A(T-2) (.cu:787-790) with the ring restore
  Px_b = Pz_b = ring/2 (.cu:269-270, 805) reproduces U_t = P(T-1-t) at every
  cell and step;
* brute-force imaging (imaging): the gradient equals
  sum_t Lambda_t lap8(P(T-1-t)), Lambda being a forward run that injects the
  time-flipped adjoint source at the receivers (.cu:793-823), lap8 the
  9-point stencil whose weights are fixed by polynomial exactness;
* mirror symmetry without damping (mirror), before edge reflections arrive;
* arrival times against the dispersion relation of the staggered leapfrog
  pair, sin(w/2) = (v/2) K(k), K(k) = 2 sum a_j sin((2j-1) k/2), with
  dt != dx (arrival).

test_each_check_catches_its_mutant builds, per check, a copy of the oracle
with one quirk flipped (the copy lives in the test's temporary directory;
oracle/ is never changed) and asserts that the check fails on it.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests import np_scheme as NP

ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
A = np.array(NP.C, dtype=np.float64)   # a1..a4 with their signs (pinned in test_oracle_scheme)
a1, a2, a3, a4 = A
HAS_GCC = shutil.which("gcc") is not None


def _rng_src(rng, n, T):
    s = rng.standard_normal((n, T))
    s[:, 0] = 0.7 + rng.random(n)        # s[0] nonzero: the scheme must never read it
    return s


def _every(Z, X):
    zz, xx = np.mgrid[0:Z, 0:X]
    return xx.ravel().astype(np.int32), zz.ravel().astype(np.int32)


def _pts(*xz):
    return (np.array([p[0] for p in xz], np.int32), np.array([p[1] for p in xz], np.int32))


# --------------------------------------------------------------------------
# the checks (each takes an oracle library handle from O.load)
# --------------------------------------------------------------------------

def check_first_samples(L):
    X, Z, pml = 48, 44, 14
    dxv, dtv = 1.0, 0.75          # dt exact in FP32 (constvec is FP32)
    r = dtv / dxv
    rng = np.random.default_rng(11)
    cq = 0.1 + 0.15 * rng.random((Z, X))
    d_x, d_z = O.calc_pml(X, Z, pml, R=1e-3)
    T = 3
    # the stencils of the first two steps must stay inside the box 4 <= x < X-4
    cases = [(8, 9, d_x, d_z), (X - 9, Z - 9, d_x, d_z),  # PML corners
             (22, 20, d_x, d_z),                          # interior
             (21, 19, 0 * d_x, 0 * d_z)]                  # undamped, neighbours sampled
    for x0, z0, dx_, dz_ in cases:
        s = _rng_src(rng, 1, T)
        nb = [(x0, z0), (x0 - 1, z0), (x0 + 1, z0), (x0, z0 - 1), (x0, z0 + 1)]
        rec = O.forward(cq, dx_, dz_, _pts((x0, z0)), s, _pts(*nb), T,
                        constvec=(dxv, dxv, dtv), precision="f64", _lib=L)["rec"]
        s1, s2 = s[0, 1], s[0, 2]
        sc = abs(s1) + abs(s2) + 1
        assert np.all(rec[:, 0] == 0)
        assert abs(rec[0, 1] - s1) <= 1e-14 * sc, "rec[1] at the source"
        assert np.all(rec[1:, 1] == 0)
        dxt, dzt = dx_[z0, x0 + 1], dz_[z0 + 1, x0]
        m = cq[z0, x0]
        want = s1 * (1 - (dxt + dzt) * dtv / 2) - 4 * m * r * r * s1 * (A ** 2).sum() + s1 + s2
        assert abs(rec[0, 2] - want) <= 1e-13 * sc, "rec[2] at the source"
        if not dx_.any():
            auto = a1 * a1 - 2 * (a1 * a2 + a2 * a3 + a3 * a4)
            for k, (x, z) in enumerate(nb[1:], 1):
                w = cq[z, x] * r * r * s1 * auto
                assert abs(rec[k, 2] - w) <= 1e-13 * sc, f"rec[2] at neighbour {(x, z)}"


def check_damping_placement(L):
    X, Z, T = 40, 40, 4
    x0, z0 = 20, 20
    dtv = 0.625                   # exact in FP32 (constvec is FP32)
    rng = np.random.default_rng(12)
    cq = 0.1 + 0.15 * rng.random((Z, X))
    s = _rng_src(rng, 1, T)
    s1 = s[0, 1]
    delta = 0.3
    zero = np.zeros((Z, X))
    for axis in ("x", "z"):
        d = np.zeros((Z, X))
        if axis == "x":
            d[:, x0] = delta             # shifted lookup: acts on cells x0-1
            rx, args = (x0 - 2, z0), (d, zero)
        else:
            d[z0, :] = delta             # acts on rows z0-1
            rx, args = (x0, z0 - 2), (zero, d)
        run = lambda dx_, dz_: O.forward(cq, dx_, dz_, _pts((x0, z0)), s, _pts(rx), T,
                                         constvec=(1.0, 1.0, dtv), precision="f64",
                                         _lib=L)["rec"][0]
        diff = run(*args) - run(zero, zero)
        want = -cq[rx[1], rx[0]] * a1 * a2 * s1 * dtv ** 2 * delta * dtv
        assert abs(diff[2]) <= 1e-15 * abs(s1), f"damping acts too early ({axis})"
        assert abs(diff[3] - want) <= 1e-9 * abs(want), f"A-step damping cell ({axis})"


def _recon_problem(seed=13, X=40, Z=36, T=70, pml=6):
    rng = np.random.default_rng(seed)
    cq = 0.08 + 0.12 * rng.random((Z, X))
    src_mask = np.zeros((Z, X), bool)
    src_mask[pml + 3, [12, 27]] = True
    rx_mask = np.zeros((Z, X), bool)
    rx_mask[pml + 2, [9, 16, 23, 30]] = True
    src = _rng_src(rng, 2, T)
    ring = O.point_order(O.revert_mask(X, Z, pml, 4, src_mask))
    return dict(X=X, Z=Z, T=T, cq=cq, src_xz=O.point_order(src_mask), src=src,
                rx_xz=O.point_order(rx_mask), ring_xz=ring, rng=rng)


def _history(L, p, d_x, d_z, src_xz, src):
    """P = Px + Pz after every step t (row t; row 0 = the zero initial state)."""
    rec = O.forward(p["cq"], d_x, d_z, src_xz, src, _every(p["Z"], p["X"]), p["T"],
                    precision="f64", _lib=L)["rec"]
    return rec.T.reshape(p["T"], p["Z"], p["X"])


def check_reconstruction(L):
    p = _recon_problem()
    Z, X, T = p["Z"], p["X"], p["T"]
    zero = np.zeros((Z, X))
    P = _history(L, p, zero, zero, p["src_xz"], p["src"])
    fw = O.forward(p["cq"], zero, zero, p["src_xz"], p["src"], p["rx_xz"], T,
                   ring_xz=p["ring_xz"], precision="f64", want_seed=True, _lib=L)
    adj = p["rng"].standard_normal(fw["rec"].shape)
    _, U, _ = O.adjoint(p["cq"], zero, zero, p["rx_xz"], adj, p["ring_xz"], fw["ring"],
                        fw["seed"], precision="f64", want_hist=True, _lib=L)
    sc = np.abs(P).max()
    assert sc > 0
    for t in range(T):
        e = np.abs(U[t] - P[T - 1 - t]).max()
        assert e <= 1e-11 * sc, f"U_{t} != P({T - 1 - t}): {e / sc:.2e}"


def lap8_weights():
    return np.array([-9.0, 128.0, -1008.0, 8064.0, -14350.0, 8064.0, -1008.0, 128.0, -9.0]) / 5040


def _lap8(u):
    """9-point second differences along x and z, operands outside the
    interior box 4 <= x < X-4, 4 <= z < Z-4 read as 0."""
    Z, X = u.shape
    v = np.zeros((Z + 8, X + 8))
    v[8:Z, 8:X] = u[4:Z - 4, 4:X - 4]
    w = lap8_weights()
    out = np.zeros((Z, X))
    for j, wj in enumerate(w):
        out += wj * v[4:Z + 4, j:j + X] + wj * v[j:j + Z, 4:X + 4]
    return out


def check_imaging(L):
    p = _recon_problem(seed=14, T=60)
    Z, X, T = p["Z"], p["X"], p["T"]
    zero = np.zeros((Z, X))
    P = _history(L, p, zero, zero, p["src_xz"], p["src"])
    fw = O.forward(p["cq"], zero, zero, p["src_xz"], p["src"], p["rx_xz"], T,
                   ring_xz=p["ring_xz"], precision="f64", want_seed=True, _lib=L)
    adj = p["rng"].standard_normal(fw["rec"].shape)
    g = O.adjoint(p["cq"], zero, zero, p["rx_xz"], adj, p["ring_xz"], fw["ring"], fw["seed"],
                  precision="f64", _lib=L)
    # Lambda: the forward scheme driven by the time-flipped adjoint source at the receivers
    Lam = _history(L, p, zero, zero, p["rx_xz"], adj[:, ::-1].copy())
    Lam[0] = 0.0
    want = sum(Lam[t] * _lap8(P[T - 1 - t]) for t in range(T))
    assert np.abs(want).max() > 0
    assert np.abs(g - want).max() <= 1e-10 * np.abs(want).max(), "imaging condition"


def check_mirror(L):
    X, Z, T = 121, 81, 150
    xc, zc = 60, 40
    rng = np.random.default_rng(15)
    half = 0.12 + 0.08 * rng.random((Z, xc + 1))
    cq = np.concatenate([half, half[:, :-1][:, ::-1]], axis=1)      # cq[:, x] = cq[:, X-1-x]
    assert np.array_equal(cq, cq[:, ::-1])
    s = _rng_src(rng, 1, T)
    s[0] = np.convolve(s[0], np.hanning(15), mode="same")            # smooth, broadband
    pts = [(xc + j, z) for j in (3, 10) for z in (zc, zc - 9)]
    pts += [(xc - j, z) for j in (3, 10) for z in (zc, zc - 9)]
    zero = np.zeros((Z, X))
    rec = O.forward(cq, zero, zero, _pts((xc, zc)), s, _pts(*pts), T, precision="f64",
                    _lib=L)["rec"]
    n = len(pts) // 2
    sc = np.abs(rec).max()
    assert sc > 0
    assert np.abs(rec[:n] - rec[n:]).max() <= 1e-10 * sc, "mirror symmetry"


def discrete_group_velocity(v, k):
    """Group velocity (cells per step) along x of the staggered leapfrog pair
    at wavenumber k (radians per cell): w(k) = 2 arcsin(v K(k) / 2)."""
    j = np.arange(1, 5)
    K = 2 * np.sum(A * np.sin((2 * j - 1) * k / 2))
    dK = np.sum(A * (2 * j - 1) * np.cos((2 * j - 1) * k / 2))
    return v * dK / np.sqrt(1 - (v * K / 2) ** 2)


def check_arrival(L):
    N, T = 201, 300
    c, dtv = 0.5, 0.8
    v = c * dtv                                    # cells per step
    f0 = 0.02                                      # cycles per step: 20 cells per wavelength
    t = np.arange(T)
    a = (np.pi * f0 * (t - 60.0)) ** 2
    s = ((1 - 2 * a) * np.exp(-a))[None, :]
    cq = np.full((N, N), c * c)
    zero = np.zeros((N, N))
    xc = N // 2
    r1, r2 = 40, 70
    rec = O.forward(cq, zero, zero, _pts((xc, xc)), s, _pts((xc + r1, xc), (xc + r2, xc)), T,
                    constvec=(1.0, 1.0, dtv), precision="f64", _lib=L)["rec"]
    # the direct pulse only (edge reflections reach r2 after ~365 steps)
    cc = np.correlate(rec[1], rec[0], mode="full")
    i = int(np.argmax(cc))
    y0, y1, y2 = cc[i - 1], cc[i], cc[i + 1]
    lag = i - (T - 1) + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)
    k0 = 2 * np.pi * f0 / v
    want = (r2 - r1) / discrete_group_velocity(v, k0)
    assert abs(lag - want) <= 3e-3 * want, f"arrival lag {lag:.3f} vs {want:.3f}"


CHECKS = {f.__name__[6:]: f for f in (check_first_samples, check_damping_placement,
                                      check_reconstruction, check_imaging, check_mirror,
                                      check_arrival)}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_oracle_passes(name):
    CHECKS[name](O.lib())


def test_lap8_weights_are_the_eighth_order_second_difference():
    w = lap8_weights()
    j = np.arange(-4, 5, dtype=np.float64)
    for p in range(0, 10):
        want = 2.0 if p == 2 else 0.0
        assert abs((w * j ** p).sum() - want) < 1e-9 * (1 + 4.0 ** p)
    assert np.array_equal(w, w[::-1])


def test_group_velocity_formula_is_the_derivative_of_the_dispersion_relation():
    v = 0.4
    j = np.arange(1, 5)
    w = lambda k: 2 * np.arcsin(v * np.sum(A * np.sin((2 * j - 1) * k / 2)))
    for k in (0.05, 0.3, 0.9):
        h = 1e-6
        assert abs((w(k + h) - w(k - h)) / (2 * h) - discrete_group_velocity(v, k)) < 1e-8
    # long waves travel at v (first-derivative consistency of the pair)
    assert abs(discrete_group_velocity(v, 1e-6) - v) < 1e-6


# --------------------------------------------------------------------------
# mutants: one quirk flipped each, in a copy of the oracle source
# --------------------------------------------------------------------------

IMPL = "oracle_pml_impl.h"
MUTANTS = {
    # 1/2 + 1/2 split of the running sum into Px and Pz
    "full_injection": ([("(AR)integ[k] / (AR)2", "(AR)integ[k]", 2)], ["first_samples"]),
    # the injected value is the running sum, not the sample
    "no_running_sum": ([("integ[k] = (ST)((AR)integ[k] + (AR)s[", "integ[k] = (ST)((AR)s[", 1)],
                       ["first_samples"]),
    # sample t (never s[0]) at step t
    "reads_s0": ([("(flip ? T - 1 - t : t)", "(flip ? T - 1 - t : t - 1)", 1)], ["first_samples"]),
    # damping looked up at the cell itself, P step / A step, x / z
    "p_damp_x_unshifted": ([("(:505-506) */\n            const AR dxx = (AR)dxm[i + 1];",
                             "(:505-506) */\n            const AR dxx = (AR)dxm[i];", 1)],
                           ["first_samples", "damping_placement"]),
    "p_damp_z_unshifted": ([("const AR dzz = (AR)dzm[i + X];\n            const AR m",
                             "const AR dzz = (AR)dzm[i];\n            const AR m", 1)],
                           ["first_samples", "damping_placement"]),
    "a_damp_x_unshifted": ([("0));\n            const AR dxx = (AR)dxm[i + 1];",
                             "0));\n            const AR dxx = (AR)dxm[i];", 1)],
                           ["damping_placement"]),
    "a_damp_z_unshifted": ([("dxm[i + 1];\n            const AR dzz = (AR)dzm[i + X];\n"
                             "            if (!revert) {\n                Axc",
                             "dxm[i + 1];\n            const AR dzz = (AR)dzm[i];\n"
                             "            if (!revert) {\n                Axc", 1)],
                           ["damping_placement"]),
    # D+ in the P step (both half-steps forward-staggered)
    "p_step_dplus": ([("(Axp, X, Z, x, z, 1)", "(Axp, X, Z, x, z, 0)", 1),
                      ("(Azp, X, Z, x, z, 1)", "(Azp, X, Z, x, z, 0)", 1)],
                     ["first_samples", "mirror"]),
    # D- in the A step
    "a_step_dminus": ([("(Pxc, X, Z, x, z, 0)", "(Pxc, X, Z, x, z, 1)", 2),
                       ("(Pzc, X, Z, x, z, 0)", "(Pzc, X, Z, x, z, 1)", 2)],
                      ["first_samples", "mirror"]),
    # dt/dx of the A step dropped
    "a_step_no_dtdx": ([("dt / dx * (dPxdx + dPzdx)", "(dPxdx + dPzdx)", 2),
                        ("dt / dx * (dPxdz + dPzdz)", "(dPxdz + dPzdz)", 2)],
                       ["first_samples", "arrival"]),
    # a staggered literal mistyped (deriv_macros.h:87-94)
    "literal_c2": ([("#define C2 7.975260e-02", "#define C2 7.957260e-02", 1)], ["first_samples"]),
    # reverse seeds: P from slot (T-1)%2, A from slot (T-2)%2
    "seed_a_slot": ([("Ax + ((T - 2) % 2) * XZ", "Ax + ((T - 1) % 2) * XZ", 1),
                     ("Az + ((T - 2) % 2) * XZ", "Az + ((T - 1) % 2) * XZ", 1)],
                    ["reconstruction", "imaging"]),
    "seed_p_slot": ([("Px + ((T - 1) % 2) * XZ", "Px + ((T - 2) % 2) * XZ", 1),
                     ("Pz + ((T - 1) % 2) * XZ", "Pz + ((T - 2) % 2) * XZ", 1)],
                    ["reconstruction", "imaging"]),
    # ring restore writes half of the saved P into each split field, sample T-1-t
    "ring_full_value": ([("ring[(size_t)k * T + T - 1 - t] / (AR)2)",
                          "ring[(size_t)k * T + T - 1 - t])", 1)],
                        ["reconstruction", "imaging"]),
    "ring_time_shift": ([("ring[(size_t)k * T + T - 1 - t] / (AR)2)",
                          "ring[(size_t)k * T + T - t] / (AR)2)", 1)],
                        ["reconstruction", "imaging"]),
    # adjoint source injected without the time flip
    "adjoint_not_flipped": ([("adj, integ, t, 1, -1)", "adj, integ, t, 0, -1)", 1)], ["imaging"]),
    # imaging with the adjoint field of the previous step
    "imaging_lambda_lag": ([("L[i] = (ST)((AR)Pxf[cur * XZ + i] + (AR)Pzf[cur * XZ + i]);",
                             "L[i] = (ST)((AR)Pxf[prv * XZ + i] + (AR)Pzf[prv * XZ + i]);", 1)],
                           ["imaging"]),
    # Laplacian without its z part
    "lap_x_only": ([("    return lx + lz;", "    return lx + (AR)0 * lz;", 1)], ["imaging"]),
}


def _build_mutant(tmp, edits):
    for f in ("oracle_pml.c", IMPL):
        shutil.copy(os.path.join(ORACLE_DIR, f), tmp)
    for old, new, count in edits:
        fn = os.path.join(tmp, "oracle_pml.c" if old.startswith("#define") else IMPL)
        s = open(fn).read()
        assert s.count(old) == count, f"mutation site {old!r} found {s.count(old)} times"
        open(fn, "w").write(s.replace(old, new))
    so = os.path.join(tmp, "libmutant.so")
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
                           "-std=gnu11", "-w", "-o", so, os.path.join(tmp, "oracle_pml.c")])
    return O.load(so)


@pytest.mark.skipif(not HAS_GCC, reason="gcc builds the mutant copies")
@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_each_check_catches_its_mutant(mutant, tmp_path):
    edits, caught_by = MUTANTS[mutant]
    L = _build_mutant(str(tmp_path), edits)
    for name in caught_by:
        with pytest.raises(AssertionError):
            CHECKS[name](L)


def test_every_check_has_a_mutant():
    used = {c for _, caught in MUTANTS.values() for c in caught}
    assert used == set(CHECKS)
