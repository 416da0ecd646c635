"""CPU oracle for the Otfwi 2D acoustic FWI hot path.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py import this module, and only as the checker.
The product package (otfwi_amd) never imports it.

Contents, each a restatement of one reference function:

* :func:`calc_pml`       -- ``Cuda_interface_PML._calcPML`` (cuda_interface_pml.py:116-132)
* :func:`point_order`    -- ``np.where`` point ordering (cuda_interface_pml.py:40-53)
* :func:`revert_mask`    -- the boundary-saving ring of ``init_regression``
                            (cuda_interface_pml.py:134-148)
* :func:`forward`        -- ``cuda_simulate`` / phase 1 of ``cuda_grad_ext``
                            (regression_PML_single.cu:618-702, 745-769), via the C library
* :func:`adjoint`        -- phase 2 of ``cuda_grad_ext`` (regression_PML_single.cu:783-836)
* :func:`l2`             -- ``adjoint_functions.L2`` (adjoint_functions.py:8-11)
* :func:`wasserstein`    -- ``adjoint_functions.wasserstein`` (adjoint_functions.py:25-67)
* :func:`wasserstein_quad`, :func:`wasserstein_experimental`, :func:`wasserstein_pm`,
  :func:`wasserstein_yang` -- the further misfits of adjoint_functions.py:70-266
* :func:`hilbert`        -- ``scipy.signal.hilbert`` as adjoint_functions.py:15-16 calls it
* :func:`quad_envelope`  -- ``adjoint_functions.quad_envelope`` (adjoint_functions.py:14-22)
* :func:`cpml_forward`   -- ``cuda_simulate2`` of the second-order C-PML scheme
                            (regression_CPML.cu:427-501), via the C library
* :func:`gradient`       -- ``Cuda_interface_PML.calc_grad`` (cuda_interface_pml.py:208-249)
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def load(path):
    """ctypes handle of a library built from oracle_pml.c, with the argument
    types of its entry points set."""
    L = ctypes.CDLL(path)
    fp = ctypes.c_void_p
    i = ctypes.c_int
    for sfx in ("_ref", "_f64"):
        f = getattr(L, "orc_forward" + sfx)
        f.argtypes = [i, i, i, fp, fp, fp, fp, i, fp, fp, fp, i, i, fp, fp, fp,
                      i, fp, fp, fp, fp]
        f.restype = i
        a = getattr(L, "orc_adjoint" + sfx)
        a.argtypes = [i, i, i, fp, fp, fp, fp, i, fp, fp, fp, i, fp, fp, fp, fp, fp, fp, fp]
        a.restype = i
    L.orc_num_threads.restype = i
    L.orc_set_threads.argtypes = [i]
    return L


def lib():
    """Load (building on demand when a compiler is present) liboracle_pml.so."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_pml.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", _HERE])
        _LIB = load(path)
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------
# host-side pieces of Cuda_interface_PML
# --------------------------------------------------------------------------

def calc_pml(X, Z, pml_size, R=1e-5, vmax=0.5):
    """Quadratic split-field damping, cuda_interface_pml.py:116-132.

    d0 = 3*Vmax*ln(1/R) / (2*L^3), profile d0*x^2 on x = linspace(0, L, L);
    the top rows / left columns get it reversed, the bottom / right forward.
    Returns dense (Z, X) float64 arrays d_x (varies along x) and d_z.
    """
    L = pml_size
    d0 = 3.0 * vmax * np.log(1.0 / R) / (2.0 * L ** 3)
    prof = d0 * np.linspace(0.0, L, L) ** 2
    d_x = np.zeros((Z, X))
    d_z = np.zeros((Z, X))
    d_z[:L, :] = prof[::-1, None]
    d_z[Z - L:, :] = prof[:, None]
    d_x[:, X - L:] = prof[None, :]
    d_x[:, :L] = prof[None, ::-1]
    return d_x, d_z


def point_order(mask):
    """Row-major point list of a boolean (Z, X) mask -> (x, z) int32 arrays."""
    zz, xx = np.nonzero(np.asarray(mask))
    return xx.astype(np.int32), zz.astype(np.int32)


def revert_mask(X, Z, pml_size, prec_deriv, source_mask):
    """Boundary ring of init_regression, cuda_interface_pml.py:134-148:
    four bands of width prec_deriv+1 just inside the PML plus every source
    cell."""
    m = prec_deriv + 1
    L = pml_size
    mask = np.zeros((Z, X), dtype=bool)
    mask[L:L + m, L:X - L] = True
    mask[Z - L - m:Z - L, L:X - L] = True
    mask[L:Z - L, L:L + m] = True
    mask[L:Z - L, X - L - m:X - L] = True
    mask |= np.asarray(source_mask, dtype=bool)
    return mask


# --------------------------------------------------------------------------
# propagator (C library)
# --------------------------------------------------------------------------

def _prep(dtype, *arrs):
    return [None if a is None else np.ascontiguousarray(a, dtype=dtype) for a in arrs]


def forward(cquad, d_x, d_z, src_xz, src, rx_xz, T, idx=-1, ring_xz=None,
            constvec=(1.0, 1.0, 1.0), precision="ref", want_seed=False, _lib=None):
    """Forward run.  precision "ref" = FP32 state with FP64 derivatives (the
    reference's arithmetic), "f64" = FP64 throughout.  _lib: a handle from
    load() to run instead of liboracle_pml.so.

    Returns dict(rec=(n_rx,T), ring=(n_rev,T) or None, seed=(4,Z,X) or None).
    """
    dt = np.float32 if precision == "ref" else np.float64
    Z, X = cquad.shape
    cq, dxm, dzm, s = _prep(dt, cquad, d_x, d_z, src)
    cv = np.ascontiguousarray(constvec, dtype=np.float32)
    sx, sz = _prep(np.int32, *src_xz)
    rx, rz = _prep(np.int32, *rx_xz)
    ns, nr = len(sx), len(rx)
    rec = np.zeros((nr, T), dt)
    if ring_xz is not None:
        vx, vz = _prep(np.int32, *ring_xz)
        nv = len(vx)
        ring = np.zeros((nv, T), dt)
    else:
        vx = vz = ring = None
        nv = 0
    seed = np.zeros((4, Z, X), dt) if want_seed else None
    L = _lib or lib()
    fn = L.orc_forward_ref if precision == "ref" else L.orc_forward_f64
    rc = fn(X, Z, T, _p(cq), _p(cv), _p(dxm), _p(dzm), ns, _p(sx), _p(sz), _p(s), idx,
            nr, _p(rx), _p(rz), _p(rec), nv, _p(vx), _p(vz), _p(ring), _p(seed))
    if rc != 0:
        raise RuntimeError(f"oracle forward failed ({rc})")
    return dict(rec=rec, ring=ring, seed=seed)


def adjoint(cquad, d_x, d_z, rx_xz, adj, ring_xz, ring, seed,
            constvec=(1.0, 1.0, 1.0), precision="ref", want_hist=False, _lib=None):
    """Phase 2: reverse reconstruction + adjoint run + imaging -> grad (Z,X).
    With want_hist, returns (grad, U (T,Z,X), Lambda (T,Z,X)): the
    reconstructed forward field Px_b+Pz_b and the adjoint field Px_f+Pz_f of
    every reverse step t, as the imaging condition reads them."""
    dt = np.float32 if precision == "ref" else np.float64
    Z, X = cquad.shape
    T = adj.shape[1]
    cq, dxm, dzm, a, rg, sd = _prep(dt, cquad, d_x, d_z, adj, ring, seed)
    cv = np.ascontiguousarray(constvec, dtype=np.float32)
    rx, rz = _prep(np.int32, *rx_xz)
    vx, vz = _prep(np.int32, *ring_xz)
    grad = np.zeros((Z, X), dt)
    uh = np.zeros((T, Z, X), dt) if want_hist else None
    lh = np.zeros((T, Z, X), dt) if want_hist else None
    L = _lib or lib()
    fn = L.orc_adjoint_ref if precision == "ref" else L.orc_adjoint_f64
    rc = fn(X, Z, T, _p(cq), _p(cv), _p(dxm), _p(dzm), len(rx), _p(rx), _p(rz), _p(a),
            len(vx), _p(vx), _p(vz), _p(rg), _p(sd), _p(grad), _p(uh), _p(lh))
    if rc != 0:
        raise RuntimeError(f"oracle adjoint failed ({rc})")
    return (grad, uh, lh) if want_hist else grad


# --------------------------------------------------------------------------
# misfits (adjoint_functions.py)
# --------------------------------------------------------------------------

def l2(sim, obs):
    """adjoint_functions.L2 (:8-11): (sum((sim-obs)^2), sim-obs)."""
    r = sim - obs
    return np.sum(r ** 2), r


def _interp_linear(xs, ys, q):
    """scipy interp1d(kind='linear', fill_value='extrapolate') semantics:
    stable (mergesort) sort of xs, searchsorted-left, index clipped to
    [1, n-1], then slope*(q-x_lo)+y_lo with slope=(y_hi-y_lo)/(x_hi-x_lo)."""
    xs = np.asarray(xs, dtype=np.float64)
    ys = np.asarray(ys)
    if ys.dtype.kind != "f":
        ys = ys.astype(np.float64)
    # y_hi - y_lo is formed in ys's own dtype (float32 data keeps it), the
    # rest in float64, as interp1d does
    order = np.argsort(xs, kind="mergesort")
    xs = xs[order]
    ys = ys[order]
    i = np.searchsorted(xs, q).clip(1, len(xs) - 1)
    x_lo, x_hi = xs[i - 1], xs[i]
    y_lo, y_hi = ys[i - 1], ys[i]
    with np.errstate(divide="ignore", invalid="ignore"):
        slope = (y_hi - y_lo) / (x_hi - x_lo)
        return slope * (q - x_lo) + y_lo


def _cumtrapz0(y, d):
    """[0, cumsum(d*(y[1:]+y[:-1])/2.0)] (scipy cumulative_trapezoid, initial=0)."""
    return np.concatenate([[0.0], np.cumsum(d * (y[1:] + y[:-1]) / 2.0)])


def wasserstein_trace(f0, g0):
    """One receiver trace of adjoint_functions.wasserstein (:33-64); FP64.
    Returns (w2_contribution, adjoint_source_row)."""
    T = len(f0)
    t = np.linspace(0, 1, T)
    dt = t[1] - t[0]
    f0 = np.asarray(f0, dtype=np.float64)
    g0 = np.asarray(g0, dtype=np.float64)
    # affine shift taken from the observed trace only (:38-41)
    c = max(-1.1 * np.min(g0), 0.1 * np.max(g0))
    g = g0 + c
    f = f0 + c
    g = g / np.sum(g * dt)
    f = f / np.sum(f * dt)
    G = _cumtrapz0(g, np.diff(t))
    F = _cumtrapz0(f, np.diff(t))
    Fi = _interp_linear(t, F, t)
    q = _interp_linear(G, t, Fi)              # G^-1(F(t))
    F1 = _interp_linear(F, t, t)
    G1 = _interp_linear(G, t, t)
    D = 2 * _cumtrapz0(t - q, 1.0)            # unit spacing (:61)
    adj = _interp_linear(t, D, t) - _interp_linear(t, D, np.array([1.0]))[0]
    w = 0.0
    for v in np.abs(F1 - G1) ** 2:            # builtin sequential sum (:64)
        w += v
    return w, adj


def wasserstein_exact_trace(f0, g0):
    """Value of wasserstein_trace and its exact derivative with respect to the
    simulated trace f0 (the build's exact-adjoint mode; the reference's own
    adjoint source, adjoint_functions.py:58-62, is a different expression).
    Follows :33-64 forward, then differentiates F^-1(t_k) on its
    interpolation bracket, the trapezoid CDF and the normalization.  Needs a
    monotone CDF of f0 + c.  FP64."""
    T = len(f0)
    t = np.linspace(0, 1, T)
    dt = t[1] - t[0]
    f0 = np.asarray(f0, dtype=np.float64)
    g0 = np.asarray(g0, dtype=np.float64)
    c = max(-1.1 * np.min(g0), 0.1 * np.max(g0))
    g = g0 + c
    h = f0 + c
    g = g / np.sum(g * dt)
    S = np.sum(h * dt)
    f = h / S
    G = _cumtrapz0(g, np.diff(t))
    F = _cumtrapz0(f, np.diff(t))
    if np.any(np.diff(F) < 0):
        raise ValueError("CDF of sim + c is not monotone")
    F1 = _interp_linear(F, t, t)
    G1 = _interp_linear(G, t, t)
    w = 0.0
    for v in np.abs(F1 - G1) ** 2:
        w += v
    e = F1 - G1
    i = np.clip(np.searchsorted(F, t, side="left"), 1, T - 1)
    lo, hi = i - 1, i
    d2 = (F[hi] - F[lo]) ** 2
    dts = t[hi] - t[lo]
    Fb = np.zeros(T)
    np.add.at(Fb, lo, 2 * e * dts * (t - F[hi]) / d2)
    np.add.at(Fb, hi, -2 * e * dts * (t - F[lo]) / d2)
    suf = np.cumsum(Fb[::-1])[::-1]                    # suf[j] = sum_{i >= j} Fb[i]
    d = np.diff(t)
    fb = np.zeros(T)
    fb[1:] += 0.5 * d * suf[1:]
    fb[:-1] += 0.5 * d * suf[1:]
    rho = fb / S - dt * np.dot(fb, h) / S ** 2
    return w, rho


def wasserstein_exact(sim, obs):
    """Per-trace exact W2 derivative: (W2 sum over traces, dW2/dsim (n_rx, T))."""
    sim = np.asarray(sim)
    obs = np.asarray(obs)
    out = np.zeros(sim.shape)
    W2 = 0.0
    for s in range(sim.shape[0]):
        w, out[s] = wasserstein_exact_trace(sim[s], obs[s])
        W2 += w
    return W2, out


def wasserstein(sim, obs):
    """adjoint_functions.wasserstein (:25-67): per-trace W2, summed in trace
    order; returns (W2, adjoint source (n_rx, T) float64)."""
    sim = np.asarray(sim)
    obs = np.asarray(obs)
    n, T = sim.shape
    out = np.zeros((n, T))
    W2 = 0
    for s in range(n):
        w, out[s] = wasserstein_trace(sim[s], obs[s])
        W2 += w
    return W2, out


def _trace_sum(v):
    """Python's builtin sum over an array (sequential), as the reference's
    misfit accumulations use it."""
    w = 0
    for x in v:
        w += x
    return w


def wasserstein_quad(sim, obs):
    """adjoint_functions.wasserstein_quad (:70-104), per trace: densities
    sim^2 and obs^2 (in the traces' own dtypes) shifted by 0.05 max(obs^2),
    normalized by their rectangle sums, CDFs cumsum(x dt); adjoint
    dist(t) - dist(1) with dist = interp1d(t, 2 cumsum(t - G^-1(F(t))) dt);
    misfit sum (F^-1(t) - G^-1(t))^2."""
    sim, obs = np.asarray(sim), np.asarray(obs)
    n, T = sim.shape
    t = np.linspace(0, 1, T)
    dt = t[1] - t[0]
    W2 = 0
    out = np.zeros((n, T))
    for s in range(n):
        g = obs[s, :] ** 2
        f = sim[s, :] ** 2
        c = np.max(g) * 0.05
        g += c
        f += c
        g /= np.sum(g * dt)
        f /= np.sum(f * dt)
        G = np.cumsum(g * dt)
        F = np.cumsum(f * dt)
        q = _interp_linear(G, t, _interp_linear(t, F, t))
        D = 2 * np.cumsum(t - q) * dt
        out[s] = _interp_linear(t, D, t) - _interp_linear(t, D, np.array([1.0]))[0]
        W2 += _trace_sum(np.abs(_interp_linear(F, t, t) - _interp_linear(G, t, t)) ** 2)
    return W2, out


def wasserstein_experimental(sim, obs):
    """adjoint_functions.wasserstein_experimental (:107-141), per trace: the
    shift of wasserstein (from obs), rectangle-rule CDFs; adjoint
    Y(t) - Y(1) with Y = t ((t - q)^2 - 2 t q), q = G^-1(F(t)); misfit
    sum (F^-1(t) - G^-1(t))^2 F^-1(t)."""
    sim, obs = np.asarray(sim), np.asarray(obs)
    n, T = sim.shape
    t = np.linspace(0, 1, T)
    dt = t[1] - t[0]
    W2 = 0
    out = np.zeros((n, T))
    for s in range(n):
        f0 = np.concatenate([np.zeros(0), sim[s, :]])
        g0 = np.concatenate([np.zeros(0), obs[s, :]])
        c = np.maximum(-1.1 * np.min(g0), 0.1 * np.max(g0))
        g = g0 + c
        f = f0 + c
        g /= np.sum(g * dt)
        f /= np.sum(f * dt)
        G = np.cumsum(g * dt)
        F = np.cumsum(f * dt)
        q = _interp_linear(G, t, _interp_linear(t, F, t))
        Y = t * ((t - q) ** 2 - 2 * t * q)
        out[s] = _interp_linear(t, Y, t) - _interp_linear(t, Y, np.array([1.0]))[0]
        F1 = _interp_linear(F, t, t)
        W2 += _trace_sum(np.abs(F1 - _interp_linear(G, t, t)) ** 2 * F1)
    return W2, out


def wasserstein_pm(sim, obs):
    """adjoint_functions.wasserstein_pm (:144-223), per trace: positive parts
    max(0, x) (traces' own dtypes) shifted by 0.01 max(max(0, obs)) and
    normalized; the reference's negative-part CDFs are built from the
    positive parts (G_n = cumsum(g_p dt), F_n = cumsum(f_p dt), :180-183), so
    both halves equal the positive one: adjoint dist(t) - dist(1) with
    dist = cumsum(t - q) dt + cumsum(t - q) dt, misfit twice
    sum (F_p^-1(t) - G_p^-1(t))^2.  (The intermediate TT, DD and I of the
    reference do not reach its outputs.)"""
    sim, obs = np.asarray(sim), np.asarray(obs)
    n, T = sim.shape
    t = np.linspace(0, 1, T)
    dt = t[1] - t[0]
    W2 = 0
    out = np.zeros((n, T))
    for s in range(n):
        g_p = np.maximum(0, obs[s, :])
        f_p = np.maximum(0, sim[s, :])
        c_p = np.max(g_p) * 0.01
        g_p += c_p
        f_p += c_p
        g_p /= np.sum(g_p * dt)
        f_p /= np.sum(f_p * dt)
        G_p = np.cumsum(g_p * dt)
        F_p = np.cumsum(f_p * dt)
        q = _interp_linear(G_p, t, _interp_linear(t, F_p, t))
        D = np.cumsum(t - q) * dt + np.cumsum(t - q) * dt
        out[s] = _interp_linear(t, D, t) - _interp_linear(t, D, np.array([t[-1]]))[0]
        e = _trace_sum(np.abs(_interp_linear(F_p, t, t) - _interp_linear(G_p, t, t)) ** 2)
        W2 += e + e
    return W2, out


def wasserstein_yang(sim, obs):
    """adjoint_functions.wasserstein_yang (:226-266), per trace: shift
    c = -1.1 min(obs) (no floor), normalized densities g, f, rectangle-rule
    CDFs, q = G^-1(F(t)) (interp1d sorts a decreasing G stably), r = t - q,
    w = -2 f(t) / g(q); adjoint (U diag(w) + diag(r)) r with U the upper
    triangle of ones, i.e. s_k = sum_{j>k} w_j r_j + (w_k + r_k) r_k; misfit
    sum (t - G^-1(F))^2 f."""
    sim, obs = np.asarray(sim), np.asarray(obs)
    n, T = sim.shape
    t = np.linspace(0, 1, T)
    dt = t[1] - t[0]
    W2 = 0
    out = np.zeros((n, T))
    U = np.triu(np.ones((T, T)))
    for s in range(n):
        c = -1.1 * np.min(obs[s, :])
        g = obs[s, :] + c
        f = sim[s, :] + c
        g /= np.sum(g * dt)
        f /= np.sum(f * dt)
        G = np.cumsum(g * dt)
        F = np.cumsum(f * dt)
        q = _interp_linear(G, t, _interp_linear(t, F, t))
        r = t - q
        w = -2 * _interp_linear(t, f, t) / _interp_linear(t, g, q)
        # U @ diag(w) has the entries U_jk w_k exactly (one nonzero product each)
        out[s] = (U * w[None, :] + np.diag(r)) @ r
        W2 += _trace_sum(np.abs(t - _interp_linear(G, t, F)) ** 2 * f)
    return W2, out


def hilbert(x):
    """Analytic signal along the last axis, as scipy.signal.hilbert forms it
    (scipy 1.15, signal/_signaltools.py): ifft(fft(x) * h) with h = 1 at k = 0
    (and k = N/2 for even N), 2 for 0 < k < N/2, 0 above, h in the spectrum's
    dtype.  numpy's FFT, like scipy.fft, keeps float32 input in single
    precision (complex64)."""
    x = np.asarray(x)
    N = x.shape[-1]
    Xf = np.fft.fft(x, axis=-1)
    h = np.zeros(N, dtype=Xf.dtype)
    if N % 2 == 0:
        h[0] = h[N // 2] = 1
        h[1:N // 2] = 2
    else:
        h[0] = 1
        h[1:(N + 1) // 2] = 2
    return np.fft.ifft(Xf * h, axis=-1)


def quad_envelope(sim, obs):
    """adjoint_functions.quad_envelope (:14-22): E = |hilbert(sim)|^2 -
    |hilbert(obs)|^2 per sample, misfit sum E^2 over the whole (n_rx, T)
    array, adjoint source sim E - Im(hilbert(E Im(hilbert(sim)))).  The
    operands keep the dtypes the reference gives them (sim float64 from the
    callback, obs as stored)."""
    hilb_obs = hilbert(obs)
    hilb_sim = hilbert(sim)
    temp = np.abs(hilb_sim) ** 2 - np.abs(hilb_obs) ** 2
    adj = sim * temp - np.imag(hilbert(temp * np.imag(hilb_sim)))
    return np.sum(temp ** 2), adj


def illumination(cquad, d_x, d_z, src_xz, src, T, shots, stride, precision="ref"):
    """Source-side illumination of the build's plain inversion step (no
    reference counterpart): sum over shots and over the steps t = 1..T-1 with
    t % stride == 0 of stride * P(t)^2, where P(t) = Px + Pz after the source
    injection of step t (the field the reference records,
    regression_PML_single.cu:295-309).  The fields come from forward() with a
    receiver on every cell.  FP64 sum."""
    Z, X = cquad.shape
    zz, xx = np.mgrid[0:Z, 0:X]
    every = (xx.ravel().astype(np.int32), zz.ravel().astype(np.int32))
    ill = np.zeros(Z * X)
    ts = [t for t in range(1, T) if t % stride == 0]
    for idx in shots:
        rec = forward(cquad, d_x, d_z, src_xz, src, every, T, idx=idx,
                      precision=precision)["rec"].astype(np.float64)
        ill += stride * (rec[:, ts] ** 2).sum(axis=1)
    return ill.reshape(Z, X)


# --------------------------------------------------------------------------
# whole gradient, as Cuda_interface_PML.calc_grad drives it
# --------------------------------------------------------------------------

def gradient(cquad, d_x, d_z, src_xz, src, rx_xz, ring_xz, obs, T, misfit,
             shots=(-1,), constvec=(1.0, 1.0, 1.0), precision="ref"):
    """Per shot: forward with ring save, host misfit + adjoint source (FP64,
    rounded to the state precision as adj_wrapper does,
    cuda_interface_pml.py:238-249), reverse/adjoint run.  Returns
    (grad mean over shots, mse mean over shots, per-shot records)."""
    dt = np.float32 if precision == "ref" else np.float64
    grads, mses, recs = [], [], []
    for k, idx in enumerate(shots):
        fw = forward(cquad, d_x, d_z, src_xz, src, rx_xz, T, idx=idx, ring_xz=ring_xz,
                     constvec=constvec, precision=precision, want_seed=True)
        o = obs[k] if obs.ndim == 3 else obs
        mse, adj = misfit(fw["rec"].astype(np.float64), o)
        g = adjoint(cquad, d_x, d_z, rx_xz, adj.astype(dt), ring_xz, fw["ring"], fw["seed"],
                    constvec=constvec, precision=precision)
        grads.append(g.astype(np.float64))
        mses.append(float(mse))
        recs.append(fw["rec"])
    return np.mean(grads, axis=0), float(np.mean(mses)), np.array(recs)


def cpml_forward(p, idx=-1):
    """One run of the second-order C-PML scheme (regression_CPML.cu:427-501,
    oracle_cpml.c) for the problem dict p (keys as cpml_problem in
    tests/problems.py builds them); returns rec (n_rx, T) float32."""
    L = lib()
    f = L.orc_cpml_forward
    f.restype = None
    X, Z, T = p["X"], p["Z"], p["T"]
    i32 = lambda a: np.ascontiguousarray(a, np.int32)
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    sx, sz, rx, rz = i32(p["sx"]), i32(p["sz"]), i32(p["rx"]), i32(p["rz"])
    arrs = [f32(p[k]) for k in ("cquad", "kappa", "rho_hx", "rho_hz", "K_x", "a_x", "b_x",
                                "K_xh", "a_xh", "b_xh", "K_z", "a_z", "b_z", "K_zh", "a_zh",
                                "b_zh", "src", "initial")]
    rec = np.zeros((len(rx), T), np.float32)
    c = ctypes
    f(c.c_int(X), c.c_int(Z), c.c_int(T), c.c_int(len(sx)), c.c_int(len(rx)), c.c_int(p["ppe"]),
      _p(sx), _p(sz), _p(rx), _p(rz), c.c_float(p["dx"]), c.c_float(p["dz"]), c.c_float(p["dt2"]),
      c.c_float(p["qt_pi"]), *[_p(a) for a in arrs], c.c_int(idx), _p(rec))
    return rec
