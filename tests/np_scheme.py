"""Independent, vectorized NumPy restatement of the reference scheme in FP64,
written from regression_PML_single.cu:260-846 and deriv_macros.h:84-136
separately from the C oracle, for brute-force cross-checks on tiny grids.

Only tests import this module.
"""
from __future__ import annotations

import numpy as np

C = (1.196289e+00, -7.975260e-02, 9.570313e-03, -6.975446e-04)
B = 4


def _interior(Z, X):
    m = np.zeros((Z, X), bool)
    m[B:Z - B, B:X - B] = True
    return m


def _shift(a, k, axis):
    """out[..., i, ...] = a[..., i + k, ...], zero where i + k is off the grid."""
    out = np.zeros_like(a)
    n = a.shape[axis]
    src = [slice(None)] * a.ndim
    dst = [slice(None)] * a.ndim
    if k >= 0:
        src[axis] = slice(k, n)
        dst[axis] = slice(0, n - k)
    else:
        src[axis] = slice(0, n + k)
        dst[axis] = slice(-k, n)
    out[tuple(dst)] = a[tuple(src)]
    return out


def d_minus(a, axis, inside):
    """sum_k c_k (a[i+k-1] - a[i-k]) with a read as 0 off the interior."""
    a = np.where(inside, a, 0.0)
    return sum(c * (_shift(a, k - 1, axis) - _shift(a, -k, axis)) for k, c in enumerate(C, 1))


def d_plus(a, axis, inside):
    """sum_k c_k (a[i+k] - a[i+1-k]) with a read as 0 off the interior."""
    a = np.where(inside, a, 0.0)
    return sum(c * (_shift(a, k, axis) - _shift(a, 1 - k, axis)) for k, c in enumerate(C, 1))


def lap8(u, inside):
    u = np.where(inside, u, 0.0)
    w = (-9.0, 128.0, -1008.0, 8064.0, -14350.0, 8064.0, -1008.0, 128.0, -9.0)
    lx = sum(wj * _shift(u, j - 4, 1) for j, wj in enumerate(w)) / 5040.0
    lz = sum(wj * _shift(u, j - 4, 0) for j, wj in enumerate(w)) / 5040.0
    return lx + lz


def _damp(d_x, d_z):
    """Shifted damping d_x[z][x+1], d_z[z+1][x] (off-grid reads unused)."""
    dx = np.zeros_like(d_x)
    dz = np.zeros_like(d_z)
    dx[:, :-1] = d_x[:, 1:]
    dz[:-1, :] = d_z[1:, :]
    return dx, dz


class State:
    def __init__(self, Z, X):
        self.Px = np.zeros((Z, X))
        self.Pz = np.zeros((Z, X))
        self.Ax = np.zeros((Z, X))
        self.Az = np.zeros((Z, X))


def p_step(s, m, dx, dz, inside, revert=False):
    dAx = d_minus(s.Ax, 1, inside)
    dAz = d_minus(s.Az, 0, inside)
    if revert:
        px, pz = s.Px - m * dAx, s.Pz - m * dAz
    else:
        px, pz = s.Px * (1 - dx) + m * dAx, s.Pz * (1 - dz) + m * dAz
    s.Px = np.where(inside, px, s.Px)
    s.Pz = np.where(inside, pz, s.Pz)


def a_step(s, dx, dz, inside, revert=False):
    gx = d_plus(s.Px, 1, inside) + d_plus(s.Pz, 1, inside)
    gz = d_plus(s.Px, 0, inside) + d_plus(s.Pz, 0, inside)
    if revert:
        ax, az = s.Ax - gx, s.Az - gz
    else:
        ax, az = s.Ax * (1 - dx) + gx, s.Az * (1 - dz) + gz
    s.Ax = np.where(inside, ax, s.Ax)
    s.Az = np.where(inside, az, s.Az)


def forward(cq, d_x, d_z, src_xz, src, rx_xz, T, idx=-1, ring_xz=None):
    """Unit dx = dt = 1.  Returns (rec, ring, seed) like oracle.forward."""
    Z, X = cq.shape
    inside = _interior(Z, X)
    dx, dz = _damp(np.asarray(d_x, float), np.asarray(d_z, float))
    m = np.asarray(cq, float)
    s = State(Z, X)
    sx, sz = src_xz
    rx, rz = rx_xz
    integ = np.zeros(len(sx))
    rec = np.zeros((len(rx), T))
    ring = np.zeros((len(ring_xz[0]), T)) if ring_xz is not None else None
    prev_A = None
    for t in range(1, T):
        prev_A = (s.Ax.copy(), s.Az.copy())
        p_step(s, m, dx, dz, inside)
        for n in range(len(sx)):
            if idx != -1 and idx != n:
                continue
            integ[n] += src[n][t]
            s.Px[sz[n], sx[n]] += integ[n] / 2
            s.Pz[sz[n], sx[n]] += integ[n] / 2
        if ring is not None:
            ring[:, t] = (s.Px + s.Pz)[ring_xz[1], ring_xz[0]]
        rec[:, t] = (s.Px + s.Pz)[rz, rx]
        a_step(s, dx, dz, inside)
    seed = np.stack([s.Px, s.Pz, prev_A[0], prev_A[1]])
    return rec, ring, seed


def adjoint(cq, d_x, d_z, rx_xz, adj, ring_xz, ring, seed):
    Z, X = cq.shape
    T = adj.shape[1]
    inside = _interior(Z, X)
    dx, dz = _damp(np.asarray(d_x, float), np.asarray(d_z, float))
    m = np.asarray(cq, float)
    lam = State(Z, X)
    ub = State(Z, X)
    ub.Px, ub.Pz, ub.Ax, ub.Az = (np.array(a, float) for a in seed)
    rx, rz = rx_xz
    vx, vz = ring_xz
    J = np.zeros(len(rx))
    grad = np.zeros((Z, X))
    for t in range(T):
        if t > 0:
            p_step(lam, m, dx, dz, inside)
            p_step(ub, m, dx, dz, inside, revert=True)
            for n in range(len(rx)):
                J[n] += adj[n][T - 1 - t]
                lam.Px[rz[n], rx[n]] += J[n] / 2
                lam.Pz[rz[n], rx[n]] += J[n] / 2
            ub.Px[vz, vx] = ring[:, T - 1 - t] / 2
            ub.Pz[vz, vx] = ring[:, T - 1 - t] / 2
            a_step(lam, dx, dz, inside)
            a_step(ub, dx, dz, inside, revert=True)
        grad += (lam.Px + lam.Pz) * lap8(ub.Px + ub.Pz, inside)
    return grad


def exact_gradient(cq, d_x, d_z, src_xz, src, rx_xz, obs, T, idx=-1, rho_fn=None):
    """Exact discrete adjoint of forward() (unit spacing): dJ/d(cq) for
    J = misfit(rec), rho_fn(rec) -> (J, dJ/drec); default L2 sum (rec-obs)^2.
    Stores every A field (tiny grids only)."""
    Z, X = cq.shape
    inside = _interior(Z, X)
    dx, dz = _damp(np.asarray(d_x, float), np.asarray(d_z, float))
    m = np.asarray(cq, float)
    s = State(Z, X)
    sx, sz = src_xz
    rx, rz = rx_xz
    integ = np.zeros(len(sx))
    rec = np.zeros((len(rx), T))
    Ax = [np.zeros((Z, X))]
    Az = [np.zeros((Z, X))]
    for t in range(1, T):
        p_step(s, m, dx, dz, inside)
        for n in range(len(sx)):
            if idx != -1 and idx != n:
                continue
            integ[n] += src[n][t]
            s.Px[sz[n], sx[n]] += integ[n] / 2
            s.Pz[sz[n], sx[n]] += integ[n] / 2
        rec[:, t] = (s.Px + s.Pz)[rz, rx]
        a_step(s, dx, dz, inside)
        Ax.append(s.Ax.copy())
        Az.append(s.Az.copy())
    if rho_fn is None:
        J = float(((rec - obs) ** 2).sum())
        rho = 2.0 * (rec - obs)
    else:
        J, rho = rho_fn(rec)
    u = np.zeros((Z, X))
    w = np.zeros((Z, X))
    a = np.zeros((Z, X))
    b = np.zeros((Z, X))
    g = np.zeros((Z, X))
    kx, kz = 1.0 - dx, 1.0 - dz
    rin = inside[rz, rx]
    for t in range(T - 1, 0, -1):
        dA = d_minus(a, 1, inside) + d_minus(b, 0, inside)
        u = np.where(inside, kx * u - dA, 0.0)
        w = np.where(inside, kz * w - dA, 0.0)
        u[rz[rin], rx[rin]] += rho[rin, t]
        w[rz[rin], rx[rin]] += rho[rin, t]
        g += np.where(inside, u * d_minus(Ax[t - 1], 1, inside) + w * d_minus(Az[t - 1], 0, inside),
                      0.0)
        a = np.where(inside, kx * a - d_plus(m * u, 1, inside), 0.0)
        b = np.where(inside, kz * b - d_plus(m * w, 0, inside), 0.0)
    return J, g, rec
