"""Synthetic workloads: normalized velocity models, Ricker sources and
transducer arrays for the benchmark configurations.

Units follow the reference's convention (cuda_interface_pml.py:58, 99;
setups/setup1_OCNDT.py:22-28): the propagator runs with dx = dz = dt = 1 and
``c`` is the Courant number v*dt/dx.  At CFL 0.5 with v_max = 5900 m/s,
dt = 0.5*dx/5900, so c(5900) = 0.5.

A source or receiver "on the top row" is placed on row ``pml`` (the first
undamped row): rows z < 4 are never updated by the scheme and the PML rows
attenuate, so a transducer there would be dead or damped.
"""
from __future__ import annotations

import math

import numpy as np

V_BG = 5900.0          # background (steel) velocity, m/s
V_MAX = 5900.0
CFL = 0.5


def courant(v, v_max=V_MAX, cfl=CFL):
    """Normalized velocity c = v*dt/dx at dt = cfl*dx/v_max."""
    return np.asarray(v, dtype=np.float64) * cfl / v_max


def ricker(f0, dt, T, t0=None):
    """Ricker wavelet (1 - 2 (pi f (t-t0))^2) exp(-(pi f (t-t0))^2) sampled at
    t = k*dt, k = 0..T-1; t0 defaults to 1.2/f0."""
    if t0 is None:
        t0 = 1.2 / f0
    t = np.arange(T) * dt - t0
    a = (math.pi * f0 * t) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def linear_array(N, n_elem, z, pitch=None, x_first=None, x_last=None):
    """Boolean (N, N) mask with n_elem points on row z.  Either a pitch (in
    cells, array centred) or the end columns are given."""
    mask = np.zeros((N, N), dtype=bool)
    if pitch is not None:
        span = (n_elem - 1) * pitch
        x0 = (N - 1 - span) // 2
        xs = x0 + pitch * np.arange(n_elem)
    else:
        xs = np.round(np.linspace(x_first, x_last, n_elem)).astype(int)
    mask[z, xs] = True
    return mask, xs


def draw_defects(c, rng, pml, n_min=1, n_max=5, semi_min=6.0, semi_max=60.0,
                 v_lo=1480.0, v_hi=3200.0, z_frac=0.2):
    """Paint 1..5 circles/ellipses (probability 1/2 each) into the normalized
    model ``c`` (modified in place); semi-axes uniform in cells, velocity
    uniform in [v_lo, v_hi] m/s, centres uniform in the lower 80 %."""
    Z, X = c.shape
    zz, xx = np.mgrid[0:Z, 0:X]
    k = int(rng.integers(n_min, n_max + 1))
    for _ in range(k):
        a = rng.uniform(semi_min, semi_max)
        b = a if rng.random() < 0.5 else rng.uniform(semi_min, semi_max)
        v = rng.uniform(v_lo, v_hi)
        cz = rng.uniform(z_frac * Z, Z - pml)
        cx = rng.uniform(pml, X - pml)
        inside = ((xx - cx) / a) ** 2 + ((zz - cz) / b) ** 2 <= 1.0
        c[inside] = courant(v)
    return c


def config1(seed=0):
    """Plain single shot: 256x256, dx = 0.1 mm, background 5900 m/s with one
    1480 m/s circular inclusion of radius 1.5 mm (15 cells) at a seeded
    position, 20-cell PML, 5 MHz Ricker at the top centre, 32 receivers along
    the top row, CFL 0.5, 1500 steps."""
    N, T, pml = 256, 1500, 20
    dx = 1e-4
    dt = CFL * dx / V_MAX
    rng = np.random.default_rng(seed)
    c_bg = np.full((N, N), float(courant(V_BG)))
    c_true = c_bg.copy()
    r = 15
    cz = rng.uniform(pml + r + 40, N - pml - r - 5)
    cx = rng.uniform(pml + r + 5, N - pml - r - 5)
    zz, xx = np.mgrid[0:N, 0:N]
    c_true[(xx - cx) ** 2 + (zz - cz) ** 2 <= r * r] = float(courant(1480.0))
    src_mask = np.zeros((N, N), dtype=bool)
    src_mask[pml, N // 2] = True
    rx_mask, _ = linear_array(N, 32, pml, x_first=pml, x_last=N - pml - 1)
    src = ricker(5e6, dt, T)[None, :]
    return dict(X=N, Z=N, T=T, pml=pml, R=1e-3, c_true=c_true, c_bg=c_bg,
                src_mask=src_mask, rx_mask=rx_mask, src=src, dx=dx, dt=dt)


def config2(model_seed=0, N=1024, T=4000, n_elem=64, pitch=12, pml=20):
    """Full-matrix capture: N x N, dx = 0.05 mm, 64-element array (pitch
    0.6 mm = 12 cells) on the top row, each element fires in turn and all
    receive; 1-5 random defects; CFL 0.5; 20-cell PML."""
    dx = 5e-5
    dt = CFL * dx / V_MAX
    rng = np.random.default_rng(model_seed)
    c_bg = np.full((N, N), float(courant(V_BG)))
    c_true = draw_defects(c_bg.copy(), rng, pml)
    mask, xs = linear_array(N, n_elem, pml, pitch=pitch)
    w = ricker(5e6, dt, T)
    src = np.tile(w, (n_elem, 1))
    return dict(X=N, Z=N, T=T, pml=pml, R=1e-3, c_true=c_true, c_bg=c_bg,
                src_mask=mask, rx_mask=mask.copy(), src=src, dx=dx, dt=dt,
                n_shots=n_elem)


C3_SIZES = (512, 768, 1024, 1536, 2048)


def config3_T(N, t_scale=1.0):
    """Steps of an N x N model of the uneven batch: 2500 at 512^2 to 6000 at
    2048^2, linear in the grid depth (times t_scale)."""
    return int(round((2500 + 3500 * (N - 512) / 1536) * t_scale))


def config3_pitch(N, n_elem=32, pml=20):
    """Array pitch (cells) of the uneven batch's 32-element arrays: half the
    spacing that would span the undamped width less 40 cells."""
    return max(2, (N - 2 * pml - 40) // (n_elem - 1) // 2)


def config3(batch_seed=0, n_models=16, t_scale=1.0, n_elem=32):
    """Uneven batch of independent inversions: n_models grid sizes drawn from
    {512, 768, 1024, 1536, 2048}^2 (seeded), each with dx = 0.05 mm, a
    32-element array on the top row (32 shots), config-2 defects (model seed
    1000 + index) and config3_T(N) steps.  Returns a list of per-model dicts
    in the form of config2."""
    rng = np.random.default_rng(batch_seed)
    sizes = rng.choice(C3_SIZES, size=n_models)
    return [config2(model_seed=1000 + i, N=int(N), T=config3_T(int(N), t_scale), n_elem=n_elem,
                    pitch=config3_pitch(int(N), n_elem))
            for i, N in enumerate(sizes)]


def config4(model_seed=0, N=2048, T=8000, n_elem=16, pitch=24, pml=20):
    """Differentiation with checkpointing: N x N at dx = 0.025 mm (CFL 0.5),
    a 16-element array at pitch 24 cells (0.6 mm) on the top row, config-2
    defects at this dx (semi-axes 0.3-3 mm = 12-120 cells), T = 8000."""
    dx = 2.5e-5
    dt = CFL * dx / V_MAX
    rng = np.random.default_rng(model_seed)
    c_bg = np.full((N, N), float(courant(V_BG)))
    c_true = draw_defects(c_bg.copy(), rng, pml, semi_min=12.0, semi_max=120.0)
    mask, xs = linear_array(N, n_elem, pml, pitch=pitch)
    w = ricker(5e6, dt, T)
    src = np.tile(w, (n_elem, 1))
    return dict(X=N, Z=N, T=T, pml=pml, R=1e-3, c_true=c_true, c_bg=c_bg,
                src_mask=mask, rx_mask=mask.copy(), src=src, dx=dx, dt=dt,
                n_shots=n_elem)


def learn_config(model_seed=0, N=256, T=1200, n_elem=32, pitch=6, pml=20, n_max=3,
                 semi_min=6.0, semi_max=30.0):
    """Full-matrix capture block for the inversion comparisons and the learned
    preconditioner's training pairs: the 1024^2 full-matrix-capture block
    scaled to N x N at the same dx = 0.05 mm and 5 MHz (so the same cells per
    wavelength), an n_elem array at pitch 0.3 mm on the top row, 1..n_max
    circles/ellipses with semi-axes 0.3-1.5 mm, c uniform in [1480, 3200]
    m/s in the lower 80 %, T long enough for the two-way travel to the
    bottom."""
    dx = 5e-5
    dt = CFL * dx / V_MAX
    rng = np.random.default_rng(model_seed)
    c_bg = np.full((N, N), float(courant(V_BG)))
    c_true = draw_defects(c_bg.copy(), rng, pml, n_max=n_max, semi_min=semi_min,
                          semi_max=semi_max)
    mask, xs = linear_array(N, n_elem, pml, pitch=pitch)
    w = ricker(5e6, dt, T)
    src = np.tile(w, (n_elem, 1))
    return dict(X=N, Z=N, T=T, pml=pml, R=1e-3, c_true=c_true, c_bg=c_bg,
                src_mask=mask, rx_mask=mask.copy(), src=src, dx=dx, dt=dt,
                n_shots=n_elem)


def cpml_profiles(n, L, dt, d0, alpha_max, kmax=1.0, half=False):
    """C-PML coefficients along one axis of n cells with an L-cell layer at
    each end: d = d0 q^2, K = 1 + (kmax - 1) q^2, alpha = alpha_max (1 - q)
    inside the layer (q: depth into it, 0..1, at the cell or, with half, at
    the half point i + 1/2), b = exp(-(d/K + alpha) dt),
    a = d (b - 1) / (K (d + K alpha)).  Returns (K, a, b) float32."""
    pos = np.arange(n, dtype=np.float64) + (0.5 if half else 0.0)
    depth = np.maximum(np.maximum(L - pos, pos - (n - 1 - L)), 0.0) / L
    d = d0 * depth ** 2
    K = 1.0 + (kmax - 1.0) * depth ** 2
    alpha = alpha_max * (1.0 - depth) * (depth > 0)
    b = np.exp(-(d / K + alpha) * dt)
    a = np.where(d > 1e-9, d * (b - 1.0) / (K * (d + K * alpha) + 1e-300), 0.0)
    return K.astype(np.float32), a.astype(np.float32), b.astype(np.float32)


def cpml_config(X, Z, T, n_src, n_rx, L=20, c0=0.4, f0=0.05, seed=0, ppe=1):
    """Second-order C-PML setup (the inputs of regression_CPML.cu's
    init_memory_sim) in normalized units (dx = dz = dt = 1): velocity c0 with
    a slower blob, density 1 with a denser blob (half-point densities as
    arithmetic means), sources and receivers on row L + 2, Ricker wavelets of
    f0 cycles per sample, reflection 1e-3 over the layer."""
    rng = np.random.default_rng(seed)
    zz, xx = np.mgrid[0:Z, 0:X]
    c = np.full((Z, X), c0)
    cz, cx = rng.uniform(0.5 * Z, 0.7 * Z), rng.uniform(0.35 * X, 0.65 * X)
    c[(xx - cx) ** 2 + (zz - cz) ** 2 <= (0.1 * min(X, Z)) ** 2] = 0.7 * c0
    rho = np.ones((Z, X))
    rho[(xx - 0.5 * X) ** 2 + (zz - 0.4 * Z) ** 2 <= (0.08 * min(X, Z)) ** 2] = 1.5
    rho_hx = rho.copy()
    rho_hx[:, :-1] = 0.5 * (rho[:, :-1] + rho[:, 1:])
    rho_hz = rho.copy()
    rho_hz[:-1, :] = 0.5 * (rho[:-1, :] + rho[1:, :])
    d0 = 3.0 * c0 * np.log(1e3) / (2.0 * L)
    prof = {}
    for ax, n in (("x", X), ("z", Z)):
        for half in (False, True):
            K, a, b = cpml_profiles(n, L, 1.0, d0, np.pi * f0, 1.0, half)
            h = "h" if half else ""
            prof[f"K_{ax}{h}"], prof[f"a_{ax}{h}"], prof[f"b_{ax}{h}"] = K, a, b
    sx = np.round(np.linspace(L + 6, X - L - 7, n_src)).astype(np.int32)
    rx = np.round(np.linspace(L + 3, X - L - 4, n_rx)).astype(np.int32)
    w = ricker(f0, 1.0, T, t0=1.2 / f0)
    src = np.tile(w, (n_src, 1)) * (1.0 + 0.1 * np.arange(n_src))[:, None]
    return dict(X=X, Z=Z, T=T, ppe=ppe, dx=1.0, dz=1.0, dt=1.0, dt2=1.0, qt_pi=4 * np.pi,
                sx=sx, sz=np.full(n_src, L + 2, np.int32), rx=rx,
                rz=np.full(n_rx, L + 2, np.int32), cquad=(c ** 2).astype(np.float32),
                kappa=(rho * c ** 2).astype(np.float32), rho=rho.astype(np.float32),
                rho_hx=rho_hx.astype(np.float32), rho_hz=rho_hz.astype(np.float32),
                src=src.astype(np.float32), initial=np.zeros((2, Z, X), np.float32), **prof)
