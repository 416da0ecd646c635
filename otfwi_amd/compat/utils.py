"""Array, pulse and noise helpers of the reference's utils.py (utils.py:7-29),
with the plotting helpers left out (they need matplotlib and a display)."""
import math

import numpy as np

__all__ = ["add_array", "pulse", "add_noise"]


def add_array(xzm, center, n_elements, pitch, dx, dz, angle=0, elem_size=1):
    """Mark a linear array of n_elements at `pitch` (m) centred at center =
    (z, x) (m) in the boolean mask xzm; returns the (2, n) index array."""
    idx = np.zeros((2, n_elements), dtype=int)
    idx[0, :] = center[0] / dx
    length = (n_elements - 1) * pitch / dx + elem_size
    idx[1, :] = np.round(center[1] / dx + np.arange(-length / 2, length / 2, pitch / dx)).astype(int)
    xzm[idx[0, :], idx[1, :]] = True
    return idx


def pulse(t, f0):
    """Gaussian-windowed cosine of 90 % bandwidth centred at 2/f0."""
    bw = 0.9 * f0
    t0 = 2 / f0
    alpha = -(math.pi * bw / 2) ** 2 / math.log(math.sqrt(2) / 2)
    return np.exp(-alpha * (t - t0) ** 2) * np.cos(2 * np.pi * f0 * (t - t0))


def add_noise(u, db):
    """u plus white Gaussian noise at -db dB of its mean power."""
    s = np.sqrt(np.mean(u ** 2) * 10 ** (-db / 10))
    return u + s * np.random.randn(*u.shape)
