"""Shape rasterizers of the reference's geometry.py (geometry.py:8-48), without
scikit-image.

`star_idx` builds the same 10-vertex star (geometry.py:8-24) and rasterizes it
the way skimage.draw.polygon does: the pixels (r, c), 0 <= r < shape[0],
0 <= c < shape[1], whose integer coordinates fall inside the polygon by the
even-odd rule, scanned row-major.  `draw_rectangle` and `draw_circle` keep the
reference's index arithmetic (geometry.py:27-48).
"""
import numpy as np

__all__ = ["star_idx", "draw_rectangle", "draw_circle", "polygon"]


def polygon(r, c, shape=None):
    """(rr, cc) of the integer points inside the polygon with vertex rows r
    and columns c (even-odd rule), limited to `shape` when given."""
    r = np.asarray(r, np.float64)
    c = np.asarray(c, np.float64)
    r0, r1 = int(np.floor(r.min())), int(np.ceil(r.max()))
    c0, c1 = int(np.floor(c.min())), int(np.ceil(c.max()))
    if shape is not None:
        r0, c0 = max(r0, 0), max(c0, 0)
        r1, c1 = min(r1, shape[0] - 1), min(c1, shape[1] - 1)
    if r1 < r0 or c1 < c0:
        return np.zeros(0, np.intp), np.zeros(0, np.intp)
    rr, cc = np.mgrid[r0:r1 + 1, c0:c1 + 1]
    inside = np.zeros(rr.shape, bool)
    n = len(r)
    for i in range(n):
        ya, xa = r[i], c[i]
        yb, xb = r[i - 1], c[i - 1]
        crosses = (ya > rr) != (yb > rr)
        with np.errstate(divide="ignore", invalid="ignore"):
            xint = xa + (rr - ya) * (xb - xa) / (yb - ya)
        inside ^= crosses & (cc < xint)
    return rr[inside].astype(np.intp), cc[inside].astype(np.intp)


def star_idx(shape, altura, centro):
    """Five-pointed star of height `altura` centred at `centro` (geometry.py:8-24)."""
    external = np.deg2rad(np.arange(90, 90 - 360, -360 / 5) - 36)
    internal = np.deg2rad(np.arange(90, 90 - 360, -360 / 5))
    xx = np.zeros(10)
    zz = np.zeros(10)
    centro = np.asarray(centro)
    inner = np.sin(np.deg2rad(22)) * altura
    for i in range(5):
        xx[2 * i] = inner * np.sin(internal[i]) + centro[0]
        zz[2 * i] = inner * np.cos(internal[i]) + centro[1]
        xx[2 * i + 1] = altura * np.sin(external[i]) + centro[0]
        zz[2 * i + 1] = altura * np.cos(external[i]) + centro[1]
    return polygon(xx, zz, shape)


def draw_rectangle(target, altura, largura, centro=None, value=1):
    """Fill target[z0:z1, x0:x1] = value for the box of height `altura` and
    width `largura` centred at centro = (x, z) (geometry.py:27-38)."""
    shape = target.shape
    if centro is None:
        centro = [int(shape[0] / 2), int(shape[1] / 2)]
    x0 = int(centro[0] - largura / 2)
    z0 = int(centro[1] - altura / 2)
    x1 = int(centro[0] + largura / 2)
    z1 = int(centro[1] + altura / 2)
    target[z0:z1, x0:x1] = value


def draw_circle(target, diametro, centro=None, value=1):
    """Fill the disc of diameter `diametro` centred at centro = (x, z)
    (geometry.py:41-48, including its use of shape[1] / shape[0] for the
    z / x ranges)."""
    shape = target.shape
    if centro is None:
        centro = [int(shape[0] / 2), int(shape[1] / 2)]
    r = int(round(diametro // 2))
    x = int(round(centro[0]))
    z = int(round(centro[1]))
    v_z, v_x = np.ogrid[-z:shape[1] - z, -x:shape[0] - x]
    target[v_x ** 2 + v_z ** 2 <= r ** 2] = value
