"""skimage-free drop-ins for the reference's geometry.py / utils.py helpers
(otfwi_amd/compat), pinned by closed forms: the shoelace area of the star,
lattice-point counts of discs and boxes, the array pitch, the pulse peak."""
import os
import sys

import numpy as np

COMPAT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "otfwi_amd",
                      "compat")


def _mods():
    if COMPAT not in sys.path:
        sys.path.insert(0, COMPAT)
    import geometry
    import utils
    return geometry, utils


def test_polygon_square_and_triangle():
    geometry, _ = _mods()
    rr, cc = geometry.polygon([0.5, 0.5, 4.5, 4.5], [0.5, 4.5, 4.5, 0.5])
    assert len(rr) == 16 and set(rr) == {1, 2, 3, 4} and set(cc) == {1, 2, 3, 4}
    # row-major scan order
    assert list(zip(rr, cc))[:4] == [(1, 1), (1, 2), (1, 3), (1, 4)]
    rr, cc = geometry.polygon([0, 20, 0], [0, 0, 20], shape=(10, 30))
    assert rr.max() <= 9 and (rr >= 0).all()


def test_star_area_matches_shoelace():
    geometry, _ = _mods()
    shape = (400, 400)
    rr, cc = geometry.star_idx(shape, 120.0, (200.0, 210.0))
    ext = np.deg2rad(np.arange(90, 90 - 360, -360 / 5) - 36)
    inn = np.deg2rad(np.arange(90, 90 - 360, -360 / 5))
    xs, zs = [], []
    for i in range(5):
        a = np.sin(np.deg2rad(22)) * 120.0
        xs += [a * np.sin(inn[i]) + 200, 120 * np.sin(ext[i]) + 200]
        zs += [a * np.cos(inn[i]) + 210, 120 * np.cos(ext[i]) + 210]
    xs, zs = np.array(xs), np.array(zs)
    area = 0.5 * abs(np.dot(xs, np.roll(zs, 1)) - np.dot(zs, np.roll(xs, 1)))
    perim = np.hypot(np.diff(np.r_[xs, xs[0]]), np.diff(np.r_[zs, zs[0]])).sum()
    assert abs(len(rr) - area) < perim
    assert len(set(zip(rr, cc))) == len(rr)


def test_draw_rectangle_and_circle():
    geometry, _ = _mods()
    a = np.zeros((100, 120))
    geometry.draw_rectangle(a, 30, 50, centro=(60, 40), value=2)
    assert a.sum() == 2 * 30 * 50
    assert a[25:55, 35:85].min() == 2
    b = np.zeros((101, 101))
    geometry.draw_circle(b, 41, centro=(50, 50), value=1)
    r = 20
    zz, xx = np.mgrid[-50:51, -50:51]
    assert b.sum() == ((xx ** 2 + zz ** 2) <= r * r).sum()


def test_setup_style_calls():
    """The calls of the reference's setups (setup1_OCNDT.py:57-90) on these modules."""
    geometry, utils = _mods()
    Nx = Nz = 500
    gmask = np.full((Nz, Nx), 0).astype(int)
    geometry.draw_rectangle(gmask, 0.5 * Nz, 0.5 * Nx, centro=(Nx // 2, Nz // 2))
    assert gmask.sum() == 250 * 250
    params = np.ones((Nz, Nx))
    geometry.draw_rectangle(params, 0.3 * Nz, 0.3 * Nx, value=3.0)
    assert (params == 3.0).sum() == 150 * 150
    xzm = np.full((Nz, Nx), False)
    idx = utils.add_array(xzm, (20e-3, 75e-3), 16, 0.6e-3, 3e-4, 3e-4)
    assert xzm.sum() == 16 and np.all(np.diff(idx[1]) == 2)
    t = np.arange(0, 10e-6, 25e-9)
    s = utils.pulse(t, 1e6)
    assert abs(s.max() - 1.0) < 1e-3 and abs(t[np.argmax(s)] - 2e-6) < 25e-9
    u = np.sin(np.arange(1000) / 10.0)
    np.random.seed(0)
    n = utils.add_noise(u, 20) - u
    assert abs(np.mean(n ** 2) / np.mean(u ** 2) - 0.01) < 0.002
