"""Host-side logic of the drop-in wrapper (no GPU): PML profile, point
ordering, boundary ring, and the C-ABI export set."""
import ctypes
import os
import re

import numpy as np

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pml_profile_closed_form():
    # cuda_interface_pml.py:116-132 with R = 1e-5, Vmax = 0.5, L = 20:
    # d0 = 3*0.5*ln(1e5)/(2*20^3); edge value d0*L^2 = 0.4317...
    d_x, d_z = O.calc_pml(64, 48, 20)
    edge = 3 * 0.5 * np.log(1e5) / (2 * 20)
    assert abs(d_x[0, 0] - edge) < 1e-12 and abs(edge - 0.4317) < 1e-4
    assert d_x[5, 0] == d_x[0, 0] and d_x[0, 19] == 0 and d_x[0, 63] == d_x[0, 0]
    assert d_z[0, 7] == d_z[0, 0] and abs(d_z[0, 0] - edge) < 1e-12
    assert d_z[19, 0] == 0 and d_z[47, 3] == d_z[0, 0]
    # the profile is d0 * x^2 on linspace(0, L, L)
    x = np.linspace(0, 20, 20)
    assert np.allclose(d_x[0, 44:], 3 * 0.5 * np.log(1e5) / (2 * 20 ** 3) * x ** 2, rtol=0, atol=0)


def test_wrapper_profile_equals_oracle():
    from otfwi_amd.cuda_interface_pml import Cuda_interface_PML

    class Probe(Cuda_interface_PML):
        def __init__(self, X, Z, L, R):   # host part only, no library call
            self.X, self.Z, self.PML_size, self.R, self.Vmax = X, Z, L, R, 0.5

    for R in (1e-5, 1e-3):
        a = Probe(70, 50, 12, R)._calcPML()
        b = O.calc_pml(70, 50, 12, R=R)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_ring_size_config1():
    # 256^2, PML 20, prec_deriv 4 (margin 5): 4 bands minus corners = 4220 points
    src = np.zeros((256, 256), bool)
    src[20, 128] = True
    mask = O.revert_mask(256, 256, 20, 4, src)
    assert mask.sum() == 4220          # the source cell lies in the top band
    x, z = O.point_order(mask)
    assert np.all(np.diff(z * 256 + x) > 0)   # np.where row-major order


def test_point_order_row_major():
    m = np.zeros((5, 6), bool)
    m[3, 1] = m[1, 4] = m[1, 2] = True
    x, z = O.point_order(m)
    assert list(zip(z, x)) == [(1, 2), (1, 4), (3, 1)]


def header_functions(header="otfwi_pml.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = "\n".join(line for line in txt.split("\n") if not line.strip().startswith("typedef"))
    names = re.findall(r"^\s*(?:const\s+)?(?:long\s+long|unsigned|[a-z_]+)\s*\*?\s*"
                       r"([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt, re.M)
    return sorted(set(n for n in names if n not in ("typedef",)))


def test_library_exports_every_declared_symbol():
    from otfwi_amd import _lib

    names = header_functions()
    assert "otfwi_gradient" in names and "cuda_grad_ext" in names and len(names) >= 25
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTED) == names


def test_reference_abi_symbols_present():
    from otfwi_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in ("init_memory_sim", "init_memory_reg", "setCquad", "cuda_simulate", "cuda_grad_ext",
              "def_python_adj"):
        assert hasattr(lib, n)


def test_plain_direction_properties():
    from otfwi_amd import inversion

    rng = np.random.default_rng(0)
    Z, X = 40, 50
    g = rng.standard_normal((Z, X))
    ill = rng.uniform(0.5, 2.0, (Z, X))
    c = rng.uniform(0.2, 0.5, (Z, X))
    mask = inversion.array_mask(Z, X, 5, 12, taper=4)
    d = inversion.plain_direction(g, ill, c, mask)
    assert np.isclose(np.abs(d).max(), 1.0)
    assert np.all(d[mask == 0] == 0)
    inside = mask > 0
    # a descent direction: opposite sign to the gradient wherever it acts
    assert np.all(np.sign(d[inside]) == -np.sign(g[inside]))
    # positive rescaling of the gradient or of the illumination leaves it unchanged
    assert np.allclose(inversion.plain_direction(7.0 * g, ill, c, mask), d)
    assert np.allclose(inversion.plain_direction(g, 3.0 * ill, c, mask), d)
    # preconditioning: with a uniform gradient the step is largest where the
    # illumination is weakest
    d1 = inversion.plain_direction(np.ones((Z, X)), ill, np.ones((Z, X)), np.ones((Z, X)))
    k = np.unravel_index(np.argmax(np.abs(d1)), d1.shape)
    assert ill[k] == ill.min()


def test_array_mask_shape():
    from otfwi_amd import inversion

    m = inversion.array_mask(64, 64, 8, 20, taper=5)
    assert m[:20].max() == 0 and m[:, :8].max() == 0 and m[-8:].max() == 0
    assert m[30, 30] == 1.0
    assert np.all(np.diff(m[20:26, 30]) > 0)


def test_device_misfit_kinds():
    """The drop-in keeps L2, W2 and the squared-envelope misfit on the device
    (by marker or by name from an adjoint_functions module) and any other
    callable on the host callback path."""
    from otfwi_amd import _lib
    from otfwi_amd import adjoint_functions as AF
    from otfwi_amd.cuda_interface_pml import _device_kind
    from otfwi_amd.model import PMLModel  # noqa: F401  (import only)

    assert [_device_kind(f) for f in (AF.L2, AF.wasserstein, AF.quad_envelope)] == ["l2", "w2", "env"]
    assert _device_kind(lambda s, o: (0.0, s)) is None
    assert (_lib.MISFIT_L2, _lib.MISFIT_W2, _lib.MISFIT_ENV) == (0, 1, 2)


def test_cpml_library_exports_every_declared_symbol():
    """libotfwi_cpml.so loads and exports every function include/otfwi_cpml.h
    declares (no GPU call)."""
    from otfwi_amd import cpml

    names = header_functions("otfwi_cpml.h")
    assert "init_memory_sim" in names and "otfwi_cpml_forward" in names
    lib = ctypes.CDLL(cpml.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(cpml.EXPORTED) == names


def test_unet_library_exports_every_declared_symbol():
    """libotfwi_unet.so loads and exports every function include/otfwi_unet.h
    declares, and the ctypes table binds exactly those (no GPU call)."""
    from otfwi_amd import _lib

    names = header_functions("otfwi_unet.h")
    assert "otfwi_unet_conv3x3" in names and "otfwi_unet_set_conv_tiles" in names
    lib = ctypes.CDLL(_lib.UNET_LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.UNET_EXPORTED) == names


def test_fmc_array_mirror_axis():
    """The mirror augmentation of the learned preconditioner's training maps
    x -> xs.min() + xs.max() - x; on both full-matrix-capture geometries it
    maps the array's source and receiver cells onto each other, so a mirrored
    training pair is the pair of the mirrored model."""
    from otfwi_amd import configs, learned

    for kw in (learned.GEOMETRIES["fmc256"], learned.GEOMETRIES["fmc1024"]):
        p = configs.learn_config(0, **kw)
        src, rx = np.asarray(p["src_mask"]), np.asarray(p["rx_mask"])
        xs = np.where(src.any(0))[0]
        W = src.shape[1]
        cols = (xs.min() + xs.max() - np.arange(W)) % W
        assert np.array_equal(src[:, cols], src) and np.array_equal(rx[:, cols], rx)
        assert src.sum() == kw.get("n_elem", 32)
