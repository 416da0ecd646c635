"""Second-order C-PML scheme on the GPU (libotfwi_cpml.so) against the C
oracle of regression_CPML.cu: traces within relative L2 1e-5 (FP32 state,
derivatives summed in double as the reference does; the two differ in
multiply-add contraction only), batching, source selection and the
reference's own C ABI."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from otfwi_amd import cpml
from tests.problems import cpml_problem, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("X,Z,T,shots", [(96, 80, 300, [0, 2, -1]), (97, 61, 240, [1]),
                                         (200, 150, 500, [2, 0])])
def test_traces_match_oracle(gpu, X, Z, T, shots):
    p = cpml_problem(X=X, Z=Z, T=T, seed=X + Z)
    if X == 97:   # nonzero initial state
        rng = np.random.default_rng(1)
        p["initial"] = (1e-2 * rng.standard_normal((2, Z, X))).astype(np.float32)
    m = cpml.CPMLModel.from_problem(p)
    rec = m.forward(shots)
    for k, s in enumerate(shots):
        ref = O.cpml_forward(p, idx=s)
        e = rel(rec[k], ref)
        print(f"{X}x{Z} T={T} shot {s}: rel {e:.2e}")
        assert np.abs(ref).max() > 1e-3 and e < 1e-5


def test_batch_equals_single_and_ppe(gpu):
    p = cpml_problem(T=260, seed=3)
    m = cpml.CPMLModel.from_problem(p)
    both = m.forward([2, 0, 1])
    for k, s in enumerate([2, 0, 1]):
        assert np.array_equal(both[k], m.forward([s])[0])
    # pixel_per_element != 1: every source fires whatever the index
    q = dict(p, ppe=3)
    mq = cpml.CPMLModel.from_problem(q)
    assert np.array_equal(mq.forward([1])[0], m.forward([-1])[0])


def test_long_run_stable(gpu):
    p = cpml_problem(T=1500, seed=6)
    rec = cpml.CPMLModel.from_problem(p).forward([-1])[0]
    ref = O.cpml_forward(p)
    assert np.isfinite(rec).all() and rel(rec, ref) < 1e-5
    assert np.abs(rec[:, -300:]).max() < 0.05 * np.abs(rec).max()


def test_reference_abi(gpu):
    """init_memory_sim (43 arguments) + cuda_simulate2 + setCquad + set_source
    as regression_CPML.cu exports them."""
    p = cpml_problem(T=220, seed=8)
    X, Z, T = p["X"], p["Z"], p["T"]
    lib = cpml.load()
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    i32 = lambda a: np.ascontiguousarray(a, np.int32)
    keep = []

    def ptr(a):
        keep.append(a)
        return a.ctypes.data_as(ctypes.c_void_p)

    zx, zz = np.zeros(X, np.float32), np.zeros(Z, np.float32)   # d, alpha: unused
    prof = []
    for ax, z0 in (("x", zx), ("z", zz)):
        for h in ("", "h"):
            prof += [z0, p[f"K_{ax}{h}"], z0, p[f"a_{ax}{h}"], p[f"b_{ax}{h}"]]
    rec = ctypes.POINTER(ctypes.c_float)()
    ns, nm = len(p["sx"]), len(p["rx"])
    lib.init_memory_sim(ctypes.c_int(X), ctypes.c_int(Z), ctypes.c_int(T), ctypes.c_int(ns),
                        ctypes.c_int(nm), ctypes.c_int(1), ptr(i32(p["sx"])), ptr(i32(p["sz"])),
                        ptr(i32(p["rx"])), ptr(i32(p["rz"])), ctypes.c_float(1.0),
                        ctypes.c_float(1.0), ctypes.c_float(1.0), ctypes.c_float(1.0),
                        ctypes.c_float(p["qt_pi"]), ptr(f32(p["cquad"])), ptr(f32(p["kappa"])),
                        ptr(f32(p["rho"])), ptr(f32(p["rho_hx"])), ptr(f32(p["rho_hz"])),
                        *[ptr(f32(a)) for a in prof], ptr(f32(p["src"])),
                        ptr(f32(p["initial"])), ctypes.byref(rec))
    assert rec
    lib.cuda_simulate2(0, 1)
    out = np.ctypeslib.as_array(rec, shape=(nm * T,)).reshape(nm, T).copy()
    assert rel(out, O.cpml_forward(p, idx=1)) < 1e-5
    assert np.array_equal(out, cpml.CPMLModel.from_problem(p).forward([1])[0])
    cq2 = f32(0.9 * p["cquad"])
    lib.setCquad(ptr(cq2))
    lib.cuda_simulate2(0, 1)
    out2 = np.ctypeslib.as_array(rec, shape=(nm * T,)).reshape(nm, T).copy()
    assert rel(out2, O.cpml_forward(dict(p, cquad=cq2), idx=1)) < 1e-5
    sx2 = i32(p["sx"] + 3)
    lib.set_source(ctypes.c_int(ns), ptr(sx2), ptr(i32(p["sz"])))
    lib.cuda_simulate2(0, 0)
    out3 = np.ctypeslib.as_array(rec, shape=(nm * T,)).reshape(nm, T).copy()
    assert rel(out3, O.cpml_forward(dict(p, cquad=cq2, sx=sx2), idx=0)) < 1e-5


def test_rejects_points_outside_grid(gpu):
    p = cpml_problem(T=50)
    with pytest.raises(Exception):
        cpml.CPMLModel.from_problem(dict(p, sx=np.array([p["X"] + 1], np.int32),
                                         sz=np.array([5], np.int32), src=p["src"][:1]))
