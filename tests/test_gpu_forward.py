"""Forward propagator on the GPU against the CPU oracle.

Oracle: oracle/oracle_pml.c (restatement of regression_PML_single.cu:447-702).
Bars: FP32 traces within rel-L2 1e-5 of the reference-precision oracle; the
FP64 build within 1e-11 of the FP64 oracle (same scheme, same inputs).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.problems import config1_problem, make_model, rel, small_problem

pytestmark = pytest.mark.gpu


def oracle_rec(p, idx, precision="ref", cq="cq_true"):
    return O.forward(p[cq], p["d_x"], p["d_z"], p["src_xz"], p["src"], p["rx_xz"], p["T"],
                     idx=idx, precision=precision)["rec"]


def test_small_fp32_vs_oracle(gpu):
    p = small_problem(seed=1)
    m = make_model(p)
    rec = m.forward([0, 1, -1])
    for k, idx in enumerate([0, 1, -1]):
        ref = oracle_rec(p, idx)
        e = rel(rec[k], ref)
        print(f"shot {idx}: rel-L2 {e:.3e}")
        assert e < 1e-5
        assert np.all(rec[k][:, 0] == 0)


def test_small_fp64_vs_oracle_f64(gpu):
    p = small_problem(seed=2)
    m = make_model(p, precision=8)
    rec = m.forward([1])
    ref = oracle_rec(p, 1, precision="f64")
    e = rel(rec[0], ref)
    print(f"fp64 rel-L2 {e:.3e}")
    assert e < 1e-6   # traces come back rounded to FP32


@pytest.mark.parametrize("th", [16, 32, 64])
def test_tile_heights_bitwise(gpu, th):
    p = small_problem(X=150, Z=130, T=160, seed=4)
    base = make_model(p, th_fwd=32).forward([0, 1])
    other = make_model(p, th_fwd=th).forward([0, 1])
    assert np.array_equal(base, other)


def test_batched_equals_single(gpu):
    p = small_problem(seed=5, n_src=3)
    m = make_model(p)
    both = m.forward([2, 0, 1])
    for k, idx in enumerate([2, 0, 1]):
        assert np.array_equal(both[k], m.forward([idx])[0])
    m2 = make_model(p, max_batch=2)
    assert np.array_equal(m2.forward([2, 0, 1]), both)


def test_config1_vs_oracle(gpu):
    p = config1_problem(seed=0)
    m = make_model(p)
    rec = m.forward([-1])[0]
    ref = oracle_rec(p, -1)
    e = rel(rec, ref)
    f64 = oracle_rec(p, -1, precision="f64")
    print(f"config1: vs reference-precision oracle {e:.3e}; vs FP64 oracle {rel(rec, f64):.3e}; "
          f"oracle ref vs FP64 {rel(ref, f64):.3e}")
    assert e < 1e-5
    m64 = make_model(p, precision=8)
    r64 = m64.forward([-1])[0]
    print(f"config1 FP64 build vs FP64 oracle {rel(r64, f64):.3e}")
    assert rel(r64, f64) < 1e-6


def test_non_multiple_grid(gpu):
    # X not a multiple of the 64-wide tile, Z not a multiple of the tile height
    p = small_problem(X=101, Z=77, T=150, seed=6)
    rec = make_model(p).forward([0])
    assert rel(rec[0], oracle_rec(p, 0)) < 1e-5


def test_zero_source_gives_zero(gpu):
    p = small_problem(seed=7)
    p["src"] = np.zeros_like(p["src"])
    rec = make_model(p).forward([-1])
    assert np.all(rec == 0)
