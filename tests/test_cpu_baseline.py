"""The CPU baseline (oracle/libcpu_mpfr.so: system MPFR / MPC in the kernels'
operation order, the contender of demos/bench3_cpu.cpp) computes the same
values as the oracle, so the GPU/CPU comparisons in bench.py compare equal
work.  Zeros are compared by value: MPFR signs the zero of a product with a
zero factor, the reference's types give +0."""
import numpy as np
import pytest

import mpc_cuda_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.skipif(not O.cpu_available(), reason="MPFR/MPC not installed: no CPU baseline")


@pytest.mark.parametrize("pb,cplx", [(64, False), (256, False), (512, True), (1024, True), (96, False)])
def test_cpu_axpy_equals_oracle(pb, cplx):
    n = 3000
    a, x, y = (M.random_array(pb, c, s, -16, 16, cplx).pair() for c, s in ((1, 1), (n, 2), (n, 3)))
    got, ms = O.cpu_axpy(pb, a, x, y, cplx)
    ref = (O.axpy_complex if cplx else O.axpy_real)(pb, a, x, y)
    assert ms >= 0.0
    assert O.mismatches(got, ref, ignore_zero_sign=True) == 0


@pytest.mark.parametrize("mode,pb,lo,hi", [("perop", 256, -4, 4), ("perop", 512, -4, 4), ("exact", 256, -4, 4),
                                           ("exact", 256, -80, 80), ("cperop", 128, -8, 8)])
def test_cpu_gemm_rows_equal_oracle(mode, pb, lo, hi):
    m, n, k = 12, 9, 70
    cplx = mode == "cperop"
    A = M.random_array(pb, m * k, 5, lo, hi, cplx).pair()
    B = M.random_array(pb, k * n, 6, lo, hi, cplx).pair()
    rows = (3, 8)
    got, ms = O.cpu_gemm(pb, m, n, k, A, B, mode, rows)
    ref = O.gemm(pb, m, n, k, A, B, mode, rows=rows)
    assert O.mismatches(got, ref, ignore_zero_sign=True) == 0


def test_cpu_versions_and_threads():
    v = O.cpu().cpu_versions().decode()
    assert v.startswith("MPFR ") and "MPC " in v
    assert O.cpu().cpu_num_threads() >= 1
