"""Parity at the workload sizes: the HIP path against the oracle on the full
arrays where the oracle finishes in seconds, and on documented samples plus
size-independent properties where it would not."""
import numpy as np
import pytest

import mpc_cuda_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def same(gpu_arr, oracle_pair):
    return O.mismatches(gpu_arr.pair(), oracle_pair, ignore_zero_sign=False)


@pytest.mark.parametrize("pb,n", [(256, 1 << 20), (64, 1 << 20), (128, 1 << 20), (512, 1 << 20), (2048, 1 << 18)])
def test_axpy_real_full_size(gpu, pb, n):
    """Real AXPY at n = 2^20 (2^18 at pb 2048 to keep the oracle to seconds),
    exponents in [-8, 8]: every output bit-exact against the oracle."""
    a = M.random_real(pb, 1, 101, -8, 8)
    x = M.random_real(pb, n, 102, -8, 8)
    y = M.random_real(pb, n, 103, -8, 8)
    r = M.axpy(a.cuda(), x.cuda(), y.cuda())
    assert same(r, O.axpy_real(pb, a.pair(), x.pair(), y.pair())) == 0


def test_axpy_complex_1024_full_size(gpu):
    """cu_fcomplex<1024> AXPY at n = 2^18, Re/Im exponents in [-16, 16]."""
    pb, n = 1024, 1 << 18
    a = M.random_complex(pb, 1, 201, -16, 16)
    x = M.random_complex(pb, n, 202, -16, 16)
    y = M.random_complex(pb, n, 203, -16, 16)
    r = M.axpy(a.cuda(), x.cuda(), y.cuda())
    assert same(r, O.axpy_complex(pb, a.pair(), x.pair(), y.pair())) == 0


def test_gemv_512_full_size(gpu):
    """Per-op GEMV, A 4096 x 4096 at pb 512, exponents in [-4, 4]: all 4096
    outputs bit-exact against the oracle's k-ordered fold."""
    pb, m, k = 512, 4096, 4096
    A = M.random_real(pb, m * k, 301, -4, 4)
    x = M.random_real(pb, k, 302, -4, 4)
    r = M.gemv(A.cuda(), x.cuda(), m, k)
    assert same(r, O.gemm(pb, m, 1, k, A.pair(), x.pair(), "perop")) == 0


def test_gemm_perop_256_full_size_sampled(gpu):
    """Per-op GEMM 1024^3 at pb 256: the oracle recomputes a sample of rows
    (first, last, and 6 spread rows = 8192 outputs, every column) and the GPU
    result must match those rows bit for bit; every output must be a
    normalised nonzero value."""
    pb, m, n, k = 256, 1024, 1024, 1024
    A = M.random_real(pb, m * k, 401, -4, 4)
    B = M.random_real(pb, k * n, 402, -4, 4)
    C = M.gemm(A.cuda(), B.cuda(), m, n, k, "perop").numpy()
    Ap, Bp = A.pair(), B.pair()
    N = M.nlimbs(pb)
    for r0 in (0, 137, 300, 511, 700, 901, 1000, m - 1):
        ref = O.gemm(pb, m, n, k, Ap, Bp, "perop", rows=(r0, r0 + 1))
        sl = slice(r0 * n, (r0 + 1) * n)
        assert np.array_equal(C.limbs[:, sl], ref[0][:, sl]), r0
        assert np.array_equal(C.se[sl], ref[1][sl]), r0
    assert np.all(C.limbs[N - 1] >> np.uint64(63) == 1)
