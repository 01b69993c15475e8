"""Per-operation GEMM (C[i,j] = fold over k of RNDN(acc + RNDN(A[i,k] B[k,j])),
acc from +0, the order of tools/linalg_fast.cu:53-68) on ragged shapes: tiles
cut by m, n and k that are not multiples of the 16 x 16 LDS tile, single rows
and columns, k below and above one staged block, zero entries, and the wide
precisions that run the untiled kernel.  Bit-exact against the oracle."""
import numpy as np
import pytest

import mpc_cuda_amd as M
from mpc_cuda_amd.arrays import MPArray
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def same(gpu_arr, oracle_pair):
    return O.mismatches(gpu_arr.pair(), oracle_pair, ignore_zero_sign=False)


@pytest.mark.parametrize("pb,m,n,k", [
    (256, 33, 17, 40),     # ragged in every dimension
    (256, 16, 16, 16),     # one whole tile, one staged block
    (256, 1, 50, 7),       # one row, k below one block
    (256, 40, 1, 33),      # one column (the GEMV shape through the GEMM entry)
    (64, 35, 29, 70),
    (96, 18, 19, 23),      # sub-limb precision (low storage word zero)
    (128, 20, 31, 17),
    (512, 17, 18, 65),     # widest precision on the tiled kernel
    (1024, 9, 10, 21),     # untiled kernel
])
def test_gemm_perop_ragged_vs_oracle(gpu, pb, m, n, k):
    A = M.random_real(pb, m * k, 31 + pb + m, -6, 6)
    B = M.random_real(pb, k * n, 32 + pb + n, -6, 6)
    r = M.gemm(A.cuda(), B.cuda(), m, n, k, "perop")
    assert same(r, O.gemm(pb, m, n, k, A.pair(), B.pair(), "perop")) == 0


def test_gemm_perop_zero_entries_and_cancellation(gpu):
    """Zero entries inside and at the edges of staged blocks, a row of zeros
    (its outputs are +0), and B's column 1 the negation of column 0."""
    pb, m, n, k = 256, 21, 19, 37
    A = M.random_real(pb, m * k, 41, -4, 4)
    B = M.random_real(pb, k * n, 42, -4, 4)
    Al, As = A.limbs.copy(), A.se.copy()
    Bl, Bs = B.limbs.copy(), B.se.copy()
    for idx in (0, 15, 16, 17, 5 * k + 31, m * k - 1):
        Al[:, idx] = 0
        As[idx] = M.ZERO_EXP * 2
    Al[:, 3 * k:4 * k] = 0
    As[3 * k:4 * k] = M.ZERO_EXP * 2
    for kk in range(k):
        Bl[:, kk * n + 1] = Bl[:, kk * n]
        Bs[kk * n + 1] = Bs[kk * n] ^ 1
    A2, B2 = MPArray(pb, Al, As), MPArray(pb, Bl, Bs)
    r = M.gemm(A2.cuda(), B2.cuda(), m, n, k, "perop")
    assert same(r, O.gemm(pb, m, n, k, A2.pair(), B2.pair(), "perop")) == 0
    h = r.numpy()
    assert np.all((h.se[3 * n:4 * n] >> 1) == M.ZERO_EXP)


@pytest.mark.parametrize("pb,m,n,k", [
    (256, 33, 17, 40),     # ragged in every dimension
    (256, 1, 20, 5),
    (64, 18, 35, 33),
    (96, 17, 16, 19),      # sub-limb precision
    (512, 16, 17, 18),     # widest precision on the tiled kernel
    (1024, 7, 9, 11),      # untiled kernel
])
def test_gemm_complex_perop_ragged_vs_oracle(gpu, pb, m, n, k):
    """Complex per-op GEMM (acc = acc + A[i,k] B[k,j], MPC_RNDNN products and
    sums, k ascending; demos/matmul_mpc.cu:33-52), bit for bit against the
    oracle, zero-sign fields included."""
    A = M.random_complex(pb, m * k, 51 + pb + m, -6, 6)
    B = M.random_complex(pb, k * n, 52 + pb + n, -6, 6)
    r = M.gemm(A.cuda(), B.cuda(), m, n, k, "perop")
    assert same(r, O.gemm(pb, m, n, k, A.pair(), B.pair(), "cperop")) == 0


def test_gemm_complex_perop_256_sampled(gpu):
    """pb 256, 512 x 384 x 1024 (a large shape, k over 64 staged blocks):
    rows 0, 255 and 511 recomputed by the oracle, bit for bit."""
    pb, m, n, k = 256, 512, 384, 1024
    A = M.random_complex(pb, m * k, 61, -4, 4)
    B = M.random_complex(pb, k * n, 62, -4, 4)
    C = M.gemm(A.cuda(), B.cuda(), m, n, k, "perop").numpy()
    Ap, Bp = A.pair(), B.pair()
    for r0 in (0, 255, m - 1):
        ref = O.gemm(pb, m, n, k, Ap, Bp, "cperop", rows=(r0, r0 + 1))
        sl = slice(r0 * n, (r0 + 1) * n)
        for part in (0, 1):
            assert np.array_equal(C.limbs[part][:, sl], ref[0][part][:, sl]), (r0, part)
            assert np.array_equal(C.se[part][sl], ref[1][part][sl]), (r0, part)
