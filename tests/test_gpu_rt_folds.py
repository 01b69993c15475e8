"""Run-time precision folds on the device (mpgfx_rt_gemv, mpgfx_rt_gemm, the
dot product as a one-row GEMV), real and complex, at precisions that are not
compiled in and some that are: value for value against the oracle's per-op
GEMM, and bit for bit, zero signs included, against MPFR (mpfr_mul +
mpfr_add from +0, demos/matmul_mpfr.cu) and MPC (mpc_mul + mpc_add,
demos/matmul_mpc.cu) where oracle/libcpu_mpfr.so is built."""
import numpy as np
import pytest

import mpc_cuda_amd as M
from mpc_cuda_amd.arrays import MPArray, nlimbs
from oracle import oracle as O
from test_rt_oracle import cplx

pytestmark = pytest.mark.gpu
PRECS = [2, 24, 53, 100, 256, 333, 640, 1000, 1024, 2048]


def dev(pair, p, c=False):
    return MPArray(p, np.ascontiguousarray(pair[0]), np.ascontiguousarray(pair[1]), c).cuda()


def mat(p, rows, cols, seed, zeros=True):
    """rows x cols values (row-major) with +0 / -0 entries sprinkled in"""
    n = rows * cols
    l, se = O.fill_ref(p, n, seed, 20)
    if zeros and n > 3:
        rng = np.random.default_rng(seed)
        for idx in rng.choice(n, size=max(1, n // 11), replace=False):
            l[:, idx] = 0
            se[idx] = (O.ZERO_EXP << 1) | int(rng.integers(0, 2))
    return np.ascontiguousarray(l), np.ascontiguousarray(se)


def cancelling(p, m, n, seed):
    """A (m x 2), B (2 x n) with A[:,0] = A[:,1] and B[0,:] = -B[1,:]: every
    sum is RNDN(t) + RNDN(-t), an exact zero (+0 under RNDN)"""
    a = O.fill_ref(p, m, seed, 10)
    b = O.fill_ref(p, n, seed + 1, 10)
    A = (np.ascontiguousarray(np.repeat(a[0], 2, axis=1)), np.ascontiguousarray(np.repeat(a[1], 2)))
    Bl = np.ascontiguousarray(np.concatenate([b[0], b[0]], 1))
    Bse = np.ascontiguousarray(np.concatenate([b[1], b[1] ^ 1]))
    return A, (Bl, Bse)


def check(p, got, m, n, k, A, B, c):
    mode = "cperop" if c else "perop"
    assert O.mismatches(got.pair(), O.gemm(p, m, n, k, A, B, mode), ignore_zero_sign=True) == 0
    if O.cpu_available():
        want, _ = O.cpu_gemm(p, m, n, k, A, B, mode, (0, m))
        assert O.mismatches(got.pair(), want, ignore_zero_sign=False) == 0


@pytest.mark.parametrize("p", PRECS)
def test_rt_gemm_real(gpu, p):
    m, n, k = 5, 300, 13                      # n ragged against the 256-lane rows
    A, B = mat(p, m, k, 0x31 + p), mat(p, k, n, 0x32 + p)
    check(p, M.rt_gemm(dev(A, p), dev(B, p), m, n, k), m, n, k, A, B, False)
    A, B = cancelling(p, 3, 70, 0x33 + p)
    C = M.rt_gemm(dev(A, p), dev(B, p), 3, 70, 2)
    check(p, C, 3, 70, 2, A, B, False)
    assert all(int(s) == (O.ZERO_EXP << 1) for s in C.numpy().se)        # +0 everywhere


@pytest.mark.parametrize("p", PRECS)
def test_rt_gemv_and_dot_real(gpu, p):
    m, k = 300, 21
    A, x = mat(p, m, k, 0x41 + p), mat(p, k, 1, 0x42 + p)
    check(p, M.rt_gemv(dev(A, p), dev(x, p), m, k), m, 1, k, A, x, False)
    k = 700
    u, v = mat(p, 1, k, 0x43 + p, zeros=False), mat(p, k, 1, 0x44 + p, zeros=False)
    check(p, M.rt_dot(dev(u, p), dev(v, p)), 1, 1, k, u, v, False)


def cmat(p, rows, cols, seed):
    n = rows * cols
    return cplx(mat(p, 2 * n, 1, seed), n)


@pytest.mark.parametrize("p", PRECS)
def test_rt_gemm_and_gemv_complex(gpu, p):
    m, n, k = 3, 140, 9                       # 2 lanes per element: ragged again
    A, B = cmat(p, m, k, 0x51 + p), cmat(p, k, n, 0x52 + p)
    check(p, M.rt_gemm(dev(A, p, True), dev(B, p, True), m, n, k), m, n, k, A, B, True)
    m, k = 200, 11
    A, x = cmat(p, m, k, 0x53 + p), cmat(p, k, 1, 0x54 + p)
    check(p, M.rt_gemv(dev(A, p, True), dev(x, p, True), m, k), m, 1, k, A, x, True)
    u, v = cmat(p, 1, 300, 0x55 + p), cmat(p, 300, 1, 0x56 + p)
    check(p, M.rt_dot(dev(u, p, True), dev(v, p, True)), 1, 1, 300, u, v, True)


@pytest.mark.parametrize("p", [256, 1024])
def test_rt_gemm_matches_fixed_precision_perop(gpu, p):
    """same fold, same order: the compiled per-op GEMM agrees value for value
    (the compiled path keeps the reference's +0 rule for zeros)"""
    m, n, k = 9, 130, 17
    A, B = M.random_real(p, m * k, 3, -10, 10).cuda(), M.random_real(p, k * n, 4, -10, 10).cuda()
    assert O.mismatches(M.rt_gemm(A, B, m, n, k).pair(), M.gemm(A, B, m, n, k, "perop").pair()) == 0
    x = M.random_real(p, k, 5, -10, 10).cuda()
    assert O.mismatches(M.rt_gemv(A, x, m, k).pair(), M.gemv(A, x, m, k).pair()) == 0
    cA, cB = M.random_complex(p, m * k, 6, -10, 10).cuda(), M.random_complex(p, k * n, 7, -10, 10).cuda()
    assert O.mismatches(M.rt_gemm(cA, cB, m, n, k).pair(), M.gemm(cA, cB, m, n, k, "perop").pair()) == 0


def test_rt_fold_empty_sum_and_shapes(gpu):
    p = 333
    A = dev(O.fill_ref(p, 12, 1, 5), p)
    out = M.rt_gemm(M.empty(p, 0, device="cuda"), M.empty(p, 0, device="cuda"), 3, 4, 0)
    assert out.n == 12 and all(int(s) == (O.ZERO_EXP << 1) for s in out.numpy().se)
    assert not out.numpy().limbs.any()
    with pytest.raises(ValueError):
        M.rt_gemv(A, A, 3, 4)
    with pytest.raises(ValueError):
        M.rt_gemm(A, A, 3, 4, 5)
    c = M.rt_gemv(M.empty(p, 0, True, device="cuda"), M.empty(p, 0, True, device="cuda"), 5, 0)
    assert c.n == 5 and all(int(s) == (O.ZERO_EXP << 1) for s in c.numpy().se.reshape(-1))
