"""The complex per-op GEMV (mpgfx_gemv_complex: the products inside the
cooperative fold, the real and imaginary parts of a row as two chains;
reference demos/matvec_mpc.cu:42-67) against the oracle's complex per-op GEMM
with n = 1 (cu_cmul then cu_cadd per step, k ascending, acc from +0), bit for
bit including the sign field of zeros: every compiled precision, shapes around
the fold's blocks (fewer rows than a wavefront's chains, k below, at and above
a block of lane-group steps, k = 0), wide exponent spreads, products the short
products cannot decide (they take the exact path inside the fold), and the
same result through gemm(n = 1)."""
import numpy as np
import pytest

import mpc_cuda_amd as M
from mpc_cuda_amd.arrays import MPArray, nlimbs
from oracle import oracle as O
from mpcases import pack, unpack

pytestmark = pytest.mark.gpu
PBS = [32, 64, 96, 128, 160, 192, 256, 288, 320, 384, 512, 1024, 1056, 2048]
ZERO = (1, -0x7FFFFFFFFFFFFF, 0, True)


def same(gpu_arr, oracle_pair):
    return O.mismatches(gpu_arr.pair(), oracle_pair, ignore_zero_sign=False)


def cplx(re, im, pb):
    return MPArray(pb, np.ascontiguousarray(np.stack([re.limbs, im.limbs])),
                   np.ascontiguousarray(np.stack([re.se, im.se])), True)


@pytest.mark.parametrize("pb", PBS)
def test_gemv_complex_vs_oracle(gpu, pb):
    m, k = (37, 70) if pb <= 1056 else (9, 33)
    A = M.random_complex(pb, m * k, 501 + pb, -30, 30)
    x = M.random_complex(pb, k, 502 + pb, -30, 30)
    y = M.gemv(A.cuda(), x.cuda(), m, k)
    assert same(y, O.gemm(pb, m, 1, k, A.pair(), x.pair(), "cperop")) == 0
    y2 = M.gemm(A.cuda(), x.cuda(), m, 1, k, "perop")
    assert same(y2, y.pair()) == 0


@pytest.mark.parametrize("pb", [64, 256, 512, 1024])
@pytest.mark.parametrize("mm,kk", [(1, 1), (2, 3), (3, 16), (5, 17), (9, 33), (4, 100), (3, 0)])
def test_gemv_complex_small_and_ragged(gpu, pb, mm, kk):
    A = M.random_complex(pb, max(mm * kk, 1), 511 + pb + kk, -60, 60)
    x = M.random_complex(pb, max(kk, 1), 512 + pb + mm, -60, 60)
    if kk == 0:
        A = M.empty(pb, 0, True, device="cpu")
        x = M.empty(pb, 0, True, device="cpu")
        y = M.gemv(A.cuda(), x.cuda(), mm, kk)
        h = y.pair()
        assert np.all(h[0] == 0) and np.all(h[1] == (-0x7FFFFFFFFFFFFF) * 2)
        return
    y = M.gemv(A.cuda(), x.cuda(), mm, kk)
    assert same(y, O.gemm(pb, mm, 1, kk, A.pair(), x.pair(), "cperop")) == 0


@pytest.mark.parametrize("pb", [512, 1024, 1056])
def test_gemv_complex_undecided_products(gpu, pb):
    """rows whose products have a real part on a rounding midpoint, exactly
    zero, or cancelled to a few bits (the short products cannot decide them),
    mixed into random ones, and zero matrix entries"""
    top = 1 << (pb - 1)
    v = lambda m_, e, s=1: (s, e, m_ << (64 * nlimbs(pb) - pb), False)
    m, k = 12, 40
    R = M.random_complex(pb, m * k, 530 + pb, -20, 20)
    Ar, Ai = unpack(R.real_part()), unpack(R.imag_part())
    X = M.random_complex(pb, k, 531 + pb, -20, 20)
    Xr, Xi = unpack(X.real_part()), unpack(X.imag_part())
    Xr[5], Xi[5] = v(top | 1, 0), v(top, -pb, -1)
    Xr[6], Xi[6] = v(top, 7), v(top | 1, 7)
    Xr[7], Xi[7] = v(top, 7), v(top | 3, 7)
    for r in range(m):
        for j in (5, 6, 7):
            Ar[r * k + j], Ai[r * k + j] = v(top | 1, 0), v(top, 0)
        Ar[r * k + 11] = ZERO
        if r % 3 == 0:
            Ar[r * k + 12], Ai[r * k + 12] = ZERO, ZERO
    A = cplx(pack(Ar, pb), pack(Ai, pb), pb)
    x = cplx(pack(Xr, pb), pack(Xi, pb), pb)
    y = M.gemv(A.cuda(), x.cuda(), m, k)
    assert same(y, O.gemm(pb, m, 1, k, A.pair(), x.pair(), "cperop")) == 0


def test_gemv_complex_4096_512(gpu):
    """the GEMV config's shape at pb 512 (complex), 64 rows checked"""
    pb, m, k = 512, 4096, 4096
    A = M.device_random(pb, m * k, 541, -4, 4, cplx=True)
    x = M.device_random(pb, k, 542, -4, 4, cplx=True)
    y = M.gemv(A, x, m, k)
    rows = 64
    Ar = (np.ascontiguousarray(A.limbs[..., :rows * k].cpu().numpy().view(np.uint64)),
          np.ascontiguousarray(A.se[..., :rows * k].cpu().numpy()))
    want = O.gemm(pb, rows, 1, k, Ar, x.pair(), "cperop")
    got = y.pair()
    got = (np.ascontiguousarray(got[0][:, :, :rows]), np.ascontiguousarray(got[1][:, :rows]))
    assert O.mismatches(got, want, ignore_zero_sign=False) == 0
