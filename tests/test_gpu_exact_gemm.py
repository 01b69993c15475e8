"""Exact-dot GEMM (INT8-MFMA slice products, one rounding per output) against
the MPFR golden fixtures (exact products + mpfr_sum) and the oracle's exact
sum, including the VALU wide-accumulator path for exponent spreads beyond the
MFMA window."""
import numpy as np
import pytest

import mpc_cuda_amd as M
from mpc_cuda_amd.arrays import MPArray
from oracle import oracle as O
import goldens

pytestmark = pytest.mark.gpu


def dev(pair, pb, cplx=False):
    return MPArray(pb, np.ascontiguousarray(pair[0]), np.ascontiguousarray(pair[1]), cplx).cuda()


@pytest.mark.parametrize("pb", goldens.PBS)
def test_exact_gemm_vs_golden(gpu, pb):
    g = goldens.load(pb)
    m, n, k = (int(v) for v in g["dims"])
    for tag in ("g4", "g80"):
        A, B = goldens.pair(g, tag + "A"), goldens.pair(g, tag + "B")
        r = M.gemm(dev(A, pb), dev(B, pb), m, n, k, "exact")
        assert O.mismatches(r.pair(), goldens.pair(g, tag + "exact"), ignore_zero_sign=True) == 0, tag
        assert O.mismatches(r.pair(), O.gemm(pb, m, n, k, A, B, "exact"), ignore_zero_sign=False) == 0, tag


@pytest.mark.parametrize("pb,m,n,k,lo,hi", [
    (256, 33, 17, 100, -4, 4),      # ragged tiles and k not a multiple of 64
    (256, 40, 40, 64, -80, 80),     # wide spread, still inside the MFMA window
    (64, 50, 30, 300, -20, 20),
    (512, 20, 24, 129, -4, 4),
    (96, 18, 19, 70, -40, 40),      # sub-limb precision (SB = 32)
    (1024, 16, 16, 65, -16, 16),
    (64, 70, 40, 4200, -4, 4),      # k beyond one int32 accumulation chunk (4096)
    (64, 544, 512, 64, -4, 4),      # more 32x32 tiles (272) than workgroups (256)
])
def test_exact_gemm_vs_oracle(gpu, pb, m, n, k, lo, hi):
    A = M.random_real(pb, m * k, 11 + pb + k, lo, hi)
    B = M.random_real(pb, k * n, 12 + pb + k, lo, hi)
    r = M.gemm(A.cuda(), B.cuda(), m, n, k, "exact")
    assert O.mismatches(r.pair(), O.gemm(pb, m, n, k, A.pair(), B.pair(), "exact"), ignore_zero_sign=False) == 0


def test_exact_gemm_cancellation_and_zeros(gpu):
    """Rows whose products cancel exactly give +0; zero entries and an all-zero
    row are skipped; a tiny term that survives a huge cancellation is exact."""
    pb, m, n, k = 256, 4, 3, 6
    A = M.random_real(pb, m * k, 5, -4, 4)
    B = M.random_real(pb, k * n, 6, -4, 4)
    Al, Ase = A.limbs.copy(), A.se.copy()
    # row 0: [x, x] against column pairs that negate -> exact zero per column
    Al[:, 1] = Al[:, 0]; Ase[1] = Ase[0] ^ 1
    Bl, Bse = B.limbs.copy(), B.se.copy()
    for j in range(n):
        Bl[:, 1 * n + j] = Bl[:, 0 * n + j]; Bse[1 * n + j] = Bse[0 * n + j]
    Ase[2:6] = M.ZERO_EXP * 2; Al[:, 2:6] = 0          # rest of row 0 zero
    Ase[6:12] = M.ZERO_EXP * 2; Al[:, 6:12] = 0        # row 1 all zero
    A2 = MPArray(pb, Al, Ase)
    B2 = MPArray(pb, Bl, Bse)
    r = M.gemm(A2.cuda(), B2.cuda(), m, n, k, "exact")
    ref = O.gemm(pb, m, n, k, A2.pair(), B2.pair(), "exact")
    assert O.mismatches(r.pair(), ref, ignore_zero_sign=False) == 0
    h = r.numpy()
    assert np.all((h.se[:2 * n] >> 1) == M.ZERO_EXP)


def test_exact_gemm_fallback_extreme_spread(gpu):
    """Exponent spreads of thousands of bits exceed the MFMA window (64 digits)
    and take the VALU wide-accumulator path; results stay exact."""
    pb, m, n, k = 128, 6, 5, 40
    A = M.random_real(pb, m * k, 21, -3000, 3000)
    B = M.random_real(pb, k * n, 22, -3000, 3000)
    r = M.gemm(A.cuda(), B.cuda(), m, n, k, "exact")
    assert O.mismatches(r.pair(), O.gemm(pb, m, n, k, A.pair(), B.pair(), "exact"), ignore_zero_sign=False) == 0


@pytest.mark.parametrize("lo,hi", [(-4, 4), (-80, 80)])
def test_exact_gemm_256_full_size_sampled(gpu, lo, hi):
    """The 1024^3 pb-256 configuration (both exponent sets): the oracle's exact
    sum recomputes 8 full rows (8192 outputs) and they must match bit for bit;
    the per-op GEMM of the same inputs differs from the exact one somewhere."""
    pb, m, n, k = 256, 1024, 1024, 1024
    A = M.random_real(pb, m * k, 501 + hi, lo, hi)
    B = M.random_real(pb, k * n, 502 + hi, lo, hi)
    C = M.gemm(A.cuda(), B.cuda(), m, n, k, "exact").numpy()
    Ap, Bp = A.pair(), B.pair()
    for r0 in (0, 77, 255, 400, 640, 777, 999, m - 1):
        ref = O.gemm(pb, m, n, k, Ap, Bp, "exact", rows=(r0, r0 + 1))
        sl = slice(r0 * n, (r0 + 1) * n)
        assert np.array_equal(C.limbs[:, sl], ref[0][:, sl]), r0
        assert np.array_equal(C.se[sl], ref[1][sl]), r0


@pytest.mark.parametrize("pb", goldens.PBS)
def test_complex_exact_gemm_vs_golden(gpu, pb):
    """Complex exact GEMM (each component one exact sum, one rounding): by value
    against MPFR's mpfr_sum fixture, bit for bit against the oracle."""
    g = goldens.load(pb)
    m, n, k = (int(v) for v in g["dims"])
    A, B = goldens.pair(g, "cgA"), goldens.pair(g, "cgB")
    r = M.gemm(dev(A, pb, True), dev(B, pb, True), m, n, k, "exact")
    assert O.mismatches(r.pair(), goldens.pair(g, "cgexact"), ignore_zero_sign=True) == 0
    assert O.mismatches(r.pair(), O.gemm(pb, m, n, k, A, B, "cexact"), ignore_zero_sign=False) == 0


@pytest.mark.parametrize("pb,m,n,k,lo,hi", [(256, 40, 33, 300, -4, 4), (96, 17, 20, 64, -60, 60), (512, 8, 9, 70, -8, 8)])
def test_complex_exact_gemm_vs_oracle(gpu, pb, m, n, k, lo, hi):
    A = M.random_complex(pb, m * k, 91 + pb, lo, hi)
    B = M.random_complex(pb, k * n, 92 + pb, lo, hi)
    r = M.gemm(A.cuda(), B.cuda(), m, n, k, "exact")
    assert O.mismatches(r.pair(), O.gemm(pb, m, n, k, A.pair(), B.pair(), "cexact"), ignore_zero_sign=False) == 0


def test_complex_exact_gemm_cancellation(gpu):
    """ai = ar and bi = br: the real parts cancel exactly (+0); then one bit of
    one imaginary entry changed: near-total cancellation."""
    pb, m, n, k = 256, 7, 6, 33
    A = M.random_complex(pb, m * k, 95, -20, 20)
    B = M.random_complex(pb, k * n, 96, -20, 20)
    Al, As = (np.array(v) for v in A.pair())
    Bl, Bs = (np.array(v) for v in B.pair())
    Al[1], As[1] = Al[0], As[0]
    Bl[1], Bs[1] = Bl[0], Bs[0]
    for flip in (False, True):
        if flip:
            Bl[1][0] ^= np.uint64(1 << 40)
        a = (np.ascontiguousarray(Al), np.ascontiguousarray(As))
        b = (np.ascontiguousarray(Bl), np.ascontiguousarray(Bs))
        r = M.gemm(dev(a, pb, True), dev(b, pb, True), m, n, k, "exact")
        assert O.mismatches(r.pair(), O.gemm(pb, m, n, k, a, b, "cexact"), ignore_zero_sign=False) == 0, flip


@pytest.mark.parametrize("k", [480000, 600000], ids=["mfma_at_limit", "beyond_slot_headroom"])
def test_exact_gemm_long_k_constant_digits(gpu, k):
    """Constant significands whose bytes are 0x7F (the largest positive digit,
    127, in every digit position) and equal exponents put the most into every
    carry-save slot per k term.  Up to about 490,000 terms (pb 256) the MFMA
    path's int64 slots hold the sum; beyond, the call takes the wide
    accumulator.  Closed form: C = RNDN(k * a * b)."""
    import bigint_ref as R
    pb, N = 256, 4
    Mv = int("80" + "7f" * 31, 16)                     # top bit set, then 0x7F bytes
    limbs = np.array([(Mv >> (64 * l)) & ((1 << 64) - 1) for l in range(N)], np.uint64)
    A = MPArray(pb, np.ascontiguousarray(np.repeat(limbs[:, None], k, 1)), np.full(k, 3 * 2, np.int64))
    B = MPArray(pb, np.ascontiguousarray(np.repeat(limbs[:, None], k, 1)), np.full(k, 5 * 2 + 1, np.int64))
    r = M.gemm(A.cuda(), B.cuda(), 1, 1, k, "exact")
    want = R.round_exact(-k * Mv * Mv, 3 + 5 - 2 * 64 * N, pb)
    got = M.arrays.decode(r.pair()[0][:, 0], r.pair()[1][0], pb)
    assert (got[0], got[1], got[2]) == (want[0], want[1], want[2])
