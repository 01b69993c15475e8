"""The complex AXPY with its products on INT8 MFMA (mpc_cuda_amd/csrc/axpy_mfma.h)
run in a host emulation of its wavefront (tests/axmf_emul: the kernel's digit
recoding, A-tile bytes, word assembly, carry sweep and rounding helpers, with
the lanes, the LDS buffers and v_mfma_i32_16x16x64_i8 emulated), against the
oracle's complex AXPY (cu_cmul + cu_cadd, MPC_RNDNN per operation), bit for bit
including the sign field of zeros.  Operands: random values with exponent
spreads up to +-200, zero parts of a, x and y, and the products the short
window cannot decide -- a real part on a rounding midpoint that only the low
product bits move off, exact cancellation (+0) and cancellation of all but a
few bits -- which must take the exact products (the emulator counts them).
The real AXPY with its product on MFMA (k_axpy_real_mfma, pb 1024 and 2048)
likewise against the oracle's real AXPY: random values, products just above a
rounding midpoint (which must take the exact product), exact products, zero
operands, all-ones significands."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O
from mpcases import pack, unpack
from mpc_cuda_amd.arrays import MPArray, nlimbs
import mpc_cuda_amd as M

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "axmf_emul", "libaxmf_emul.so")
u64p, i64p = ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int64)
ZERO = (1, -0x7FFFFFFFFFFFFF, 0, True)


@pytest.fixture(scope="module")
def emul():
    if not os.path.exists(LIB):
        pytest.fail("tests/axmf_emul/libaxmf_emul.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB)
    L.axmf_emul_axpy.restype = ctypes.c_long
    L.axmf_emul_axpy.argtypes = [ctypes.c_int, ctypes.c_int64] + [u64p, i64p] * 4
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


def run_emul(L, pb, a, x, y):
    n = x[1].shape[1]
    zl = np.zeros_like(x[0])
    zs = np.zeros_like(x[1])
    nodd = L.axmf_emul_axpy(pb, n, _p(a[0], u64p), _p(a[1], i64p), _p(x[0], u64p), _p(x[1], i64p),
                            _p(y[0], u64p), _p(y[1], i64p), _p(zl, u64p), _p(zs, i64p))
    assert nodd >= 0
    return (zl, zs), nodd


def cplx(re, im, pb):
    return MPArray(pb, np.ascontiguousarray(np.stack([re.limbs, im.limbs])),
                   np.ascontiguousarray(np.stack([re.se, im.se])), True)


@pytest.mark.parametrize("pb", [512, 1024, 2048])
@pytest.mark.parametrize("spread", [16, 200])
def test_random_vs_oracle(emul, pb, spread):
    n = 32 * (12 if pb < 2048 else 4)
    a = M.random_complex(pb, 1, 11 + pb + spread, -spread, spread)
    x = M.random_complex(pb, n, 12 + pb + spread, -spread, spread)
    y = M.random_complex(pb, n, 13 + pb + spread, -spread, spread)
    got, nodd = run_emul(emul, pb, a.pair(), x.pair(), y.pair())
    want = O.axpy_complex(pb, a.pair(), x.pair(), y.pair())
    assert O.mismatches(got, want, ignore_zero_sign=False) == 0
    assert nodd == 0                         # random products never need the exact path


@pytest.mark.parametrize("pb", [512, 1024])
def test_zero_parts_vs_oracle(emul, pb):
    """zero real or imaginary parts of a, x and y, and whole zero elements"""
    n = 64
    rnd = M.random_complex(pb, n, 70 + pb, -20, 20)
    Xr, Xi = unpack(rnd.real_part()), unpack(rnd.imag_part())
    Yr, Yi = list(reversed(Xi)), list(reversed(Xr))
    for k in range(0, n, 5):
        Xr[k] = ZERO
    for k in range(1, n, 7):
        Xi[k] = ZERO
    for k in range(2, n, 3):
        Yr[k] = ZERO
    for k in range(3, n, 11):
        Xr[k], Xi[k], Yr[k], Yi[k] = ZERO, ZERO, ZERO, ZERO
    x = cplx(pack(Xr, pb), pack(Xi, pb), pb)
    y = cplx(pack(Yr, pb), pack(Yi, pb), pb)
    for ar_zero, ai_zero in ((True, False), (False, True), (True, True)):
        A = M.random_complex(pb, 1, 80 + pb, -5, 5)
        Ar, Ai = unpack(A.real_part()), unpack(A.imag_part())
        if ar_zero:
            Ar = [ZERO]
        if ai_zero:
            Ai = [ZERO]
        a = cplx(pack(Ar, pb), pack(Ai, pb), pb)
        got, _ = run_emul(emul, pb, a.pair(), x.pair(), y.pair())
        want = O.axpy_complex(pb, a.pair(), x.pair(), y.pair())
        assert O.mismatches(got, want, ignore_zero_sign=False) == 0


@pytest.mark.parametrize("pb", [512, 1024, 2048])
def test_undecided_products_take_the_exact_path(emul, pb):
    """a = (2^(pb-1)+1, 2^(pb-1)) 2^-pb; per element a real part on a
    midpoint, exactly zero, or cancelled to a few bits; y zero there"""
    top = 1 << (pb - 1)
    v = lambda m, e, s=1: (s, e, m << (64 * nlimbs(pb) - pb), False)
    n = 96
    rnd = M.random_complex(pb, n, 900 + pb, -20, 20)
    Rr, Ri = unpack(rnd.real_part()), unpack(rnd.imag_part())
    Xr, Xi, Yr, Yi = list(Rr), list(Ri), list(Ri), list(Rr)
    planted = 0
    for k in (1, 40, 75, 90):
        Xr[k], Xi[k] = v(top | 1, 0), v(top, -pb, -1)           # Re on a midpoint
        Xr[k + 1], Xi[k + 1] = v(top, 7), v(top | 1, 7)         # Re = +0
        Xr[k - 1], Xi[k - 1] = v(top, 7), v(top | 3, 7)         # Re cancelled to a few bits
        for j in (k - 1, k, k + 1):
            Yr[j], Yi[j] = ZERO, ZERO
        planted += 3
    a = cplx(pack([v(top | 1, 0)], pb), pack([v(top, 0)], pb), pb)
    x = cplx(pack(Xr, pb), pack(Xi, pb), pb)
    y = cplx(pack(Yr, pb), pack(Yi, pb), pb)
    got, nodd = run_emul(emul, pb, a.pair(), x.pair(), y.pair())
    want = O.axpy_complex(pb, a.pair(), x.pair(), y.pair())
    assert O.mismatches(got, want, ignore_zero_sign=False) == 0
    assert nodd >= planted


@pytest.mark.parametrize("pb", [512, 1024])
def test_extreme_significands_vs_oracle(emul, pb):
    """all-ones, single-bit and alternating significands (digit recoding carries
    through every byte; products at both ends of [2^(2pb-2), 2^(2pb)))"""
    top, ones = 1 << (pb - 1), (1 << pb) - 1
    pats = [ones, top, top | 1, ones ^ 1, int("10" * (pb // 2), 2), int("01" * (pb // 2), 2) | top,
            top | ((1 << (pb // 2)) - 1), ones ^ (top >> 1), int(("7f" * (pb // 8)), 16) | top,
            int(("80" * (pb // 8)), 16)]
    v = lambda m, e, s=1: (s, e, m << (64 * nlimbs(pb) - pb), False)
    n = 32 * 4
    Xr = [v(pats[k % len(pats)], k % 7 - 3, 1 if k % 3 else -1) for k in range(n)]
    Xi = [v(pats[(k // len(pats)) % len(pats)], 2 - k % 5, -1 if k % 4 else 1) for k in range(n)]
    Yr = [v(pats[(k + 3) % len(pats)], k % 9 - 4) for k in range(n)]
    Yi = [v(pats[(k + 5) % len(pats)], 4 - k % 9, -1) for k in range(n)]
    x = cplx(pack(Xr, pb), pack(Xi, pb), pb)
    y = cplx(pack(Yr, pb), pack(Yi, pb), pb)
    for ar, ai in ((ones, top), (top, ones), (pats[4], pats[5]), (pats[8], pats[9])):
        a = cplx(pack([v(ar, 1)], pb), pack([v(ai, -1, -1)], pb), pb)
        got, _ = run_emul(emul, pb, a.pair(), x.pair(), y.pair())
        want = O.axpy_complex(pb, a.pair(), x.pair(), y.pair())
        assert O.mismatches(got, want, ignore_zero_sign=False) == 0


# ---------------------------------------------------------------- real AXPY

@pytest.fixture(scope="module")
def emul_real(emul):
    emul.axmf_emul_axpy_real.restype = ctypes.c_long
    emul.axmf_emul_axpy_real.argtypes = [ctypes.c_int, ctypes.c_int64] + [u64p, i64p] * 4
    return emul


def run_real(L, pb, a, x, y):
    n = x[1].shape[0]
    zl = np.zeros_like(x[0])
    zs = np.zeros_like(x[1])
    nodd = L.axmf_emul_axpy_real(pb, n, _p(a[0], u64p), _p(a[1], i64p), _p(x[0], u64p), _p(x[1], i64p),
                                 _p(y[0], u64p), _p(y[1], i64p), _p(zl, u64p), _p(zs, i64p))
    assert nodd >= 0
    return (zl, zs), nodd


@pytest.mark.parametrize("pb", [1024, 2048])
@pytest.mark.parametrize("spread", [8, 300])
def test_real_random_vs_oracle(emul_real, pb, spread):
    n = 64 * (6 if pb < 2048 else 3)
    a = M.random_real(pb, 1, 21 + pb + spread, -spread, spread)
    x = M.random_real(pb, n, 22 + pb + spread, -spread, spread)
    y = M.random_real(pb, n, 23 + pb + spread, -spread, spread)
    got, nodd = run_real(emul_real, pb, a.pair(), x.pair(), y.pair())
    assert O.mismatches(got, O.axpy_real(pb, a.pair(), x.pair(), y.pair()), ignore_zero_sign=False) == 0
    assert nodd == 0


@pytest.mark.parametrize("pb", [1024, 2048])
def test_real_midpoints_zeros_extremes_vs_oracle(emul_real, pb):
    """products just above a rounding midpoint ((2^(pb-1) + 2^(pb/2) + 1)
    (2^(pb-1) + 2^(pb/2-2) + 1)), exact products of powers of two and of
    all-ones significands, zero x and y, and products that cancel y"""
    top, ones = 1 << (pb - 1), (1 << pb) - 1
    v = lambda m, e, s=1: (s, e, m << (64 * nlimbs(pb) - pb), False)
    n = 128
    rnd = M.random_real(pb, n, 60 + pb, -30, 30)
    X = unpack(rnd)
    Y = list(reversed(X))
    a_m = top + (1 << (pb // 2)) + 1
    for k in range(0, n, 6):
        X[k] = v(top + (1 << (pb // 2 - 2)) + 1, 3)             # a x near a midpoint
    for k in range(1, n, 9):
        X[k] = v(top, -2)                                        # a x exact (a times a power of two)
    for k in range(2, n, 11):
        X[k] = ZERO
    for k in range(3, n, 7):
        Y[k] = ZERO
    for k in range(4, n, 13):
        X[k] = v(ones, 1, -1)
    a = pack([v(a_m, 2)], pb)
    x, y = pack(X, pb), pack(Y, pb)
    got, nodd = run_real(emul_real, pb, a.pair(), x.pair(), y.pair())
    assert O.mismatches(got, O.axpy_real(pb, a.pair(), x.pair(), y.pair()), ignore_zero_sign=False) == 0
    assert nodd >= len(range(0, n, 6)) - len([k for k in range(0, n, 6) if k % 11 == 2])
    for a_val in (ZERO, v(ones, -5, -1), v(top, 0)):
        a = pack([a_val], pb)
        got, _ = run_real(emul_real, pb, a.pair(), x.pair(), y.pair())
        assert O.mismatches(got, O.axpy_real(pb, a.pair(), x.pair(), y.pair()), ignore_zero_sign=False) == 0
