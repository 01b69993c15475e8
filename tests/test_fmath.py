"""Division and square root (cu_fdiv / cu_sqrt, include/mpc_cuda/cu_fmath.cuh;
reference include/mpc_cuda/cu_fmath.cuh:53-100): faithful results of Newton
refinement at pb + 128 bits.  Pins of the oracle's restatement
(oracle/mpfixed.c or_fdiv / or_fsqrt):

* the reference itself (its header compiled into oracle/_ref): bit-exact on
  the reference generator's values and on the structured edge cases;
* MPFR's correctly rounded mpfr_div / mpfr_sqrt: within one unit in the last
  place (the reference's own accuracy statement, cu_fmath.cuh:3-11), and
  exact on quotients and roots that are representable;
* the header's host build (tests/host/libcufp_host.so) bit-exact with the
  oracle.  CPU only."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O
import goldens
from mpcases import edge_arrays, pack, unpack
from mpc_cuda_amd.arrays import MPArray, nlimbs

PBS = goldens.PBS
HOSTLIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host", "libcufp_host.so")


def operands(pb, n=400, seed=1):
    """reference-generator values plus the edge cases; divisors never zero
    (the reference has no defined result for a zero divisor)"""
    a = O.fill_ref(pb, n, 0x1357 + pb + seed, 30)
    b = O.fill_ref(pb, n, 0x2468 + pb + seed, 30)
    ea, eb, _, _ = edge_arrays(pb, seed=seed + 2)
    A = (np.ascontiguousarray(np.concatenate([a[0], ea.limbs], 1)), np.ascontiguousarray(np.concatenate([a[1], ea.se])))
    B = (np.ascontiguousarray(np.concatenate([b[0], eb.limbs], 1)), np.ascontiguousarray(np.concatenate([b[1], eb.se])))
    zb = (B[1] >> 1) == O.ZERO_EXP
    B[1][zb] = 2 * 3
    B[0][:, zb] = 0
    B[0][-1, zb] = np.uint64(1 << 63)
    return A, B


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("pb", PBS)
def test_oracle_div_sqrt_vs_reference_build(pb):
    A, B = operands(pb)
    assert O.mismatches(O.elementwise_real(pb, "div", A, B), O.elementwise_real(pb, "div", A, B, impl="ref"),
                        ignore_zero_sign=False) == 0
    assert O.mismatches(O.sqrt_real(pb, A), O.sqrt_real(pb, A, impl="ref"), ignore_zero_sign=False) == 0


def _ulps(pb, x, y):
    """distance between two lists of nonzero same-sign values in units of the
    smaller one's last place"""
    out = []
    for (sx, ex, Mx, zx), (sy, ey, My, zy) in zip(x, y):
        assert not zx and not zy and sx == sy
        sh = 64 * nlimbs(pb) - pb
        mx, my = Mx >> sh, My >> sh
        e0 = min(ex, ey)
        out.append(abs((mx << (ex - e0)) - (my << (ey - e0))))
    return np.array(out)


@pytest.mark.skipif(not O.pin_available(), reason="MPFR driver (oracle/build/pin_mpfr) not built")
@pytest.mark.parametrize("pb", PBS)
def test_oracle_div_sqrt_faithful_vs_mpfr(pb):
    A, B = operands(pb, n=300, seed=7)
    # drop zero dividends / non-positive radicands: the fixed type returns +0 there
    keep_a = (A[1] >> 1) != O.ZERO_EXP
    A = (np.ascontiguousarray(A[0][:, keep_a]), np.ascontiguousarray(A[1][keep_a]))
    B = (np.ascontiguousarray(B[0][:, keep_a]), np.ascontiguousarray(B[1][keep_a]))
    n = A[1].shape[0]
    ours = O.elementwise_real(pb, "div", A, B)
    mp = O.pin("div", pb, [A, B], ((nlimbs(pb), n), (n,)), [n])
    wrap = lambda p: unpack(MPArray(pb, p[0], p[1]))
    d = _ulps(pb, wrap(ours), wrap(mp))
    assert d.max() <= 1 and (d == 0).mean() > 0.5
    pos = (A[1] & 1) == 0
    X = (np.ascontiguousarray(A[0][:, pos]), np.ascontiguousarray(A[1][pos]))
    m = X[1].shape[0]
    d = _ulps(pb, wrap(O.sqrt_real(pb, X)), wrap(O.pin("sqrt", pb, [X], ((nlimbs(pb), m), (m,)), [m])))
    assert d.max() <= 1 and (d == 0).mean() > 0.5


@pytest.mark.parametrize("pb", [64, 256, 1024])
def test_exact_quotients_and_roots(pb):
    """Representable results come out exact: 6/3 = 2, 1/4, -7/(-7) = 1,
    (2^(pb-1)+1)/1; sqrt(4) = 2, sqrt(1/4) = 1/2, sqrt(2^-200) = 2^-100,
    sqrt(9/16) = 3/4."""
    top = 1 << (pb - 1)
    v = lambda m, e, s=1: (s, e, m << (64 * nlimbs(pb) - pb), False)
    num = [v(3 << (pb - 2), 3), v(top, 1), v(top | (top >> 1) | (top >> 2), 3, -1), v(top | 1, 1)]
    den = [v(3 << (pb - 2), 2), v(top, 3), v(top | (top >> 1) | (top >> 2), 3, -1), v(top, 1)]
    q = unpack(MPArray(pb, *O.elementwise_real(pb, "div", pack(num, pb).pair(), pack(den, pb).pair())))
    assert q == [v(top, 2), v(top, -1), v(top, 1), v(top | 1, 1)]
    rad = [v(top, 3), v(top, -1), v(top, -199), v(9 << (pb - 4), 0)]
    r = unpack(MPArray(pb, *O.sqrt_real(pb, pack(rad, pb).pair())))
    assert r == [v(top, 2), v(top, 0), v(top, -99), v(3 << (pb - 2), 0)]
    # zero and negative radicands, zero dividend and divisor give +0
    z = (1, O.ZERO_EXP, 0, True)
    r = unpack(MPArray(pb, *O.sqrt_real(pb, pack([z, v(top, 3, -1)], pb).pair())))
    assert [x[3] for x in r] == [True, True] and all(x[0] == 1 for x in r)
    q = unpack(MPArray(pb, *O.elementwise_real(pb, "div", pack([z, v(top, 1)], pb).pair(),
                                               pack([v(top, 1), z], pb).pair())))
    assert [x[3] for x in q] == [True, True]


@pytest.mark.parametrize("pb", PBS)
def test_header_host_build_vs_oracle(pb):
    if not os.path.exists(HOSTLIB):
        pytest.fail("tests/host/libcufp_host.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(HOSTLIB)
    u64p, i64p = ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int64)
    L.host_elementwise_real.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64] + [u64p, i64p] * 3
    L.host_sqrt_real.argtypes = [ctypes.c_int, ctypes.c_int64] + [u64p, i64p] * 2
    A, B = operands(pb, n=300, seed=11)
    n = A[1].shape[0]
    P = lambda a, t: a.ctypes.data_as(t)
    ol, os_ = O.empty_real(pb, n)
    assert L.host_elementwise_real(pb, 3, n, P(A[0], u64p), P(A[1], i64p), P(B[0], u64p), P(B[1], i64p),
                                   P(ol, u64p), P(os_, i64p)) == 0
    assert O.mismatches((ol, os_), O.elementwise_real(pb, "div", A, B), ignore_zero_sign=False) == 0
    ol, os_ = O.empty_real(pb, n)
    assert L.host_sqrt_real(pb, n, P(A[0], u64p), P(A[1], i64p), P(ol, u64p), P(os_, i64p)) == 0
    assert O.mismatches((ol, os_), O.sqrt_real(pb, A), ignore_zero_sign=False) == 0
