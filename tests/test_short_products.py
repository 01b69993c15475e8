"""The short-product paths -- of the complex product component RNDN(a*b -/+
c*d) (core::fmm_short, include/mpc_cuda/cu_fcomplex.cuh) and of the real
product RNDN(a*b) (core::fmul_short, cu_freal.cuh), by word columns below 16
words and by limb columns above -- against the exact paths (core::fmm_exact, core::fmul_exact), both from
the g++ build of the headers (tests/host/libcufp_host.so): wherever the short
path does not flag a lane as undecided, its result is the exact one, bit for
bit; the cases its error bound cannot decide (exact and massive cancellation,
a sum on a rounding midpoint that only the dropped low product bits move off)
are flagged.  CPU only."""
import ctypes
import os

import numpy as np
import pytest

from mpc_cuda_amd import gen as G
from mpc_cuda_amd.arrays import nlimbs
from mpcases import pack

HOSTLIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host", "libcufp_host.so")
PBS = [256, 288, 320, 384, 512, 1024, 1056, 2048]


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HOSTLIB):
        pytest.fail("tests/host/libcufp_host.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(HOSTLIB)
    L.host_fmm_check.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int] + \
        [ctypes.c_void_p] * 5
    L.host_fmul_check.argtypes = [ctypes.c_int, ctypes.c_int64] + [ctypes.c_void_p] * 7
    return L


def run(L, pb, ops, sub):
    """ops: four host real MPArrays of equal n -> (short, exact, amb)"""
    n = ops[0].n
    N = nlimbs(pb)
    l = np.ascontiguousarray(np.stack([o.limbs for o in ops]))
    s = np.ascontiguousarray(np.stack([o.se for o in ops]))
    sl, el = np.zeros((N, n), np.uint64), np.zeros((N, n), np.uint64)
    ss, es = np.zeros(n, np.int64), np.zeros(n, np.int64)
    amb = np.zeros(n, np.uint8)
    assert L.host_fmm_check(pb, n, l.ctypes.data, s.ctypes.data, int(sub), sl.ctypes.data, ss.ctypes.data,
                            el.ctypes.data, es.ctypes.data, amb.ctypes.data) == 0
    return (sl, ss), (el, es), amb.astype(bool)


def agree_where_decided(short, exact, amb):
    ok = ~amb
    return np.array_equal(short[0][:, ok], exact[0][:, ok]) and np.array_equal(short[1][ok], exact[1][ok])


@pytest.mark.parametrize("pb", PBS)
@pytest.mark.parametrize("sub", [0, 1])
def test_short_path_exact_on_random_operands(host, pb, sub):
    n = 3000
    ops = [G.random_real(pb, n, 1000 * pb + 10 * sub + k, -40, 40) for k in range(4)]
    short, exact, amb = run(host, pb, ops, sub)
    assert agree_where_decided(short, exact, amb)
    assert amb.sum() == 0          # random operands sit far from the midpoints


@pytest.mark.parametrize("pb", PBS)
def test_cancellation_is_flagged(host, pb):
    n = 200
    a = G.random_real(pb, n, 7 + pb, -8, 8)
    b = G.random_real(pb, n, 8 + pb, -8, 8)
    # exact cancellation: a*b - a*b = 0
    short, exact, amb = run(host, pb, [a, b, a, b], 1)
    assert amb.all()
    # massive cancellation: d differs from b in its last bit
    N = nlimbs(pb)
    d_l = b.limbs.copy()
    d_l[0] ^= np.uint64(1 << (64 * N - pb))
    d = type(b)(pb, d_l, b.se.copy())
    short, exact, amb = run(host, pb, [a, b, a, d], 1)
    assert amb.all()
    assert agree_where_decided(short, exact, amb)


@pytest.mark.parametrize("pb", PBS)
def test_midpoint_moved_by_dropped_bits_is_flagged(host, pb):
    """(2^(pb-1)+1)^2 2^-2pb + 2^(-pb-2): the short product drops the final
    +1 of the square, which leaves the sum exactly on a midpoint (rounded to
    even: down), while the exact sum lies above it (rounded up).  The two
    paths differ, so the short path must flag the lane."""
    top = 1 << (pb - 1)
    v = lambda m, e, s=1: (s, e, m << (64 * nlimbs(pb) - pb), False)
    a = pack([v(top | 1, 0)], pb)
    c = pack([v(top, 0)], pb)
    d = pack([v(top, -pb)], pb)
    short, exact, amb = run(host, pb, [a, a, c, d], 0)
    assert not (np.array_equal(short[0], exact[0]) and np.array_equal(short[1], exact[1]))
    assert amb.all()
    # the same sum reached as a difference: a*a - c*(-d)
    dn = pack([v(top, -pb, -1)], pb)
    short, exact, amb = run(host, pb, [a, a, c, dn], 1)
    assert amb.all()


def run_mul(L, pb, a, b):
    n = a.n
    N = nlimbs(pb)
    l = np.ascontiguousarray(np.stack([a.limbs, b.limbs]))
    s = np.ascontiguousarray(np.stack([a.se, b.se]))
    sl, el = np.zeros((N, n), np.uint64), np.zeros((N, n), np.uint64)
    ss, es = np.zeros(n, np.int64), np.zeros(n, np.int64)
    amb = np.zeros(n, np.uint8)
    assert L.host_fmul_check(pb, n, l.ctypes.data, s.ctypes.data, sl.ctypes.data, ss.ctypes.data,
                             el.ctypes.data, es.ctypes.data, amb.ctypes.data) == 0
    return (sl, ss), (el, es), amb.astype(bool)


@pytest.mark.parametrize("pb", PBS)
def test_short_real_product(host, pb):
    """Random products: never flagged, always exact.  (2^(pb-1) + 2^(pb/2) + 1)
    (2^(pb-1) + 2^(pb/2-2) + 1): the kept columns put the product exactly on a
    midpoint with an even last bit (rounded down), the dropped low columns
    lift it above (rounded up): flagged, and the two paths differ."""
    a = G.random_real(pb, 3000, 77 + pb, -40, 40)
    b = G.random_real(pb, 3000, 78 + pb, -40, 40)
    short, exact, amb = run_mul(host, pb, a, b)
    assert amb.sum() == 0 and agree_where_decided(short, exact, amb)
    top = 1 << (pb - 1)
    v = lambda m, e, s=1: (s, e, m << (64 * nlimbs(pb) - pb), False)
    x = pack([v(top + (1 << (pb // 2)) + 1, 0)], pb)
    y = pack([v(top + (1 << (pb // 2 - 2)) + 1, 0)], pb)
    short, exact, amb = run_mul(host, pb, x, y)
    assert amb.all()
    assert not (np.array_equal(short[0], exact[0]) and np.array_equal(short[1], exact[1]))
