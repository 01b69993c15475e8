"""Parity of the HIP library (through the C ABI) with the oracle and with the
MPFR/MPC golden fixtures.  Bit-exact: every limb and the sign/exponent word,
including the sign field of zeros, against the oracle (a restatement of the
reference); zeros compared by value against MPFR fixtures (the reference's own
tests do the same: tools/test_freal.cu:73)."""
import numpy as np
import torch
import pytest

import mpc_cuda_amd as M
from mpc_cuda_amd.arrays import MPArray
from oracle import oracle as O
import goldens
from mpcases import edge_arrays

pytestmark = pytest.mark.gpu
PBS = goldens.PBS


def dev(pair, pb, cplx=False):
    return MPArray(pb, np.ascontiguousarray(pair[0]), np.ascontiguousarray(pair[1]), cplx).cuda()


def host_pair(x):
    return x.pair()


def same(gpu_arr, oracle_pair):
    return O.mismatches(host_pair(gpu_arr), oracle_pair, ignore_zero_sign=False)


def same_value(gpu_arr, golden_pair):
    return O.mismatches(host_pair(gpu_arr), golden_pair, ignore_zero_sign=True)


@pytest.mark.parametrize("pb", PBS)
def test_elementwise_real_vs_golden_and_oracle(gpu, pb):
    g = goldens.load(pb)
    a, b = goldens.pair(g, "a"), goldens.pair(g, "b")
    da, db = dev(a, pb), dev(b, pb)
    for op in ("mul", "add", "sub"):
        r = M.elementwise(op, da, db)
        assert same_value(r, goldens.pair(g, op)) == 0, op
        assert same(r, O.elementwise_real(pb, op, a, b)) == 0, op
    # the reference's generator and spread mix, plus edge cases
    n = 30000 if pb <= 1056 else 8000
    l, s = O.fill_ref(pb, 2 * n, 0x1234 + pb, -1)
    ea, eb, _, _ = edge_arrays(pb, seed=13)
    a = (np.concatenate([l[:, 0::2], ea.limbs], 1), np.concatenate([s[0::2], ea.se]))
    b = (np.concatenate([l[:, 1::2], eb.limbs], 1), np.concatenate([s[1::2], eb.se]))
    a = tuple(np.ascontiguousarray(v) for v in a)
    b = tuple(np.ascontiguousarray(v) for v in b)
    da, db = dev(a, pb), dev(b, pb)
    for op in ("mul", "add", "sub"):
        assert same(M.elementwise(op, da, db), O.elementwise_real(pb, op, a, b)) == 0, op


@pytest.mark.parametrize("pb", PBS)
def test_elementwise_complex_vs_golden_and_oracle(gpu, pb):
    g = goldens.load(pb)
    ca, cb = goldens.pair(g, "ca"), goldens.pair(g, "cb")
    da, db = dev(ca, pb, True), dev(cb, pb, True)
    for op in ("mul", "add", "sub"):
        r = M.elementwise(op, da, db)
        assert same_value(r, goldens.pair(g, "c" + op)) == 0, op
        assert same(r, O.elementwise_complex(pb, op, ca, cb)) == 0, op
    n = 6000 if pb <= 1056 else 2000
    x = M.random_complex(pb, n, 77 + pb, -40, 40)
    y = M.random_complex(pb, n, 78 + pb, -40, 40)
    dx, dy = x.cuda(), y.cuda()
    for op in ("mul", "add", "sub"):
        assert same(M.elementwise(op, dx, dy), O.elementwise_complex(pb, op, x.pair(), y.pair())) == 0, op


@pytest.mark.parametrize("pb", PBS)
def test_axpy_vs_golden_and_oracle(gpu, pb):
    g = goldens.load(pb)
    r = M.axpy(dev(goldens.pair(g, "ax"), pb), dev(goldens.pair(g, "xx"), pb), dev(goldens.pair(g, "yy"), pb))
    assert same_value(r, goldens.pair(g, "axpy")) == 0
    r = M.axpy(dev(goldens.pair(g, "cax"), pb, True), dev(goldens.pair(g, "cxx"), pb, True),
               dev(goldens.pair(g, "cyy"), pb, True))
    assert same_value(r, goldens.pair(g, "caxpy")) == 0
    # ragged length (not a multiple of the workgroup), oracle bit-exact
    n = 4099
    a = M.random_real(pb, 1, 5, -8, 8)
    x = M.random_real(pb, n, 6, -8, 8)
    y = M.random_real(pb, n, 7, -8, 8)
    assert same(M.axpy(a.cuda(), x.cuda(), y.cuda()), O.axpy_real(pb, a.pair(), x.pair(), y.pair())) == 0
    ca = M.random_complex(pb, 1, 8, -16, 16)
    cx = M.random_complex(pb, 1031, 9, -16, 16)
    cy = M.random_complex(pb, 1031, 10, -16, 16)
    assert same(M.axpy(ca.cuda(), cx.cuda(), cy.cuda()), O.axpy_complex(pb, ca.pair(), cx.pair(), cy.pair())) == 0


@pytest.mark.parametrize("pb", [512, 1024, 1056])
def test_axpy_complex_chunk_loop(gpu, pb):
    """Wide precisions run the complex AXPY as a loop over 32-element chunks
    per wavefront with the next chunk copied into LDS ahead of use: three
    chunks per wavefront plus a 17-element tail, oracle bit-exact."""
    n = 3 * 1024 * 32 + 17
    ca = M.random_complex(pb, 1, 30 + pb, -16, 16)
    cx = M.random_complex(pb, n, 31 + pb, -16, 16)
    cy = M.random_complex(pb, n, 32 + pb, -16, 16)
    assert same(M.axpy(ca.cuda(), cx.cuda(), cy.cuda()), O.axpy_complex(pb, ca.pair(), cx.pair(), cy.pair())) == 0


def _cplx_from(re, im, pb):
    return MPArray(pb, np.ascontiguousarray(np.stack([re.limbs, im.limbs])), np.ascontiguousarray(np.stack([re.se, im.se])),
                   True)


@pytest.mark.parametrize("pb", [512, 1024, 1056, 2048])
def test_complex_products_the_short_path_leaves_open(gpu, pb):
    """Complex products whose rounding the short products cannot decide, mixed
    into random ones: a real part on a midpoint that only the low product bits
    move off ((2^(pb-1)+1)^2 2^-2pb + 2^(-pb-2)), exact cancellation (+0) and
    cancellation of all but a few bits.  Through the elementwise product and
    the AXPY (inside whole 32-element chunks and in the tail), oracle
    bit-exact."""
    from mpcases import pack
    from mpc_cuda_amd.arrays import nlimbs
    top = 1 << (pb - 1)
    v = lambda m, e, s=1: (s, e, m << (64 * nlimbs(pb) - pb), False)
    zero = (1, -0x7FFFFFFFFFFFFF, 0, True)
    n = 100
    rnd = M.random_complex(pb, n, 900 + pb, -20, 20)
    rr, ri = rnd.real_part(), rnd.imag_part()
    from mpcases import unpack
    Rr, Ri = unpack(rr), unpack(ri)
    # elementwise: a*b per element
    Ar, Ai, Br, Bi = list(Rr), list(Ri), list(reversed(Rr)), list(reversed(Ri))
    for k in (1, 40, 75, 98):
        Ar[k], Br[k], Ai[k], Bi[k] = v(top | 1, 0), v(top | 1, 0), v(top, 0), v(top, -pb, -1)     # Re on a midpoint
        Ar[k + 1], Ai[k + 1], Br[k + 1], Bi[k + 1] = v(top | 5, 3), v(top | 5, 3), v(top | 9, -2), v(top | 9, -2)  # Re = +0
        Ar[k - 1], Ai[k - 1], Br[k - 1], Bi[k - 1] = v(top | 5, 3), v(top | 5, 3), v(top | 9, -2), v(top | 11, -2)
    a = _cplx_from(pack(Ar, pb), pack(Ai, pb), pb)
    b = _cplx_from(pack(Br, pb), pack(Bi, pb), pb)
    assert same(M.elementwise("mul", a.cuda(), b.cuda()), O.elementwise_complex(pb, "mul", a.pair(), b.pair())) == 0
    # AXPY with a = (2^(pb-1)+1, 2^(pb-1)) 2^-pb: x picks the same three cases
    ca = _cplx_from(pack([v(top | 1, 0)], pb), pack([v(top, 0)], pb), pb)
    Xr, Xi, Yr, Yi = list(Rr), list(Ri), list(Ri), list(Rr)
    for k in (1, 40, 75, 98):
        Xr[k], Xi[k] = v(top | 1, 0), v(top, -pb, -1)           # Re = a_r x_r - a_i x_i on a midpoint
        Xr[k + 1], Xi[k + 1] = v(top, 7), v(top | 1, 7)         # Re = +0
        Xr[k - 1], Xi[k - 1] = v(top, 7), v(top | 3, 7)         # Re = -2^(pb-1) 2^(14-2pb)
        for j in (k - 1, k, k + 1):
            Yr[j], Yi[j] = zero, zero
    x = _cplx_from(pack(Xr, pb), pack(Xi, pb), pb)
    y = _cplx_from(pack(Yr, pb), pack(Yi, pb), pb)
    assert same(M.axpy(ca.cuda(), x.cuda(), y.cuda()), O.axpy_complex(pb, ca.pair(), x.pair(), y.pair())) == 0


@pytest.mark.parametrize("pb", [512, 1024, 1056, 2048])
def test_real_products_near_a_midpoint(gpu, pb):
    """Real products just above a rounding midpoint,
    (2^(pb-1) + 2^(pb/2) + 1)(2^(pb-1) + 2^(pb/2-2) + 1) -- the case the short
    product (used at pb 512) cannot decide from its kept columns -- mixed into
    random ones, through the elementwise product, the AXPY and the GEMV;
    oracle bit-exact."""
    from mpcases import pack, unpack
    from mpc_cuda_amd.arrays import nlimbs
    top = 1 << (pb - 1)
    v = lambda m, e, s=1: (s, e, m << (64 * nlimbs(pb) - pb), False)
    n = 300
    X, Y = unpack(M.random_real(pb, n, 40 + pb, -20, 20)), unpack(M.random_real(pb, n, 41 + pb, -20, 20))
    for k in (0, 37, 150, 299):
        X[k], Y[k] = v(top + (1 << (pb // 2)) + 1, 3, -1), v(top + (1 << (pb // 2 - 2)) + 1, -2)
    x, y = pack(X, pb), pack(Y, pb)
    assert same(M.elementwise("mul", x.cuda(), y.cuda()), O.elementwise_real(pb, "mul", x.pair(), y.pair())) == 0
    a = pack([v(top + (1 << (pb // 2)) + 1, 1)], pb)
    Xa = list(X)
    for k in (5, 64, 200):
        Xa[k] = v(top + (1 << (pb // 2 - 2)) + 1, 2)
    xa = pack(Xa, pb)
    assert same(M.axpy(a.cuda(), xa.cuda(), y.cuda()), O.axpy_real(pb, a.pair(), xa.pair(), y.pair())) == 0
    m, k = 4, 75
    A = pack(X[:m * k], pb)
    xv = pack(Y[:k], pb)
    g = M.gemv(A.cuda(), xv.cuda(), m, k)
    assert same(g, O.gemm(pb, m, 1, k, A.pair(), xv.pair(), "perop")) == 0


@pytest.mark.parametrize("pb", PBS)
def test_gemm_gemv_perop_vs_golden_and_oracle(gpu, pb):
    g = goldens.load(pb)
    m, n, k = (int(v) for v in g["dims"])
    for tag in ("g4", "g80"):
        A, B = goldens.pair(g, tag + "A"), goldens.pair(g, tag + "B")
        r = M.gemm(dev(A, pb), dev(B, pb), m, n, k, "perop")
        assert same_value(r, goldens.pair(g, tag + "perop")) == 0, tag
        assert same(r, O.gemm(pb, m, n, k, A, B, "perop")) == 0, tag
    r = M.gemm(dev(goldens.pair(g, "cgA"), pb, True), dev(goldens.pair(g, "cgB"), pb, True), m, n, k)
    assert same_value(r, goldens.pair(g, "cgperop")) == 0
    # GEMV: y = A x, with a ragged row count
    mm, kk = 67, 45
    A = M.random_real(pb, mm * kk, 3 + pb, -4, 4)
    x = M.random_real(pb, kk, 4 + pb, -4, 4)
    r = M.gemv(A.cuda(), x.cuda(), mm, kk)
    assert same(r, O.gemm(pb, mm, 1, kk, A.pair(), x.pair(), "perop")) == 0


@pytest.mark.parametrize("pb", PBS)
def test_fold_edge_cases(gpu, pb):
    """The GEMV fold adds RNDN(RNDN(+0 + a*1) + b*1) = RNDN(a + b) for every
    structured edge pair (ties, carry-out, cancellation to +0 and to one bit,
    far operands with sticky borrow, zero operands): bit-exact against the
    oracle's k-ordered fold, whose additions are pinned against MPFR."""
    ea, eb, _, _ = edge_arrays(pb, seed=29)
    rows = ea.n
    A = MPArray(pb, np.ascontiguousarray(np.stack([ea.limbs, eb.limbs], 2).reshape(ea.limbs.shape[0], 2 * rows)),
                np.ascontiguousarray(np.stack([ea.se, eb.se], 1).reshape(2 * rows)), False)
    N = M.nlimbs(pb)
    one_l = np.zeros((N, 2), np.uint64)
    one_l[N - 1, :] = np.uint64(1 << 63)
    x = MPArray(pb, one_l, np.array([2, 2], np.int64), False)      # 1.0, 1.0
    r = M.gemv(A.cuda(), x.cuda(), rows, 2)
    assert same(r, O.gemm(pb, rows, 1, 2, A.pair(), x.pair(), "perop")) == 0


@pytest.mark.parametrize("pb", PBS)
def test_gemv_wide_exponents(gpu, pb):
    """Per-op GEMV with exponents in [-150, 150]: far alignments across
    several limbs, sticky-only additions, and cancellations in the fold."""
    mm, kk = 130, 37
    A = M.random_real(pb, mm * kk, 71 + pb, -150, 150)
    x = M.random_real(pb, kk, 72 + pb, -150, 150)
    r = M.gemv(A.cuda(), x.cuda(), mm, kk)
    assert same(r, O.gemm(pb, mm, 1, kk, A.pair(), x.pair(), "perop")) == 0


def test_gemv_partial_wavefront(gpu):
    """GEMV of 200 rows at pb 128 (16 rows per wavefront: 12 full wavefronts
    and one of 8 rows) and k = 33 (two full blocks of 16 steps and one step)."""
    pb, mm, kk = 128, 200, 33
    A = M.random_real(pb, mm * kk, 51, -4, 4)
    x = M.random_real(pb, kk, 52, -4, 4)
    ref = O.gemm(pb, mm, 1, kk, A.pair(), x.pair(), "perop")
    r = M.gemv(A.cuda(), x.cuda(), mm, kk)
    torch.cuda.synchronize()
    assert same(r, ref) == 0


def test_empty_and_single_element(gpu):
    pb = 256
    a = M.random_real(pb, 1, 1).cuda()
    e = M.empty(pb, 0)
    r = M.axpy(a, e, e)
    assert r.n == 0
    x = M.random_real(pb, 1, 2)
    y = M.random_real(pb, 1, 3)
    r = M.axpy(a, x.cuda(), y.cuda())
    assert same(r, O.axpy_real(pb, a.pair(), x.pair(), y.pair())) == 0
    # k = 0 GEMM gives +0 everywhere (acc starts at +0)
    A = M.empty(pb, 0)
    C = M.gemm(A, A, 3, 2, 0)
    h = C.numpy()
    assert np.all(h.limbs == 0) and np.all(h.se == M.ZERO_EXP * 2)


@pytest.mark.parametrize("pb,cplx", [(256, False), (1024, False), (96, True), (512, True)])
def test_pack_unpack_roundtrip(gpu, pb, cplx):
    import ctypes
    import torch
    n = 1000
    x = M.random_array(pb, n, 5, -30, 30, cplx).cuda()
    N = M.nlimbs(pb)
    aos = torch.zeros(((2 if cplx else 1) * n * (2 + N),), dtype=torch.int64, device="cuda")
    L = M.lib()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.mpgfx_unpack(pb, int(cplx), n, *x.ptrs(), aos.data_ptr(), s) == 0
    # AoS struct: {int sign; <pad>; long exp; cu_limb m[N]}
    h = aos.cpu().numpy().reshape(-1, 2 + N)
    xr = x.numpy()
    re_se = xr.se[0] if cplx else xr.se
    first = h[0::2] if cplx else h
    assert np.all(first[:, 1] == (re_se >> 1))
    sign = first[:, 0].astype(np.int64) & 0xFFFFFFFF
    assert np.all(np.where(re_se & 1, sign == 0xFFFFFFFF, sign == 1))
    y = M.empty(pb, n, cplx)
    assert L.mpgfx_pack(pb, int(cplx), n, aos.data_ptr(), *y.ptrs(), s) == 0
    assert M.equal(x, y) == 0


@pytest.mark.parametrize("pb", [32, 64, 96, 128, 192, 256, 320, 384, 512, 1024])
@pytest.mark.parametrize("mm,kk", [(1, 1), (2, 3), (5, 16), (3, 17), (9, 33), (4, 100)])
def test_gemv_small_and_ragged(gpu, pb, mm, kk):
    """GEMV shapes around the fold's blocks: fewer rows than a wavefront's
    groups, k below, at and just above one block of lane-group steps"""
    A = M.random_real(pb, mm * kk, 91 + pb + kk, -40, 40)
    x = M.random_real(pb, kk, 92 + pb + mm, -40, 40)
    r = M.gemv(A.cuda(), x.cuda(), mm, kk)
    assert same(r, O.gemm(pb, mm, 1, kk, A.pair(), x.pair(), "perop")) == 0
