"""The oracle (oracle/mpfixed.c) pinned before it is trusted:
  1. against the committed MPFR/MPC golden fixtures,
  2. against MPFR/MPC live, on the reference's own test generator and seeds
     (tools/test_freal.cu:92-102, tools/test_fcomplex.cu:197-206),
  3. against the reference implementation compiled from its own header
     (oracle/_ref), bit for bit including the sign field of zeros,
  4. against an independent exact big-integer rounding (tests/bigint_ref.py)
     on structured edge cases, with planted mistakes showing the cases are
     discriminating."""
import numpy as np
import pytest

from oracle import oracle as O
import bigint_ref as BI
import goldens
from mpcases import edge_arrays, unpack, pack

PBS = goldens.PBS


@pytest.mark.parametrize("pb", PBS)
def test_oracle_vs_golden(pb):
    g = goldens.load(pb)
    a, b = goldens.pair(g, "a"), goldens.pair(g, "b")
    for op in ("mul", "add", "sub"):
        assert O.mismatches(O.elementwise_real(pb, op, a, b), goldens.pair(g, op)) == 0, op
    ca, cb = goldens.pair(g, "ca"), goldens.pair(g, "cb")
    for op in ("mul", "add", "sub"):
        assert O.mismatches(O.elementwise_complex(pb, op, ca, cb), goldens.pair(g, "c" + op)) == 0, op
    r = O.axpy_real(pb, goldens.pair(g, "ax"), goldens.pair(g, "xx"), goldens.pair(g, "yy"))
    assert O.mismatches(r, goldens.pair(g, "axpy")) == 0
    r = O.axpy_complex(pb, goldens.pair(g, "cax"), goldens.pair(g, "cxx"), goldens.pair(g, "cyy"))
    assert O.mismatches(r, goldens.pair(g, "caxpy")) == 0
    m, n, k = (int(v) for v in g["dims"])
    for tag in ("g4", "g80"):
        A, B = goldens.pair(g, tag + "A"), goldens.pair(g, tag + "B")
        for mode in ("perop", "exact"):
            assert O.mismatches(O.gemm(pb, m, n, k, A, B, mode), goldens.pair(g, tag + mode)) == 0, (tag, mode)
    r = O.gemm(pb, m, n, k, goldens.pair(g, "cgA"), goldens.pair(g, "cgB"), "cperop")
    assert O.mismatches(r, goldens.pair(g, "cgperop")) == 0
    r = O.gemm(pb, m, n, k, goldens.pair(g, "cgA"), goldens.pair(g, "cgB"), "cexact")
    assert O.mismatches(r, goldens.pair(g, "cgexact")) == 0


@pytest.mark.skipif(not O.pin_available(), reason="GMP/MPFR/MPC not available (oracle/build/pin_mpfr)")
@pytest.mark.parametrize("pb", [64, 96, 256, 512])
def test_oracle_complex_exact_gemm_vs_mpfr(pb):
    """Complex exact GEMM (each component one exact sum, one rounding) against
    MPFR's mpfr_sum over the exact products, on wide exponent spreads and on
    inputs whose real parts cancel exactly (ai = ar, bi = br: Re = 0) or almost
    (bi = br with the last limb changed)."""
    from mpc_cuda_amd import gen as G
    m, n, k = 6, 5, 19
    A = G.random_complex(pb, m * k, 61 + pb, -80, 80)
    B = G.random_complex(pb, k * n, 62 + pb, -80, 80)
    cases = [(A.pair(), B.pair())]
    Al, As = (np.array(v) for v in A.pair())
    Bl, Bs = (np.array(v) for v in B.pair())
    Al[1], As[1] = Al[0], As[0]
    Bl[1], Bs[1] = Bl[0], Bs[0]
    cases.append(((Al.copy(), As.copy()), (Bl.copy(), Bs.copy())))
    Bl[1][0] ^= np.uint64(1 << 40)                  # inside the kept bits at every pb here
    cases.append(((Al, As), (Bl, Bs)))
    for ca, cb in cases:
        ca = tuple(np.ascontiguousarray(v) for v in ca)
        cb = tuple(np.ascontiguousarray(v) for v in cb)
        ref = O.pin("cgemm_exact", pb, [ca, cb], ((2, O.nlimbs(pb), m * n), (2, m * n)), [m, n, k])
        assert O.mismatches(O.gemm(pb, m, n, k, ca, cb, "cexact"), ref) == 0


REF_SEEDS = {32: 0x1111, 64: 0x2222, 96: 0x3333, 128: 0x4444, 160: 0x5555, 256: 0x6666, 288: 0x7777,
             512: 0x8888, 1024: 0x9999, 1056: 0xAAAA, 2048: 0xBBBB}


def _ref_pairs(pb, M, seed):
    """A[i], B[i] drawn alternately from one xorshift stream with the spread
    mix (i&3) ? 40 : pb+200, as tools/test_freal.cu:55 does."""
    l, s = O.fill_ref(pb, 2 * M, seed, -1)
    # fill_ref applies the spread mix by value index; re-draw to match pair index
    a = (np.ascontiguousarray(l[:, 0::2]), np.ascontiguousarray(s[0::2]))
    b = (np.ascontiguousarray(l[:, 1::2]), np.ascontiguousarray(s[1::2]))
    return a, b


@pytest.mark.skipif(not O.pin_available(), reason="GMP/MPFR/MPC not available (oracle/build/pin_mpfr)")
@pytest.mark.parametrize("pb", sorted(REF_SEEDS))
def test_oracle_vs_mpfr_reference_generator(pb):
    M = 20000 if pb <= 1056 else 8000
    a, b = _ref_pairs(pb, M, REF_SEEDS[pb])
    for op in ("mul", "add", "sub"):
        ref = O.pin(op, pb, [a, b], ((O.nlimbs(pb), M), (M,)), [M])
        assert O.mismatches(O.elementwise_real(pb, op, a, b), ref) == 0, op
    Mc = M // 4
    def cpl(x):
        return (np.ascontiguousarray(np.stack([x[0][:, :Mc], x[0][:, Mc:2 * Mc]])),
                np.ascontiguousarray(np.stack([x[1][:Mc], x[1][Mc:2 * Mc]])))
    ca, cb = cpl(a), cpl(b)
    for op in ("mul", "add", "sub"):
        ref = O.pin("c" + op, pb, [ca, cb], ((2, O.nlimbs(pb), Mc), (2, Mc)), [Mc])
        assert O.mismatches(O.elementwise_complex(pb, op, ca, cb), ref) == 0, op


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("pb", PBS)
def test_oracle_vs_reference_build(pb):
    """Bit-exact, zero-sign fields included, against the reference's own code."""
    M = 6000
    a, b = _ref_pairs(pb, M, 0xC0FFEE + pb)
    ea, eb, _, _ = edge_arrays(pb, seed=3)
    a = (np.ascontiguousarray(np.concatenate([a[0], ea.limbs], 1)), np.ascontiguousarray(np.concatenate([a[1], ea.se])))
    b = (np.ascontiguousarray(np.concatenate([b[0], eb.limbs], 1)), np.ascontiguousarray(np.concatenate([b[1], eb.se])))
    for op in ("mul", "add", "sub"):
        o = O.elementwise_real(pb, op, a, b)
        r = O.elementwise_real(pb, op, a, b, impl="ref")
        assert O.mismatches(o, r, ignore_zero_sign=False) == 0, op
    n = a[1].shape[0] // 2
    ca = (np.ascontiguousarray(np.stack([a[0][:, :n], b[0][:, n:2 * n]])), np.ascontiguousarray(np.stack([a[1][:n], b[1][n:2 * n]])))
    cb = (np.ascontiguousarray(np.stack([b[0][:, :n], a[0][:, n:2 * n]])), np.ascontiguousarray(np.stack([b[1][:n], a[1][n:2 * n]])))
    for op in ("mul", "add", "sub"):
        o = O.elementwise_complex(pb, op, ca, cb)
        r = O.elementwise_complex(pb, op, ca, cb, impl="ref")
        assert O.mismatches(o, r, ignore_zero_sign=False) == 0, op
    # AXPY and per-op GEMM orders
    x = O.fill_ref(pb, 512, 5 + pb, 8)
    y = O.fill_ref(pb, 512, 6 + pb, 8)
    al = O.fill_ref(pb, 1, 7 + pb, 8)
    assert O.mismatches(O.axpy_real(pb, al, x, y), O.axpy_real(pb, al, x, y, impl="ref"), False) == 0
    A = O.fill_ref(pb, 6 * 40, 8 + pb, 4)
    B = O.fill_ref(pb, 40 * 3, 9 + pb, 4)
    assert O.mismatches(O.gemm(pb, 6, 3, 40, A, B), O.gemm(pb, 6, 3, 40, A, B, impl="ref"), False) == 0


EDGE_PBS = [32, 64, 96, 160, 256, 512, 1056]


@pytest.mark.parametrize("pb", EDGE_PBS)
def test_oracle_vs_bigint_edge_cases(pb):
    A, B, la, lb = edge_arrays(pb, seed=11, n_random=48)
    for op, f in (("mul", BI.mul), ("add", BI.add), ("sub", BI.sub)):
        got = unpack(_wrap(pb, O.elementwise_real(pb, op, A.pair(), B.pair())))
        want = [f(x, y, pb) for x, y in zip(la, lb)]
        assert got == want, op


def test_oracle_vs_bigint_complex_and_dots():
    for pb in (64, 96, 256):
        A, B, la, lb = edge_arrays(pb, seed=5, n_random=32)
        n = len(la) // 2
        ca = _cplx(pb, la[:n], lb[n:2 * n])
        cb = _cplx(pb, lb[:n], la[n:2 * n])
        got = O.elementwise_complex(pb, "mul", ca, cb)
        gre = unpack(_wrap(pb, (np.ascontiguousarray(got[0][0]), np.ascontiguousarray(got[1][0]))))
        gim = unpack(_wrap(pb, (np.ascontiguousarray(got[0][1]), np.ascontiguousarray(got[1][1]))))
        for i in range(n):
            wr, wi = BI.cmul(la[i], lb[n + i], lb[i], la[n + i], pb)
            assert gre[i] == wr and gim[i] == wi, (pb, i)
        # dots: per-op fold and exact sum, one row of length n
        Ar, Br = pack(la[:n], pb), pack(lb[:n], pb)
        per = unpack(_wrap(pb, O.gemm(pb, 1, 1, n, Ar.pair(), Br.pair(), "perop")))
        ex = unpack(_wrap(pb, O.gemm(pb, 1, 1, n, Ar.pair(), Br.pair(), "exact")))
        assert per[0] == BI.fold_dot(la[:n], lb[:n], pb)
        assert ex[0] == BI.sum_products(list(zip(la[:n], lb[:n])), pb)


def test_edge_cases_catch_planted_rounding_mistakes():
    """Each plausible rounding mistake, planted in a copy of the big-integer
    rounding, disagrees with the oracle on the edge-case set."""
    import bigint_ref

    def make(mutation):
        def round_exact(num, e2, pb):
            if num == 0:
                return (1, BI.ZERO_EXP, 0, True)
            s = -1 if num < 0 else 1
            a = abs(num)
            bl = a.bit_length()
            if bl > pb:
                sh = bl - pb
                q, r = a >> sh, a & ((1 << sh) - 1)
                half = 1 << (sh - 1)
                if mutation == "half_away":
                    up = r >= half
                elif mutation == "truncate":
                    up = False
                elif mutation == "no_sticky":
                    up = (r >> (sh - 1)) & 1 and ((q & 1) == 1)
                else:
                    up = r > half or (r == half and (q & 1))
                if up:
                    q += 1
                if q.bit_length() > pb:
                    q >>= 1
                    sh += 1 if mutation != "carry_exp" else 0
                a, e2 = q, e2 + sh
            exp = e2 + a.bit_length()
            M = a << (64 * BI.nlimbs(pb) - a.bit_length())
            return (s, exp, M, False)
        return round_exact

    pb = 64
    A, B, la, lb = edge_arrays(pb, seed=2, n_random=64)
    got = {op: unpack(_wrap(pb, O.elementwise_real(pb, op, A.pair(), B.pair()))) for op in ("mul", "add", "sub")}
    orig = bigint_ref.round_exact
    try:
        for mutation in ("half_away", "truncate", "no_sticky", "carry_exp"):
            bigint_ref.round_exact = make(mutation)
            diff = 0
            for op, f in (("mul", BI.mul), ("add", BI.add), ("sub", BI.sub)):
                diff += sum(g != f(x, y, pb) for g, x, y in zip(got[op], la, lb))
            assert diff > 0, mutation
    finally:
        bigint_ref.round_exact = orig


def _wrap(pb, pair):
    from mpc_cuda_amd.arrays import MPArray
    return MPArray(pb, pair[0], pair[1], False)


def _cplx(pb, re, im):
    r, i = pack(re, pb), pack(im, pb)
    return (np.ascontiguousarray(np.stack([r.limbs, i.limbs])), np.ascontiguousarray(np.stack([r.se, i.se])))
