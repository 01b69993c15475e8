"""The drop-in template API (include/mpc_cuda/cu_freal.cuh, cu_fcomplex.cuh)
exercised the way the reference's users call it: user code in
tests/api/api_cases.h -- AoS kern<PB> kernels (tools/test_freal.cu:41-48,
tools/test_fcomplex.cu), the doubles-in/doubles-out AXPY of
demos/sample_fixed.cu:20-30, and every public function and member -- built
twice: by g++ for the host (CPU tests) and by hipcc for gfx950 (GPU tests).
Every result is checked bit for bit (sign field included) against exact
closed forms: Python integer arithmetic rounded once to nearest-even
(tests/bigint_ref.py), exact comparisons, Python's correctly rounded
integer-to-float division for to_double, round-half-even for cu_lround.
"""
import random

import pytest

import api_harness as H
import bigint_ref as R
from mpcases import edge_pairs

KINDS = [pytest.param("host", id="host"), pytest.param("dev", id="gpu", marks=pytest.mark.gpu)]
ZERO = (1, H.ZERO_EXP, 0, True)


@pytest.fixture(scope="module", params=KINDS)
def lib(request):
    if request.param == "dev":
        import torch
        if not torch.cuda.is_available():
            pytest.fail("GPU test run without a visible GPU")
    return H.Lib(request.param)


def operands(pb, n=300, seed=0):
    A, B = edge_pairs(pb, seed=seed + 3, n_random=40)
    RA = H.random_vals(pb, n, 1000 * pb + seed)
    RB = H.random_vals(pb, n, 2000 * pb + seed)
    # near operands (cancellation) and far operands (sticky paths)
    for i in range(0, n, 3):
        s, e, M, z = RA[i]
        if not z:
            RB[i] = (-s, e + random.Random(i).choice([0, 0, 1, -1, -pb, -2 * pb]), M if RB[i][3] else RB[i][2], False)
    return A + RA, B + RB


@pytest.mark.parametrize("pb", H.PBS)
def test_struct_sizes(lib, pb):
    """AoS layouts of the reference (cu_freal.cuh:75-77, cu_fcomplex.cuh:25-28)"""
    N = H.nlimbs(pb)
    assert lib.L.api_sizeof_real(pb) == 16 + 8 * N == H.real_dtype(pb).itemsize
    assert lib.L.api_sizeof_wterm(pb) == 24 + 16 * N == H.wterm_dtype(pb).itemsize


@pytest.mark.parametrize("pb", H.PBS)
@pytest.mark.parametrize("named", [True, False], ids=["functions", "operators"])
def test_arith_kernel(lib, pb, named):
    """kern<PB> of tools/test_freal.cu:41-48 (cu_fmul / cu_fadd / cu_fsub) and
    the same with operators * + -"""
    A, B = operands(pb)
    mu, ad, su = lib.arith(pb, A, B, named)
    for i, (a, b) in enumerate(zip(A, B)):
        assert H.same(mu[i], R.mul(a, b, pb)), (i, "mul")
        assert H.same(ad[i], R.add(a, b, pb)), (i, "add")
        assert H.same(su[i], H.sub_ref(a, b, pb)), (i, "sub")


@pytest.mark.parametrize("pb", H.PBS)
def test_sample_fixed_axpy(lib, pb):
    """demos/sample_fixed.cu:20-30: cu_freal<PB> fa = a; r = fa * x + y with
    operators; (double) r.  Each step rounded to pb bits, the last to double."""
    rng = random.Random(pb)
    x = [1.0 + i * 0.25 for i in range(8)] + [rng.uniform(-1, 1) * 2.0 ** rng.randint(-60, 60) for _ in range(200)]
    y = [2.0 - i * 0.1 for i in range(8)] + [rng.uniform(-1, 1) * 2.0 ** rng.randint(-60, 60) for _ in range(200)]
    x += [1.0, -1.0, 0.0, 3.0]
    y += [-1.5, 1.5, 0.0, -4.5]
    a = 1.5
    out = lib.axpy_double(pb, a, x, y)
    fa = H.double_to_val(a, pb)
    for i in range(len(x)):
        t = R.mul(fa, H.double_to_val(x[i], pb), pb)
        r = R.add(t, H.double_to_val(y[i], pb), pb)
        assert out[i] == H.to_double_ref(r, pb), i
    if pb >= 64:
        assert out[0] == 3.5 and out[7] == 1.5 * 2.75 + (2.0 - 0.7)


@pytest.mark.parametrize("pb", H.PBS)
def test_utilities(lib, pb):
    """cu_addmag, cu_submag, cu_neg, cu_abs, cu_scale2, cu_cmpmag, cu_cmp,
    cu_sgn, cu_lround (cu_freal.cuh:215-277, 294-356)"""
    A, B = operands(pb, seed=5)
    rng = random.Random(7 + pb)
    # integers, ties (k + 1/2), values near 2^62 and beyond 2^63 for cu_lround
    for k in range(60):
        v = rng.choice([rng.randint(-10, 10), rng.randint(-2 ** 40, 2 ** 40)])
        num, e2 = 2 * v + rng.choice([1, -1, 0]), -1
        A.append(H.rndn(num, e2, pb) if num else ZERO)
        B.append(H.random_vals(pb, 1, k)[0])
    for e in (61, 62, 63, 64, 70):
        A.append(H.random_vals(pb, 1, e, e, e, 0)[0])
        B.append(A[-1])
    r = lib.utils(pb, A, B)
    for i, (a, b) in enumerate(zip(A, B)):
        s = -1 if i & 1 else 1
        na, ea = H.value(a, pb)
        nb, eb = H.value(b, pb)
        mag = R.round_exact(abs(na) * 2 ** (ea - min(ea, eb)) + abs(nb) * 2 ** (eb - min(ea, eb)), min(ea, eb), pb) \
            if (na or nb) else ZERO
        exp_am = (s, mag[1], mag[2], mag[3])
        assert H.same(r["addmag"][i], exp_am), (i, "addmag")
        cm = H.mag_cmp(a, b, pb)
        big, small = (a, b) if cm >= 0 else (b, a)
        nbig, ebig = H.value(big, pb)
        nsm, esm = H.value(small, pb)
        e0 = min(ebig, esm)
        d = abs(nbig) * 2 ** (ebig - e0) - abs(nsm) * 2 ** (esm - e0)
        rd = R.round_exact(d, e0, pb)
        exp_sm = rd if rd[3] else (s, rd[1], rd[2], rd[3])
        assert H.same(r["submag"][i], exp_sm), (i, "submag")
        zero_a = a[3]
        assert H.same(r["neg"][i], a if zero_a else (-a[0], a[1], a[2], a[3])), (i, "neg")
        assert H.same(r["abs"][i], a if zero_a else (1, a[1], a[2], a[3])), (i, "abs")
        k = (i % 41) - 20
        assert H.same(r["scale2"][i], a if zero_a else (a[0], a[1] + k, a[2], a[3])), (i, "scale2")
        assert r["cmpmag"][i] == cm, (i, "cmpmag")
        assert r["cmp"][i] == H.frac_cmp(a, b, pb), (i, "cmp")
        assert r["sgn"][i] == (0 if zero_a else a[0]), (i, "sgn")
        assert r["lround"][i] == H.lround_ref(a, pb), (i, "lround", a[1])


@pytest.mark.parametrize("pb", H.PBS)
def test_convert_and_from_si(lib, pb):
    """cu_convert<TO,FROM> to 32, 64 and 1024 bits (RNDN, exact when
    widening) and cu_from_si"""
    A = H.random_vals(pb, 250, 31 + pb) + edge_pairs(pb, seed=2, n_random=0)[0]
    rng = random.Random(pb)
    V = [0, 1, -1, 2 ** 63 - 1, -2 ** 63, 2 ** 32 + 1, -(2 ** 33) - 3, (1 << 40) | 1]
    V += [rng.randint(-2 ** 63, 2 ** 63 - 1) >> rng.randint(0, 62) for _ in range(len(A) - len(V))]
    r = lib.convert(pb, A, V)
    for i, a in enumerate(A):
        num, e2 = H.value(a, pb)
        for to in (32, 64, 1024):
            want = H.rndn(num, e2, to) if num else ZERO
            assert H.same(r[to][i], want), (i, to)
        want = H.rndn(V[i], 0, pb) if V[i] else ZERO
        assert H.same(r["si"][i], want), (i, "from_si", V[i])


@pytest.mark.parametrize("pb", H.PBS)
def test_doubles(lib, pb):
    """from_double (and the double constructor): exact injection, RNDN below
    53 bits, subnormals included; to_double (and the cast): RNDN to the
    nearest double, once, subnormal and overflowing values included"""
    rng = random.Random(pb + 99)
    D = [0.0, -0.0, 1.0, -0.75, 3.625, 2.0 ** -1074, -(2.0 ** -1060) * 3, 2.0 ** 1023 * 1.75, 1e-310, 1e308]
    D += [rng.uniform(-1, 1) * 2.0 ** rng.randint(-1080, 1020) for _ in range(300)]
    A = H.random_vals(pb, len(D) - 12, 5 + pb, -1100, 1100)
    A += [H.random_vals(pb, 1, 77 + e, e, e, 0)[0] for e in (-1021, -1022, -1023, -1073, -1074, -1075, 1024, 1025)]
    N = H.nlimbs(pb)
    half_min = (1, -1074, 1 << (64 * N - 1), False)          # exactly 2^-1075: ties to 0
    just_above = (1, -1074, (1 << (64 * N - 1)) | (1 << (64 * N - pb)), False)
    max_mid = (1, 1024, ((1 << 53) - 1) << (64 * N - 53), False)
    A += [half_min, just_above, max_mid, ZERO][: len(D) - len(A)]
    while len(A) < len(D):
        A.append(ZERO)
    fd, back, td = lib.doubles(pb, D, A)
    for i, d in enumerate(D):
        want = H.double_to_val(d, pb)
        assert H.same(fd[i], want), (i, d)
        assert back[i] == H.to_double_ref(want, pb), (i, d)
        assert td[i] == H.to_double_ref(A[i], pb), (i, A[i][1])
    if pb >= 64:
        assert back[4] == 3.625 and back[5] == 2.0 ** -1074


@pytest.mark.parametrize("pb", H.PBS)
def test_bit_helpers(lib, pb):
    """cu_wbit and cu_wsticky_below (cu_freal.cuh:52-64) on significand limbs"""
    N = H.nlimbs(pb)
    A = H.random_vals(pb, 400, 3 + pb, zeros=0.0)
    rng = random.Random(pb)
    IX = [rng.randint(-3, 64 * N + 3) for _ in A]
    IX[:6] = [0, 1, 64 * N - 1, 64 * N, -1, 64 * N - pb]
    wb, ws = lib.bits(pb, A, IX)
    for i, (a, ix) in enumerate(zip(A, IX)):
        M = a[2]
        assert wb[i] == ((M >> ix) & 1 if 0 <= ix < 64 * N else 0), (i, ix)
        assert ws[i] == (1 if ix > 0 and (M & ((1 << min(ix, 64 * N)) - 1)) else 0), (i, ix)


@pytest.mark.parametrize("pb", H.PBS)
def test_products(lib, pb):
    """cu_umul (64 x 64 -> 128) and cu_mul_full<PB> (N x N -> 2N limbs)"""
    A = H.random_vals(pb, 300, 11 + pb, zeros=0.0)
    B = H.random_vals(pb, 300, 12 + pb, zeros=0.0)
    A[0] = (1, 0, (1 << (64 * H.nlimbs(pb))) - 1 - ((1 << (64 * H.nlimbs(pb) - pb)) - 1), False)
    B[0] = A[0]
    umul, full = lib.products(pb, A, B)
    mask = (1 << 64) - 1
    for i, (a, b) in enumerate(zip(A, B)):
        assert umul[i] == (a[2] & mask) * (b[2] & mask), i
        assert full[i] == a[2] * b[2], i


@pytest.mark.parametrize("pb", H.PBS)
def test_finalize(lib, pb):
    """cu_freal<PB>::finalize (cu_freal.cuh:83-138): round a LEN-limb
    magnitude with its top bit at `msb`, worth 2^Ehi, and round / sticky
    material below bit 0 (not consulted when fewer than PB bits remain)"""
    rng = random.Random(17 + pb)
    N = H.nlimbs(pb)
    L = 2 * N + 3
    D, MSB, EHI, RB, SB, SG, ints = [], [], [], [], [], [], []
    for t in range(400):
        kind = t % 5
        msb = {0: L * 64 - 1, 1: rng.randint(0, L * 64 - 1), 2: pb - 1, 3: pb, 4: rng.randint(0, pb + 40)}[kind]
        msb = min(msb, L * 64 - 1)
        x = (1 << msb) | rng.getrandbits(msb) if msb > 0 else 1
        if t % 7 == 0 and msb >= pb:           # a tie or an all-ones run below the kept bits
            keep = msb - pb + 1
            x = (x >> keep << keep) | (1 << (keep - 1)) if keep >= 1 else x
        if t == 1:
            x, msb = 0, -1
        rb, sb = rng.randint(0, 1), rng.choice([0, 0, 1, rng.getrandbits(64)])
        sg = rng.choice([1, -1])
        D.append([(x >> (64 * k)) & ((1 << 64) - 1) for k in range(L)])
        MSB.append(msb); EHI.append(rng.randint(-100, 100)); RB.append(rb); SB.append(sb); SG.append(sg)
        ints.append(x)
    out = lib.finalize(pb, L, D, MSB, EHI, RB, SB, SG)
    for i in range(len(MSB)):
        msb, x = MSB[i], ints[i]
        if msb < 0:
            want = ZERO
        elif msb + 1 - pb >= 0:
            # extended value: x, then the round bit, then a sticky bit
            num = (x << 2) | (RB[i] << 1) | (1 if SB[i] else 0)
            want = H.rndn(SG[i] * num, EHI[i] - msb - 2, pb)
        else:
            want = H.rndn(SG[i] * x, EHI[i] - msb, pb)
        assert H.same(out[i], want), (i, msb)


@pytest.mark.parametrize("pb", H.PBS)
def test_wide_terms(lib, pb):
    """cu_mk_term, cu_wcmp, cu_round_sum2 (cu_fcomplex.cuh:30-88), cu_fmma,
    cu_fmms (91-96): exact normalised products, their comparison, and the
    exact sum (difference) rounded once"""
    A, B = operands(pb, n=150, seed=1)
    C, D = operands(pb, n=150, seed=2)
    # planted exact cancellation of the two products: c*d = -a*b
    for i in range(0, len(A), 9):
        C[i], D[i] = (-A[i][0], A[i][1], A[i][2], A[i][3]), B[i]
    t1, t2, wc, rs, fa, fs = lib.terms(pb, A, B, C, D)
    N = H.nlimbs(pb)
    for i in range(len(A)):
        for t, (x, y) in ((t1, (A[i], B[i])), (t2, (C[i], D[i]))):
            if x[3] or y[3]:
                want = (1, 0, 0, True)
            else:
                P = x[2] * y[2]
                sh = 0 if P >> (128 * N - 1) else 1
                want = (x[0] * y[0], x[1] + y[1] - sh, P << sh, False)
            assert tuple(t[i]) == want, (i, "mk_term")
        k1, k2 = (t1[i][1], t1[i][2]), (t2[i][1], t2[i][2])
        assert wc[i] == (k1 > k2) - (k1 < k2), (i, "wcmp")
        assert H.same(fa[i], R.sum_products([(A[i], B[i]), (C[i], D[i])], pb)), (i, "fmma")
        nC = (-C[i][0], C[i][1], C[i][2], C[i][3])
        assert H.same(fs[i], R.sum_products([(A[i], B[i]), (nC, D[i])], pb)), (i, "fmms")
        assert H.same(rs[i], fa[i]), (i, "round_sum2")


@pytest.mark.parametrize("pb", H.PBS)
def test_complex_kernel(lib, pb):
    """complex kern<PB> (tools/test_fcomplex.cu): cu_cmul / cu_cadd / cu_csub
    and the operators, MPC_RNDNN per component"""
    A, B = operands(pb, n=120, seed=4)
    C, D = operands(pb, n=120, seed=6)
    ZA = list(zip(A, C))
    ZB = list(zip(B, D))
    f3, o3 = lib.cplx(pb, ZA, ZB)
    for i, ((ar, ai), (br, bi)) in enumerate(zip(ZA, ZB)):
        want = [R.cmul(ar, ai, br, bi, pb), (R.add(ar, br, pb), R.add(ai, bi, pb)),
                (H.sub_ref(ar, br, pb), H.sub_ref(ai, bi, pb))]
        for j in range(3):
            for got in (f3[3 * i + j], o3[3 * i + j]):
                assert H.same(got[0], want[j][0]) and H.same(got[1], want[j][1]), (i, j)


@pytest.mark.parametrize("pb", H.PBS)
def test_complex_from_doubles(lib, pb):
    """cu_fcomplex(double, double), from_doubles, real_d / imag_d and
    z = a*x + y; the known answer of demos/sample_fixed.cu:61-66 is
    (1.5-0.25i)*(1+0.5i)+(2-1i) = (3.625, -0.5)"""
    rng = random.Random(pb + 5)
    rows = [[1.5, -0.25, 1.0, 0.5, 2.0, -1.0]]
    rows += [[rng.uniform(-1, 1) * 2.0 ** rng.randint(-30, 30) for _ in range(6)] for _ in range(150)]
    out, Z = lib.cplx_doubles(pb, rows)
    for i, r in enumerate(rows):
        a = [H.double_to_val(v, pb) for v in r]
        pr = R.cmul(a[0], a[1], a[2], a[3], pb)
        z = (R.add(pr[0], a[4], pb), R.add(pr[1], a[5], pb))
        assert H.same(Z[i][0], z[0]) and H.same(Z[i][1], z[1]), i
        assert out[i, 0] == H.to_double_ref(z[0], pb) and out[i, 1] == H.to_double_ref(z[1], pb), i
    if pb >= 64:
        assert out[0, 0] == 3.625 and out[0, 1] == -0.5
