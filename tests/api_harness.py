"""Harness for the template-API tests (test infrastructure).

Loads tests/api/libapi_host.so (g++ build) or tests/api/libapi_dev.so (hipcc
build for gfx950) -- the same user code of tests/api/api_cases.h, written
against include/mpc_cuda.cuh the way the reference's programs use it -- and
converts between Python value tuples (sign, exp, M, zero), with value =
sign * M * 2^(exp - 64N) as in tests/bigint_ref.py, and the reference's AoS
struct layouts:
  cu_freal<PB>    {int sign; long exp; cu_limb m[N];}             16 + 8N bytes
  cu_wterm<PB>    {int sign; long exp; bool zero; cu_limb M[2N];}  24 + 16N bytes
  cu_fcomplex<PB> {cu_freal<PB> re, im;}
"""
import ctypes
import os
import random

import numpy as np

import bigint_ref as R

HERE = os.path.dirname(os.path.abspath(__file__))
PBS = [32, 64, 96, 256, 512, 1024]
ZERO_EXP = R.ZERO_EXP

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64


def nlimbs(pb):
    return (pb + 63) // 64


def real_dtype(pb):
    return np.dtype([("sign", "<i4"), ("pad", "<i4"), ("exp", "<i8"), ("m", "<u8", (nlimbs(pb),))])


def wterm_dtype(pb):
    return np.dtype([("sign", "<i4"), ("pad", "<i4"), ("exp", "<i8"), ("zero", "u1"), ("pad2", "u1", (7,)),
                     ("M", "<u8", (2 * nlimbs(pb),))])


def cplx_dtype(pb):
    return np.dtype([("re", real_dtype(pb)), ("im", real_dtype(pb))])


def to_aos(vals, pb, arr=None):
    """list of (sign, exp, M, zero) -> structured array (or fill `arr`)"""
    N = nlimbs(pb)
    a = np.zeros(len(vals), real_dtype(pb)) if arr is None else arr
    for i, (s, e, M, z) in enumerate(vals):
        a["sign"][i] = s
        a["exp"][i] = ZERO_EXP if z else e
        for l in range(N):
            a["m"][i, l] = (M >> (64 * l)) & 0xFFFFFFFFFFFFFFFF
    return a


def from_aos(a, pb):
    N = nlimbs(pb)
    out = []
    for i in range(a.shape[0]):
        M = 0
        for l in range(N - 1, -1, -1):
            M = (M << 64) | int(a["m"][i, l])
        e = int(a["exp"][i])
        out.append((int(a["sign"][i]), e, M, e == ZERO_EXP))
    return out


def value(v, pb):
    """exact (numerator, exponent2) of a value tuple"""
    return R.to_exact(v, pb)


def same(got, want):
    """bit-level equality of two value tuples (sign field included)"""
    return tuple(got) == tuple(want)


# ---------------------------------------------------------------- library

class Lib:
    def __init__(self, kind):
        path = os.path.join(HERE, "api", f"libapi_{kind}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (run __graft_entry__.build())")
        self.kind = kind
        L = ctypes.CDLL(path)
        L.api_sizeof_real.argtypes = [ctypes.c_int]
        L.api_sizeof_wterm.argtypes = [ctypes.c_int]
        L.api_arith.argtypes = [ctypes.c_int, ctypes.c_int, _i64] + [_vp] * 5
        L.api_axpy_double.argtypes = [ctypes.c_int, _i64, ctypes.c_double] + [_vp] * 3
        L.api_utils.argtypes = [ctypes.c_int, _i64] + [_vp] * 9
        L.api_convert.argtypes = [ctypes.c_int, _i64] + [_vp] * 6
        L.api_doubles.argtypes = [ctypes.c_int, _i64] + [_vp] * 5
        L.api_bits.argtypes = [ctypes.c_int, _i64] + [_vp] * 4
        L.api_products.argtypes = [ctypes.c_int, _i64] + [_vp] * 4
        L.api_finalize.argtypes = [ctypes.c_int, _i64, ctypes.c_int] + [_vp] * 7
        L.api_terms.argtypes = [ctypes.c_int, _i64] + [_vp] * 10
        L.api_cplx.argtypes = [ctypes.c_int, _i64] + [_vp] * 4
        L.api_cplx_doubles.argtypes = [ctypes.c_int, _i64] + [_vp] * 3
        self.L = L

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{self.kind} {what}: status {rc}")

    def arith(self, pb, A, B, named=True):
        n = len(A)
        a, b = to_aos(A, pb), to_aos(B, pb)
        outs = [np.zeros(n, real_dtype(pb)) for _ in range(3)]
        self._chk(self.L.api_arith(pb, int(named), n, a.ctypes.data, b.ctypes.data,
                                   *[o.ctypes.data for o in outs]), "arith")
        return [from_aos(o, pb) for o in outs]

    def axpy_double(self, pb, a, x, y):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        out = np.zeros_like(x)
        self._chk(self.L.api_axpy_double(pb, len(x), float(a), x.ctypes.data, y.ctypes.data, out.ctypes.data),
                  "axpy_double")
        return out

    def utils(self, pb, A, B):
        n = len(A)
        a, b = to_aos(A, pb), to_aos(B, pb)
        outs = [np.zeros(n, real_dtype(pb)) for _ in range(5)]
        cmp = np.zeros(3 * n, np.int32)
        lr = np.zeros(n, np.int64)
        self._chk(self.L.api_utils(pb, n, a.ctypes.data, b.ctypes.data, *[o.ctypes.data for o in outs],
                                   cmp.ctypes.data, lr.ctypes.data), "utils")
        am, sm, ng, ab, sc = [from_aos(o, pb) for o in outs]
        return dict(addmag=am, submag=sm, neg=ng, abs=ab, scale2=sc, cmpmag=cmp[0::3].tolist(),
                    cmp=cmp[1::3].tolist(), sgn=cmp[2::3].tolist(), lround=lr.tolist())

    def convert(self, pb, A, V):
        n = len(A)
        a = to_aos(A, pb)
        v = np.ascontiguousarray(V, np.int64)
        c32, c64, c1024 = (np.zeros(n, real_dtype(p)) for p in (32, 64, 1024))
        si = np.zeros(n, real_dtype(pb))
        self._chk(self.L.api_convert(pb, n, a.ctypes.data, v.ctypes.data, c32.ctypes.data, c64.ctypes.data,
                                     c1024.ctypes.data, si.ctypes.data), "convert")
        return {32: from_aos(c32, 32), 64: from_aos(c64, 64), 1024: from_aos(c1024, 1024), "si": from_aos(si, pb)}

    def doubles(self, pb, D, A):
        n = len(D)
        d = np.ascontiguousarray(D, np.float64)
        a = to_aos(A, pb)
        fd = np.zeros(n, real_dtype(pb))
        back, td = np.zeros(n), np.zeros(n)
        self._chk(self.L.api_doubles(pb, n, d.ctypes.data, a.ctypes.data, fd.ctypes.data, back.ctypes.data,
                                     td.ctypes.data), "doubles")
        return from_aos(fd, pb), back, td

    def bits(self, pb, A, IX):
        n = len(A)
        a = to_aos(A, pb)
        ix = np.ascontiguousarray(IX, np.int64)
        wb, ws = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
        self._chk(self.L.api_bits(pb, n, a.ctypes.data, ix.ctypes.data, wb.ctypes.data, ws.ctypes.data), "bits")
        return wb.tolist(), ws.tolist()

    def products(self, pb, A, B):
        n, N = len(A), nlimbs(pb)
        a, b = to_aos(A, pb), to_aos(B, pb)
        hl = np.zeros(2 * n, np.uint64)
        r = np.zeros(2 * N * n, np.uint64)
        self._chk(self.L.api_products(pb, n, a.ctypes.data, b.ctypes.data, hl.ctypes.data, r.ctypes.data),
                  "products")
        umul = [(int(hl[2 * i]) << 64) | int(hl[2 * i + 1]) for i in range(n)]
        full = []
        for i in range(n):
            x = 0
            for k in range(2 * N - 1, -1, -1):
                x = (x << 64) | int(r[2 * N * i + k])
            full.append(x)
        return umul, full

    def finalize(self, pb, length, D, MSB, EHI, RB, SB, SG):
        n = len(MSB)
        d = np.ascontiguousarray(np.asarray(D, np.uint64).reshape(n * length))
        arrs = [np.ascontiguousarray(MSB, np.int32), np.ascontiguousarray(EHI, np.int64),
                np.ascontiguousarray(RB, np.uint64), np.ascontiguousarray(SB, np.uint64),
                np.ascontiguousarray(SG, np.int32)]
        out = np.zeros(n, real_dtype(pb))
        self._chk(self.L.api_finalize(pb, n, length, d.ctypes.data, *[x.ctypes.data for x in arrs],
                                      out.ctypes.data), "finalize")
        return from_aos(out, pb)

    def terms(self, pb, A, B, C, D):
        n = len(A)
        ins = [to_aos(x, pb) for x in (A, B, C, D)]
        t1, t2 = np.zeros(n, wterm_dtype(pb)), np.zeros(n, wterm_dtype(pb))
        wc = np.zeros(n, np.int32)
        outs = [np.zeros(n, real_dtype(pb)) for _ in range(3)]
        self._chk(self.L.api_terms(pb, n, *[x.ctypes.data for x in ins], t1.ctypes.data, t2.ctypes.data,
                                   wc.ctypes.data, *[o.ctypes.data for o in outs]), "terms")
        rs, fa, fs = [from_aos(o, pb) for o in outs]
        return _terms_of(t1, pb), _terms_of(t2, pb), wc.tolist(), rs, fa, fs

    def cplx(self, pb, A, B):
        """A, B: lists of (re, im) value-tuple pairs"""
        n = len(A)
        a, b = _to_caos(A, pb), _to_caos(B, pb)
        f3, o3 = np.zeros(3 * n, cplx_dtype(pb)), np.zeros(3 * n, cplx_dtype(pb))
        self._chk(self.L.api_cplx(pb, n, a.ctypes.data, b.ctypes.data, f3.ctypes.data, o3.ctypes.data), "cplx")
        return _from_caos(f3, pb), _from_caos(o3, pb)

    def cplx_doubles(self, pb, D6):
        d = np.ascontiguousarray(D6, np.float64).reshape(-1)
        n = d.shape[0] // 6
        out = np.zeros(2 * n)
        z = np.zeros(n, cplx_dtype(pb))
        self._chk(self.L.api_cplx_doubles(pb, n, d.ctypes.data, out.ctypes.data, z.ctypes.data), "cplx_doubles")
        return out.reshape(n, 2), _from_caos(z, pb)


def _terms_of(t, pb):
    WN = 2 * nlimbs(pb)
    out = []
    for i in range(t.shape[0]):
        M = 0
        for l in range(WN - 1, -1, -1):
            M = (M << 64) | int(t["M"][i, l])
        out.append((int(t["sign"][i]), int(t["exp"][i]), M, bool(t["zero"][i])))
    return out


def _to_caos(Z, pb):
    a = np.zeros(len(Z), cplx_dtype(pb))
    re = to_aos([z[0] for z in Z], pb)
    im = to_aos([z[1] for z in Z], pb)
    a["re"] = re
    a["im"] = im
    return a


def _from_caos(a, pb):
    return list(zip(from_aos(np.ascontiguousarray(a["re"]), pb), from_aos(np.ascontiguousarray(a["im"]), pb)))


# ---------------------------------------------------------------- inputs

def random_vals(pb, n, seed, emin=-40, emax=40, zeros=0.05):
    """random full-significand values (top bit set, low 64N-pb bits clear)"""
    rng = random.Random(seed)
    N = nlimbs(pb)
    out = []
    for _ in range(n):
        if rng.random() < zeros:
            out.append((1, ZERO_EXP, 0, True))
            continue
        m = (1 << (pb - 1)) | rng.getrandbits(pb - 1)
        out.append((rng.choice([1, -1]), rng.randint(emin, emax), m << (64 * N - pb), False))
    return out


# ---------------------------------------------------------------- closed forms

def rndn(num, e2, pb):
    return R.round_exact(num, e2, pb)


def add_ref(x, y, pb):
    """cu_fadd: exact sum rounded once; x + 0 and 0 + x return the operand"""
    return R.add(x, y, pb)


def sub_ref(x, y, pb):
    s, e, M, z = y
    return R.add(x, (-s, e, M, z), pb)


def mul_ref(x, y, pb):
    r = R.mul(x, y, pb)
    return r


def frac_cmp(x, y, pb):
    nx, ex = value(x, pb)
    ny, ey = value(y, pb)
    e = min(ex, ey)
    a, b = nx << (ex - e), ny << (ey - e)
    return (a > b) - (a < b)


def mag_cmp(x, y, pb):
    ax = (1, x[1], x[2], x[3])
    ay = (1, y[1], y[2], y[3])
    return frac_cmp(ax, ay, pb)


def lround_ref(x, pb):
    """nearest integer, ties to even; saturated beyond 2^63"""
    from fractions import Fraction
    num, e2 = value(x, pb)
    if num == 0:
        return 0
    if x[1] > 63:
        return -(1 << 63) if x[0] < 0 else (1 << 63) - 1
    f = Fraction(num) * (Fraction(2) ** e2)
    return round(f)


def to_double_ref(x, pb):
    """RNDN to the nearest double (correct rounding, subnormals included)"""
    from fractions import Fraction
    num, e2 = value(x, pb)
    if num == 0:
        return 0.0
    f = Fraction(num) * (Fraction(2) ** e2)
    try:
        return float(f)
    except OverflowError:
        return float("inf") if num > 0 else float("-inf")


def double_to_val(d, pb):
    """from_double: exact (RNDN below 53 bits); both zeros give +0"""
    if d == 0.0:
        return (1, ZERO_EXP, 0, True)
    n, den = float(d).as_integer_ratio()
    e2 = -(den.bit_length() - 1)
    return rndn(n, e2, pb)
