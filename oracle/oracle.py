"""TEST INFRASTRUCTURE ONLY: Python bindings to the oracle.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import
this module.  It wraps

* ``oracle/liboracle.so`` -- the C restatement of the reference arithmetic
  (oracle/mpfixed.c),
* ``oracle/_ref/libref.so`` -- the reference's own header types compiled from
  /root/reference (optional; absent where the reference is not mounted),
* ``oracle/build/pin_mpfr`` -- the MPFR/MPC semantics driver (optional),
* ``oracle/libcpu_mpfr.so`` -- the CPU baseline of bench.py: MPFR/MPC in the
  kernels' operation order, OpenMP over the host's cores (optional).

Arrays use the structure-of-arrays layout shared with the HIP library: a real
array of n values at precision pb is ``limbs`` (uint64, shape (N, n)) plus
``se`` (int64, shape (n,), se = exp*2 + (sign<0)); a complex array is the real
array of its n real parts followed by the real array of its n imaginary parts,
stored as limbs of shape (2, N, n) and se of shape (2, n).
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ZERO_EXP = -0x7FFFFFFFFFFFFF

_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)


def nlimbs(pb):
    return (pb + 63) // 64


def _p(a, kind):
    assert a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous"
    return a.ctypes.data_as(kind)


_LIB = None
_REF = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        arr2 = [_u64p, _i64p]
        L.or_elementwise_real.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64] + arr2 * 3
        L.or_elementwise_complex.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64] + arr2 * 3
        L.or_axpy_real.argtypes = [ctypes.c_int, ctypes.c_int64] + arr2 * 4
        L.or_axpy_complex.argtypes = [ctypes.c_int, ctypes.c_int64] + arr2 * 4
        for f in (L.or_gemm_perop, L.or_gemm_exact, L.or_cgemm_perop):
            f.argtypes = [ctypes.c_int] + [ctypes.c_int64] * 3 + arr2 * 3 + [ctypes.c_int64] * 2
        L.or_sqrt_real.argtypes = [ctypes.c_int, ctypes.c_int64] + arr2 * 2
        L.or_fill_ref.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int, _u64p, _i64p]
        L.or_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libref.so"))


def ref():
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libref.so is not built")
        L = ctypes.CDLL(path)
        arr2 = [_u64p, _i64p]
        L.ref_elementwise_real.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64] + arr2 * 3
        L.ref_elementwise_complex.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64] + arr2 * 3
        L.ref_axpy_real.argtypes = [ctypes.c_int, ctypes.c_int64] + arr2 * 4
        L.ref_axpy_complex.argtypes = [ctypes.c_int, ctypes.c_int64] + arr2 * 4
        L.ref_sqrt_real.argtypes = [ctypes.c_int, ctypes.c_int64] + arr2 * 2
        L.ref_gemm_perop.argtypes = [ctypes.c_int] + [ctypes.c_int64] * 3 + arr2 * 3 + [ctypes.c_int64] * 2
        L.ref_supported.argtypes = [ctypes.c_int]
        L.ref_num_threads.restype = ctypes.c_int
        _REF = L
    return _REF


_CPU = None


def cpu_available():
    return os.path.exists(os.path.join(HERE, "libcpu_mpfr.so"))


def cpu():
    """oracle/libcpu_mpfr.so: timed MPFR/MPC loops (the CPU baseline)."""
    global _CPU
    if _CPU is None:
        path = os.path.join(HERE, "libcpu_mpfr.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/libcpu_mpfr.so is not built (make -C oracle cpu)")
        L = ctypes.CDLL(path)
        arr2 = [_u64p, _i64p]
        for f in (L.cpu_axpy_real, L.cpu_axpy_complex):
            f.argtypes = [ctypes.c_int, ctypes.c_int64] + arr2 * 4
            f.restype = ctypes.c_double
        for f in (L.cpu_gemm_perop, L.cpu_cgemm_perop, L.cpu_gemm_exact):
            f.argtypes = [ctypes.c_int] + [ctypes.c_int64] * 3 + arr2 * 3 + [ctypes.c_int64] * 2
            f.restype = ctypes.c_double
        L.cpu_num_threads.restype = ctypes.c_int
        L.cpu_set_reps.argtypes = [ctypes.c_int]
        L.cpu_versions.restype = ctypes.c_char_p
        _CPU = L
    return _CPU


def cpu_axpy(pb, a, x, y, cplx, reps=1):
    """MPFR (mpfr_mul + mpfr_add) or MPC (mpc_mul + mpc_add) AXPY over the
    whole arrays: (result, milliseconds of the timed loop, averaged over
    `reps` runs of it)"""
    n = x[1].shape[-1]
    cpu().cpu_set_reps(reps)
    ol, ose = (empty_complex if cplx else empty_real)(pb, n)
    f = cpu().cpu_axpy_complex if cplx else cpu().cpu_axpy_real
    ms = f(pb, n, _p(a[0], _u64p), _p(a[1], _i64p), _p(x[0], _u64p), _p(x[1], _i64p),
           _p(y[0], _u64p), _p(y[1], _i64p), _p(ol, _u64p), _p(ose, _i64p))
    return (ol, ose), ms


def cpu_gemm(pb, m, n, k, A, B, mode, rows):
    """MPFR / MPC GEMM on output rows rows=(r0, r1): mode perop | cperop |
    exact; (result with the other rows zero, milliseconds of the timed loop)"""
    r0, r1 = rows
    cpu().cpu_set_reps(1)
    Cl, Cse = (empty_complex if mode == "cperop" else empty_real)(pb, m * n)
    f = {"perop": cpu().cpu_gemm_perop, "cperop": cpu().cpu_cgemm_perop, "exact": cpu().cpu_gemm_exact}[mode]
    ms = f(pb, m, n, k, _p(A[0], _u64p), _p(A[1], _i64p), _p(B[0], _u64p), _p(B[1], _i64p),
           _p(Cl, _u64p), _p(Cse, _i64p), r0, r1)
    return (Cl, Cse), ms


def pin_available():
    return os.path.exists(os.path.join(HERE, "build", "pin_mpfr"))


# ---------------------------------------------------------------- arrays
def empty_real(pb, n):
    return np.zeros((nlimbs(pb), n), np.uint64), np.zeros((n,), np.int64)


def empty_complex(pb, n):
    return np.zeros((2, nlimbs(pb), n), np.uint64), np.zeros((2, n), np.int64)


def fill_ref(pb, n, seed, spread):
    """The reference test generator (tools/test_freal.cu:15-27): xorshift64
    (13,7,17); spread < 0 selects test_freal.cu:55's mix."""
    limbs, se = empty_real(pb, n)
    lib().or_fill_ref(pb, n, seed, spread, _p(limbs, _u64p), _p(se, _i64p))
    return limbs, se


OPS = {"mul": 0, "add": 1, "sub": 2, "div": 3}


def elementwise_real(pb, op, a, b, impl="oracle"):
    (al, ase), (bl, bse) = a, b
    n = ase.shape[0]
    ol, ose = empty_real(pb, n)
    args = (pb, OPS[op], n, _p(al, _u64p), _p(ase, _i64p), _p(bl, _u64p), _p(bse, _i64p),
            _p(ol, _u64p), _p(ose, _i64p))
    if impl == "oracle":
        lib().or_elementwise_real(*args)
    else:
        assert ref().ref_elementwise_real(*args) == 0
    return ol, ose


def sqrt_real(pb, x, impl="oracle"):
    """cu_sqrt per element (faithful; +0 for zero and negative values)."""
    xl, xse = x
    n = xse.shape[0]
    ol, ose = empty_real(pb, n)
    args = (pb, n, _p(xl, _u64p), _p(xse, _i64p), _p(ol, _u64p), _p(ose, _i64p))
    if impl == "oracle":
        lib().or_sqrt_real(*args)
    else:
        assert ref().ref_sqrt_real(*args) == 0
    return ol, ose


def elementwise_complex(pb, op, a, b, impl="oracle"):
    (al, ase), (bl, bse) = a, b
    n = ase.shape[1]
    ol, ose = empty_complex(pb, n)
    args = (pb, OPS[op], n, _p(al, _u64p), _p(ase, _i64p), _p(bl, _u64p), _p(bse, _i64p),
            _p(ol, _u64p), _p(ose, _i64p))
    if impl == "oracle":
        lib().or_elementwise_complex(*args)
    else:
        assert ref().ref_elementwise_complex(*args) == 0
    return ol, ose


def axpy_real(pb, a, x, y, impl="oracle"):
    n = x[1].shape[0]
    ol, ose = empty_real(pb, n)
    args = (pb, n, _p(a[0], _u64p), _p(a[1], _i64p), _p(x[0], _u64p), _p(x[1], _i64p),
            _p(y[0], _u64p), _p(y[1], _i64p), _p(ol, _u64p), _p(ose, _i64p))
    if impl == "oracle":
        lib().or_axpy_real(*args)
    else:
        assert ref().ref_axpy_real(*args) == 0
    return ol, ose


def axpy_complex(pb, a, x, y, impl="oracle"):
    n = x[1].shape[1]
    ol, ose = empty_complex(pb, n)
    args = (pb, n, _p(a[0], _u64p), _p(a[1], _i64p), _p(x[0], _u64p), _p(x[1], _i64p),
            _p(y[0], _u64p), _p(y[1], _i64p), _p(ol, _u64p), _p(ose, _i64p))
    if impl == "oracle":
        lib().or_axpy_complex(*args)
    else:
        assert ref().ref_axpy_complex(*args) == 0
    return ol, ose


def gemm(pb, m, n, k, A, B, mode="perop", rows=None, impl="oracle"):
    """C = A*B with A m x k, B k x n (row-major element order).  mode: perop |
    exact | cperop (complex per-op) | cexact (complex, exact per component).  rows=(r0, r1) limits the computed rows;
    the other rows of the result stay zero."""
    r0, r1 = rows if rows is not None else (0, m)
    if mode == "cexact":
        return _cgemm_exact(pb, m, n, k, A, B, rows)
    if mode == "cperop":
        Cl, Cse = empty_complex(pb, m * n)
    else:
        Cl, Cse = empty_real(pb, m * n)
    args = (pb, m, n, k, _p(A[0], _u64p), _p(A[1], _i64p), _p(B[0], _u64p), _p(B[1], _i64p),
            _p(Cl, _u64p), _p(Cse, _i64p), r0, r1)
    if impl == "ref":
        assert mode == "perop"
        assert ref().ref_gemm_perop(*args) == 0
    elif mode == "perop":
        lib().or_gemm_perop(*args)
    elif mode == "exact":
        lib().or_gemm_exact(*args)
    elif mode == "cperop":
        lib().or_cgemm_perop(*args)
    else:
        raise ValueError(mode)
    return Cl, Cse


def _cgemm_exact(pb, m, n, k, A, B, rows):
    """Complex exact GEMM: Re C = RNDN(sum_k ar*br - ai*bi), Im C = RNDN(sum_k
    ar*bi + ai*br), each component one exact sum rounded once.  Computed as two
    real exact GEMMs over 2k terms: Re = [Ar | Ai] [Br ; -Bi], Im = [Ar | Ai]
    [Bi ; Br] (negation flips the sign bit)."""
    (Al, As), (Bl, Bs) = A, B
    N = nlimbs(pb)
    A2l = np.concatenate([Al[0].reshape(N, m, k), Al[1].reshape(N, m, k)], axis=2).reshape(N, 2 * m * k)
    A2s = np.concatenate([As[0].reshape(m, k), As[1].reshape(m, k)], axis=1).reshape(2 * m * k)
    Bre = (np.concatenate([Bl[0], Bl[1]], axis=1), np.concatenate([Bs[0], Bs[1] ^ 1]))
    Bim = (np.concatenate([Bl[1], Bl[0]], axis=1), np.concatenate([Bs[1], Bs[0]]))
    A2 = (np.ascontiguousarray(A2l), np.ascontiguousarray(A2s))
    re = gemm(pb, m, n, 2 * k, A2, tuple(np.ascontiguousarray(v) for v in Bre), "exact", rows)
    im = gemm(pb, m, n, 2 * k, A2, tuple(np.ascontiguousarray(v) for v in Bim), "exact", rows)
    return np.stack([re[0], im[0]]), np.stack([re[1], im[1]])


# ---------------------------------------------------------------- MPFR driver
def pin(op, pb, arrays, out_shape, dims):
    """Run oracle/build/pin_mpfr OP PB on the given SoA arrays; returns the
    result as (limbs, se) with out_shape = (limb shape, se shape)."""
    exe = os.path.join(HERE, "build", "pin_mpfr")
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            for limbs, se in arrays:
                f.write(np.ascontiguousarray(limbs).tobytes())
                f.write(np.ascontiguousarray(se).tobytes())
        subprocess.run([exe, op, str(pb), fin, fout] + [str(d) for d in dims], check=True)
        raw = np.fromfile(fout, np.uint64)
    ls, ss = out_shape
    nl = int(np.prod(ls))
    return raw[:nl].reshape(ls).copy(), raw[nl:].view(np.int64).reshape(ss).copy()


def canon_zero_sign(limbs, se):
    """Zero compares equal to zero whatever its sign field (the reference's
    tests compare with mpfr_equal_p / mpc_cmp: tools/test_freal.cu:73)."""
    se = se.copy()
    z = (se >> 1) == ZERO_EXP
    se[z] = ZERO_EXP * 2
    return limbs, se


def mismatches(x, y, ignore_zero_sign=True):
    (xl, xs), (yl, ys) = x, y
    if ignore_zero_sign:
        xl, xs = canon_zero_sign(xl, xs)
        yl, ys = canon_zero_sign(yl, ys)
    bad = xs != ys
    bad = bad | np.any(xl != yl, axis=-2)
    return int(np.count_nonzero(bad))
