"""The MPFR value struct that mpgfx_import_mpfr / mpgfx_export_mpfr read and
write (mpc_cuda_amd.interop.MpfrStruct, mpc_cuda_amd/csrc/interop.cpp) pinned
against the MPFR library itself: for values set through MPFR's own API, the
struct fields read back as precision, sign, MPFR exponent and the significand
limbs of the same value in the MP-array format; MPFR's zero / NaN / infinity
exponents are the constants interop.cpp uses.  CPU only (needs libmpfr, here
under /opt/conda)."""
import ctypes
import os

import numpy as np
import pytest

from mpc_cuda_amd import gen as G
from mpc_cuda_amd.interop import MpfrStruct
from mpc_cuda_amd.arrays import nlimbs

LIB = "/opt/conda/lib/libmpfr.so"
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libmpfr not available")
LONG_MAX = 2 ** 63 - 1


def _mpfr():
    m = ctypes.CDLL(LIB)
    p = ctypes.POINTER(MpfrStruct)
    m.mpfr_init2.argtypes = [p, ctypes.c_long]
    m.mpfr_set_str.argtypes = [p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    m.mpfr_get_exp.argtypes = [p]
    m.mpfr_get_exp.restype = ctypes.c_long
    m.mpfr_signbit.argtypes = [p]
    m.mpfr_set_zero.argtypes = [p, ctypes.c_int]
    m.mpfr_set_nan.argtypes = [p]
    m.mpfr_set_inf.argtypes = [p, ctypes.c_int]
    m.mpfr_clear.argtypes = [p]
    return m


@pytest.mark.parametrize("pb", [64, 96, 256, 1024])
def test_struct_fields_match_mpfr(pb):
    m = _mpfr()
    N = nlimbs(pb)
    a = G.random_real(pb, 16, 7 + pb, -300, 300)
    for i in range(a.n):
        x = MpfrStruct()
        m.mpfr_init2(ctypes.byref(x), pb)
        M = 0
        for l in range(N):
            M |= int(a.limbs[l, i]) << (64 * l)
        e = int(a.se[i]) >> 1
        neg = int(a.se[i]) & 1
        s = ("-" if neg else "") + "0x" + format(M, "x") + "p" + str(e - 64 * N)
        assert m.mpfr_set_str(ctypes.byref(x), s.encode(), 0, 0) == 0      # exact at precision pb
        assert x.prec == pb
        assert x.sign == (-1 if neg else 1) and bool(m.mpfr_signbit(ctypes.byref(x))) == bool(neg)
        assert x.exp == m.mpfr_get_exp(ctypes.byref(x)) == e
        assert [x.d[l] for l in range(N)] == [int(a.limbs[l, i]) for l in range(N)]
        m.mpfr_clear(ctypes.byref(x))


def test_special_exponents():
    m = _mpfr()
    x = MpfrStruct()
    m.mpfr_init2(ctypes.byref(x), 128)
    m.mpfr_set_zero(ctypes.byref(x), -1)
    assert x.exp == 0 - LONG_MAX and x.sign == -1
    m.mpfr_set_nan(ctypes.byref(x))
    assert x.exp == 1 - LONG_MAX
    m.mpfr_set_inf(ctypes.byref(x), 1)
    assert x.exp == 2 - LONG_MAX
    m.mpfr_clear(ctypes.byref(x))
