"""Host arrays of MPFR numbers <-> device MP arrays (mpgfx_import_mpfr /
mpgfx_export_mpfr).  `MpfrStruct` is MPFR's public value struct
(__mpfr_struct in mpfr.h: precision, sign, exponent, limb pointer); a C caller
passes its mpfr_t / mpc_t arrays directly, this module is the same boundary
for Python callers that hold such arrays (for example through ctypes)."""
import ctypes

from . import _lib
from .arrays import empty


class MpfrStruct(ctypes.Structure):
    _fields_ = [("prec", ctypes.c_long), ("sign", ctypes.c_int), ("exp", ctypes.c_long),
                ("d", ctypes.POINTER(ctypes.c_uint64))]


MPFR_STRUCT_SIZE = ctypes.sizeof(MpfrStruct)


def import_mpfr(pb, n, addr, stride=MPFR_STRUCT_SIZE, im_addr=None):
    """Device MP array of the n MPFR values at host address `addr` (structs
    `stride` bytes apart); with im_addr, a complex array whose imaginary parts
    are the structs at im_addr (for an mpc_t array: addr = &z[0].re,
    im_addr = &z[0].im, stride = sizeof(mpc_t))."""
    cplx = im_addr is not None
    out = empty(pb, n, cplx)
    L = _lib.lib()
    (rl, rs) = (out.limbs[0], out.se[0]) if cplx else (out.limbs, out.se)
    _lib.check("mpgfx_import_mpfr", L.mpgfx_import_mpfr(pb, n, addr, stride, rl.data_ptr(), rs.data_ptr(), None))
    if cplx:
        _lib.check("mpgfx_import_mpfr", L.mpgfx_import_mpfr(pb, n, im_addr, stride, out.limbs[1].data_ptr(),
                                                            out.se[1].data_ptr(), None))
    return out


def export_mpfr(arr, addr, stride=MPFR_STRUCT_SIZE, im_addr=None):
    """Write a device MP array into n MPFR structs at `addr` (precision pb,
    limb buffers allocated); complex arrays also need im_addr."""
    L = _lib.lib()
    if arr.cplx:
        if im_addr is None:
            raise ValueError("complex array: im_addr required")
        parts = [(arr.limbs[0], arr.se[0], addr), (arr.limbs[1], arr.se[1], im_addr)]
    else:
        parts = [(arr.limbs, arr.se, addr)]
    for l, s, a in parts:
        _lib.check("mpgfx_export_mpfr", L.mpgfx_export_mpfr(arr.pb, arr.n, l.data_ptr(), s.data_ptr(), a, stride, None))
