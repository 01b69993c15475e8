"""mpgfx_import_mpfr / mpgfx_export_mpfr on the device: arrays of MPFR value
structs (layout pinned against libmpfr in test_interop_layout.py) built from
seeded MP arrays import bit-for-bit, export back bit-for-bit, mpc_t-strided
complex arrays split into real and imaginary parts, zeros keep their sign, and
NaN / infinity / precision mismatches are rejected."""
import ctypes

import numpy as np
import pytest

from mpc_cuda_amd import gen as G
from mpc_cuda_amd import _lib
from mpc_cuda_amd.arrays import MPArray, nlimbs, ZERO_EXP
from mpc_cuda_amd.interop import MpfrStruct, import_mpfr, export_mpfr, MPFR_STRUCT_SIZE

pytestmark = pytest.mark.gpu
LONG_MAX = 2 ** 63 - 1


def _structs(pb, count, stride_structs=1):
    """count value slots, each `stride_structs` structs apart, every struct
    with its own zeroed limb buffer of nlimbs(pb) limbs."""
    N = nlimbs(pb)
    arr = (MpfrStruct * (count * stride_structs))()
    bufs = []
    for j in range(count * stride_structs):
        b = (ctypes.c_uint64 * N)()
        bufs.append(b)
        arr[j].prec, arr[j].sign, arr[j].exp = pb, 1, 0
        arr[j].d = ctypes.cast(b, ctypes.POINTER(ctypes.c_uint64))
    return arr, bufs


def _fill(arr, k, host, i):
    """struct arr[k] := value i of the host real MP array."""
    N = nlimbs(host.pb)
    w = int(host.se[i])
    e, neg = w >> 1, w & 1
    arr[k].sign = -1 if neg else 1
    if e == ZERO_EXP:
        arr[k].exp = -LONG_MAX
        return
    arr[k].exp = e
    for l in range(N):
        arr[k].d[l] = int(host.limbs[l, i])


def _read(arr, k, pb):
    N = nlimbs(pb)
    neg = 1 if arr[k].sign < 0 else 0
    if arr[k].exp == -LONG_MAX:
        return [0] * N, ZERO_EXP * 2 + neg
    return [arr[k].d[l] for l in range(N)], arr[k].exp * 2 + neg


def _with_zeros(h):
    se = h.se.copy()
    limbs = h.limbs.copy()
    se[3] = ZERO_EXP * 2
    se[5] = ZERO_EXP * 2 + 1
    limbs[:, 3] = 0
    limbs[:, 5] = 0
    return MPArray(h.pb, limbs, se)


@pytest.mark.parametrize("pb", [64, 96, 256, 1024, 2048])
def test_import_export_real(pb):
    n = 37
    h = _with_zeros(G.random_real(pb, n, 100 + pb, -2000, 2000))
    arr, bufs = _structs(pb, n)
    for i in range(n):
        _fill(arr, i, h, i)
    d = import_mpfr(pb, n, ctypes.addressof(arr))
    back = d.numpy()
    assert np.array_equal(back.limbs, h.limbs) and np.array_equal(back.se, h.se)

    out, obufs = _structs(pb, n)
    export_mpfr(d, ctypes.addressof(out))
    for i in range(n):
        lw, s = _read(out, i, pb)
        assert s == int(h.se[i])
        assert lw == [int(v) for v in h.limbs[:, i]]
        assert out[i].prec == pb


@pytest.mark.parametrize("pb", [128, 512])
def test_import_export_complex_mpc_stride(pb):
    """mpc_t = {mpfr_t re, mpfr_t im}: stride 2 structs, im at offset 1."""
    n = 21
    h = G.random_complex(pb, n, 7 + pb, -100, 100)
    re, im = h.real_part(), h.imag_part()
    arr, bufs = _structs(pb, n, 2)
    for i in range(n):
        _fill(arr, 2 * i, re, i)
        _fill(arr, 2 * i + 1, im, i)
    base = ctypes.addressof(arr)
    d = import_mpfr(pb, n, base, 2 * MPFR_STRUCT_SIZE, im_addr=base + MPFR_STRUCT_SIZE)
    assert d.cplx
    back = d.numpy()
    assert np.array_equal(back.limbs, h.limbs) and np.array_equal(back.se, h.se)

    out, obufs = _structs(pb, n, 2)
    ob = ctypes.addressof(out)
    export_mpfr(d, ob, 2 * MPFR_STRUCT_SIZE, im_addr=ob + MPFR_STRUCT_SIZE)
    for i in range(n):
        for part, src in ((0, re), (1, im)):
            lw, s = _read(out, 2 * i + part, pb)
            assert s == int(src.se[i]) and lw == [int(v) for v in src.limbs[:, i]]


def test_rejects_nan_inf_and_precision():
    pb, n = 256, 4
    h = G.random_real(pb, n, 5)
    L = _lib.lib()
    good, gbufs = _filled(pb, h)
    d = import_mpfr(pb, n, ctypes.addressof(good))
    for bad in ("nan", "inf", "prec"):
        arr, bufs = _filled(pb, h)
        if bad == "nan":
            arr[2].exp = 1 - LONG_MAX
        elif bad == "inf":
            arr[2].exp = 2 - LONG_MAX
        else:
            arr[2].prec = pb + 64
        l, s = d.ptrs()
        assert L.mpgfx_import_mpfr(pb, n, ctypes.addressof(arr), MPFR_STRUCT_SIZE, l, s, None) == 1
        if bad == "prec":
            assert L.mpgfx_export_mpfr(pb, n, l, s, ctypes.addressof(arr), MPFR_STRUCT_SIZE, None) == 1
    # a stride shorter than the struct is rejected; n = 0 is a no-op
    l, s = d.ptrs()
    arr, bufs = _filled(pb, h)
    assert L.mpgfx_import_mpfr(pb, n, ctypes.addressof(arr), 16, l, s, None) == 1
    assert L.mpgfx_import_mpfr(pb, 0, None, MPFR_STRUCT_SIZE, None, None, None) == 0


def _filled(pb, h):
    arr, bufs = _structs(pb, h.n)
    for i in range(h.n):
        _fill(arr, i, h, i)
    return arr, bufs


def test_import_runtime_precision_axpy_export():
    """An mpfr_t array of precision 1000 (not compiled in) imports, goes
    through the run-time precision AXPY, and exports; the exported structs
    hold the oracle's values."""
    import mpc_cuda_amd as M
    from oracle import oracle as O
    p, n = 1000, 77
    a, x, y = G.random_real(p, 1, 3), G.random_real(p, n, 4, -30, 30), G.random_real(p, n, 5, -30, 30)
    sa, ba = _filled(p, a)
    sx, bx = _filled(p, x)
    sy, by = _filled(p, y)
    da, dx, dy = (import_mpfr(p, v.n, ctypes.addressof(s)) for v, s in ((a, sa), (x, sx), (y, sy)))
    r = M.rt_axpy(da, dx, dy)
    out, obufs = _structs(p, n)
    export_mpfr(r, ctypes.addressof(out))
    want = O.axpy_real(p, a.pair(), x.pair(), y.pair())
    for i in range(n):
        lw, s = _read(out, i, p)
        assert s == int(want[1][i]) and lw == [int(v) for v in want[0][:, i]]
