"""Host-side calls into the HIP library on device MP arrays.

These mirror the reference's operations on cu_fp::cu_freal<PB> /
cu_fcomplex<PB> (tkouya/mpc_cuda include/mpc_cuda/cu_freal.cuh,
cu_fcomplex.cuh, demos/bench_3way.cu, tools/linalg_fast.cu), each launched
asynchronously on the current torch stream.  Arguments must be device
MPArrays of matching precision; results are new device MPArrays.
"""
import ctypes

from . import _lib
from .arrays import MPArray, empty


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(x):
    if not isinstance(x, MPArray) or not x.on_device:
        raise TypeError("expected a device MPArray (use .cuda())")
    return x


def _same(pb, *xs):
    for x in xs:
        if x.pb != pb:
            raise ValueError("precision mismatch")


def _out(out, pb, n, cplx, like):
    """a caller's output array, checked against the result's precision, length
    and kind (a short or mismatched array would be written out of bounds), or
    a new device array"""
    if out is None:
        return empty(pb, n, cplx, device=like.se.device)
    _dev(out)
    if out.pb != pb or out.n != n or out.cplx != cplx:
        raise ValueError(f"output array is pb={out.pb} n={out.n} cplx={out.cplx}, "
                         f"the result is pb={pb} n={n} cplx={cplx}")
    if not (out.limbs.is_contiguous() and out.se.is_contiguous()):
        raise ValueError("output array must be contiguous")
    return out


_OPS = {"mul": _lib.MUL, "add": _lib.ADD, "sub": _lib.SUB, "div": _lib.DIV}


def elementwise(op, a, b, out=None):
    """out = a OP b elementwise (real: cu_fmul/cu_fadd/cu_fsub, and cu_fdiv for
    "div" -- faithful, Newton at pb + 128 bits; complex: cu_cmul/cu_cadd/cu_csub)."""
    _dev(a), _dev(b)
    _same(a.pb, b)
    if a.cplx != b.cplx or a.n != b.n:
        raise ValueError("shape mismatch")
    if a.cplx and op == "div":
        raise ValueError("complex division is not provided")
    out = _out(out, a.pb, a.n, a.cplx, a)
    L = _lib.lib()
    f = L.mpgfx_elementwise_complex if a.cplx else L.mpgfx_elementwise_real
    _lib.check(f.__name__, f(a.pb, _OPS[op], a.n, *a.ptrs(), *b.ptrs(), *out.ptrs(), _stream()))
    return out


def sqrt(x, out=None):
    """out = cu_sqrt(x) elementwise for a real array (faithful; +0 for zero and
    negative values)."""
    _dev(x)
    if x.cplx:
        raise ValueError("complex square root is not provided")
    out = _out(out, x.pb, x.n, False, x)
    L = _lib.lib()
    _lib.check("mpgfx_sqrt_real", L.mpgfx_sqrt_real(x.pb, x.n, *x.ptrs(), *out.ptrs(), _stream()))
    return out


def rt_elementwise(op, a, b, out=None):
    """Run-time precision (any a.pb in 1..2048): out = a OP b, OP in mul/add/sub,
    RNDN at a.pb with MPFR's zero signs."""
    _dev(a), _dev(b)
    if a.pb != b.pb or a.n != b.n or a.cplx or b.cplx:
        raise ValueError("shape mismatch")
    out = _out(out, a.pb, a.n, False, a)
    L = _lib.lib()
    _lib.check("mpgfx_rt_elementwise_real",
               L.mpgfx_rt_elementwise_real(a.pb, _OPS[op], a.n, *a.ptrs(), *b.ptrs(), *out.ptrs(), _stream()))
    return out


def rt_axpy(a, x, y, out=None):
    """Run-time precision AXPY (any x.pb in 1..2048), real or complex: RNDN /
    MPC_RNDNN per operation, as mpfr_mul + mpfr_add / mpc_mul + mpc_add."""
    _dev(a), _dev(x), _dev(y)
    if not (a.pb == x.pb == y.pb) or a.n != 1 or x.n != y.n or not (a.cplx == x.cplx == y.cplx):
        raise ValueError("shape mismatch")
    out = _out(out, x.pb, x.n, x.cplx, x)
    L = _lib.lib()
    f = L.mpgfx_rt_axpy_complex if x.cplx else L.mpgfx_rt_axpy_real
    _lib.check(f.__name__, f(x.pb, x.n, *a.ptrs(), *x.ptrs(), *y.ptrs(), *out.ptrs(), _stream()))
    return out


def rt_gemv(A, x, m, k, out=None):
    """Run-time precision GEMV y = A x (any A.pb in 1..2048), real or complex:
    the per-op fold with mpfr_mul + mpfr_add (mpc_mul + mpc_add) at A.pb."""
    _dev(A), _dev(x)
    if A.pb != x.pb or A.n != m * k or x.n != k or A.cplx != x.cplx:
        raise ValueError("shape mismatch")
    out = _out(out, A.pb, m, A.cplx, A)
    _lib.check("mpgfx_rt_gemv", _lib.lib().mpgfx_rt_gemv(A.pb, int(A.cplx), m, k, *A.ptrs(), *x.ptrs(), *out.ptrs(),
                                                         _stream()))
    return out


def rt_dot(x, y, out=None):
    """Run-time precision dot product (the per-op fold), a GEMV of one row."""
    return rt_gemv(x, y, 1, x.n, out=out)


def rt_gemm(A, B, m, n, k, out=None):
    """Run-time precision GEMM C = A B, per-op fold per element at A.pb."""
    _dev(A), _dev(B)
    if A.pb != B.pb or A.n != m * k or B.n != k * n or A.cplx != B.cplx:
        raise ValueError("shape mismatch")
    out = _out(out, A.pb, m * n, A.cplx, A)
    _lib.check("mpgfx_rt_gemm", _lib.lib().mpgfx_rt_gemm(A.pb, int(A.cplx), m, n, k, *A.ptrs(), *B.ptrs(),
                                                         *out.ptrs(), _stream()))
    return out


def axpy(a, x, y, out=None):
    """out = RNDN(RNDN(a*x) + y) with a scalar (an MPArray of one value)."""
    _dev(a), _dev(x), _dev(y)
    _same(x.pb, a, y)
    if a.n != 1 or x.n != y.n or not (a.cplx == x.cplx == y.cplx):
        raise ValueError("shape mismatch")
    out = _out(out, x.pb, x.n, x.cplx, x)
    L = _lib.lib()
    f = L.mpgfx_axpy_complex if x.cplx else L.mpgfx_axpy_real
    _lib.check(f.__name__, f(x.pb, x.n, *a.ptrs(), *x.ptrs(), *y.ptrs(), *out.ptrs(), _stream()))
    return out


def gemv(A, x, m, k, out=None):
    """y = A x with per-operation rounding in k order (A: m*k values; real,
    or complex with MPC_RNDNN products and sums)."""
    _dev(A), _dev(x)
    _same(A.pb, x)
    if A.n != m * k or x.n != k or A.cplx != x.cplx:
        raise ValueError("shape mismatch")
    out = _out(out, A.pb, m, A.cplx, A)
    L = _lib.lib()
    if A.cplx:
        _lib.check("mpgfx_gemv_complex", L.mpgfx_gemv_complex(A.pb, m, k, *A.ptrs(), *x.ptrs(), *out.ptrs(), _stream()))
    else:
        _lib.check("mpgfx_gemv_real", L.mpgfx_gemv_real(A.pb, m, k, *A.ptrs(), *x.ptrs(), *out.ptrs(), _stream()))
    return out


def gemm(A, B, m, n, k, mode="perop", out=None, workspace=None):
    """C = A B (A: m*k, B: k*n values, row-major).  mode 'perop' (the
    reference's matmul order) or 'exact' (exact sum, one rounding per real
    value or per complex component)."""
    _dev(A), _dev(B)
    _same(A.pb, B)
    if A.n != m * k or B.n != k * n or A.cplx != B.cplx:
        raise ValueError("shape mismatch")
    out = _out(out, A.pb, m * n, A.cplx, A)
    L = _lib.lib()
    if A.cplx:
        if mode == "perop":
            _lib.check("mpgfx_gemm_complex",
                       L.mpgfx_gemm_complex(A.pb, m, n, k, *A.ptrs(), *B.ptrs(), *out.ptrs(), _stream()))
        elif mode == "exact":
            _lib.check("mpgfx_gemm_complex_exact",
                       L.mpgfx_gemm_complex_exact(A.pb, m, n, k, *A.ptrs(), *B.ptrs(), *out.ptrs(), _stream()))
        else:
            raise ValueError(mode)
        return out
    md = {"perop": _lib.GEMM_PEROP, "exact": _lib.GEMM_EXACT}[mode]
    ws_ptr, ws_bytes = None, 0
    need = gemm_workspace_size(A.pb, mode, m, n, k) if mode == "exact" else 0
    if workspace is None and mode == "exact":
        workspace = _workspace(A.se.device, need)
    if workspace is not None:
        ws_ptr, ws_bytes = workspace.data_ptr(), workspace.numel() * workspace.element_size()
    _lib.check("mpgfx_gemm_real",
               L.mpgfx_gemm_real(A.pb, md, m, n, k, *A.ptrs(), *B.ptrs(), *out.ptrs(),
                                 ws_ptr, ws_bytes, _stream()))
    if mode == "exact" and k > 0:
        # asynchronous mode: the outcome lands in the workspace's first word
        _LAST_WS[str(A.se.device)] = workspace if ws_bytes >= need else None
    return out


_WS = {}
_LAST_WS = {}


def exact_status(device="cuda"):
    """Outcome of the last real exact GEMM gemm() ran on `device`: with a
    workspace of gemm_workspace_size() bytes (what gemm() passes) the call is
    asynchronous and its outcome is the first 8 bytes of that workspace: 0, or
    EUNSUPPORTED when the exponent spread needs more than 2^22 bits of
    accumulator (C is then +0).  Waits for the stream."""
    import torch
    key = str(torch.device(device) if not isinstance(device, torch.device) else device)
    if key == "cuda":
        key = "cuda:%d" % torch.cuda.current_device()
    t = _LAST_WS.get(key)
    if t is None:
        return _lib.OK
    return int(t[:8].view(torch.int64).item())


def _workspace(device, nbytes):
    """a per-device workspace tensor kept between calls and grown on demand
    (torch's caching allocator), so repeated exact GEMMs do not allocate"""
    import torch
    key = str(device)
    t = _WS.get(key)
    if t is None or t.numel() < nbytes:
        _WS.pop(key, None)
        t = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)
        _WS[key] = t
    return t


def gemm_workspace_size(pb, mode, m, n, k):
    md = {"perop": _lib.GEMM_PEROP, "exact": _lib.GEMM_EXACT}[mode]
    b = ctypes.c_size_t(0)
    _lib.check("mpgfx_gemm_workspace_size", _lib.lib().mpgfx_gemm_workspace_size(pb, md, m, n, k, ctypes.byref(b)))
    return b.value


def supported_pb(pb):
    return bool(_lib.lib().mpgfx_supported_pb(pb))
