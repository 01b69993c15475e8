"""ctypes binding of the mpgfx C ABI (include/mpgfx.h) built as
mpc_cuda_amd/libmpgfx.so.  There is no CPU fallback: if the library is absent
or fails to load, every call raises."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpgfx.so")

OK, EINVAL, EUNSUPPORTED, EHIP, ENOMEM = 0, 1, 2, 3, 4
STATUS = {OK: "OK", EINVAL: "EINVAL", EUNSUPPORTED: "EUNSUPPORTED", EHIP: "EHIP", ENOMEM: "ENOMEM"}
MUL, ADD, SUB, DIV = 0, 1, 2, 3
GEMM_PEROP, GEMM_EXACT = 0, 1

# every symbol include/mpgfx.h declares, with its ctypes signature
_p = ctypes.c_void_p
_i = ctypes.c_int
_l = ctypes.c_int64
SIGNATURES = {
    "mpgfx_version": (_i, []),
    "mpgfx_supported_pb": (_i, [_i]),
    "mpgfx_last_error": (ctypes.c_char_p, []),
    "mpgfx_elementwise_real": (_i, [_i, _i, _l] + [_p] * 6 + [_p]),
    "mpgfx_elementwise_complex": (_i, [_i, _i, _l] + [_p] * 6 + [_p]),
    "mpgfx_sqrt_real": (_i, [_i, _l] + [_p] * 4 + [_p]),
    "mpgfx_rt_elementwise_real": (_i, [_i, _i, _l] + [_p] * 6 + [_p]),
    "mpgfx_rt_axpy_real": (_i, [_i, _l] + [_p] * 8 + [_p]),
    "mpgfx_rt_axpy_complex": (_i, [_i, _l] + [_p] * 8 + [_p]),
    "mpgfx_rt_gemv": (_i, [_i, _i, _l, _l] + [_p] * 6 + [_p]),
    "mpgfx_rt_gemm": (_i, [_i, _i, _l, _l, _l] + [_p] * 6 + [_p]),
    "mpgfx_axpy_real": (_i, [_i, _l] + [_p] * 8 + [_p]),
    "mpgfx_axpy_complex": (_i, [_i, _l] + [_p] * 8 + [_p]),
    "mpgfx_gemv_real": (_i, [_i, _l, _l] + [_p] * 6 + [_p]),
    "mpgfx_gemv_complex": (_i, [_i, _l, _l] + [_p] * 6 + [_p]),
    "mpgfx_gemm_workspace_size": (_i, [_i, _i, _l, _l, _l, ctypes.POINTER(ctypes.c_size_t)]),
    "mpgfx_gemm_real": (_i, [_i, _i, _l, _l, _l] + [_p] * 6 + [_p, ctypes.c_size_t, _p]),
    "mpgfx_gemm_complex": (_i, [_i, _l, _l, _l] + [_p] * 6 + [_p]),
    "mpgfx_gemm_complex_exact": (_i, [_i, _l, _l, _l] + [_p] * 6 + [_p]),
    "mpgfx_gemm_exact_noalloc": (_i, [_i, _l, _l, _l] + [_p] * 6 + [_p]),
    "mpgfx_import_mpfr": (_i, [_i, _l, _p, ctypes.c_size_t, _p, _p, _p]),
    "mpgfx_export_mpfr": (_i, [_i, _l, _p, _p, _p, ctypes.c_size_t, _p]),
    "mpgfx_pack": (_i, [_i, _i, _l, _p, _p, _p, _p]),
    "mpgfx_fill_random": (_i, [_i, _l, _l, ctypes.c_uint64, _l, _l, _p, _p, _p]),
    "mpgfx_batch_create": (_i, [_p, _l, _i, ctypes.c_char_p, _p, ctypes.POINTER(_p)]),
    "mpgfx_batch_run": (_i, [_p, _p]),
    "mpgfx_batch_info": (_i, [_p, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(_i)]),
    "mpgfx_batch_destroy": (_i, [_p]),
    "mpgfx_unpack": (_i, [_i, _i, _l, _p, _p, _p, _p]),
}

_LIB = None


class MPGFXError(RuntimeError):
    def __init__(self, fn, code, detail=""):
        super().__init__(f"{fn} failed: {STATUS.get(code, code)} {detail}".strip())
        self.code = code


def lib():
    """Load libmpgfx.so (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `make -C mpc_cuda_amd/csrc -j8`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(fn, code):
    if code != OK:
        detail = ""
        if code == EHIP:
            detail = lib().mpgfx_last_error().decode(errors="replace")
        raise MPGFXError(fn, code, detail)
    return code
