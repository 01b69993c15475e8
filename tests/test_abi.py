"""The C-ABI library loads, exports every symbol include/mpgfx.h declares, and
rejects bad arguments on the host before any GPU work (no compute calls)."""
import ctypes
import os
import re

import pytest

import mpc_cuda_amd
from mpc_cuda_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "mpgfx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mpgfx_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = mpc_cuda_amd.lib()
    syms = declared_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.SIGNATURES), "ctypes signatures out of sync with include/mpgfx.h"


def test_version_and_supported_precisions():
    L = mpc_cuda_amd.lib()
    assert L.mpgfx_version() >= 100
    for pb in (32, 64, 96, 128, 160, 192, 256, 288, 320, 384, 512, 1024, 1056, 2048):
        assert L.mpgfx_supported_pb(pb) == 1, pb
    for pb in (0, -32, 31, 33, 4096, 100):
        assert L.mpgfx_supported_pb(pb) == 0, pb


def test_argument_errors_return_before_any_launch():
    L = mpc_cuda_amd.lib()
    p = ctypes.c_void_p(0x1000)   # never dereferenced: every call below fails validation first
    # negative sizes
    assert L.mpgfx_axpy_real(256, -1, p, p, p, p, p, p, p, p, None) == _lib.EINVAL
    assert L.mpgfx_gemv_real(512, -4, 4, p, p, p, p, p, p, None) == _lib.EINVAL
    assert L.mpgfx_gemm_real(256, 0, 4, -1, 4, p, p, p, p, p, p, None, 0, None) == _lib.EINVAL
    # unsupported precision
    assert L.mpgfx_axpy_real(100, 8, p, p, p, p, p, p, p, p, None) == _lib.EUNSUPPORTED
    assert L.mpgfx_elementwise_complex(4096, 0, 8, p, p, p, p, p, p, None) == _lib.EUNSUPPORTED
    # bad op / mode
    assert L.mpgfx_elementwise_real(256, 7, 8, p, p, p, p, p, p, None) == _lib.EINVAL
    assert L.mpgfx_elementwise_complex(256, _lib.DIV, 8, p, p, p, p, p, p, None) == _lib.EINVAL
    assert L.mpgfx_sqrt_real(256, -1, p, p, p, p, None) == _lib.EINVAL
    assert L.mpgfx_sqrt_real(256, 8, p, None, p, p, None) == _lib.EINVAL
    assert L.mpgfx_sqrt_real(100, 8, p, p, p, p, None) == _lib.EUNSUPPORTED
    assert L.mpgfx_rt_axpy_real(0, 8, p, p, p, p, p, p, p, p, None) == _lib.EINVAL
    assert L.mpgfx_rt_axpy_complex(4096, 8, p, p, p, p, p, p, p, p, None) == _lib.EUNSUPPORTED
    assert L.mpgfx_rt_elementwise_real(100, _lib.DIV, 8, p, p, p, p, p, p, None) == _lib.EINVAL
    assert L.mpgfx_rt_axpy_real(100, 8, p, None, p, p, p, p, p, p, None) == _lib.EINVAL
    assert L.mpgfx_gemm_real(256, 5, 4, 4, 4, p, p, p, p, p, p, None, 0, None) == _lib.EINVAL
    # null pointers
    assert L.mpgfx_axpy_complex(256, 8, None, p, p, p, p, p, p, p, None) == _lib.EINVAL
    assert L.mpgfx_pack(256, 0, 8, None, p, p, None) == _lib.EINVAL
    # empty problems are no-ops
    assert L.mpgfx_axpy_real(256, 0, None, None, None, None, None, None, None, None, None) == _lib.OK
    assert L.mpgfx_gemm_real(256, 0, 0, 4, 4, None, None, None, None, None, None, None, 0, None) == _lib.OK
    b = ctypes.c_size_t(123)
    assert L.mpgfx_gemm_workspace_size(256, 0, 8, 8, 8, ctypes.byref(b)) == _lib.OK and b.value == 0
    assert L.mpgfx_gemm_workspace_size(256, 0, 8, 8, 8, None) == _lib.EINVAL


def test_python_wrappers_refuse_host_arrays():
    x = mpc_cuda_amd.random_real(256, 8, 1)
    with pytest.raises(TypeError):
        mpc_cuda_amd.axpy(x, x, x)
