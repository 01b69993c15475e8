"""mpgfx_gemm_exact_noalloc -- the exact GEMM on the library's static device
scratch, with no allocation -- against mpgfx_gemm_real(MPGFX_GEMM_EXACT), bit
for bit, on consecutive calls (the second smaller than the first, the scratch
keeping the first call's contents), on the wide-accumulator path, and a call
too large for the scratch, which is refused with MPGFX_ENOMEM.  Called through
the C ABI."""
import ctypes

import pytest

import mpc_cuda_amd as M
from mpc_cuda_amd import _lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def noalloc(pb, A, B, m, n, k, C):
    import torch
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return _lib.lib().mpgfx_gemm_exact_noalloc(pb, m, n, k, *A.ptrs(), *B.ptrs(), *C.ptrs(), s)


def one(pb, m, n, k, lo, hi, seed):
    import torch
    A = M.device_random(pb, m * k, seed, lo, hi)
    B = M.device_random(pb, k * n, seed + 1, lo, hi)
    ref = M.gemm(A, B, m, n, k, "exact")
    C = M.empty(pb, m * n)
    C.limbs.fill_(-1)
    C.se.fill_(12345)
    rc = noalloc(pb, A, B, m, n, k, C)
    torch.cuda.synchronize()
    assert rc == _lib.OK, _lib.lib().mpgfx_last_error()
    assert M.equal(C, ref) == 0, (pb, m, n, k)
    return A, B, C


def test_noalloc_exact_gemm_matches_exact_entry(gpu):
    print("noalloc: pb 256 1024^3, then 200x150x300, then pb 512 33x17x100 spread 3000", flush=True)
    one(256, 1024, 1024, 1024, -4, 4, 11)                    # the headline GEMM shape
    A, B, C = one(256, 200, 150, 300, -4, 4, 21)            # smaller, on the scratch the first call left
    assert O.mismatches(C.pair(), O.gemm(256, 200, 150, 300, A.pair(), B.pair(), "exact"),
                        ignore_zero_sign=False) == 0
    A, B, C = one(512, 33, 17, 100, -3000, 3000, 31)        # wide-accumulator path
    assert O.mismatches(C.pair(), O.gemm(512, 33, 17, 100, A.pair(), B.pair(), "exact"),
                        ignore_zero_sign=False) == 0


def test_noalloc_refuses_beyond_scratch(gpu):
    import torch
    pb, d = 256, 2048                                        # digit planes alone need 285 MB
    A = M.device_random(pb, d * d, 41, -4, 4)
    B = M.device_random(pb, d * d, 42, -4, 4)
    C = M.empty(pb, d * d)
    rc = noalloc(pb, A, B, d, d, d, C)
    torch.cuda.synchronize()
    assert rc == _lib.ENOMEM
    assert b"workspace" in _lib.lib().mpgfx_last_error()
    one(256, 64, 64, 64, -4, 4, 51)                          # the entry point still works afterwards
