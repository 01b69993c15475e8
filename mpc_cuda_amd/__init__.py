"""mpc_cuda_amd -- MI355X (gfx950) fixed-precision multiple-precision arithmetic.

The compute path is the HIP library ``libmpgfx.so`` (C ABI in include/mpgfx.h,
sources in mpc_cuda_amd/csrc/), built on the drop-in C++ template types
cu_fp::cu_freal<PB> / cu_fcomplex<PB> of include/mpc_cuda/.  This package is
the Python host side: device MP arrays (torch tensors as HIP allocations), a
seeded input generator, and thin wrappers over the C ABI.  There is no CPU
fallback: every operation runs the HIP kernels or raises.
"""
from . import _lib
from .arrays import MPArray, empty, equal, nlimbs, ZERO_EXP
from .gen import random_real, random_complex, random_array, device_random
from . import batch
from . import interop
from .ops import (elementwise, sqrt, rt_elementwise, rt_axpy, rt_gemv, rt_dot, rt_gemm, axpy, gemv, gemm,
                  gemm_workspace_size, exact_status, supported_pb)
from ._lib import MPGFXError, lib

__all__ = ["sqrt", "rt_elementwise", "rt_axpy", "rt_gemv", "rt_dot", "rt_gemm", "MPArray", "empty", "equal", "nlimbs", "ZERO_EXP", "random_real", "random_complex",
           "random_array", "device_random", "batch", "interop", "elementwise", "axpy", "gemv", "gemm", "gemm_workspace_size",
           "exact_status", "supported_pb", "MPGFXError", "lib"]
