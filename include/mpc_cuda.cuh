/* mpc_cuda.cuh -- umbrella header for the fixed-precision types (drop-in for
 * the fixed-precision part of tkouya/mpc_cuda include/mpc_cuda.cuh:42,47).
 * The runtime-precision cu_mpfr/cu_mpc layer of the reference is not part of
 * this library. */
#ifndef MPGFX_MPC_CUDA_CUH
#define MPGFX_MPC_CUDA_CUH
#include "mpc_cuda/cu_freal.cuh"
#include "mpc_cuda/cu_fcomplex.cuh"
#endif
