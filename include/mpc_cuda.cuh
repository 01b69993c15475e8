/* mpc_cuda.cuh -- umbrella header for the fixed-precision types (drop-in for
 * the fixed-precision part of tkouya/mpc_cuda include/mpc_cuda.cuh:42,47,55;
 * of cu_fmath.cuh, division and square root).
 * The runtime-precision cu_mpfr/cu_mpc layer of the reference is not part of
 * this library. */
#ifndef MPGFX_MPC_CUDA_CUH
#define MPGFX_MPC_CUDA_CUH
#include "mpc_cuda/cu_freal.cuh"
#include "mpc_cuda/cu_fcomplex.cuh"
#include "mpc_cuda/cu_fmath.cuh"
#endif
