/* common.h -- shared device helpers of the mpgfx HIP library (gfx950).
 *
 * MP arrays are structure-of-arrays (see include/mpgfx.h): limb l of value i
 * at limbs[l*n + i], one int64 sign/exponent word per value.  A wavefront of
 * 64 lanes reading value i..i+63 therefore reads 512 contiguous bytes per limb
 * plane: fully coalesced 8-byte loads.
 */
#ifndef MPGFX_COMMON_H
#define MPGFX_COMMON_H

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include "mpc_cuda.cuh"
#include "mpgfx.h"

namespace mpgfx {

using cu_fp::cu_freal;
using cu_fp::cu_fcomplex;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;

template <int PB>
__device__ __forceinline__ cu_freal<PB> ld (const u64 *__restrict__ l, const i64 *__restrict__ se,
                                            i64 stride, i64 i)
{
  cu_freal<PB> r;
#pragma unroll
  for (int k = 0; k < cu_freal<PB>::N; k++) r.m[k] = __builtin_nontemporal_load (l + (i64) k * stride + i);
  i64 w = __builtin_nontemporal_load (se + i);
  r.sign = (w & 1) ? -1 : 1;
  r.exp = (long) (w >> 1);
  return r;
}
/* cached load (operands re-read by other lanes / waves) */
template <int PB>
__device__ __forceinline__ cu_freal<PB> ldc (const u64 *__restrict__ l, const i64 *__restrict__ se,
                                             i64 stride, i64 i)
{
  cu_freal<PB> r;
#pragma unroll
  for (int k = 0; k < cu_freal<PB>::N; k++) r.m[k] = l[(i64) k * stride + i];
  i64 w = se[i];
  r.sign = (w & 1) ? -1 : 1;
  r.exp = (long) (w >> 1);
  return r;
}
template <int PB>
__device__ __forceinline__ void st (const cu_freal<PB> &r, u64 *__restrict__ l, i64 *__restrict__ se,
                                    i64 stride, i64 i)
{
#pragma unroll
  for (int k = 0; k < cu_freal<PB>::N; k++) __builtin_nontemporal_store (r.m[k], l + (i64) k * stride + i);
  __builtin_nontemporal_store ((i64) ((u64) (i64) r.exp << 1) | (r.sign < 0 ? 1 : 0), se + i);
}

/* host-side error bookkeeping */
void set_error (const char *where, hipError_t e);
void set_error_text (const char *msg);
int launch_status (const char *where);

inline int nlimbs (int pb) { return (pb + 63) / 64; }

/* lanes per number of the cooperative folds: the power of two >= N + 2 (a
 * guard limb, N limbs, a carry limb), at least 4 */
__host__ __device__ constexpr int coop_lanes (int pb)
{
  return ((pb + 63) / 64 + 2 <= 4) ? 4 : ((pb + 63) / 64 + 2 <= 8) ? 8 : ((pb + 63) / 64 + 2 <= 16) ? 16
       : ((pb + 63) / 64 + 2 <= 32) ? 32 : 64;
}

} // namespace mpgfx

/* precisions compiled into the elementwise / AXPY / GEMV / GEMM kernels */
#define MPGFX_PB_LIST(X) X(32) X(64) X(96) X(128) X(160) X(192) X(256) X(288) X(320) \
                         X(384) X(512) X(1024) X(1056) X(2048)

#endif
