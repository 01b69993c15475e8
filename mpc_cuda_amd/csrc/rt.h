/* rt.h -- run-time precision arithmetic for the mpgfx_rt_* entry points: the
 * role of the reference's runtime-precision MPFR / MPC path on the device
 * (cu_mpfr_mul / cu_mpfr_add in demos/axpy_mpfr.cu:36-50, cu_mpc_mul /
 * cu_mpc_add in demos/axpy_mpc.cu:40-53), for any precision p up to
 * 32 CW bits, with MPFR's RNDN results.
 *
 * A value of precision p is stored as an MP array (include/mpgfx.h) with
 * nl = ceil(p/64) limbs, MPFR's left-justified significand with the low
 * 64 nl - p bits zero.  In registers it is held in a container of CW 32-bit
 * words (CW a power of two >= 2 nl), the significand at the top and the words
 * below it zero, so the word engine of cu_fp_core.cuh runs unchanged on
 * compile-time sizes and only the final rounding (round_top_p) works at the
 * run-time bit position.  Zero signs follow MPFR for real products and sums
 * (mpfr_mul: the product of the signs; mpfr_add in RNDN: -0 only for -0 + -0);
 * a complex component that cancels exactly is +0.
 */
#ifndef MPGFX_RT_H
#define MPGFX_RT_H

#include "common.h"

namespace mpgfx {

template <int CW>
struct rt_val
{
  int sign;
  i64 exp;
  bool zero;
  cu_fp::core::u32 w[CW];
};

/* value i of an SoA array of precision p (nl limbs) */
template <int CW>
__device__ __forceinline__ rt_val<CW> rt_ld (const u64 *__restrict__ l, const i64 *__restrict__ se, i64 n, i64 i,
                                             int nl)
{
  rt_val<CW> r;
#pragma unroll
  for (int k = 0; k < CW / 2; k++)
    {
      const int li = k - CW / 2 + nl;            /* container limb k holds storage limb li */
      const u64 v = (li >= 0) ? l[(i64) li * n + i] : 0ull;
      r.w[2 * k] = (u32) v;
      r.w[2 * k + 1] = (u32) (v >> 32);
    }
  const i64 w = se[i];
  r.sign = (w & 1) ? -1 : 1;
  r.exp = w >> 1;
  r.zero = r.exp == MPGFX_ZERO_EXP;
  return r;
}

template <int CW>
__device__ __forceinline__ void rt_st (const rt_val<CW> &r, u64 *__restrict__ l, i64 *__restrict__ se, i64 n, i64 i,
                                       int nl)
{
#pragma unroll
  for (int k = 0; k < CW / 2; k++)
    {
      const int li = k - CW / 2 + nl;
      if (li >= 0) l[(i64) li * n + i] = r.zero ? 0ull : (((u64) r.w[2 * k + 1] << 32) | r.w[2 * k]);
    }
  const i64 e = r.zero ? (i64) MPGFX_ZERO_EXP : r.exp;
  se[i] = (i64) ((u64) e << 1) | (r.sign < 0 ? 1 : 0);
}

template <int CW>
__device__ __forceinline__ rt_val<CW> rt_zero (int sign)
{
  rt_val<CW> r;
#pragma unroll
  for (int k = 0; k < CW; k++) r.w[k] = 0;
  r.sign = sign;
  r.exp = MPGFX_ZERO_EXP;
  r.zero = true;
  return r;
}

/* exact product, normalised: 2CW words, top bit set; value = M 2^(e - 64 CW) */
template <int CW>
struct rt_term
{
  int sign;
  i64 exp;
  bool zero;
  cu_fp::core::u32 M[2 * CW];
};

template <int CW>
__device__ __forceinline__ rt_term<CW> rt_mk_term (const rt_val<CW> &a, const rt_val<CW> &b)
{
  using namespace cu_fp::core;
  rt_term<CW> t;
  u32 P[2 * CW];
  mul_words<CW> (P, a.w, b.w);
  const u32 s = (P[2 * CW - 1] >> 31) ^ 1u;
#pragma unroll
  for (int i = 2 * CW - 1; i >= 0; i--) t.M[i] = fshl (P[i], (i >= 1) ? P[i - 1] : 0u, s);
  t.zero = a.zero || b.zero;
  t.sign = a.sign * b.sign;
  t.exp = t.zero ? 0 : a.exp + b.exp - (i64) s;
  return t;
}

/* RNDN_p(a b) (mpfr_mul) */
template <int CW>
__device__ __forceinline__ rt_val<CW> rt_mul (const rt_val<CW> &a, const rt_val<CW> &b, int p)
{
  using namespace cu_fp::core;
  if (a.zero || b.zero) return rt_zero<CW> (a.sign * b.sign);
  const rt_term<CW> t = rt_mk_term<CW> (a, b);
  rt_val<CW> r;
  const u32 c = round_top_p<CW, 2 * CW> (r.w, t.M, p, 0u);
  r.sign = t.sign;
  r.exp = t.exp + (i64) c;
  r.zero = false;
  return r;
}

/* RNDN_p(a + b) or RNDN_p(a - b) (mpfr_add / mpfr_sub) */
template <int CW>
__device__ __forceinline__ rt_val<CW> rt_add (const rt_val<CW> &a, const rt_val<CW> &b, int p, bool sub)
{
  using namespace cu_fp::core;
  const int sb = sub ? -b.sign : b.sign;
  if (a.zero && b.zero) return rt_zero<CW> (a.sign == sb ? a.sign : 1);
  const rres<CW> q = add_round<CW, CW> (a.sign, a.exp, a.zero, a.w, sb, b.exp, b.zero, b.w, p);
  rt_val<CW> r;
#pragma unroll
  for (int k = 0; k < CW; k++) r.w[k] = q.k[k];
  r.sign = q.sign;
  r.exp = q.exp;
  r.zero = q.zero;
  return r;
}

/* RNDN_p(a b - c d) if sub, else RNDN_p(a b + c d), from the exact products
 * (one component of mpc_mul) */
template <int CW>
__device__ __forceinline__ rt_val<CW> rt_fmm (const rt_val<CW> &a, const rt_val<CW> &b, const rt_val<CW> &c,
                                              const rt_val<CW> &d, int p, bool sub)
{
  using namespace cu_fp::core;
  const rt_term<CW> t1 = rt_mk_term<CW> (a, b), t2 = rt_mk_term<CW> (c, d);
  const rres<CW> q = add_round<CW, 2 * CW> (t1.zero ? 1 : t1.sign, t1.exp, t1.zero, t1.M,
                                            t2.zero ? 1 : (sub ? -t2.sign : t2.sign), t2.exp, t2.zero, t2.M, p);
  rt_val<CW> r;
#pragma unroll
  for (int k = 0; k < CW; k++) r.w[k] = q.k[k];
  r.sign = q.sign;
  r.exp = q.exp;
  r.zero = q.zero;
  return r;
}

/* host side: one table per container size (rt_kernels.hip, -DMPGFX_RT_CW) */
struct rt_ops
{
  int cw;
  int (*ew) (int op, int p, i64 n, const u64 *, const i64 *, const u64 *, const i64 *, u64 *, i64 *, hipStream_t);
  int (*axpy_real) (int p, i64 n, const u64 *, const i64 *, const u64 *, const i64 *, const u64 *, const i64 *, u64 *,
                    i64 *, hipStream_t);
  int (*axpy_cplx) (int p, i64 n, const u64 *, const i64 *, const u64 *, const i64 *, const u64 *, const i64 *, u64 *,
                    i64 *, hipStream_t);
  int (*gemv) (int cplx, int p, i64 m, i64 k, const u64 *, const i64 *, const u64 *, const i64 *, u64 *, i64 *,
               hipStream_t);
  int (*gemm) (int cplx, int p, i64 m, i64 n, i64 k, const u64 *, const i64 *, const u64 *, const i64 *, u64 *, i64 *,
               hipStream_t);
};
const rt_ops *rt_ops_cw_2 ();
const rt_ops *rt_ops_cw_4 ();
const rt_ops *rt_ops_cw_8 ();
const rt_ops *rt_ops_cw_16 ();
const rt_ops *rt_ops_cw_32 ();
const rt_ops *rt_ops_cw_64 ();

/* run-time precisions served: 1 .. MPGFX_RT_MAXPREC bits */
constexpr int MPGFX_RT_MAXPREC = 2048;

} // namespace mpgfx

#endif
