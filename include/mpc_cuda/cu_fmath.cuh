/* cu_fmath.cuh -- fixed-precision division and square root of
 * cu_fp::cu_freal<PB>, for AMD CDNA4 (gfx950) device code and for the host.
 * Drop-in for cu_fdiv / cu_sqrt of tkouya/mpc_cuda
 * include/mpc_cuda/cu_fmath.cuh:53-100.
 *
 * Faithful, not correctly rounded: the operands are widened to W = PB +
 * CU_FGUARD bits, a reciprocal (reciprocal square root) seeded from a double
 * is refined by Newton steps at W bits with every multiply and subtract RNDN,
 * and the W-bit result is rounded to PB.  The seed, the recurrences, the step
 * count and the guard width are the reference's, which makes every result
 * bit-identical to it; the W-bit arithmetic is the word engine of
 * cu_fp_core.cuh (cu_fmul / cu_fsub of cu_freal<W>).
 *
 * Arguments outside the reference's domain: a zero divisor gives +0 here (the
 * reference seeds its Newton step with 1/0 and has no defined result), and the
 * square root of zero or of a negative value is +0 (as in the reference).
 */
#ifndef MPGFX_CU_FMATH_CUH
#define MPGFX_CU_FMATH_CUH

#include <cmath>
#include "mpc_cuda/cu_freal.cuh"

#ifndef CU_FGUARD
#define CU_FGUARD 128            /* guard bits of the working precision (multiple of 32) */
#endif

namespace cu_fp {

/* Newton steps to reach W bits from a 52-bit seed, counting the seed itself
 * (one plus the number of doublings of 52 up to W); the loops below run one
 * more (cu_fmath.cuh:25-26, 62, 82) */
__host__ __device__ static inline constexpr int cu_newton_iters (int W)
{
  int steps = 1;
  for (long bits = 52; bits < W; bits *= 2) steps++;
  return steps;
}

/* 1/b at precision W: m = |b| scaled into [1/2, 1), z0 = double(1/m),
 * z <- z (2 - m z) (cu_fmath.cuh:53-66) */
template <int W>
__host__ __device__ static cu_freal<W> cu_recip_w (cu_freal<W> b)
{
  const int sgn = b.sign;
  const long e = b.exp;
  cu_freal<W> m = b;
  m.sign = 1;
  m.exp = 0;
  cu_freal<W> z = cu_freal<W>::from_double (1.0 / m.to_double ());
  const cu_freal<W> two = cu_from_si<W> (2);
  const int steps = cu_newton_iters (W) + 1;
  for (int s = 0; s < steps; s++)
    z = cu_fmul<W> (z, cu_fsub<W> (two, cu_fmul<W> (m, z)));
  z.sign = sgn;
  z.exp -= e;
  return z;
}

/* a / b = RNDN_PB(RNDN_W(a * (1/b)_W)) (cu_fmath.cuh:68-75) */
template <int PB>
__host__ __device__ static cu_freal<PB> cu_fdiv (const cu_freal<PB> &a, const cu_freal<PB> &b)
{
  if (a.is_zero () || b.is_zero ())
    {
      cu_freal<PB> r;
      r.set_zero ();
      return r;
    }
  constexpr int W = PB + CU_FGUARD;
  const cu_freal<W> aw = cu_convert<W> (a), bw = cu_convert<W> (b);
  return cu_convert<PB> (cu_fmul<W> (aw, cu_recip_w<W> (bw)));
}

/* sqrt(x), x > 0, at precision W: x = m 2^(2t) with m in [1/2, 2), y0 =
 * double(1/sqrt(m)), y <- y (3 - m y^2) / 2, sqrt(x) = (m y) 2^t
 * (cu_fmath.cuh:78-92) */
template <int W>
__host__ __device__ static cu_freal<W> cu_sqrt_w (cu_freal<W> x)
{
  const long e = x.exp;
  cu_freal<W> m = x;
  m.sign = 1;
  long t;
  if (e & 1L)
    {
      m.exp = 1;
      t = (e - 1) / 2;
    }
  else
    {
      m.exp = 0;
      t = e / 2;
    }
  cu_freal<W> y = cu_freal<W>::from_double (1.0 / sqrt (m.to_double ()));
  const cu_freal<W> three = cu_from_si<W> (3);
  const int steps = cu_newton_iters (W) + 1;
  for (int s = 0; s < steps; s++)
    {
      const cu_freal<W> y2 = cu_fmul<W> (y, y);
      y = cu_scale2<W> (cu_fmul<W> (y, cu_fsub<W> (three, cu_fmul<W> (m, y2))), -1);
    }
  cu_freal<W> r = cu_fmul<W> (m, y);
  r.exp += t;
  return r;
}

/* sqrt(x) = RNDN_PB(sqrt_W(x)); +0 for zero and negative x (cu_fmath.cuh:93-100) */
template <int PB>
__host__ __device__ static cu_freal<PB> cu_sqrt (const cu_freal<PB> &x)
{
  if (x.is_zero () || x.sign < 0)
    {
      cu_freal<PB> r;
      r.set_zero ();
      return r;
    }
  constexpr int W = PB + CU_FGUARD;
  return cu_convert<PB> (cu_sqrt_w<W> (cu_convert<W> (x)));
}

} // namespace cu_fp

#endif
