/* tests/host/cufp_host.cpp -- host (g++) instantiation of the drop-in template
 * API in include/mpc_cuda/, exported with the SoA C signatures the oracle uses,
 * so tests can compare the header's host path with the oracle bit for bit.
 * Test harness only: the library's entry points run on the GPU. */
#include "mpc_cuda.cuh"
#include <cstdint>
using namespace cu_fp;

namespace {
template <int PB> cu_freal<PB> ld (const uint64_t *l, const int64_t *se, int64_t n, int64_t i)
{
  cu_freal<PB> r;
  for (int k = 0; k < cu_freal<PB>::N; k++) r.m[k] = l[(int64_t) k * n + i];
  r.sign = (se[i] & 1) ? -1 : 1;
  r.exp = (long) (se[i] >> 1);
  return r;
}
template <int PB> void st (const cu_freal<PB> &r, uint64_t *l, int64_t *se, int64_t n, int64_t i)
{
  for (int k = 0; k < cu_freal<PB>::N; k++) l[(int64_t) k * n + i] = r.m[k];
  se[i] = (int64_t) ((uint64_t) (int64_t) r.exp << 1) | (r.sign < 0 ? 1 : 0);
}
template <int PB> void ew_real (int op, int64_t n, const uint64_t *al, const int64_t *ase,
                                const uint64_t *bl, const int64_t *bse, uint64_t *ol, int64_t *ose)
{
#pragma omp parallel for
  for (int64_t i = 0; i < n; i++)
    {
      cu_freal<PB> a = ld<PB> (al, ase, n, i), b = ld<PB> (bl, bse, n, i);
      st<PB> (op == 0 ? a * b : (op == 1 ? a + b : (op == 2 ? a - b : cu_fdiv<PB> (a, b))), ol, ose, n, i);
    }
}
template <int PB> void ew_cplx (int op, int64_t n, const uint64_t *al, const int64_t *ase,
                                const uint64_t *bl, const int64_t *bse, uint64_t *ol, int64_t *ose)
{
  const int64_t o = (int64_t) cu_freal<PB>::N * n;
#pragma omp parallel for
  for (int64_t i = 0; i < n; i++)
    {
      cu_fcomplex<PB> a (ld<PB> (al, ase, n, i), ld<PB> (al + o, ase + n, n, i));
      cu_fcomplex<PB> b (ld<PB> (bl, bse, n, i), ld<PB> (bl + o, bse + n, n, i));
      cu_fcomplex<PB> r = op == 0 ? a * b : (op == 1 ? a + b : a - b);
      st<PB> (r.re, ol, ose, n, i);
      st<PB> (r.im, ol + o, ose + n, n, i);
    }
}
template <int PB> void sqrt_real (int64_t n, const uint64_t *xl, const int64_t *xse, uint64_t *ol, int64_t *ose)
{
#pragma omp parallel for
  for (int64_t i = 0; i < n; i++) st<PB> (cu_sqrt<PB> (ld<PB> (xl, xse, n, i)), ol, ose, n, i);
}
template <int PB> double conv (double d, int *exp_out)
{
  cu_freal<PB> x (d);
  *exp_out = (int) x.exp;
  return (double) x;
}
}

#define PB_LIST(X) X(32) X(64) X(96) X(128) X(160) X(192) X(256) X(288) X(320) X(384) \
                   X(512) X(1024) X(1056) X(2048)
extern "C" int host_elementwise_real (int pb, int op, int64_t n, const uint64_t *al, const int64_t *ase,
                                      const uint64_t *bl, const int64_t *bse, uint64_t *ol, int64_t *ose)
{
#define C(P) if (pb == P) { ew_real<P> (op, n, al, ase, bl, bse, ol, ose); return 0; }
  PB_LIST (C)
#undef C
  return -1;
}
extern "C" int host_elementwise_complex (int pb, int op, int64_t n, const uint64_t *al, const int64_t *ase,
                                         const uint64_t *bl, const int64_t *bse, uint64_t *ol, int64_t *ose)
{
#define C(P) if (pb == P) { ew_cplx<P> (op, n, al, ase, bl, bse, ol, ose); return 0; }
  PB_LIST (C)
#undef C
  return -1;
}
extern "C" int host_sqrt_real (int pb, int64_t n, const uint64_t *xl, const int64_t *xse, uint64_t *ol, int64_t *ose)
{
#define C(P) if (pb == P) { sqrt_real<P> (n, xl, xse, ol, ose); return 0; }
  PB_LIST (C)
#undef C
  return -1;
}
/* round-trip a double through cu_freal<PB> (from_double / to_double) */
extern "C" double host_roundtrip_double (int pb, double d, int *exp_out)
{
#define C(P) if (pb == P) return conv<P> (d, exp_out);
  PB_LIST (C)
#undef C
  return 0.0;
}

/* one component of a complex product, RNDN(a*b -/+ c*d), by the short-product
 * path (with its `amb` flag) and by the exact path, so tests can check that
 * the short path's result is the exact one wherever it does not flag the lane.
 * Operands: four real arrays stacked, limbs (4, N, n), se (4, n). */
template <int PB> void fmm_chk (int64_t n, const uint64_t *l, const int64_t *se, int sub, uint64_t *sl,
                                int64_t *sse, uint64_t *el, int64_t *ese, uint8_t *amb)
{
  const int64_t o = (int64_t) cu_freal<PB>::N * n;
#pragma omp parallel for
  for (int64_t i = 0; i < n; i++)
    {
      cu_freal<PB> a = ld<PB> (l, se, n, i), b = ld<PB> (l + o, se + n, n, i);
      cu_freal<PB> c = ld<PB> (l + 2 * o, se + 2 * n, n, i), d = ld<PB> (l + 3 * o, se + 3 * n, n, i);
      bool f = false;
      cu_freal<PB> r = core::fmm_short<PB> (a, b, c, d, sub != 0, f);
      st<PB> (r, sl, sse, n, i);
      st<PB> (core::fmm_exact<PB> (a, b, c, d, sub != 0), el, ese, n, i);
      amb[i] = f ? 1 : 0;
    }
}
#define PB_SHORT(X) X(256) X(288) X(320) X(384) X(512) X(1024) X(1056) X(2048)
extern "C" int host_fmm_check (int pb, int64_t n, const uint64_t *l, const int64_t *se, int sub, uint64_t *sl,
                               int64_t *sse, uint64_t *el, int64_t *ese, uint8_t *amb)
{
#define C(P) if (pb == P) { fmm_chk<P> (n, l, se, sub, sl, sse, el, ese, amb); return 0; }
  PB_SHORT (C)
#undef C
  return -1;
}

/* RNDN(a*b) by the short product (with its `amb` flag) and by the exact one;
 * operands stacked: limbs (2, N, n), se (2, n) */
template <int PB> void fmul_chk (int64_t n, const uint64_t *l, const int64_t *se, uint64_t *sl, int64_t *sse,
                                 uint64_t *el, int64_t *ese, uint8_t *amb)
{
  const int64_t o = (int64_t) cu_freal<PB>::N * n;
#pragma omp parallel for
  for (int64_t i = 0; i < n; i++)
    {
      cu_freal<PB> a = ld<PB> (l, se, n, i), b = ld<PB> (l + o, se + n, n, i);
      bool f = false;
      st<PB> (core::fmul_short<PB> (a, b, f), sl, sse, n, i);
      st<PB> (core::fmul_exact<PB> (a, b), el, ese, n, i);
      amb[i] = f ? 1 : 0;
    }
}
extern "C" int host_fmul_check (int pb, int64_t n, const uint64_t *l, const int64_t *se, uint64_t *sl,
                                int64_t *sse, uint64_t *el, int64_t *ese, uint8_t *amb)
{
#define C(P) if (pb == P) { fmul_chk<P> (n, l, se, sl, sse, el, ese, amb); return 0; }
  PB_SHORT (C)
#undef C
  return -1;
}
