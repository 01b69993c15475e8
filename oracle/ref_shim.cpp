/* oracle/ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * A thin C-ABI front over the reference implementation itself: it includes the
 * reference's own header-only types (include/mpc_cuda/cu_freal.cuh and
 * include/mpc_cuda/cu_fcomplex.cuh, host path: cu_freal.cuh:23-32, 46-48) from
 * /root/reference at build time and is compiled by oracle/Makefile into
 * oracle/_ref/libref.so.  No reference source is copied into this repository.
 *
 * It is used (a) to validate the C oracle in oracle/mpfixed.c bit for bit,
 * including the reference's zero-sign conventions, and (b) as the CPU baseline
 * of bench.py: the reference's own fixed-precision host code, parallelised with
 * OpenMP over host cores exactly as demos/bench3_cpu.cpp:35-85 parallelises its
 * MPFR/MPC loops.
 *
 * Arrays use the SoA layout of oracle/mpfixed.h.
 */
#include "mpc_cuda/cu_freal.cuh"
#include "mpc_cuda/cu_fcomplex.cuh"
#include "mpc_cuda/cu_fmath.cuh"
#include <cstdint>
#include <cstddef>
#include <cstdlib>
#ifdef _OPENMP
#include <omp.h>
#endif

using cu_fp::cu_freal;
using cu_fp::cu_fcomplex;

namespace {

template <int PB> inline cu_freal<PB> ld (const uint64_t *l, const int64_t *se, int64_t n, int64_t i)
{
  cu_freal<PB> r;
  for (int k = 0; k < cu_freal<PB>::N; k++) r.m[k] = l[(int64_t) k * n + i];
  int64_t w = se[i];
  r.sign = (w & 1) ? -1 : 1;
  r.exp = (long) (w >> 1);
  return r;
}
template <int PB> inline void st (const cu_freal<PB> &r, uint64_t *l, int64_t *se, int64_t n, int64_t i)
{
  for (int k = 0; k < cu_freal<PB>::N; k++) l[(int64_t) k * n + i] = r.m[k];
  se[i] = (int64_t) ((uint64_t) (int64_t) r.exp << 1) | (r.sign < 0 ? 1 : 0);
}
template <int PB> inline cu_fcomplex<PB> cld (const uint64_t *l, const int64_t *se, int64_t n, int64_t i)
{
  return cu_fcomplex<PB> (ld<PB> (l, se, n, i), ld<PB> (l + (int64_t) cu_freal<PB>::N * n, se + n, n, i));
}
template <int PB> inline void cst (const cu_fcomplex<PB> &z, uint64_t *l, int64_t *se, int64_t n, int64_t i)
{
  st<PB> (z.re, l, se, n, i);
  st<PB> (z.im, l + (int64_t) cu_freal<PB>::N * n, se + n, n, i);
}

template <int PB> void ew_real (int op, int64_t n, const uint64_t *al, const int64_t *ase,
                                const uint64_t *bl, const int64_t *bse, uint64_t *ol, int64_t *ose)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    {
      cu_freal<PB> a = ld<PB> (al, ase, n, i), b = ld<PB> (bl, bse, n, i), r;
      if (op == 0) r = cu_fp::cu_fmul<PB> (a, b);
      else if (op == 1) r = cu_fp::cu_fadd<PB> (a, b);
      else if (op == 2) r = cu_fp::cu_fsub<PB> (a, b);
      else r = cu_fp::cu_fdiv<PB> (a, b);
      st<PB> (r, ol, ose, n, i);
    }
}
template <int PB> void sqrt_real (int64_t n, const uint64_t *xl, const int64_t *xse, uint64_t *ol, int64_t *ose)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) st<PB> (cu_fp::cu_sqrt<PB> (ld<PB> (xl, xse, n, i)), ol, ose, n, i);
}
template <int PB> void ew_cplx (int op, int64_t n, const uint64_t *al, const int64_t *ase,
                                const uint64_t *bl, const int64_t *bse, uint64_t *ol, int64_t *ose)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    {
      cu_fcomplex<PB> a = cld<PB> (al, ase, n, i), b = cld<PB> (bl, bse, n, i), r;
      if (op == 0) r = cu_fp::cu_cmul<PB> (a, b);
      else if (op == 1) r = cu_fp::cu_cadd<PB> (a, b);
      else r = cu_fp::cu_csub<PB> (a, b);
      cst<PB> (r, ol, ose, n, i);
    }
}
/* demos/bench_3way.cu:44-50 on full-limb values */
template <int PB> void axpy_real (int64_t n, const uint64_t *al, const int64_t *ase,
                                  const uint64_t *xl, const int64_t *xse, const uint64_t *yl,
                                  const int64_t *yse, uint64_t *ol, int64_t *ose)
{
  const cu_freal<PB> a = ld<PB> (al, ase, 1, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    {
      cu_freal<PB> t = a * ld<PB> (xl, xse, n, i);
      st<PB> (t + ld<PB> (yl, yse, n, i), ol, ose, n, i);
    }
}
/* demos/bench_3way.cu:51-59 on full-limb values */
template <int PB> void axpy_cplx (int64_t n, const uint64_t *al, const int64_t *ase,
                                  const uint64_t *xl, const int64_t *xse, const uint64_t *yl,
                                  const int64_t *yse, uint64_t *ol, int64_t *ose)
{
  const cu_fcomplex<PB> a = cld<PB> (al, ase, 1, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    {
      cu_fcomplex<PB> t = a * cld<PB> (xl, xse, n, i);
      cst<PB> (t + cld<PB> (yl, yse, n, i), ol, ose, n, i);
    }
}
/* tools/linalg_fast.cu:53-68 on full-limb values, rows [r0, r1) */
template <int PB> void gemm_perop (int64_t m, int64_t n, int64_t k, const uint64_t *Al,
                                   const int64_t *Ase, const uint64_t *Bl, const int64_t *Bse,
                                   uint64_t *Cl, int64_t *Cse, int64_t r0, int64_t r1)
{
#pragma omp parallel for schedule(static) collapse(2)
  for (int64_t i = r0; i < r1; i++)
    for (int64_t j = 0; j < n; j++)
      {
        cu_freal<PB> acc; acc.set_zero ();
        for (int64_t kk = 0; kk < k; kk++)
          acc = cu_fp::cu_fadd<PB> (acc, cu_fp::cu_fmul<PB> (ld<PB> (Al, Ase, m * k, i * k + kk),
                                                             ld<PB> (Bl, Bse, k * n, kk * n + j)));
        st<PB> (acc, Cl, Cse, m * n, i * n + j);
      }
}

} // namespace

#define PB_LIST(X) X(32) X(64) X(96) X(128) X(160) X(192) X(256) X(288) X(320) X(384) \
                   X(512) X(1024) X(1056) X(2048)

extern "C" {

int ref_supported (int pb)
{
#define C(P) if (pb == P) return 1;
  PB_LIST (C)
#undef C
  return 0;
}

int ref_num_threads (void)
{
#ifdef _OPENMP
  return omp_get_max_threads ();
#else
  return 1;
#endif
}

int ref_elementwise_real (int pb, int op, int64_t n, const uint64_t *al, const int64_t *ase,
                          const uint64_t *bl, const int64_t *bse, uint64_t *ol, int64_t *ose)
{
#define C(P) if (pb == P) { ew_real<P> (op, n, al, ase, bl, bse, ol, ose); return 0; }
  PB_LIST (C)
#undef C
  return -1;
}
int ref_sqrt_real (int pb, int64_t n, const uint64_t *xl, const int64_t *xse, uint64_t *ol, int64_t *ose)
{
#define C(P) if (pb == P) { sqrt_real<P> (n, xl, xse, ol, ose); return 0; }
  PB_LIST (C)
#undef C
  return -1;
}
int ref_elementwise_complex (int pb, int op, int64_t n, const uint64_t *al, const int64_t *ase,
                             const uint64_t *bl, const int64_t *bse, uint64_t *ol, int64_t *ose)
{
#define C(P) if (pb == P) { ew_cplx<P> (op, n, al, ase, bl, bse, ol, ose); return 0; }
  PB_LIST (C)
#undef C
  return -1;
}
int ref_axpy_real (int pb, int64_t n, const uint64_t *al, const int64_t *ase, const uint64_t *xl,
                   const int64_t *xse, const uint64_t *yl, const int64_t *yse, uint64_t *ol,
                   int64_t *ose)
{
#define C(P) if (pb == P) { axpy_real<P> (n, al, ase, xl, xse, yl, yse, ol, ose); return 0; }
  PB_LIST (C)
#undef C
  return -1;
}
int ref_axpy_complex (int pb, int64_t n, const uint64_t *al, const int64_t *ase, const uint64_t *xl,
                      const int64_t *xse, const uint64_t *yl, const int64_t *yse, uint64_t *ol,
                      int64_t *ose)
{
#define C(P) if (pb == P) { axpy_cplx<P> (n, al, ase, xl, xse, yl, yse, ol, ose); return 0; }
  PB_LIST (C)
#undef C
  return -1;
}
int ref_gemm_perop (int pb, int64_t m, int64_t n, int64_t k, const uint64_t *Al, const int64_t *Ase,
                    const uint64_t *Bl, const int64_t *Bse, uint64_t *Cl, int64_t *Cse, int64_t r0,
                    int64_t r1)
{
#define C(P) if (pb == P) { gemm_perop<P> (m, n, k, Al, Ase, Bl, Bse, Cl, Cse, r0, r1); return 0; }
  PB_LIST (C)
#undef C
  return -1;
}

} // extern "C"
