/* tests/api/api_cases.h -- user code written against the drop-in template API
 * (include/mpc_cuda.cuh), the way the reference's own programs use it:
 *   - AoS arrays of cu_freal<PB> / cu_fcomplex<PB> in, AoS arrays out, one
 *     element per thread (tools/test_freal.cu:41-48, tools/test_fcomplex.cu
 *     kern<PB>);
 *   - doubles in, operators, doubles out (demos/sample_fixed.cu:20-30);
 *   - every public free function and member of cu_freal.cuh / cu_fcomplex.cuh.
 * Each case is one __host__ __device__ function of the element index, so the
 * same code runs in the GPU kernels of api_dev.hip and in the host loops of
 * api_host.cpp.  Test harness only. */
#ifndef API_CASES_H
#define API_CASES_H

#include "mpc_cuda.cuh"
#include <cstdint>

#ifndef __HIPCC__
#ifndef __host__
#define __host__
#define __device__
#endif
#endif
#define API_HD __host__ __device__ inline

namespace api {
using namespace cu_fp;
typedef int64_t i64;
typedef cu_fp::cu_limb u64;          /* unsigned long long, as cu_limb */

/* compiled precisions of the harness */
#define API_PBS(X) X(32) X(64) X(96) X(256) X(512) X(1024)

/* tools/test_freal.cu:41-48: mul / add / sub by the named functions */
template <int PB>
API_HD void arith (i64 i, const cu_freal<PB> *A, const cu_freal<PB> *B, cu_freal<PB> *MU, cu_freal<PB> *AD,
                   cu_freal<PB> *SU)
{
  MU[i] = cu_fmul<PB> (A[i], B[i]);
  AD[i] = cu_fadd<PB> (A[i], B[i]);
  SU[i] = cu_fsub<PB> (A[i], B[i]);
}

/* the same through the operators */
template <int PB>
API_HD void opers (i64 i, const cu_freal<PB> *A, const cu_freal<PB> *B, cu_freal<PB> *MU, cu_freal<PB> *AD,
                   cu_freal<PB> *SU)
{
  MU[i] = A[i] * B[i];
  AD[i] = A[i] + B[i];
  SU[i] = A[i] - B[i];
}

/* demos/sample_fixed.cu:20-30: doubles in, r = a*x + y with operators, double out */
template <int PB>
API_HD void axpy_double (i64 i, double a, const double *x, const double *y, double *out)
{
  cu_freal<PB> fa = a;
  cu_freal<PB> r = fa * cu_freal<PB> (x[i]) + cu_freal<PB> (y[i]);
  out[i] = (double) r;
}

/* magnitude helpers, sign utilities, comparison, integer rounding, scaling:
 * AM = cu_addmag(a, b, s), SM = cu_submag(big, small, s) with s = -1 on odd
 * indices; cmp[3 i + 0..2] = cu_cmpmag(a,b), cu_cmp(a,b), cu_sgn(a) */
template <int PB>
API_HD void utils (i64 i, const cu_freal<PB> *A, const cu_freal<PB> *B, cu_freal<PB> *AM, cu_freal<PB> *SM,
                   cu_freal<PB> *NG, cu_freal<PB> *AB, cu_freal<PB> *SC, int *cmp, long *lr)
{
  const cu_freal<PB> a = A[i], b = B[i];
  const int s = (i & 1) ? -1 : 1;
  const int cm = cu_cmpmag<PB> (a, b);
  AM[i] = cu_addmag<PB> (a, b, s);
  SM[i] = cm >= 0 ? cu_submag<PB> (a, b, s) : cu_submag<PB> (b, a, s);
  NG[i] = cu_neg<PB> (a);
  AB[i] = cu_abs<PB> (a);
  SC[i] = cu_scale2<PB> (a, (long) (i % 41) - 20);
  cmp[3 * i + 0] = cm;
  cmp[3 * i + 1] = cu_cmp<PB> (a, b);
  cmp[3 * i + 2] = cu_sgn<PB> (a);
  lr[i] = cu_lround<PB> (a);
}

/* cu_convert to 32, 64 and 1024 bits; cu_from_si */
template <int PB>
API_HD void convert (i64 i, const cu_freal<PB> *A, const long *V, cu_freal<32> *C32, cu_freal<64> *C64,
                     cu_freal<1024> *C1024, cu_freal<PB> *SI)
{
  C32[i] = cu_convert<32, PB> (A[i]);
  C64[i] = cu_convert<64, PB> (A[i]);
  C1024[i] = cu_convert<1024, PB> (A[i]);
  SI[i] = cu_from_si<PB> (V[i]);
}

/* from_double / the double constructor, to_double / the double cast */
template <int PB>
API_HD void doubles (i64 i, const double *D, const cu_freal<PB> *A, cu_freal<PB> *FD, double *back, double *TD)
{
  const cu_freal<PB> f (D[i]);
  FD[i] = cu_freal<PB>::from_double (D[i]);
  back[i] = (double) f;
  TD[i] = A[i].to_double ();
}

/* cu_wbit / cu_wsticky_below at bit positions IX[i] of A[i]'s limbs */
template <int PB>
API_HD void bits (i64 i, const cu_freal<PB> *A, const long *IX, u64 *WB, u64 *WS)
{
  WB[i] = cu_wbit (A[i].m, cu_freal<PB>::N, IX[i]);
  WS[i] = cu_wsticky_below (A[i].m, cu_freal<PB>::N, IX[i]);
}

/* cu_umul on the low limbs, cu_mul_full on the significands */
template <int PB>
API_HD void products (i64 i, const cu_freal<PB> *A, const cu_freal<PB> *B, u64 *HL, u64 *R)
{
  cu_umul (HL[2 * i + 0], HL[2 * i + 1], A[i].m[0], B[i].m[0]);
  cu_limb r[2 * cu_freal<PB>::N];
  cu_mul_full<PB> (r, A[i].m, B[i].m);
  for (int k = 0; k < 2 * cu_freal<PB>::N; k++) R[(i64) 2 * cu_freal<PB>::N * i + k] = r[k];
}

/* cu_freal<PB>::finalize on a LEN-limb buffer with round / sticky inputs */
template <int PB>
API_HD void finalize (i64 i, int len, const u64 *D, const int *MSB, const long *EHI, const u64 *RB, const u64 *SB,
                      const int *SG, cu_freal<PB> *OUT)
{
  OUT[i] = cu_freal<PB>::finalize (D + (i64) len * i, len, MSB[i], EHI[i], RB[i], SB[i], SG[i]);
}

/* wide terms: cu_mk_term, cu_wcmp, cu_round_sum2, cu_fmma, cu_fmms */
template <int PB>
API_HD void terms (i64 i, const cu_freal<PB> *A, const cu_freal<PB> *B, const cu_freal<PB> *C,
                   const cu_freal<PB> *D, cu_wterm<PB> *T1, cu_wterm<PB> *T2, int *WC, cu_freal<PB> *RS,
                   cu_freal<PB> *FA, cu_freal<PB> *FS)
{
  const cu_wterm<PB> t1 = cu_mk_term<PB> (A[i], B[i]), t2 = cu_mk_term<PB> (C[i], D[i]);
  T1[i] = t1;
  T2[i] = t2;
  WC[i] = cu_wcmp<PB> (t1, t2);
  RS[i] = cu_round_sum2<PB> (t1, t2);
  FA[i] = cu_fmma<PB> (A[i], B[i], C[i], D[i]);
  FS[i] = cu_fmms<PB> (A[i], B[i], C[i], D[i]);
}

/* tools/test_fcomplex.cu kern<PB>: complex mul / add / sub by the named
 * functions and by the operators */
template <int PB>
API_HD void cplx (i64 i, const cu_fcomplex<PB> *A, const cu_fcomplex<PB> *B, cu_fcomplex<PB> *F3,
                  cu_fcomplex<PB> *O3)
{
  F3[3 * i + 0] = cu_cmul<PB> (A[i], B[i]);
  F3[3 * i + 1] = cu_cadd<PB> (A[i], B[i]);
  F3[3 * i + 2] = cu_csub<PB> (A[i], B[i]);
  O3[3 * i + 0] = A[i] * B[i];
  O3[3 * i + 1] = A[i] + B[i];
  O3[3 * i + 2] = A[i] - B[i];
}

/* complex values from doubles: z = a*x + y (demos/sample_fixed.cu:61-66),
 * the constructors, from_doubles, real_d / imag_d */
template <int PB>
API_HD void cplx_doubles (i64 i, const double *D4, double *OUT2, cu_fcomplex<PB> *Z)
{
  const double *d = D4 + 6 * i;
  const cu_fcomplex<PB> a (d[0], d[1]);
  const cu_fcomplex<PB> x = cu_fcomplex<PB>::from_doubles (d[2], d[3]);
  const cu_freal<PB> yr = d[4], yi = d[5];
  const cu_fcomplex<PB> y (yr, yi);
  const cu_fcomplex<PB> z = a * x + y;
  Z[i] = z;
  OUT2[2 * i + 0] = z.real_d ();
  OUT2[2 * i + 1] = z.imag_d ();
}

} // namespace api
#endif
