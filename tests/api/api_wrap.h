/* tests/api/api_wrap.h -- extern "C" entry points of the API test harness,
 * shared by the GPU build (api_dev.hip: kernels, device copies) and the host
 * build (api_host.cpp: loops over host arrays).  The includer defines
 *   struct Ctx { T *in<T>(const void *h, i64 count); T *out<T>(void *h, i64 count);
 *                int finish (); };
 *   API_RUN(CASE, PB, n, args...)  -- run api::CASE<PB>(i, args...) for i < n
 * Every entry point takes host arrays and returns 0, -1 for a precision that
 * is not compiled in, or a positive HIP error code. */
#ifndef API_WRAP_H
#define API_WRAP_H

namespace api {

template <int PB>
int w_arith (int named, i64 n, const void *A, const void *B, void *MU, void *AD, void *SU)
{
  typedef cu_freal<PB> F;
  Ctx c;
  const F *a = c.in<F> (A, n), *b = c.in<F> (B, n);
  F *mu = c.out<F> (MU, n), *ad = c.out<F> (AD, n), *su = c.out<F> (SU, n);
  if (named) API_RUN (arith, PB, n, a, b, mu, ad, su);
  else API_RUN (opers, PB, n, a, b, mu, ad, su);
  return c.finish ();
}

template <int PB>
int w_axpy_double (i64 n, double a, const double *X, const double *Y, double *OUT)
{
  Ctx c;
  const double *x = c.in<double> (X, n), *y = c.in<double> (Y, n);
  double *o = c.out<double> (OUT, n);
  API_RUN (axpy_double, PB, n, a, x, y, o);
  return c.finish ();
}

template <int PB>
int w_utils (i64 n, const void *A, const void *B, void *AM, void *SM, void *NG, void *AB, void *SC, int *CMP,
             long *LR)
{
  typedef cu_freal<PB> F;
  Ctx c;
  const F *a = c.in<F> (A, n), *b = c.in<F> (B, n);
  F *am = c.out<F> (AM, n), *sm = c.out<F> (SM, n), *ng = c.out<F> (NG, n), *ab = c.out<F> (AB, n),
    *sc = c.out<F> (SC, n);
  int *cmp = c.out<int> (CMP, 3 * n);
  long *lr = c.out<long> (LR, n);
  API_RUN (utils, PB, n, a, b, am, sm, ng, ab, sc, cmp, lr);
  return c.finish ();
}

template <int PB>
int w_convert (i64 n, const void *A, const long *V, void *C32, void *C64, void *C1024, void *SI)
{
  typedef cu_freal<PB> F;
  Ctx c;
  const F *a = c.in<F> (A, n);
  const long *v = c.in<long> (V, n);
  cu_freal<32> *c32 = c.out<cu_freal<32>> (C32, n);
  cu_freal<64> *c64 = c.out<cu_freal<64>> (C64, n);
  cu_freal<1024> *c1024 = c.out<cu_freal<1024>> (C1024, n);
  F *si = c.out<F> (SI, n);
  API_RUN (convert, PB, n, a, v, c32, c64, c1024, si);
  return c.finish ();
}

template <int PB>
int w_doubles (i64 n, const double *D, const void *A, void *FD, double *BACK, double *TD)
{
  typedef cu_freal<PB> F;
  Ctx c;
  const double *d = c.in<double> (D, n);
  const F *a = c.in<F> (A, n);
  F *fd = c.out<F> (FD, n);
  double *back = c.out<double> (BACK, n), *td = c.out<double> (TD, n);
  API_RUN (doubles, PB, n, d, a, fd, back, td);
  return c.finish ();
}

template <int PB>
int w_bits (i64 n, const void *A, const long *IX, u64 *WB, u64 *WS)
{
  typedef cu_freal<PB> F;
  Ctx c;
  const F *a = c.in<F> (A, n);
  const long *ix = c.in<long> (IX, n);
  u64 *wb = c.out<u64> (WB, n), *ws = c.out<u64> (WS, n);
  API_RUN (bits, PB, n, a, ix, wb, ws);
  return c.finish ();
}

template <int PB>
int w_products (i64 n, const void *A, const void *B, u64 *HL, u64 *R)
{
  typedef cu_freal<PB> F;
  Ctx c;
  const F *a = c.in<F> (A, n), *b = c.in<F> (B, n);
  u64 *hl = c.out<u64> (HL, 2 * n), *r = c.out<u64> (R, 2 * F::N * n);
  API_RUN (products, PB, n, a, b, hl, r);
  return c.finish ();
}

template <int PB>
int w_finalize (i64 n, int len, const u64 *D, const int *MSB, const long *EHI, const u64 *RB, const u64 *SB,
                const int *SG, void *OUT)
{
  typedef cu_freal<PB> F;
  Ctx c;
  const u64 *d = c.in<u64> (D, (i64) len * n), *rb = c.in<u64> (RB, n), *sb = c.in<u64> (SB, n);
  const int *msb = c.in<int> (MSB, n), *sg = c.in<int> (SG, n);
  const long *ehi = c.in<long> (EHI, n);
  F *o = c.out<F> (OUT, n);
  API_RUN (finalize, PB, n, len, d, msb, ehi, rb, sb, sg, o);
  return c.finish ();
}

template <int PB>
int w_terms (i64 n, const void *A, const void *B, const void *C, const void *D, void *T1, void *T2, int *WC,
             void *RS, void *FA, void *FS)
{
  typedef cu_freal<PB> F;
  typedef cu_wterm<PB> T;
  Ctx c;
  const F *a = c.in<F> (A, n), *b = c.in<F> (B, n), *cc = c.in<F> (C, n), *d = c.in<F> (D, n);
  T *t1 = c.out<T> (T1, n), *t2 = c.out<T> (T2, n);
  int *wc = c.out<int> (WC, n);
  F *rs = c.out<F> (RS, n), *fa = c.out<F> (FA, n), *fs = c.out<F> (FS, n);
  API_RUN (terms, PB, n, a, b, cc, d, t1, t2, wc, rs, fa, fs);
  return c.finish ();
}

template <int PB>
int w_cplx (i64 n, const void *A, const void *B, void *F3, void *O3)
{
  typedef cu_fcomplex<PB> Z;
  Ctx c;
  const Z *a = c.in<Z> (A, n), *b = c.in<Z> (B, n);
  Z *f3 = c.out<Z> (F3, 3 * n), *o3 = c.out<Z> (O3, 3 * n);
  API_RUN (cplx, PB, n, a, b, f3, o3);
  return c.finish ();
}

template <int PB>
int w_cplx_doubles (i64 n, const double *D6, double *OUT2, void *Zo)
{
  typedef cu_fcomplex<PB> Z;
  Ctx c;
  const double *d = c.in<double> (D6, 6 * n);
  double *o = c.out<double> (OUT2, 2 * n);
  Z *z = c.out<Z> (Zo, n);
  API_RUN (cplx_doubles, PB, n, d, o, z);
  return c.finish ();
}

} // namespace api

#define API_DISPATCH(CALL) \
  switch (pb) { API_PBS (API_CASE_PB) default: return -1; }
#define API_CASE_PB(P) case P: return API_CALL_AT (P);

extern "C" {
int api_sizeof_real (int pb)
{
#define API_CALL_AT(P) (int) sizeof (cu_fp::cu_freal<P>)
  API_DISPATCH (0)
#undef API_CALL_AT
}
int api_sizeof_wterm (int pb)
{
#define API_CALL_AT(P) (int) sizeof (cu_fp::cu_wterm<P>)
  API_DISPATCH (0)
#undef API_CALL_AT
}
int api_arith (int pb, int named, api::i64 n, const void *A, const void *B, void *MU, void *AD, void *SU)
{
#define API_CALL_AT(P) api::w_arith<P> (named, n, A, B, MU, AD, SU)
  API_DISPATCH (0)
#undef API_CALL_AT
}
int api_axpy_double (int pb, api::i64 n, double a, const double *X, const double *Y, double *OUT)
{
#define API_CALL_AT(P) api::w_axpy_double<P> (n, a, X, Y, OUT)
  API_DISPATCH (0)
#undef API_CALL_AT
}
int api_utils (int pb, api::i64 n, const void *A, const void *B, void *AM, void *SM, void *NG, void *AB, void *SC,
               int *CMP, long *LR)
{
#define API_CALL_AT(P) api::w_utils<P> (n, A, B, AM, SM, NG, AB, SC, CMP, LR)
  API_DISPATCH (0)
#undef API_CALL_AT
}
int api_convert (int pb, api::i64 n, const void *A, const long *V, void *C32, void *C64, void *C1024, void *SI)
{
#define API_CALL_AT(P) api::w_convert<P> (n, A, V, C32, C64, C1024, SI)
  API_DISPATCH (0)
#undef API_CALL_AT
}
int api_doubles (int pb, api::i64 n, const double *D, const void *A, void *FD, double *BACK, double *TD)
{
#define API_CALL_AT(P) api::w_doubles<P> (n, D, A, FD, BACK, TD)
  API_DISPATCH (0)
#undef API_CALL_AT
}
int api_bits (int pb, api::i64 n, const void *A, const long *IX, api::u64 *WB, api::u64 *WS)
{
#define API_CALL_AT(P) api::w_bits<P> (n, A, IX, WB, WS)
  API_DISPATCH (0)
#undef API_CALL_AT
}
int api_products (int pb, api::i64 n, const void *A, const void *B, api::u64 *HL, api::u64 *R)
{
#define API_CALL_AT(P) api::w_products<P> (n, A, B, HL, R)
  API_DISPATCH (0)
#undef API_CALL_AT
}
int api_finalize (int pb, api::i64 n, int len, const api::u64 *D, const int *MSB, const long *EHI,
                  const api::u64 *RB, const api::u64 *SB, const int *SG, void *OUT)
{
#define API_CALL_AT(P) api::w_finalize<P> (n, len, D, MSB, EHI, RB, SB, SG, OUT)
  API_DISPATCH (0)
#undef API_CALL_AT
}
int api_terms (int pb, api::i64 n, const void *A, const void *B, const void *C, const void *D, void *T1, void *T2,
               int *WC, void *RS, void *FA, void *FS)
{
#define API_CALL_AT(P) api::w_terms<P> (n, A, B, C, D, T1, T2, WC, RS, FA, FS)
  API_DISPATCH (0)
#undef API_CALL_AT
}
int api_cplx (int pb, api::i64 n, const void *A, const void *B, void *F3, void *O3)
{
#define API_CALL_AT(P) api::w_cplx<P> (n, A, B, F3, O3)
  API_DISPATCH (0)
#undef API_CALL_AT
}
int api_cplx_doubles (int pb, api::i64 n, const double *D6, double *OUT2, void *Z)
{
#define API_CALL_AT(P) api::w_cplx_doubles<P> (n, D6, OUT2, Z)
  API_DISPATCH (0)
#undef API_CALL_AT
}
}
#endif
