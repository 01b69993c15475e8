/* api.hip -- the C ABI of the mpgfx library (declared in include/mpgfx.h):
 * argument checking, precision dispatch and error bookkeeping.  Every entry
 * point validates its arguments on the host and launches nothing on error. */
#include "common.h"
#include "dispatch.h"
#include "rt.h"
#include <cstring>

namespace mpgfx {

static thread_local char g_err[256];

void set_error (const char *where, hipError_t e)
{
  snprintf (g_err, sizeof g_err, "%s: %s", where, hipGetErrorString (e));
}

void set_error_text (const char *msg)
{
  snprintf (g_err, sizeof g_err, "%s", msg);
}

int launch_status (const char *where)
{
  hipError_t e = hipGetLastError ();
  if (e != hipSuccess) { set_error (where, e); return MPGFX_EHIP; }
  return MPGFX_OK;
}

const pb_ops *find_ops (int pb)
{
#define C(P) if (pb == P) return ops_pb_##P ();
  MPGFX_PB_LIST (C)
#undef C
  return nullptr;
}

} // namespace mpgfx

using namespace mpgfx;

#define NEED(p) do { if (!(p)) return MPGFX_EINVAL; } while (0)
#define GET_OPS(o) const pb_ops *o = find_ops (pb); if (!o) return MPGFX_EUNSUPPORTED

extern "C" int mpgfx_version (void) { return 100; }

extern "C" int mpgfx_supported_pb (int pb) { return find_ops (pb) != nullptr; }

extern "C" const char *mpgfx_last_error (void) { return g_err; }

extern "C" int mpgfx_elementwise_real (int pb, int op, int64_t n, const uint64_t *al, const int64_t *ase,
                                       const uint64_t *bl, const int64_t *bse, uint64_t *ol, int64_t *ose,
                                       void *stream)
{
  if (n < 0 || op < MPGFX_MUL || op > MPGFX_DIV) return MPGFX_EINVAL;
  GET_OPS (o);
  if (n == 0) return MPGFX_OK;
  NEED (al && ase && bl && bse && ol && ose);
  return o->ew_real (op, n, al, ase, bl, bse, ol, ose, (hipStream_t) stream);
}

extern "C" int mpgfx_sqrt_real (int pb, int64_t n, const uint64_t *xl, const int64_t *xse, uint64_t *ol,
                                int64_t *ose, void *stream)
{
  if (n < 0) return MPGFX_EINVAL;
  GET_OPS (o);
  if (n == 0) return MPGFX_OK;
  NEED (xl && xse && ol && ose);
  return o->sqrt_real (n, xl, xse, ol, ose, (hipStream_t) stream);
}

extern "C" int mpgfx_elementwise_complex (int pb, int op, int64_t n, const uint64_t *al, const int64_t *ase,
                                          const uint64_t *bl, const int64_t *bse, uint64_t *ol, int64_t *ose,
                                          void *stream)
{
  if (n < 0 || op < MPGFX_MUL || op > MPGFX_SUB) return MPGFX_EINVAL;
  GET_OPS (o);
  if (n == 0) return MPGFX_OK;
  NEED (al && ase && bl && bse && ol && ose);
  return o->ew_cplx (op, n, al, ase, bl, bse, ol, ose, (hipStream_t) stream);
}

extern "C" int mpgfx_axpy_real (int pb, int64_t n, const uint64_t *al, const int64_t *ase, const uint64_t *xl,
                                const int64_t *xse, const uint64_t *yl, const int64_t *yse, uint64_t *ol,
                                int64_t *ose, void *stream)
{
  if (n < 0) return MPGFX_EINVAL;
  GET_OPS (o);
  if (n == 0) return MPGFX_OK;
  NEED (al && ase && xl && xse && yl && yse && ol && ose);
  return o->axpy_real (n, al, ase, xl, xse, yl, yse, ol, ose, (hipStream_t) stream);
}

extern "C" int mpgfx_axpy_complex (int pb, int64_t n, const uint64_t *al, const int64_t *ase, const uint64_t *xl,
                                   const int64_t *xse, const uint64_t *yl, const int64_t *yse, uint64_t *zl,
                                   int64_t *zse, void *stream)
{
  if (n < 0) return MPGFX_EINVAL;
  GET_OPS (o);
  if (n == 0) return MPGFX_OK;
  NEED (al && ase && xl && xse && yl && yse && zl && zse);
  return o->axpy_cplx (n, al, ase, xl, xse, yl, yse, zl, zse, (hipStream_t) stream);
}

extern "C" int mpgfx_gemv_real (int pb, int64_t m, int64_t k, const uint64_t *Al, const int64_t *Ase,
                                const uint64_t *xl, const int64_t *xse, uint64_t *yl, int64_t *yse, void *stream)
{
  if (m < 0 || k < 0) return MPGFX_EINVAL;
  GET_OPS (o);
  if (m == 0) return MPGFX_OK;
  NEED (yl && yse);
  if (k > 0) NEED (Al && Ase && xl && xse);
  return o->gemv (m, k, Al, Ase, xl, xse, yl, yse, (hipStream_t) stream);
}

extern "C" int mpgfx_gemv_complex (int pb, int64_t m, int64_t k, const uint64_t *Al, const int64_t *Ase,
                                   const uint64_t *xl, const int64_t *xse, uint64_t *yl, int64_t *yse, void *stream)
{
  if (m < 0 || k < 0) return MPGFX_EINVAL;
  GET_OPS (o);
  if (m == 0) return MPGFX_OK;
  NEED (yl && yse);
  if (k > 0) NEED (Al && Ase && xl && xse);
  return o->gemm_cplx (m, 1, k, Al, Ase, xl, xse, yl, yse, (hipStream_t) stream);
}

extern "C" int mpgfx_gemm_workspace_size (int pb, int mode, int64_t m, int64_t n, int64_t k, size_t *bytes)
{
  if (!bytes || m < 0 || n < 0 || k < 0) return MPGFX_EINVAL;
  if (mode != MPGFX_GEMM_PEROP && mode != MPGFX_GEMM_EXACT) return MPGFX_EINVAL;
  if (!find_ops (pb)) return MPGFX_EUNSUPPORTED;
  *bytes = (mode == MPGFX_GEMM_EXACT) ? exact_workspace_bound (pb, m, n, k) : 0;
  return MPGFX_OK;
}

extern "C" int mpgfx_gemm_real (int pb, int mode, int64_t m, int64_t n, int64_t k, const uint64_t *Al,
                                const int64_t *Ase, const uint64_t *Bl, const int64_t *Bse, uint64_t *Cl,
                                int64_t *Cse, void *ws, size_t ws_bytes, void *stream)
{
  if (m < 0 || n < 0 || k < 0) return MPGFX_EINVAL;
  if (mode != MPGFX_GEMM_PEROP && mode != MPGFX_GEMM_EXACT) return MPGFX_EINVAL;
  GET_OPS (o);
  if (m == 0 || n == 0) return MPGFX_OK;
  NEED (Cl && Cse);
  if (k > 0) NEED (Al && Ase && Bl && Bse);
  if (mode == MPGFX_GEMM_EXACT)
    return gemm_exact (o, pb, m, n, k, Al, Ase, Bl, Bse, Cl, Cse, ws, ws_bytes, (hipStream_t) stream, false);
  return o->gemm_perop (m, n, k, Al, Ase, Bl, Bse, Cl, Cse, (hipStream_t) stream);
}

extern "C" int mpgfx_gemm_exact_noalloc (int pb, int64_t m, int64_t n, int64_t k, const uint64_t *Al,
                                         const int64_t *Ase, const uint64_t *Bl, const int64_t *Bse, uint64_t *Cl,
                                         int64_t *Cse, void *stream)
{
  if (m < 0 || n < 0 || k < 0) return MPGFX_EINVAL;
  GET_OPS (o);
  if (m == 0 || n == 0) return MPGFX_OK;
  NEED (Cl && Cse);
  if (k > 0) NEED (Al && Ase && Bl && Bse);
  void *scr = nullptr;
  const int rc = exact_noalloc_scratch (&scr);
  if (rc) return rc;
  return gemm_exact (o, pb, m, n, k, Al, Ase, Bl, Bse, Cl, Cse, scr, EXACT_NOALLOC_BYTES, (hipStream_t) stream, true);
}

extern "C" int mpgfx_gemm_complex (int pb, int64_t m, int64_t n, int64_t k, const uint64_t *Al, const int64_t *Ase,
                                   const uint64_t *Bl, const int64_t *Bse, uint64_t *Cl, int64_t *Cse, void *stream)
{
  if (m < 0 || n < 0 || k < 0) return MPGFX_EINVAL;
  GET_OPS (o);
  if (m == 0 || n == 0) return MPGFX_OK;
  NEED (Cl && Cse);
  if (k > 0) NEED (Al && Ase && Bl && Bse);
  return o->gemm_cplx (m, n, k, Al, Ase, Bl, Bse, Cl, Cse, (hipStream_t) stream);
}

extern "C" int mpgfx_gemm_complex_exact (int pb, int64_t m, int64_t n, int64_t k, const uint64_t *Al,
                                         const int64_t *Ase, const uint64_t *Bl, const int64_t *Bse, uint64_t *Cl,
                                         int64_t *Cse, void *stream)
{
  if (m < 0 || n < 0 || k < 0) return MPGFX_EINVAL;
  GET_OPS (o);
  if (m == 0 || n == 0) return MPGFX_OK;
  NEED (Cl && Cse);
  if (k > 0) NEED (Al && Ase && Bl && Bse);
  return gemm_exact_complex (o, pb, m, n, k, Al, Ase, Bl, Bse, Cl, Cse, (hipStream_t) stream);
}

extern "C" int mpgfx_pack (int pb, int cplx, int64_t n, const void *aos, uint64_t *l, int64_t *se, void *stream)
{
  if (n < 0) return MPGFX_EINVAL;
  GET_OPS (o);
  if (n == 0) return MPGFX_OK;
  NEED (aos && l && se);
  return o->pack (cplx ? 1 : 0, n, aos, l, se, (hipStream_t) stream);
}

extern "C" int mpgfx_unpack (int pb, int cplx, int64_t n, const uint64_t *l, const int64_t *se, void *aos,
                             void *stream)
{
  if (n < 0) return MPGFX_EINVAL;
  GET_OPS (o);
  if (n == 0) return MPGFX_OK;
  NEED (aos && l && se);
  return o->unpack (cplx ? 1 : 0, n, l, se, aos, (hipStream_t) stream);
}

/* ---- seeded synthetic values (same stream as mpc_cuda_amd/gen.py) ---- */
namespace mpgfx {

__device__ __forceinline__ u64 splitmix64 (u64 x)
{
  x += 0x9E3779B97F4A7C15ull;
  u64 z = x;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ u64 xs64 (u64 s)
{
  s ^= s << 13; s ^= s >> 7; s ^= s << 17;
  return s;
}

/* value i (of n, starting at `start`): stream s = splitmix64(i ^ splitmix64(seed)),
 * then xorshift64 draws: sign, nl limbs (top bit forced, low sb bits cleared),
 * exponent = draw % (emax-emin+1) + emin */
__global__ void k_fill_random (i64 n, i64 start, u64 seed, int nl, int sb, i64 emin, i64 emax, u64 *l, i64 *se)
{
  i64 i = (i64) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u64 s = splitmix64 ((u64) (start + i) ^ splitmix64 (seed));
  if (s == 0) s = 1;
  s = xs64 (s);
  i64 neg = (i64) (s & 1);
  for (int q = 0; q < nl; q++)
    {
      s = xs64 (s);
      u64 v = s;
      if (q == nl - 1) v |= 1ull << 63;
      if (q == 0 && sb) v &= ~((1ull << sb) - 1);
      l[(i64) q * n + i] = v;
    }
  s = xs64 (s);
  u64 span = (u64) (emax - emin + 1);
  i64 e = (i64) (s % span) + emin;
  se[i] = (i64) ((u64) e << 1) | neg;
}

} // namespace mpgfx

extern "C" int mpgfx_fill_random (int pb, int64_t n, int64_t start, uint64_t seed, int64_t emin, int64_t emax,
                                  uint64_t *limbs, int64_t *se, void *stream)
{
  if (n < 0 || emax < emin || pb <= 0 || pb % 32) return MPGFX_EINVAL;
  if (n == 0) return MPGFX_OK;
  if (!limbs || !se) return MPGFX_EINVAL;
  int nl = (pb + 63) / 64, sb = 64 * nl - pb;
  hipLaunchKernelGGL (mpgfx::k_fill_random, dim3 ((unsigned) ((n + 255) / 256)), dim3 (256), 0, (hipStream_t) stream, n,
                      start, seed, nl, sb, emin, emax, limbs, se);
  return launch_status ("mpgfx_fill_random");
}

/* ---------------- run-time precision (rt.h, rt_kernels.hip) ---------------- */
namespace {
const mpgfx::rt_ops *rt_get (int p)
{
  const int nl = (p + 63) / 64;
  if (nl <= 1) return mpgfx::rt_ops_cw_2 ();
  if (nl <= 2) return mpgfx::rt_ops_cw_4 ();
  if (nl <= 4) return mpgfx::rt_ops_cw_8 ();
  if (nl <= 8) return mpgfx::rt_ops_cw_16 ();
  if (nl <= 16) return mpgfx::rt_ops_cw_32 ();
  return mpgfx::rt_ops_cw_64 ();
}
int rt_check (int prec, int64_t n)
{
  if (n < 0 || prec < 1) return MPGFX_EINVAL;
  if (prec > mpgfx::MPGFX_RT_MAXPREC) return MPGFX_EUNSUPPORTED;
  return MPGFX_OK;
}
} // namespace

extern "C" int mpgfx_rt_elementwise_real (int prec, int op, int64_t n, const uint64_t *al, const int64_t *ase,
                                          const uint64_t *bl, const int64_t *bse, uint64_t *ol, int64_t *ose,
                                          void *stream)
{
  if (op < MPGFX_MUL || op > MPGFX_SUB) return MPGFX_EINVAL;
  if (int rc = rt_check (prec, n)) return rc;
  if (n == 0) return MPGFX_OK;
  NEED (al && ase && bl && bse && ol && ose);
  return rt_get (prec)->ew (op, prec, n, al, ase, bl, bse, ol, ose, (hipStream_t) stream);
}

extern "C" int mpgfx_rt_axpy_real (int prec, int64_t n, const uint64_t *al, const int64_t *ase, const uint64_t *xl,
                                   const int64_t *xse, const uint64_t *yl, const int64_t *yse, uint64_t *ol,
                                   int64_t *ose, void *stream)
{
  if (int rc = rt_check (prec, n)) return rc;
  if (n == 0) return MPGFX_OK;
  NEED (al && ase && xl && xse && yl && yse && ol && ose);
  return rt_get (prec)->axpy_real (prec, n, al, ase, xl, xse, yl, yse, ol, ose, (hipStream_t) stream);
}

extern "C" int mpgfx_rt_axpy_complex (int prec, int64_t n, const uint64_t *al, const int64_t *ase,
                                      const uint64_t *xl, const int64_t *xse, const uint64_t *yl,
                                      const int64_t *yse, uint64_t *zl, int64_t *zse, void *stream)
{
  if (int rc = rt_check (prec, n)) return rc;
  if (n == 0) return MPGFX_OK;
  NEED (al && ase && xl && xse && yl && yse && zl && zse);
  return rt_get (prec)->axpy_cplx (prec, n, al, ase, xl, xse, yl, yse, zl, zse, (hipStream_t) stream);
}

/* zero outputs of an empty sum (k = 0): +0, sign/exponent words only */
namespace {
__global__ void k_rt_zero (int64_t n, int nl, uint64_t *l, int64_t *se)
{
  const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < nl; q++) l[(int64_t) q * n + i] = 0;
  se[i] = (int64_t) ((uint64_t) MPGFX_ZERO_EXP << 1);
}
int rt_zero_fill (int prec, int64_t n, int parts, uint64_t *l, int64_t *se, hipStream_t s)
{
  const int nl = (prec + 63) / 64;
  for (int q = 0; q < parts; q++)
    hipLaunchKernelGGL (k_rt_zero, dim3 ((unsigned) ((n + 255) / 256)), dim3 (256), 0, s, n, nl, l + (int64_t) q * nl * n,
                        se + (int64_t) q * n);
  return launch_status ("rt_zero");
}
} // namespace

extern "C" int mpgfx_rt_gemv (int prec, int cplx, int64_t m, int64_t k, const uint64_t *Al, const int64_t *Ase,
                              const uint64_t *xl, const int64_t *xse, uint64_t *yl, int64_t *yse, void *stream)
{
  if (m < 0 || k < 0 || (cplx != 0 && cplx != 1)) return MPGFX_EINVAL;
  if (int rc = rt_check (prec, m)) return rc;
  if (m == 0) return MPGFX_OK;
  NEED (yl && yse);
  if (k == 0) return rt_zero_fill (prec, m, cplx ? 2 : 1, yl, yse, (hipStream_t) stream);
  NEED (Al && Ase && xl && xse);
  return rt_get (prec)->gemv (cplx, prec, m, k, Al, Ase, xl, xse, yl, yse, (hipStream_t) stream);
}

extern "C" int mpgfx_rt_gemm (int prec, int cplx, int64_t m, int64_t n, int64_t k, const uint64_t *Al,
                              const int64_t *Ase, const uint64_t *Bl, const int64_t *Bse, uint64_t *Cl, int64_t *Cse,
                              void *stream)
{
  if (m < 0 || n < 0 || k < 0 || (cplx != 0 && cplx != 1)) return MPGFX_EINVAL;
  if (int rc = rt_check (prec, m)) return rc;
  if (m == 0 || n == 0) return MPGFX_OK;
  NEED (Cl && Cse);
  if (k == 0) return rt_zero_fill (prec, m * n, cplx ? 2 : 1, Cl, Cse, (hipStream_t) stream);
  NEED (Al && Ase && Bl && Bse);
  return rt_get (prec)->gemm (cplx, prec, m, n, k, Al, Ase, Bl, Bse, Cl, Cse, (hipStream_t) stream);
}
