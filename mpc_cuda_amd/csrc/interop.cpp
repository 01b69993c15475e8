/* interop.cpp -- host-side import/export between arrays of MPFR value structs
 * and device MP arrays (mpgfx_import_mpfr / mpgfx_export_mpfr,
 * include/mpgfx.h).
 *
 * The struct is MPFR's public `__mpfr_struct` {prec, sign, exp, limb pointer}
 * (mpfr.h; LP64: long, int, long, pointer), whose special exponents are
 * 0 - EXP_MAX (zero), 1 - EXP_MAX (NaN) and 2 - EXP_MAX (infinity), EXP_MAX
 * the largest long.  A regular value's limbs and exponent are exactly the
 * storage of cu_freal<PB>, so the copy is a gather / scatter of limbs. */
#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>
#include <cstring>
#include <vector>
#include "mpgfx.h"

namespace mpgfx {
void set_error (const char *where, hipError_t e);
}

namespace {

struct mpfr_view
{
  long prec;
  int sign;
  long exp;
  uint64_t *d;
};
static_assert (sizeof (mpfr_view) == 32, "MPFR value struct on LP64");

constexpr long EXP_MAX = LONG_MAX;
constexpr long MPFR_EXP_ZERO = 0 - EXP_MAX;
constexpr long MPFR_EXP_INF = 2 - EXP_MAX;
constexpr int64_t ZERO_EXP = -0x7fffffffffffffLL;

int hip_ok (const char *where, hipError_t e)
{
  if (e != hipSuccess) { mpgfx::set_error (where, e); return MPGFX_EHIP; }
  return MPGFX_OK;
}

/* the compiled precisions, and every precision of the run-time path */
bool pb_ok (int pb) { return (pb >= 1 && pb <= 2048) || mpgfx_supported_pb (pb); }

} // namespace

extern "C" int mpgfx_import_mpfr (int pb, int64_t n, const void *x, size_t stride, uint64_t *limbs, int64_t *se,
                                  void *stream)
{
  if (n < 0 || stride < sizeof (mpfr_view)) return MPGFX_EINVAL;
  if (!pb_ok (pb)) return MPGFX_EUNSUPPORTED;
  if (n == 0) return MPGFX_OK;
  if (!x || !limbs || !se) return MPGFX_EINVAL;
  const int N = (pb + 63) / 64;
  std::vector<uint64_t> hl ((size_t) N * n);
  std::vector<int64_t> hs ((size_t) n);
  const char *base = (const char *) x;
  for (int64_t i = 0; i < n; i++)
    {
      mpfr_view v;
      std::memcpy (&v, base + (size_t) i * stride, sizeof v);
      if (v.prec != pb) return MPGFX_EINVAL;
      const bool neg = v.sign < 0;
      if (v.exp == MPFR_EXP_ZERO)
        {
          for (int l = 0; l < N; l++) hl[(size_t) l * n + i] = 0;
          hs[i] = (int64_t) ((uint64_t) ZERO_EXP << 1) | (neg ? 1 : 0);
          continue;
        }
      if (v.exp <= MPFR_EXP_INF || !v.d) return MPGFX_EINVAL;        /* NaN, infinity */
      for (int l = 0; l < N; l++) hl[(size_t) l * n + i] = v.d[l];
      hs[i] = (int64_t) ((uint64_t) v.exp << 1) | (neg ? 1 : 0);
    }
  hipStream_t s = (hipStream_t) stream;
  int rc = hip_ok ("import_mpfr copy", hipMemcpyAsync (limbs, hl.data (), hl.size () * 8, hipMemcpyHostToDevice, s));
  if (!rc) rc = hip_ok ("import_mpfr copy", hipMemcpyAsync (se, hs.data (), hs.size () * 8, hipMemcpyHostToDevice, s));
  if (!rc) rc = hip_ok ("import_mpfr sync", hipStreamSynchronize (s));
  return rc;
}

extern "C" int mpgfx_export_mpfr (int pb, int64_t n, const uint64_t *limbs, const int64_t *se, void *x, size_t stride,
                                  void *stream)
{
  if (n < 0 || stride < sizeof (mpfr_view)) return MPGFX_EINVAL;
  if (!pb_ok (pb)) return MPGFX_EUNSUPPORTED;
  if (n == 0) return MPGFX_OK;
  if (!x || !limbs || !se) return MPGFX_EINVAL;
  const int N = (pb + 63) / 64;
  char *base = (char *) x;
  for (int64_t i = 0; i < n; i++)
    {
      mpfr_view v;
      std::memcpy (&v, base + (size_t) i * stride, sizeof v);
      if (v.prec != pb || !v.d) return MPGFX_EINVAL;
    }
  std::vector<uint64_t> hl ((size_t) N * n);
  std::vector<int64_t> hs ((size_t) n);
  hipStream_t s = (hipStream_t) stream;
  int rc = hip_ok ("export_mpfr copy", hipMemcpyAsync (hl.data (), limbs, hl.size () * 8, hipMemcpyDeviceToHost, s));
  if (!rc) rc = hip_ok ("export_mpfr copy", hipMemcpyAsync (hs.data (), se, hs.size () * 8, hipMemcpyDeviceToHost, s));
  if (!rc) rc = hip_ok ("export_mpfr sync", hipStreamSynchronize (s));
  if (rc) return rc;
  for (int64_t i = 0; i < n; i++)
    {
      mpfr_view v;
      char *p = base + (size_t) i * stride;
      std::memcpy (&v, p, sizeof v);
      const int64_t w = hs[i], e = w >> 1;
      v.sign = (w & 1) ? -1 : 1;
      if (e == ZERO_EXP)
        v.exp = MPFR_EXP_ZERO;          /* MPFR leaves a zero's limbs unspecified */
      else
        {
          v.exp = (long) e;
          for (int l = 0; l < N; l++) v.d[l] = hl[(size_t) l * n + i];
        }
      std::memcpy (p, &v, sizeof v);
    }
  return MPGFX_OK;
}
