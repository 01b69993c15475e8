/* exact_gemm.hip -- host orchestration and precision-independent kernels of
 * the exact-dot GEMM (scheme in exact.h; per-precision kernels in
 * exact_pb.h).  The exponent scan has to reach the host before the digit count
 * and the MFMA kernel variant are chosen, so this mode synchronises its stream
 * once, after the scan. */
#include "common.h"
#include "dispatch.h"
#include "exact.h"

namespace mpgfx {

/* per line (row of A or column of B): largest / smallest exponent of its
 * nonzero values, and the largest spread over all lines (atomicMax) */
__global__ __launch_bounds__ (256) void k_exact_scan (i64 lines, i64 len, i64 ls, i64 ps, const i64 *__restrict__ se,
                                                      i64 *__restrict__ emax, i64 *__restrict__ emin,
                                                      unsigned long long *__restrict__ spread_max)
{
  __shared__ i64 smax[256], smin[256];
  const i64 L = blockIdx.x;
  i64 hi = INT64_MIN, lo = INT64_MAX;
  for (i64 p = threadIdx.x; p < len; p += blockDim.x)
    {
      i64 e = se[L * ls + p * ps] >> 1;
      if (e != MPGFX_ZERO_EXP)
        {
          hi = e > hi ? e : hi;
          lo = e < lo ? e : lo;
        }
    }
  smax[threadIdx.x] = hi;
  smin[threadIdx.x] = lo;
  __syncthreads ();
  for (int s = 128; s > 0; s >>= 1)
    {
      if ((int) threadIdx.x < s)
        {
          smax[threadIdx.x] = smax[threadIdx.x + s] > smax[threadIdx.x] ? smax[threadIdx.x + s] : smax[threadIdx.x];
          smin[threadIdx.x] = smin[threadIdx.x + s] < smin[threadIdx.x] ? smin[threadIdx.x + s] : smin[threadIdx.x];
        }
      __syncthreads ();
    }
  if (threadIdx.x == 0)
    {
      hi = smax[0];
      lo = smin[0];
      bool empty = (hi == INT64_MIN);
      emax[L] = empty ? EXACT_EMPTY : hi;
      emin[L] = empty ? EXACT_EMPTY : lo;
      if (!empty) atomicMax (spread_max, (unsigned long long) (hi - lo));
    }
}

static int hip_ok (const char *where, hipError_t e)
{
  if (e != hipSuccess) { set_error (where, e); return MPGFX_EHIP; }
  return MPGFX_OK;
}

/* workgroups of the MFMA kernel: one per CU (persistent over 32x32 tiles), a
 * multiple of 8 so the XCD-aware tile assignment covers every tile */
static i64 exact_grid (i64 tiles)
{
  const i64 g = tiles < 256 ? tiles : 256;
  return (g + 7) / 8 * 8;
}

/* bytes of workspace for the MFMA path with digit count up to EXACT_SMAX */
size_t exact_workspace_bound (int pb, i64 m, i64 n, i64 k)
{
  (void) pb;
  const i64 mp = pad_to (m, 32), np = pad_to (n, 32), kp = pad_to (k > 0 ? k : 1, 64);
  const i64 spad = pad_to (EXACT_SMAX, 16);              /* nG*G for S <= EXACT_SMAX, G = 16 */
  const i64 grid = exact_grid ((mp / 32) * (np / 32));
  size_t b = 0;
  b += (size_t) (2 * m + 2 * n + 8) * 8;                 /* scan outputs + spread words */
  b = (size_t) pad_to ((i64) b, 256);
  b += (size_t) (spad * (mp + np) * kp);                 /* digit planes */
  b = (size_t) pad_to ((i64) b, 256);
  b += (size_t) grid * 256 * 4 * exact_slots ((int) spad) * 8;
  return b;
}

/* digits per group: fewer passes first (a pass re-stages its planes for every
 * k step, so passes with few digits are load-latency bound), then the smaller
 * group; the padded digit count G*nG stays <= 64 (workspace bound) */
static void choose_groups (int S, int *G, int *nG)
{
  const int cand[2] = {16, 20};
  int best_passes = -1;
  for (int c : cand)
    {
      const int ng = (S + c - 1) / c;
      if (ng * c > 64) continue;
      if (best_passes < 0 || ng * ng < best_passes) { best_passes = ng * ng; *G = c; *nG = ng; }
    }
}

int gemm_exact (const pb_ops *o, int pb, i64 m, i64 n, i64 k, const u64 *Al, const i64 *Ase, const u64 *Bl,
                const i64 *Bse, u64 *Cl, i64 *Cse, void *ws, size_t ws_bytes, hipStream_t s)
{
  if (k == 0)                      /* empty sums: +0 everywhere, as the per-op path */
    return o->gemm_perop (m, n, k, Al, Ase, Bl, Bse, Cl, Cse, s);
  const i64 mp = pad_to (m, 32), np = pad_to (n, 32), kp = pad_to (k, 64);
  /* ---- scan (always needs a small buffer) ---- */
  size_t scan_bytes = (size_t) pad_to ((i64) ((2 * m + 2 * n + 8) * 8), 256);
  char *base = (char *) ws;
  char *owned = nullptr;
  size_t need_mfma = exact_workspace_bound (pb, m, n, k);
  if (!base || ws_bytes < need_mfma)
    {
      int rc = hip_ok ("gemm_exact workspace", hipMallocAsync ((void **) &owned, need_mfma, s));
      if (rc) return MPGFX_ENOMEM;
      base = owned;
    }
  i64 *emaxA = (i64 *) base, *eminA = emaxA + m, *emaxB = eminA + m, *eminB = emaxB + n;
  unsigned long long *spreads = (unsigned long long *) (eminB + n);
  int rc = MPGFX_OK;
  auto done = [&] (int r) {
    if (owned) (void) hipFreeAsync (owned, s);
    return r;
  };
  if ((rc = hip_ok ("gemm_exact memset", hipMemsetAsync (spreads, 0, 2 * sizeof (unsigned long long), s)))) return done (rc);
  if (k > 0)
    {
      hipLaunchKernelGGL (k_exact_scan, dim3 ((unsigned) m), dim3 (256), 0, s, m, k, k, (i64) 1, Ase, emaxA, eminA,
                          spreads);
      hipLaunchKernelGGL (k_exact_scan, dim3 ((unsigned) n), dim3 (256), 0, s, n, k, (i64) 1, n, Bse, emaxB, eminB,
                          spreads + 1);
      if ((rc = launch_status ("gemm_exact scan"))) return done (rc);
    }
  unsigned long long hs[2] = {0, 0};
  if (k > 0)
    {
      if ((rc = hip_ok ("gemm_exact spread copy", hipMemcpyAsync (hs, spreads, sizeof hs, hipMemcpyDeviceToHost, s))))
        return done (rc);
      if ((rc = hip_ok ("gemm_exact sync", hipStreamSynchronize (s)))) return done (rc);
    }
  exact_plan p;
  p.pb = pb; p.m = m; p.n = n; p.k = k; p.mp = mp; p.np = np; p.kp = kp;
  p.sA = (i64) hs[0]; p.sB = (i64) hs[1];
  const i64 smax = p.sA > p.sB ? p.sA : p.sB;
  p.S = (smax > 100000) ? EXACT_SMAX + 1 : exact_digits (pb, smax);
  p.fallback = p.S > EXACT_SMAX;
  if (!p.fallback)
    {
      choose_groups (p.S, &p.G, &p.nG);
      /* fragment loads use 32-bit buffer offsets within one digit group */
      if ((i64) p.G * (mp > np ? mp : np) * kp >= ((i64) 1 << 31)) return done (MPGFX_EUNSUPPORTED);
      const int spad = p.G * p.nG;
      p.slots = exact_slots (spad);
      const int grid = (int) exact_grid ((mp / 32) * (np / 32));
      char *q = base + scan_bytes;
      int8_t *DA = (int8_t *) q;
      q += pad_to ((i64) spad * mp * kp, 256);
      int8_t *DB = (int8_t *) q;
      q += pad_to ((i64) spad * np * kp, 256);
      i64 *slotbuf = (i64 *) q;
      if ((rc = o->exact_digits (m, k, k, 1, Al, Ase, m * k, emaxA, p.sA, p.S, spad, mp, kp, DA, s))) return done (rc);
      if ((rc = o->exact_digits (n, k, 1, n, Bl, Bse, k * n, emaxB, p.sB, p.S, spad, np, kp, DB, s))) return done (rc);
      rc = o->exact_mfma (p.G, p, DA, DB, emaxA, emaxB, slotbuf, grid, Cl, Cse, s);
      return done (rc);
    }
  /* ---- VALU wide-accumulator fallback, in row chunks ---- */
  const i64 spanbits = p.sA + p.sB + 2 * (i64) pb + 64 + 2;
  if (spanbits > ((i64) 1 << 22)) return done (MPGFX_EUNSUPPORTED);
  const int nw = (int) ((spanbits + 31) / 32) + 1;
  const size_t budget = (size_t) 256 << 20;
  i64 rows = (i64) (budget / ((size_t) n * nw * 8));
  if (rows < 1) rows = 1;
  if (rows > m) rows = m;
  i64 *wsp = nullptr;
  if ((rc = hip_ok ("gemm_exact fallback workspace", hipMallocAsync ((void **) &wsp, (size_t) rows * n * nw * 8, s))))
    return done (MPGFX_ENOMEM);
  for (i64 r0 = 0; r0 < m && rc == MPGFX_OK; r0 += rows)
    {
      i64 nr = (r0 + rows <= m) ? rows : m - r0;
      rc = o->exact_fallback (m, n, k, r0, nr, Al, Ase, Bl, Bse, eminA, eminB, nw, wsp, Cl, Cse, s);
    }
  (void) hipFreeAsync (wsp, s);
  return done (rc);
}

} // namespace mpgfx
