/* exact.h -- shared definitions of the exact-dot GEMM (C = RNDN(sum_k A B),
 * every product and the sum exact, one rounding per output).
 *
 * Scheme (no counterpart in the reference; semantics of mpfr_dot / mpfr_sum,
 * src/mpfr_cuda/dot.c:47-54):
 *  1. scan     per row i of A and column j of B: largest and smallest exponent
 *              of the nonzero entries; the largest spread decides the digit
 *              count S.
 *  2. digits   every entry becomes an integer X = sign * M' * 2^sh aligned to
 *              its row's (column's) largest exponent (M' = the pb-bit
 *              significand, sh = spread_max - (emax_row - e) >= 0), written as
 *              S signed 8-bit digits (two's complement bytes recoded to
 *              [-128, 127] with a carry), digit-plane-major:
 *              DA[t][i][k], DB[u][j][k] (k contiguous, padded to 64).
 *  3. MFMA     for an output tile, sum_k X_A X_B = sum_d 256^d G_d with
 *              G_d = sum_{t+u=d} sum_k DA[t][i][k] DB[u][j][k]; every (t,u)
 *              pair is an INT8 GEMM on v_mfma_i32_16x16x64_i8 accumulating
 *              exactly in int32 (|digit product| <= 2^14, <= G*4096 terms per
 *              accumulator between flushes).  Digits are processed in groups of
 *              G in {9, 11, 13, 15, 17} (chosen by padded work and pass
 *              overhead): a pass keeps B's fragments in VGPRs, streams A's,
 *              and holds the 2G-1 diagonal accumulators; each pass flushes its
 *              diagonals, shifted by 8d bits, into per-element carry-save
 *              int64 slots (32-bit steps).  The slots hold any k below
 *              2^62 / (G nG 2^38) (about 490,000 at pb 256); longer sums take
 *              the fallback below.
 *  4. round    the slots are carry-normalised into a two's complement integer,
 *              negated if negative, and rounded once to pb bits (RNDN); the
 *              exponent is emaxA_i + emaxB_j - sA - sB - 2pb + bitlength.
 *  5. fallback spreads beyond the MFMA window (S > EXACT_SMAX digits), and k
 *              beyond the slots' headroom, run on a VALU wide accumulator per
 *              output in workspace memory (64 bits of headroom for the count).
 */
#ifndef MPGFX_EXACT_H
#define MPGFX_EXACT_H

#include "common.h"

namespace mpgfx {

constexpr int EXACT_SMAX = 80;          /* digits per operand on the MFMA path */
constexpr int EXACT_SPAD_MAX = 96;      /* digits incl. the zero planes of the last group */
constexpr i64 EXACT_EMPTY = INT64_MIN;  /* row/column without nonzero entries */
constexpr int EXACT_KCHUNK = 4096;      /* int32 flush period in k (G <= 20) */
/* static scratch of mpgfx_gemm_exact_noalloc: 256 MiB per GPU (pb 256,
 * 1024^3 with exponent spreads below 8 needs 234 MB) */
constexpr size_t EXACT_NOALLOC_BYTES = (size_t) 256 << 20;

__host__ __device__ inline int exact_digits (int pb, i64 spread) { return (int) ((pb + spread + 2 + 7) / 8); }
__host__ __device__ inline i64 pad_to (i64 x, i64 m) { return (x + m - 1) / m * m; }
/* carry-save slots per output element for S digits (diagonals up to 2S-2) */
__host__ __device__ inline int exact_slots (int S) { return (8 * (2 * S - 1) + 31) / 32 + 2; }
/* slots for the largest padded digit count */
constexpr int EXACT_SLMAX = (8 * (2 * EXACT_SPAD_MAX - 1) + 31) / 32 + 2;

/* digit groups: G in {9, 11, 13, 15, 17}, nG = ceil(S/G), zero planes past S
 * up to G*nG <= EXACT_SPAD_MAX.  Cost model: padded MFMA work (G*nG)^2, times
 * (1 + 4/G) for the staging and flush overhead of a pass (a pass stages 2G
 * planes per k step for G^2 MFMAs). */
__host__ __device__ inline void exact_choose_groups (int S, int *G, int *nG)
{
  double best = -1;
  for (int c = 9; c <= 17; c += 2)
    {
      const int ng = (S + c - 1) / c;
      if (ng * c > EXACT_SPAD_MAX) continue;
      const double cost = (double) (ng * c) * (ng * c) * (1.0 + 4.0 / c);
      if (best < 0 || cost < best) { best = cost; *G = c; *nG = ng; }
    }
}

/* Headroom of the int64 carry-save slots.  Per k term, a digit position D
 * collects sum_{t+u=D} DA_t DB_u over at most spad digit pairs of magnitude
 * <= 2^14, and a slot adds 4 digit positions shifted by 0, 8, 16 and 24 bits:
 * at most spad * 2^14 * (1 + 2^8 + 2^16 + 2^24) per k.  Summed over kp terms
 * that must stay below 2^62 (the slots are signed, and the rounding kernel's
 * carry pass adds a carry); longer sums go to the wide-accumulator path. */
__host__ __device__ inline bool exact_slots_hold (int spad, i64 kp)
{
  const double per_k = (double) spad * 16384.0 * 16843009.0;
  return (double) kp * per_k < 4.611686018427388e18;        /* 2^62 */
}

/* wide-accumulator path: bits an output's exact sum can span, and the cap
 * beyond which the call is refused */
__host__ __device__ inline i64 exact_span_bits (int pb, i64 sA, i64 sB) { return sA + sB + 2 * (i64) pb + 64 + 2; }
constexpr i64 EXACT_SPAN_MAX = (i64) 1 << 22;

/* Digit pairs the asynchronous mode leaves out of the MFMA stage: digit t of
 * an A value times digit u of a B value with t + u < cut.  Digits are at most
 * 2^7 in magnitude, so per product those pairs add up to less than
 * 2^14 sum_{d<cut} (d+1) 256^d < cut 2^(15 + 8(cut-1)), and over kp terms the
 * integer the MFMA stage leaves is within 2^eb (integer units) of the exact
 * sum, eb = ceil(lg kp) + ceil(lg cut) + 15 + 8(cut-1).  The rounding stage
 * marks every element whose rounding that error could change (a value within
 * 2^eb of a rounding midpoint or of zero, or an error not below an eighth of
 * the last place), and those are recomputed exactly.  The cut is the largest
 * with eb <= pb + (sA + sB)/2 - 24: for a sum as large as a product of two
 * values at the middle of their lines' exponent ranges, the error is 24 bits
 * below the last place, so marked elements are rare.  cut = 0 (eb = -1): no
 * pair is left out. */
__host__ __device__ inline int exact_ceil_lg (i64 x)
{
  return x <= 1 ? 0 : 64 - __builtin_clzll ((unsigned long long) (x - 1));
}
/* (one lane of the plan kernel: a direct estimate ignoring the lg(cut) term,
 * then lowered while the bound exceeds the target, at most a few steps) */
__host__ __device__ inline void exact_choose_cut (int pb, i64 sA, i64 sB, i64 kp, int S, int *cut, int *eb)
{
  const i64 target = (i64) pb + (sA + sB) / 2 - 24, lk = exact_ceil_lg (kp);
  const i64 room = target - lk - 15;
  i64 d = room < 0 ? 0 : room / 8 + 1;
  if (d > 2 * (i64) S - 2) d = 2 * (i64) S - 2;
  while (d > 0 && lk + exact_ceil_lg (d) + 15 + 8 * (d - 1) > target) d--;
  *cut = (int) d;
  *eb = d > 0 ? (int) (lk + exact_ceil_lg (d) + 15 + 8 * (d - 1)) : -1;
}
/* Cse word of an element the rounding stage leaves to the exact fix-up (no
 * value has this sign/exponent word) */
constexpr i64 EXACT_FIX_MARK = INT64_MIN;

/* The plan of the asynchronous mode, decided on the device from the exponent
 * scan (k_exact_plan) and read by every later stage, which the host launches
 * for the worst case: digit planes laid out for EXACT_SPAD_MAX digits, the
 * MFMA kernel switching on G, the rounding and the wide-accumulator kernels
 * each returning at once unless the plan takes their path.  It sits at the
 * start of the caller's workspace; `status` (the first 8 bytes) is the call's
 * outcome once the stream has passed it: MPGFX_OK, or MPGFX_EUNSUPPORTED when
 * the spread needs more than EXACT_SPAN_MAX accumulator bits or more
 * accumulator memory than the workspace holds (C is then +0). */
struct exact_dplan
{
  i64 status;
  i64 sA, sB;                           /* spread maxima (bits) */
  i64 nw;                               /* wide accumulator (and the fix-up): 32-bit words per output */
  i64 lanes;                            /* wide accumulator (and the fix-up): outputs accumulated at once */
  int S, G, nG, slots;
  int fallback;                         /* 1: the wide-accumulator path */
  int cut, eb;                          /* MFMA path: diagonals left out, error bound (exact_choose_cut) */
};
constexpr size_t EXACT_DPLAN_BYTES = 256;
static_assert (sizeof (exact_dplan) <= EXACT_DPLAN_BYTES, "plan block");

/* per-launch plan, filled on the host after the scan */
struct exact_plan
{
  int pb, S, G, nG;                     /* digits, group size, groups */
  i64 m, n, k, mp, np, kp;              /* sizes, padded sizes */
  i64 sA, sB;                           /* spread maxima (bits) */
  int slots;
  bool fallback;                        /* spread beyond the MFMA window */
};

} // namespace mpgfx
#endif
