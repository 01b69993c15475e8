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

inline int exact_digits (int pb, i64 spread) { return (int) ((pb + spread + 2 + 7) / 8); }
inline i64 pad_to (i64 x, i64 m) { return (x + m - 1) / m * m; }
/* carry-save slots per output element for S digits (diagonals up to 2S-2) */
inline int exact_slots (int S) { return (8 * (2 * S - 1) + 31) / 32 + 2; }
/* slots for the largest padded digit count */
constexpr int EXACT_SLMAX = (8 * (2 * EXACT_SPAD_MAX - 1) + 31) / 32 + 2;

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
