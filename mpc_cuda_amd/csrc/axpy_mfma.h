/* axpy_mfma.h -- complex AXPY z = a*x + y (MPC_RNDNN per operation) with the
 * products on INT8 MFMA (gfx950), for precisions that are whole multiples of
 * 512 bits.  Reference: demos/bench_3way.cu:51-59, tools/axpy_fast_complex.cu:32-40
 * (one cu_cmul + cu_cadd per element); the rounding is the one of
 * cu_fcomplex.cuh:90-96 (cu_fmms / cu_fmma: each component rounded once).
 *
 * The coefficient a is the same for every element, so the significand products
 * a_q * x (q: real / imaginary part of a) of a chunk of elements form one
 * matrix product: a Toeplitz matrix of the digits of a times the digit columns
 * of the x significands.
 *
 * Digits.  A significand M (KW = PB/32 words, top bit set) is written in signed
 * base-256 digits, M = 256^ND + sum_{u<ND} s_u 256^u (ND = PB/8, s_u in
 * [-128, 127]): the bytes of M + 0x8080..80 (the carry out of the top byte is
 * always 1, as that byte is >= 0x80), each XOR 0x80.  One add-with-carry chain
 * and one XOR per word.  With a = 2^(32 KW) + A' and x = 2^(32 KW) + X',
 *     a * x = conv + 2^(32 KW) a,   conv = sum_d C_d 256^d,
 *     C_d = sum_{v + u = d} t_v s_u   over v in [0, ND] (t_ND = 1), u < ND,
 * so the MFMA computes the digit convolution of the ND + 1 digits of a with
 * the ND digits of x (|C_d| <= 128 * 2^14 = 2^21: exact in int32), and a's
 * words are added at word offset KW while the columns are carried into words.
 *
 * Short product.  add_round needs the top H = KW + 2 words of each product.
 * Only the digits d >= D0 = 4 WLO, WLO = 2 KW - H - 2, are formed: the digits
 * below are worth less than 2^21 sum_{d<D0} 256^d < 2^(8 D0 + 13), i.e. less
 * than 2^(13 - 64) units of product word WLO + 2, the lowest word of the kept
 * window.  Carrying the formed digits from word WLO upwards gives the exact
 * integer S = a*x - E, |E| < 2^(8 D0 + 13); S lies in (2^(64 KW - 3), 2^(64 KW))
 * (a*x <= (2^(32KW) - 1)^2 and >= 2^(64 KW - 2)), so the arithmetic modulo
 * 2^(64 KW) is exact and no word past the top is needed.  The window
 * T = floor(S / 2^(32 (WLO + 2))) is within 2 units of its last word of the
 * exact product: the term is handed to add_round with that error bound, which
 * flags the lanes whose rounding the bound does not decide (and a window
 * whose top two bits are clear, possible only within 2^(8 D0 + 13) of
 * 2^(64 KW - 2)); those lanes recompute the exact products (core::fmm_exact).
 *
 * MFMA layout (v_mfma_i32_16x16x64_i8, as in the exact GEMM): lane l supplies
 * row l&15 of the A tile and column l&15 of the B tile, bytes 16(l>>4) .. +15
 * of the 64-deep K block for both (the same K order for A and B, so the sum is
 * that of the logical operands), and receives rows 4(l>>4) + v, v < 4, of
 * column l&15.  Tile T of the output holds digits D0 + 16T + r (r the row), so
 * lane l's four rows are the four digits of product word WLO + 4T + (l>>4).
 * A tile (T, K block b) is formed only where it holds a nonzero Toeplitz
 * entry (0 <= d - u <= ND).
 *
 * Work per wavefront: chunks of 32 elements; lane L = 2e + p (element e, part
 * p: 0 real, 1 imaginary).  Per chunk: (1) lane L recodes x_p[e] (staged in
 * LDS) into digits in place; (2) the B tiles (x digits of both parts,
 * two 16-element column tiles, every K block) are read into registers; (3)
 * pass q = 0 multiplies by a_r, pass q = 1 by a_i: for each x part p' and
 * column tile the formed tiles are accumulated, each lane turns its four digit
 * sums into one signed 64-bit word value and writes it to LDS for the lane
 * that carries that product -- lane 2e + (p' ^ q), so that lane 2e receives
 * a_r x_r then a_i x_i (the real part) and lane 2e+1 receives a_r x_i then
 * a_i x_r (the imaginary part); the word values pass through LDS in groups of
 * three output tiles, each carried before the next is formed (while the next
 * chunk's x is copied into the staging rows); (4) each lane
 * carries its product's word values into words, adding a_q's words at offset
 * KW, and normalises the window; (5) core::add_round rounds the two terms once
 * (RNDN(p0 -/+ p1)), the lanes it leaves open take the exact products, then
 * cu_fadd adds y (loaded during pass 1).
 *
 * LDS per workgroup (8 wavefronts, two per SIMD): the A tiles of a_r and a_i
 * (2 x NF x 1 KB, built once per workgroup from a's digits), a's words, and
 * per wavefront the x staging rows and the word-value buffer.
 */
#ifndef MPGFX_AXPY_MFMA_H
#define MPGFX_AXPY_MFMA_H

#include <cstdint>
#include "mpc_cuda/cu_fp_core.cuh"

namespace mpgfx {
namespace axmf {

typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;

template <int PB>
struct cfg
{
  static_assert (PB % 512 == 0, "whole K blocks of 64 digits");
  static constexpr int KW = PB / 32;              /* words of a significand */
  static constexpr int ND = PB / 8;               /* digits of a significand */
  static constexpr int KB = ND / 64;              /* K blocks */
  static constexpr int H = KW + 2;                /* window words handed to add_round */
  static constexpr int WLO = 2 * KW - H - 2;      /* lowest product word formed */
  static constexpr int D0 = 4 * WLO;              /* lowest digit formed */
  static constexpr int NWW = 2 * KW - WLO;        /* product words formed */
  static constexpr int NT = NWW / 4;              /* 16-digit output tiles */
  static_assert (NWW % 4 == 0, "whole tiles");
  /* tile (T, b) holds a Toeplitz entry t_{d-u}, 0 <= d - u <= ND */
  static constexpr bool live (int T, int b)
  {
    return D0 + 16 * T + 15 - 64 * b >= 0 && D0 + 16 * T - (64 * b + 63) <= ND;
  }
  static constexpr int fidx (int T, int b)        /* index among the live tiles, T-major */
  {
    int k = 0;
    for (int t = 0; t < NT; t++)
      for (int c = 0; c < KB; c++)
        {
          if (t == T && c == b) return k;
          if (live (t, c)) k++;
        }
    return -1;
  }
  static constexpr int nf ()
  {
    int k = 0;
    for (int t = 0; t < NT; t++)
      for (int c = 0; c < KB; c++) k += live (t, c) ? 1 : 0;
    return k;
  }
  static constexpr int NF = nf ();
  /* padded digit string of a: byte OFF + v holds t_v (v in [0, ND]), zeros
   * elsewhere; fragment bytes are read at d - u + OFF */
  static constexpr int OFF = (64 * KB - D0 > 0) ? ((64 * KB - D0 + 3) & ~3) : 4;
  static constexpr int TPL = (OFF + D0 + 16 * NT + 3) & ~3;
  static constexpr int TPW = TPL / 4;
  static_assert (OFF + ND < TPL, "top digit inside the string");
};

/* signed digits of a significand of KW words: bytes of (M + 0x8080..80) mod
 * 2^(32 KW), each XOR 0x80 (as int8: byte - 128); the carry out (digit ND,
 * always 1 for a normalised M) is implied */
template <int KW>
CUFP_INL void recode (u32 (&d)[KW], const u32 (&w)[KW])
{
  u32 c = 0;
  cu_fp::core::sfor<0, KW> ([&] (auto J) {
    constexpr int j = decltype (J)::value;
    d[j] = cu_fp::core::addc (w[j], 0x80808080u, c) ^ 0x80808080u;
  });
}

/* dword i of a's padded digit string (cfg::TPW dwords) from its digit words */
template <int PB>
CUFP_INL u32 tp_dword (int i, const u32 (&d)[cfg<PB>::KW])
{
  typedef cfg<PB> C;
  u32 v = 0;
  cu_fp::core::sfor<0, C::KW> ([&] (auto J) {
    constexpr int j = decltype (J)::value;
    v = (i == C::OFF / 4 + j) ? d[j] : v;
  });
  v = (i == (C::OFF + C::ND) / 4) ? (1u << (8 * ((C::OFF + C::ND) & 3))) : v;
  return v;
}

CUFP_INL u32 bswap32 (u32 x)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_perm (x, x, 0x00010203u);
#else
  return __builtin_bswap32 (x);
#endif
}

/* first byte index k0 of the fragment of tile (T, b) for lane l: byte j of
 * the lane's 16 is t at string index k0 - j */
template <int PB>
CUFP_INL int frag_k0 (int T, int b, int l)
{
  typedef cfg<PB> C;
  return C::D0 + 16 * T + (l & 15) - 64 * b - 16 * (l >> 4) + C::OFF;
}

/* dword i (bytes 4i .. 4i+3) of a fragment: string bytes k0-4i .. k0-4i-3,
 * i.e. the byte reversal of the (unaligned) dword at s = k0 - 4i - 3 */
CUFP_INL u32 frag_dword (const u32 *tpw, int k0, int i)
{
  const int s = k0 - 4 * i - 3;
  const u32 lo = tpw[s >> 2], hi = tpw[(s >> 2) + 1];
#if defined(__HIP_DEVICE_COMPILE__)
  const u32 v = __builtin_amdgcn_alignbyte (hi, lo, (u32) (s & 3));
#else
  const u32 v = (u32) ((((u64) hi << 32) | lo) >> (8 * (s & 3)));
#endif
  return bswap32 (v);
}

/* value of one product word from its four digit sums (digits 4w .. 4w+3):
 * c0 + c1 2^8 + c2 2^16 + c3 2^24, |c| < 2^22 */
CUFP_INL i64 assemble (int c0, int c1, int c2, int c3)
{
  const int t1 = (int) ((u32) c3 << 8) + c2;     /* |t1| < 2^30: exact */
  const int t0 = (int) ((u32) c1 << 8) + c0;
  return (i64) ((u64) (i64) t1 << 16) + (i64) t0;
}

/* a product term in add_round's form: value = sign * M * 2^(exp - 32 H) */
template <int PB>
struct term
{
  int sign;
  i64 exp;
  bool zero;
  bool odd;          /* top two window bits clear: not decided by the bound */
  u32 M[cfg<PB>::H];
};

/* carry state of one product: the signed carry and the words P[0..H] (P[0]
 * the word below the window) */
template <int PB>
struct carry
{
  i64 c;
  u32 P[cfg<PB>::H + 1];
};

/* carry the word values of product words WLO + W0 .. WLO + W1 - 1 (V(w - W0)
 * for word WLO + w) into words, adding a_q's KW words (aw(j), word j) at word
 * offset KW; the first range starts with c = 0 */
template <int PB, int W0, int W1, typename VF, typename AF>
CUFP_INL void sweep_range (carry<PB> &st, VF &&V, AF &&aw)
{
  typedef cfg<PB> C;
  constexpr int H = C::H, NWW = C::NWW, WLO = C::WLO, KW = C::KW;
  static_assert (0 <= W0 && W0 < W1 && W1 <= NWW, "word range");
  cu_fp::core::sfor<W0, W1> ([&] (auto Wc) {
    constexpr int w = decltype (Wc)::value;
    constexpr int aw_i = WLO + w - KW;
    i64 s = V (w - W0) + st.c;
    if constexpr (aw_i >= 0) s += (i64) aw (aw_i);
    if constexpr (w >= NWW - H - 1) st.P[w - (NWW - H - 1)] = (u32) s;
    st.c = s >> 32;
  });
}

/* normalise the top H words of a swept product into the term (exp holds the
 * exponent sum on entry) */
template <int PB>
CUFP_INL void sweep_finish (term<PB> &t, const carry<PB> &st)
{
  constexpr int H = cfg<PB>::H;
  const u32 sh = (st.P[H] >> 31) ^ 1u;
  t.odd = (st.P[H] >> 30) == 0;
  cu_fp::core::sfor<0, H> ([&] (auto I) {
    constexpr int i = decltype (I)::value;
    t.M[i] = cu_fp::core::fshl (st.P[i + 1], st.P[i], sh);
  });
  t.exp -= (i64) sh;
}

/* the top H words of a swept product as they stand (top bit 31 or 30, the
 * word below dropped: less than one more unit of error), for add_round's RAW
 * form, which orders and aligns unnormalised terms itself */
template <int PB>
CUFP_INL void sweep_finish_raw (term<PB> &t, const carry<PB> &st)
{
  constexpr int H = cfg<PB>::H;
  t.odd = (st.P[H] >> 30) == 0;
  cu_fp::core::sfor<0, H> ([&] (auto I) {
    constexpr int i = decltype (I)::value;
    t.M[i] = st.P[i + 1];
  });
}

/* the whole sweep at once */
template <int PB, typename VF>
CUFP_INL void sweep (term<PB> &t, VF &&V, const u32 (&aw)[cfg<PB>::KW])
{
  carry<PB> st;
  st.c = 0;
  sweep_range<PB, 0, cfg<PB>::NWW> (st, V, [&] (int j) { return aw[j]; });
  sweep_finish<PB> (t, st);
}

/* error bound of a term for add_round: below 2 units of its last word before
 * the normalising shift (ERR = 1), in the form of core::fmm_short's bound */
constexpr int ERRB = 1 + 3;

/* RNDN of one normalised term to KW words (the real product a*x): K gets the
 * kept words, the return value is the carry out (K is then 2^(PB-1) and the
 * exponent grows by one).  The term is within 2 units of its last word of the
 * exact product (either side), so the rounding is decided unless the 64 guard
 * bits below the kept words lie within that of the rounding midpoint (or the
 * window's top two bits were clear): then `amb` is set and the exact product
 * must give the result. */
template <int PB>
CUFP_INL u32 round_term (u32 (&K)[cfg<PB>::KW], const term<PB> &t, bool &amb)
{
  constexpr int KW = cfg<PB>::KW, H = cfg<PB>::H;
  static_assert (H == KW + 2, "two guard words");
  const u64 g = ((u64) t.M[1] << 32) | t.M[0];
  const u64 d = g - (1ull << 63);                 /* wraps: |g - 2^63| <= 4 <=> d + 4 <= 8 */
  amb = (d + 4u <= 8u) || t.odd;
  return cu_fp::core::round_top<KW, H> (K, t.M, 0u, 0u);
}

} // namespace axmf
} // namespace mpgfx

#endif
