/* cu_fp_core.cuh -- word-level engine behind cu_fp::cu_freal<PB> / cu_fcomplex<PB>
 * for AMD CDNA4 (gfx950) and the host.
 *
 * Significands are handled as little-endian arrays of 32-bit words so that every
 * word product maps onto one v_mad_u64_u32 and every carry onto a v_add_co /
 * v_addc_co pair.  All loops have compile-time trip counts and every
 * data-dependent shift is a select network (log-step barrel shifter over words
 * plus a funnel shift), so the arrays stay in VGPRs and lanes of a wavefront do
 * not diverge on exponent differences.  The only branch kept is the word-level
 * normalisation after a cancellation of 32 bits or more, which random data
 * almost never takes.
 *
 * Rounding is round-to-nearest-even (MPFR RNDN).  The algorithm is the one of
 * the reference's single rounding kernel (tkouya/mpc_cuda
 * include/mpc_cuda/cu_freal.cuh:89-138): keep PB bits below the most
 * significant one, a round bit, and a sticky bit for everything lower.
 */
#ifndef MPGFX_CU_FP_CORE_CUH
#define MPGFX_CU_FP_CORE_CUH

#include <cstdint>
#include <type_traits>
#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define CUFP_INL __host__ __device__ __forceinline__
#else
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#define CUFP_INL inline __attribute__ ((always_inline))
#endif

namespace cu_fp {
namespace core {

typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;

#define CUFP_UNROLL _Pragma ("unroll")

/* compile-time loop: f(std::integral_constant<int, I>) for I in [B, E); keeps
 * every array index a constant so register arrays never fall back to scratch */
template <int B, int E, typename F>
CUFP_INL void sfor (F &&f)
{
  if constexpr (B < E)
    {
      f (std::integral_constant<int, B> ());
      sfor<B + 1, E> (f);
    }
}

/* s = a + b + c, c <- carry out (c in {0,1}); on the device a 32-bit
 * add-with-carry (v_addc_co_u32 through VCC) rather than a 64-bit add */
CUFP_INL u32 addc (u32 a, u32 b, u32 &c)
{
#if defined(__HIP_DEVICE_COMPILE__)
  u32 co;
  const u32 s = __builtin_addc (a, b, c, &co);
  c = co;
  return s;
#else
  u64 t = (u64) a + b + c;
  c = (u32) (t >> 32);
  return (u32) t;
#endif
}

/* Funnel shifts.  On the device they are v_alignbit_b32; the 64-bit form is
 * kept for the host only, because on the device the optimiser merges the
 * (hi:lo) pair of two adjacent array words into one 64-bit load, which keeps
 * the whole array out of registers. */
/* low 32 bits of (hi:lo) >> s, s in [0,31] */
CUFP_INL u32 fshr (u32 hi, u32 lo, u32 s)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbit (hi, lo, s);
#else
  return (u32) ((((u64) hi << 32) | lo) >> s);
#endif
}
/* top 32 bits of (hi:lo) << s, s in [0,31] */
CUFP_INL u32 fshl (u32 hi, u32 lo, u32 s)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return s ? __builtin_amdgcn_alignbit (hi, lo, 32u - s) : hi;
#else
  return (u32) (((((u64) hi << 32) | lo) << s) >> 32);
#endif
}

/* true if the condition holds in any active lane of the wavefront (a
 * wave-uniform branch condition: steps no lane needs are skipped by all) */
CUFP_INL bool any_lane (bool c)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_ballot_w64 (c) != 0;
#else
  return c;
#endif
}

/* Placed first in a branch that only rare data takes: the empty volatile
 * statement cannot be executed speculatively, so the compiler keeps the
 * branch (a wave-uniform jump) instead of turning its body into selects that
 * every lane would execute every time. */
#define CUFP_RARE() asm volatile ("")

CUFP_INL int clz32 (u32 x)
{
  return x ? __builtin_clz (x) : 32;
}
CUFP_INL int clz64 (u64 x)
{
  return x ? __builtin_clzll (x) : 64;
}

/* ---------------------------------------------------------------------------
 * Exact product p[2W] = a[W] * b[W] of two W-word significands.
 *
 * W < 16 (pb < 512): schoolbook on 32-bit words, by columns on the device
 * (one 32x32+64 multiply-add per word product, its carry out counted with one
 * add-with-carry; mul_words_cols) and by rows on the host; at these sizes the
 * limb split and recombination below cost more than they save (measured on
 * the pb 256 GEMM and AXPY).
 * W >= 16: the operands are re-cut into LB-bit limbs (LB = 29, or 28 for the
 * widest precisions) and the product is accumulated by columns: a column of
 * up to 63 (255) limb products plus the incoming carry cannot overflow 64 bits,
 * so every limb product is exactly one v_mad_u64_u32 with the running column
 * sum as its addend and no carry handling inside the column; each column then
 * emits one LB-bit limb and shifts its carry into the next.  The limbs are cut
 * back into 32-bit words at the end.
 * ------------------------------------------------------------------------- */
template <int W>
CUFP_INL void mul_words_rows (u32 (&p)[2 * W], const u32 (&a)[W], const u32 (&b)[W])
{
  CUFP_UNROLL
  for (int i = 0; i < 2 * W; i++) p[i] = 0;
  sfor<0, W> ([&] (auto J) {
    constexpr int j = decltype (J)::value;
    u32 c = 0;
    sfor<0, W> ([&] (auto I) {
      constexpr int i = decltype (I)::value;
      u64 t = (u64) a[i] * b[j] + p[i + j] + c;
      p[i + j] = (u32) t;
      c = (u32) (t >> 32);
    });
    p[W + j] = c;
  });
}

/* bits [LB*k0 + s0, LB*k0 + s0 + 32) of a number held as LB-bit limbs P[0..NP)
 * (s0 < LB): the word spans limbs k0, k0+1 and, when 2 LB - s0 < 32, k0+2;
 * 32-bit shifts only */
template <int LB, int NP, int k0, int s0>
CUFP_INL u32 limbs_word (const u32 *P)
{
  static_assert (s0 >= 0 && s0 < LB, "offset inside the first limb");
  u32 w = P[k0] >> s0;
  if constexpr (k0 + 1 < NP) w |= P[k0 + 1] << (LB - s0);
  if constexpr (k0 + 2 < NP && 2 * LB - s0 < 32) w |= P[k0 + 2] << (2 * LB - s0);
  return w;
}

template <int W>
struct limb_cut
{
  static constexpr int LB = (32 * W <= 29 * 63) ? 29 : 28;   /* column capacity 63 / 255 */
  static constexpr int NL = (32 * W + LB - 1) / LB;
  static constexpr u32 MASK = (1u << LB) - 1u;
  static_assert (NL <= ((LB == 29) ? 63 : 255), "column sum could overflow");
};

/* bits [LB*j, LB*j+LB) of a W-word number, for every limb j */
template <int W>
CUFP_INL void words_to_limbs (u32 (&L)[limb_cut<W>::NL], const u32 (&A)[W])
{
  typedef limb_cut<W> C;
  sfor<0, C::NL> ([&] (auto J) {
    constexpr int j = decltype (J)::value;
    constexpr int bit = C::LB * j, w = bit >> 5, s = bit & 31;
    const u32 lo = A[w];
    const u32 hi = (w + 1 < W) ? A[w + 1] : 0u;
    L[j] = fshr (hi, lo, (u32) s) & C::MASK;
  });
}

template <int W>
CUFP_INL void mul_words_limbs (u32 (&p)[2 * W], const u32 (&a)[W], const u32 (&b)[W])
{
  typedef limb_cut<W> C;
  constexpr int NL = C::NL, LB = C::LB;
  u32 La[NL], Lb[NL], P[2 * NL];
  words_to_limbs<W> (La, a);
  words_to_limbs<W> (Lb, b);
  u64 acc = 0;
  sfor<0, 2 * NL - 1> ([&] (auto K) {
    constexpr int k = decltype (K)::value;
    constexpr int i0 = (k >= NL) ? k - NL + 1 : 0, i1 = (k < NL) ? k + 1 : NL;
    sfor<i0, i1> ([&] (auto I) {
      constexpr int i = decltype (I)::value;
      acc = (u64) La[i] * Lb[k - i] + acc;
    });
    P[k] = (u32) acc & C::MASK;
    acc >>= LB;
  });
  P[2 * NL - 1] = (u32) acc;
  /* limbs -> 32-bit words */
  sfor<0, 2 * W> ([&] (auto Q) {
    constexpr int q = decltype (Q)::value;
    p[q] = limbs_word<LB, 2 * NL, (32 * q) / LB, 32 * q - LB * ((32 * q) / LB)> (P);
  });
}

/* ---------------------------------------------------------------------------
 * Short product: the top H words of a[W] * b[W] (H <= 2W - 3) from the limb
 * columns at and above K0 only, the column that holds product bit LSB =
 * 32 (2W - H), the lowest bit kept.  The columns below K0 each hold less than
 * (k+1) 2^(2 LB), so together they are worth less than
 * K0 2^(2 LB) 2^(LB (K0-1)) / (1 - 2^-LB) < 2^(LB K0 + LB + lg K0 + 1); leaving
 * out their carries and cutting the sum at bit LSB makes the result p a lower
 * bound of the true top words with an error below 2^ERR units of p's last
 * word.  When both operands have their top bit set the top word of p keeps
 * bit 31 or bit 30 set: the limbs at and above K0 of a alone, times b, are at
 * least 2^(32W-1) 2^(32W-1).
 * ------------------------------------------------------------------------- */
template <int W, int H>
struct short_cut
{
  typedef limb_cut<W> C;
  static_assert (H <= 2 * W - 3, "short product wider than the product");
  static constexpr int LSB = 32 * (2 * W - H);
  static constexpr int K0 = LSB / C::LB;
  static constexpr int lg (int x) { return x <= 1 ? 0 : 1 + lg ((x + 1) / 2); }
  static constexpr int ERR = C::LB * K0 + C::LB + lg (K0) + 2 - LSB;
};

/* W < 16: the short product by 32-bit word columns.  Columns K0 = 2W - H - 1
 * and up are computed (one column below the kept words, whose carries it
 * passes on exactly); the columns below K0 would add less than
 * sum_{k<K0} (k+1) 2^(64+32k) < (K0+1) 2^(32K0+32), so the result is a lower
 * bound of the top H words off by less than K0 + 1 units of its last word. */
template <int W, int H>
struct short_cols
{
  static_assert (H <= 2 * W - 1, "short product wider than the product");
  static constexpr int K0 = 2 * W - H - 1;
  static constexpr int lg (int x) { return x <= 1 ? 0 : 1 + lg ((x + 1) / 2); }
  static constexpr int ERR = lg (K0 + 1);
};

/* the error exponent of mul_words_short<W, H> (either form) */
template <int W, int H>
constexpr int short_err ()
{
  if constexpr (W >= 16)
    return short_cut<W, H>::ERR;
  else
    return short_cols<W, H>::ERR;
}

template <int W, int H>
CUFP_INL void mul_words_short_cols (u32 (&p)[H], const u32 (&a)[W], const u32 (&b)[W]);

template <int W, int H>
CUFP_INL void mul_words_short (u32 (&p)[H], const u32 (&a)[W], const u32 (&b)[W])
{
  if constexpr (W < 16)
    {
      mul_words_short_cols<W, H> (p, a, b);
      return;
    }
  else
    {
  typedef limb_cut<W> C;
  constexpr int NL = C::NL, LB = C::LB, K0 = short_cut<W, H>::K0;
  u32 La[NL], Lb[NL], P[2 * NL];
  words_to_limbs<W> (La, a);
  words_to_limbs<W> (Lb, b);
  u64 acc = 0;
  sfor<K0, 2 * NL - 1> ([&] (auto K) {
    constexpr int k = decltype (K)::value;
    constexpr int i0 = (k >= NL) ? k - NL + 1 : 0, i1 = (k < NL) ? k + 1 : NL;
    sfor<i0, i1> ([&] (auto I) {
      constexpr int i = decltype (I)::value;
      acc = (u64) La[i] * Lb[k - i] + acc;
    });
    P[k] = (u32) acc & C::MASK;
    acc >>= LB;
  });
  P[2 * NL - 1] = (u32) acc;
  sfor<0, H> ([&] (auto Q) {
    constexpr int q = 2 * W - H + decltype (Q)::value;
    constexpr int k0 = (32 * q) / LB, s0 = 32 * q - LB * k0;
    static_assert (k0 >= K0, "word below the computed columns");
    p[decltype (Q)::value] = limbs_word<LB, 2 * NL, k0, s0> (P);
  });
    }
}

#if defined(__HIP_DEVICE_COMPILE__)
/* acc += a*b on 64 bits, the carry out of that addition added to ovf: the
 * carry mask v_mad_u64_u32 writes is the add-with-carry's carry-in (the
 * compiler exposes neither, hence the two instructions written out; s_nop 1
 * gives the mask its wait states before it is read) */
__device__ __forceinline__ void mac_carry (u64 &acc, u32 &ovf, u32 a, u32 b)
{
  u64 t, cm, cm2;
  asm ("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(t), "=s"(cm) : "v"(a), "v"(b), "v"(acc));
  asm ("s_nop 1\n\tv_addc_co_u32_e64 %0, %1, %2, 0, %3" : "=v"(ovf), "=s"(cm2) : "v"(ovf), "s"(cm));
  acc = t;
}

/* W < 16 on the device: product scanning by columns.  Every word product is
 * one v_mad_u64_u32 into the column's 64-bit sum, its carry out counted in
 * the column's overflow word; the column's low word is emitted and the rest,
 * (sum >> 32) + overflow * 2^32, starts the next column.  No zero-extended
 * addends are built (the row form needs two moves and a 64-bit add per word
 * product for them). */
template <int W>
__device__ __forceinline__ void mul_words_cols (u32 (&p)[2 * W], const u32 (&a)[W], const u32 (&b)[W])
{
  u64 acc = 0;
  sfor<0, 2 * W - 1> ([&] (auto K) {
    constexpr int k = decltype (K)::value;
    constexpr int i0 = k >= W ? k - W + 1 : 0, i1 = k < W ? k + 1 : W;
    u32 ovf = 0;
    sfor<i0, i1> ([&] (auto I) {
      constexpr int i = decltype (I)::value;
      mac_carry (acc, ovf, a[i], b[k - i]);
    });
    p[k] = (u32) acc;
    acc = (acc >> 32) | ((u64) ovf << 32);
  });
  p[2 * W - 1] = (u32) acc;
}
#endif

template <int W, int H>
CUFP_INL void mul_words_short_cols (u32 (&p)[H], const u32 (&a)[W], const u32 (&b)[W])
{
  constexpr int K0 = short_cols<W, H>::K0, LO = 2 * W - H;
  u64 acc = 0;
  sfor<K0, 2 * W - 1> ([&] (auto K) {
    constexpr int k = decltype (K)::value;
    constexpr int i0 = k >= W ? k - W + 1 : 0, i1 = k < W ? k + 1 : W;
    u32 ovf = 0;
    sfor<i0, i1> ([&] (auto I) {
      constexpr int i = decltype (I)::value;
#if defined(__HIP_DEVICE_COMPILE__)
      mac_carry (acc, ovf, a[i], b[k - i]);
#else
      const u64 s = acc + (u64) a[i] * b[k - i];
      ovf += s < acc ? 1u : 0u;
      acc = s;
#endif
    });
    if constexpr (k >= LO) p[k - LO] = (u32) acc;
    acc = (acc >> 32) | ((u64) ovf << 32);
  });
  p[H - 1] = (u32) acc;
}

template <int W>
CUFP_INL void mul_words (u32 (&p)[2 * W], const u32 (&a)[W], const u32 (&b)[W])
{
  if constexpr (W >= 16)
    mul_words_limbs<W> (p, a, b);
  else
#if defined(__HIP_DEVICE_COMPILE__)
    mul_words_cols<W> (p, a, b);
#else
    mul_words_rows<W> (p, a, b);
#endif
}

/* ---------------------------------------------------------------------------
 * Round the LEN-word buffer D, whose top bit is bit 32*LEN-1 after a left shift
 * by s in {0,1}, to KW = PB/32 words (RNDN).  `xst` is sticky material below
 * D bit 0.  Returns the kept words in K and 1 on carry-out (K is then
 * 2^(PB-1) and the caller adds 1 to the exponent).
 * ------------------------------------------------------------------------- */
template <int KW, int LEN>
CUFP_INL u32 round_top (u32 (&K)[KW], const u32 (&D)[LEN], u32 s, u32 xst)
{
  static_assert (LEN >= KW, "buffer narrower than the precision");
  /* kept words: shifted words LEN-KW .. LEN-1 */
  CUFP_UNROLL
  for (int j = 0; j < KW; j++)
    {
      int t = LEN - KW + j;
      u32 lo = (t >= 1) ? D[t - 1] : 0u;
      K[j] = fshl (D[t], lo, s);
    }
  u32 rbit = 0, sticky = xst;
  if (LEN > KW)
    {
      const int r = LEN - KW - 1;
      u32 lo = (r >= 1) ? D[r - 1] : 0u;
      u32 rw = fshl (D[r], lo, s);
      rbit = rw >> 31;
      sticky |= rw & 0x7fffffffu;
      CUFP_UNROLL
      for (int i = 0; i < r; i++) sticky |= D[i];
    }
  u32 inc = rbit & ((sticky != 0) | (K[0] & 1u));
  u32 c = inc;
  K[0] = addc (K[0], 0u, c);
  /* the increment carries past the lowest kept word only when that word was
   * all ones: a chain of dependent add-with-carry steps just for that case */
  if (any_lane (c != 0))
    {
      CUFP_RARE ();
      CUFP_UNROLL
      for (int j = 1; j < KW; j++) K[j] = addc (K[j], 0u, c);
    }
  /* carry-out: K wrapped to 0; the value is 2^PB = 2^(PB-1) * 2 */
  K[KW - 1] |= c << 31;
  return c;
}

/* ---------------------------------------------------------------------------
 * The same rounding to the top p bits of D (0 < p <= 32 KW, a run-time
 * precision): K gets the top KW words of the result with the bits below the
 * kept p cleared.  Word indices depend on p, so every word is visited with
 * selects (no register indexing at run time).
 * ------------------------------------------------------------------------- */
template <int KW, int LEN>
CUFP_INL u32 round_top_p (u32 (&K)[KW], const u32 (&D)[LEN], int p, u32 xst)
{
  static_assert (LEN > KW, "needs a guard word below the kept ones");
  const int pos = 32 * LEN - p;                   /* bit index of the kept LSB */
  const int pw = pos >> 5, pbit = pos & 31;
  const int rw = (pos - 1) >> 5, rb = (pos - 1) & 31;
  u32 rword = 0, lword = 0, sticky = xst;
  CUFP_UNROLL
  for (int t = 0; t < LEN; t++)
    {
      rword = (t == rw) ? D[t] : rword;
      lword = (t == pw) ? D[t] : lword;
      const u32 below = (t < rw) ? 0xffffffffu : ((t == rw) ? ((1u << rb) - 1u) : 0u);
      sticky |= D[t] & below;
    }
  const u32 inc = ((rword >> rb) & 1u) & ((sticky != 0) | ((lword >> pbit) & 1u));
  u32 c = 0;
  CUFP_UNROLL
  for (int j = 0; j < KW; j++)
    {
      const int t = LEN - KW + j;
      const u32 keep = (t > pw) ? 0xffffffffu : ((t == pw) ? (0xffffffffu << pbit) : 0u);
      K[j] = addc (D[t] & keep, (t == pw) ? (inc << pbit) : 0u, c);
    }
  K[KW - 1] |= c << 31;
  return c;
}

/* ---------------------------------------------------------------------------
 * RNDN(A +/- B) for two exact operands of LW words each (top bit set unless
 * zero), value = sign * M * 2^(e - 32*LW), rounded to KW words.  The larger
 * magnitude is placed in a window of LW+2 words (one guard word below, one
 * carry word above); the smaller is shifted right by the exponent difference
 * with a barrel shifter; bits that leave the window become the sticky bit and,
 * for a subtraction, a borrow of one unit at the window's LSB (the exact floor
 * of the true difference -- cu_freal.cuh:266-270 uses the same device).
 * Output: sign, exponent, zero flag and the kept words.
 * ------------------------------------------------------------------------- */
template <int KW>
struct rres
{
  int sign;
  i64 exp;
  bool zero;
  bool amb;          /* ERRB > 0: the rounding is not decided by the bound */
  u32 k[KW];
};

/* ERRB > 0: the operands are lower or upper bounds of the true terms, off by
 * less than 2^ERRB units of the last word of the larger one (together), and
 * r.amb reports whether the true sum could round differently: that is when
 * it lies within that error of a rounding midpoint, or the computed sum is 0.
 * With 3 or 4 guard words below the kept ones (LW = KW + 2 or KW + 3) the
 * test reads the top 64 guard bits.  p > 0 rounds to p bits (a run-time precision, the
 * operands' low bits zero) rather than to KW words.
 * RAW: the operands may be unnormalised by one bit (top bit 31 or 30 of the
 * top word), their exponents those of the words as they stand; they are
 * ordered by their normalised exponents and top words, and the larger enters
 * the window one bit lower (its bottom bit in the guard word) so that the
 * smaller, shifted right by one bit more, never needs a left shift. */
template <int KW, int LW, int ERRB = 0, bool RAW = false>
CUFP_INL rres<KW> add_round (int sa, i64 ea, bool za, const u32 (&Ai)[LW],
                             int sb, i64 eb, bool zb, const u32 (&Bi)[LW], int p = 0)
{
  const int L = LW + 2;
  rres<KW> r;
  /* order by exponent, then by the top words (zero is the smallest); only
   * equal exponents and equal top words can leave the difference negative,
   * and it is negated below */
  i64 ean = ea, ebn = eb;
  u32 ta = Ai[LW - 1], tb = Bi[LW - 1];
  if constexpr (RAW)
    {
      const u32 ha = (Ai[LW - 1] >> 31) ^ 1u, hb = (Bi[LW - 1] >> 31) ^ 1u;
      ean = ea - (i64) ha;
      ebn = eb - (i64) hb;
      ta = fshl (Ai[LW - 1], Ai[LW - 2], ha);
      tb = fshl (Bi[LW - 1], Bi[LW - 2], hb);
    }
  bool age = zb || (!za && (ean > ebn || (ean == ebn && ta >= tb)));
  u32 A[LW], B[LW];
  CUFP_UNROLL
  for (int i = 0; i < LW; i++)
    {
      A[i] = age ? Ai[i] : Bi[i];
      B[i] = age ? Bi[i] : Ai[i];
    }
  int sA = age ? sa : sb, sB = age ? sb : sa;
  i64 eA = age ? ea : eb, eB = age ? eb : ea;
  bool zA = age ? za : zb, zB = age ? zb : za;
  const i64 dmax = 32 * (i64) (LW + 1);
  i64 d = zB ? dmax : (eA - eB + (RAW ? 1 : 0));
  d = d > dmax ? dmax : d;
  u32 ws = (u32) (d >> 5), bs = (u32) (d & 31);

  /* T = [guard 0][B][carry 0] shifted right by d */
  u32 T[L];
  T[0] = 0;
  CUFP_UNROLL
  for (int i = 0; i < LW; i++) T[i + 1] = B[i];
  T[L - 1] = 0;
  u32 ext = 0;
  CUFP_UNROLL
  for (int k = 0; (1 << k) <= LW + 1; k++)
    {
      const int sh = 1 << k;
      u32 on = (ws >> k) & 1u;
      if (!any_lane (on != 0)) continue;        /* step needed by no lane */
      u32 lost = 0;
      CUFP_UNROLL
      for (int i = 0; i < sh && i < L; i++) lost |= T[i];
      ext |= on ? lost : 0u;
      CUFP_UNROLL
      for (int i = 0; i < L; i++)
        {
          u32 v = (i + sh < L) ? T[i + sh] : 0u;
          T[i] = on ? v : T[i];
        }
    }
  {
    u32 mask = bs ? ((1u << bs) - 1u) : 0u;
    ext |= T[0] & mask;
    CUFP_UNROLL
    for (int i = 0; i < L; i++)
      {
        u32 hi = (i + 1 < L) ? T[i + 1] : 0u;
        T[i] = fshr (hi, T[i], bs);
      }
  }
  u32 st = (ext != 0);

  /* D = A_win + (sub ? ~T + 1 - st : T) */
  bool sub = (sA != sB) && !zB;
  u32 xm = sub ? 0xffffffffu : 0u;
  u32 c = (sub && !st) ? 1u : 0u;
  u32 D[L];
  if constexpr (RAW)
    {
      /* A one bit lower: its bottom bit in the guard word */
      D[0] = addc (A[0] << 31, T[0] ^ xm, c);
      CUFP_UNROLL
      for (int i = 0; i < LW; i++) D[i + 1] = addc (fshr (i + 1 < LW ? A[i + 1] : 0u, A[i], 1u), T[i + 1] ^ xm, c);
    }
  else
    {
      D[0] = addc (0u, T[0] ^ xm, c);
      CUFP_UNROLL
      for (int i = 0; i < LW; i++) D[i + 1] = addc (A[i], T[i + 1] ^ xm, c);
    }
  D[L - 1] = addc (0u, T[L - 1] ^ xm, c);
  /* A - B < 0: only with equal exponents (otherwise A > B), so nothing was
   * shifted out and the window holds the exact difference in two's
   * complement (top word all ones): negate it, B's sign */
  {
    const bool neg = sub && (D[L - 1] >> 31) != 0;
    if (any_lane (neg))
      {
        CUFP_RARE ();
        u32 cn = 1;
        CUFP_UNROLL
        for (int i = 0; i < L; i++)
          {
            const u32 v = addc (~D[i], 0u, cn);
            D[i] = neg ? v : D[i];
          }
        sA = neg ? sB : sA;
      }
  }

  /* Normalise into E (L-1 words, top bit in E[L-2]): with a carry into the
   * top word the window moves right by one bit (the dropped bit goes to the
   * sticky), otherwise left by the leading zeros lz. */
  const u32 cy = D[L - 1] & 1u;
  st |= cy & D[0];
  u32 E[L - 1];
  u32 lz;
  bool zres = false;
  if (!any_lane (cy == 0u && D[L - 2] == 0u))
    {
      /* no lane cancelled a whole word: one funnel shift per word, right by
       * cy from the pairs (D[i+1]:D[i]) when there is a carry or no leading
       * zero, else right by 32 - lz from the pairs (D[i]:D[i-1]) */
      lz = cy ? 0u : (u32) clz32 (D[L - 2]);
      const bool up = cy != 0u || lz == 0u;
      const u32 sa = up ? cy : 32u - lz;
      u32 lo = up ? D[0] : 0u;
      CUFP_UNROLL
      for (int i = 0; i < L - 1; i++)
        {
          const u32 hi = up ? D[i + 1] : D[i];
          E[i] = fshr (hi, lo, sa);
          lo = hi;
        }
    }
  else
    {
      /* some lane cancelled a whole word (rare): carry shift, a full scan
       * of the leading zeros, word steps, then the bit shift */
      CUFP_RARE ();
      CUFP_UNROLL
      for (int i = 0; i < L - 1; i++)
        {
          u32 hi = D[i + 1];
          E[i] = cy ? fshr (hi, D[i], 1u) : D[i];
        }
      lz = 0;
      bool seen = false;
      CUFP_UNROLL
      for (int i = L - 2; i >= 0; i--)
        {
          u32 z = (u32) clz32 (E[i]);
          lz += seen ? 0u : z;
          seen = seen || (E[i] != 0);
        }
      zres = !seen;
      u32 wl = lz >> 5, bl = lz & 31u;
      CUFP_UNROLL
      for (int k = 0; (1 << k) <= L - 1; k++)
        {
          const int sh = 1 << k;
          u32 on = (wl >> k) & 1u;
          if (!any_lane (on != 0)) continue;
          CUFP_UNROLL
          for (int i = L - 2; i >= 0; i--)
            {
              u32 v = (i - sh >= 0) ? E[i - sh] : 0u;
              E[i] = on ? v : E[i];
            }
        }
      CUFP_UNROLL
      for (int i = L - 2; i >= 0; i--)
        {
          u32 lo = (i >= 1) ? E[i - 1] : 0u;
          E[i] = fshl (E[i], lo, bl);
        }
    }
  r.amb = false;
  if constexpr (ERRB > 0)
    {
      /* E's last bit is worth 2^(cy - lz - 32) units of A's last word, so
       * the error is below 2^(ERRB + 32 - cy + lz) of E's last bits; the
       * GW guard words g = E[0..GW-1] (round bit on top) are safe unless
       * |g - 2^(32 GW - 1)| <= that, tested on the top two guard words */
      constexpr int GW = L - 1 - KW;
      static_assert (GW == 3 || GW == 4, "the test reads the top two of 3 or 4 guard words");
      /* RAW: A sits one bit lower, so E's bits are worth twice as much */
      constexpr int R = RAW ? 1 : 0;
      static_assert (ERRB + 31 - R - 32 * (GW - 2) >= 0, "error position below the tested guard bits");
      const int t = ERRB + 32 - R - (int) cy + (int) lz - 32 * (GW - 2);
      const u64 x = ((u64) E[GW - 1] << 32) | E[GW - 2];
      bool amb = true;
      if (t < 63)
        {
          const u64 m = 1ull << (63 - t), q = x >> t;
          amb = (q == m) || (q + 1 == m);
        }
      r.amb = amb || zres;
    }
  /* p > 0: round to p bits (run-time precision) instead of KW words */
  u32 co = p ? round_top_p<KW, L - 1> (r.k, E, p, st) : round_top<KW, L - 1> (r.k, E, 0u, st);
  r.sign = zres ? 1 : sA;
  r.exp = eA + (RAW ? 1 : 0) + (i64) cy - (i64) lz + (i64) co;
  r.zero = zres;
  return r;
}

} // namespace core
} // namespace cu_fp

#endif
