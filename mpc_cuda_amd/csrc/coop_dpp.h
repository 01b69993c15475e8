/* coop_dpp.h -- lanes-per-number RNDN addition for the per-op folds with
 * lane groups of at most 16 (precisions up to 896 bits), moving data between
 * lanes with DPP (row shifts and row broadcasts at VALU latency) instead of
 * ds_bpermute round trips through the LDS crossbar.
 *
 * Same frame and the same result as coop.h (bit-exact with cu_fadd,
 * reference cu_freal.cuh:279-292: sign dispatch, |a|+|b| / |a|-|b| with
 * round/sticky and the sticky borrow, exact cancellation gives +0, x + 0 and
 * 0 + x return x): a number is spread over a group of G lanes (G = the power
 * of two >= N + 2), lane j = 0 a guard limb, lanes 1..N the limbs, lane N+1
 * the carry limb, sign and exponent replicated.  A group never spans two
 * 16-lane DPP rows, so every move stays inside its row; moves that would
 * cross into the neighbouring group of the same row are masked.
 *
 * One step, common case (exponents differ by less than 64 bits and no word
 * of cancellation) -- each item a few VALU instructions:
 *   order      by exponent (the limb comparison, two ballots, only when some
 *              group has equal exponents);
 *   align      B >> d within the limb: the upper neighbour by one row shift,
 *              a funnel shift; bits that fall off lane 0 are its sticky;
 *   add/sub    per-lane sums, lane 0 adding the two's complement's +1 unless
 *              bits fell off (the sticky borrow), carries into every lane at
 *              once from two ballots: carries = ((Gen << 1) + Prop) ^ Prop;
 *   normalise  from the carry lane and the top limb broadcast to the group:
 *              one bit right on a carry, left by the top limb's leading zeros
 *              otherwise, from the neighbours by row shifts;
 *   round      RNDN from the guard and first lanes broadcast to the group,
 *              the increment carried by one more pair of ballots.
 * Limb shifts of whole lanes (exponent differences of 64 bits or more, zero
 * operands) take row shifts by 1, 2, 4 and 8 lanes, and cancellations of a
 * whole limb or more a general path (a ballot for the top lane, ds_bpermute
 * moves), each only when some group of the wavefront needs it.
 *
 * The lane operations come from a policy type L (lane index, ballot, row
 * shift left / right by a constant, broadcast of a constant lane of the row,
 * quad permutation, bpermute), so the same code runs on the device
 * (DevLanes below) and in a host emulation of a 64-lane wavefront that the
 * CPU tests use. */
#ifndef MPGFX_COOP_DPP_H
#define MPGFX_COOP_DPP_H

#include <cstdint>

#if !defined(__HIPCC__) && !defined(__HIP__)
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#endif

namespace mpgfx {
namespace coopd {

typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;

constexpr i64 ZERO_EXP = -0x7fffffffffffffLL;

constexpr int lanes_for (int pb)
{
  return ((pb + 63) / 64 + 2 <= 4) ? 4 : ((pb + 63) / 64 + 2 <= 8) ? 8 : ((pb + 63) / 64 + 2 <= 16) ? 16
       : ((pb + 63) / 64 + 2 <= 32) ? 32 : ((pb + 63) / 64 + 2 <= 64) ? 64 : 0;
}

/* one lane's part of a number in the frame */
struct cv
{
  u64 m;
  i64 exp;
  int sign;
};

template <class L>
inline __host__ __device__ u64 bc64 (const L &w, u64 v, int src)
{
  const u32 lo = w.bperm (src, (u32) v), hi = w.bperm (src, (u32) (v >> 32));
  return ((u64) hi << 32) | lo;
}
/* lane j of the row reads lane j + S of the same row (0 past the row) */
template <int S, class L>
inline __host__ __device__ u64 up64 (const L &w, u64 v)
{
  return ((u64) w.template row_shl<S> ((u32) (v >> 32)) << 32) | w.template row_shl<S> ((u32) v);
}
/* lane j reads lane j - S of the same row (0 before the row) */
template <int S, class L>
inline __host__ __device__ u64 dn64 (const L &w, u64 v)
{
  return ((u64) w.template row_shr<S> ((u32) (v >> 32)) << 32) | w.template row_shr<S> ((u32) v);
}
/* lane j of the group reads lane j + 1 (G <= 16: inside the row; wider
 * groups span rows: the wavefront-wide shift), resp. lane j - 1; the caller
 * masks the group's edge */
template <int G, class L>
inline __host__ __device__ u64 nxt64 (const L &w, u64 v)
{
  if constexpr (G <= 16)
    return up64<1> (w, v);
  else
    return ((u64) w.wave_shl1 ((u32) (v >> 32)) << 32) | w.wave_shl1 ((u32) v);
}
template <int G, class L>
inline __host__ __device__ u64 prv64 (const L &w, u64 v)
{
  if constexpr (G <= 16)
    return dn64<1> (w, v);
  else
    return ((u64) w.wave_shr1 ((u32) (v >> 32)) << 32) | w.wave_shr1 ((u32) v);
}

/* lane OFF of the lane's group of G lanes */
template <int OFF, int G, class L>
inline __host__ __device__ u32 gbc32 (const L &w, u32 v)
{
  static_assert (OFF >= 0 && OFF < G, "lane inside the group");
  if constexpr (G == 64)
    return w.readlane (v, OFF);
  else if constexpr (G == 32)
    {
      const u32 x0 = w.readlane (v, OFF), x1 = w.readlane (v, 32 + OFF);
      return (w.lane () & 32) ? x1 : x0;
    }
  else if constexpr (G == 16)
    return w.template row_bcast<OFF> (v);
  else if constexpr (G == 8)
    {
      const u32 x0 = w.template row_bcast<OFF> (v), x1 = w.template row_bcast<8 + OFF> (v);
      return (w.lane () & 8) ? x1 : x0;
    }
  else
    return w.template quad_perm<OFF | (OFF << 2) | (OFF << 4) | (OFF << 6)> (v);
}
template <int OFF, int G, class L>
inline __host__ __device__ u64 gbc64 (const L &w, u64 v)
{
  return ((u64) gbc32<OFF, G> (w, (u32) (v >> 32)) << 32) | gbc32<OFF, G> (w, (u32) v);
}

inline __host__ __device__ int clz64 (u64 x)
{
  return x ? __builtin_clzll (x) : 64;
}

template <int G>
inline __host__ __device__ u64 gbits (u64 mask, int base)
{
  const u64 gm = (G == 64) ? ~0ull : ((1ull << (G & 63)) - 1);
  return (mask >> base) & gm;
}

inline __host__ __device__ cv zero ()
{
  cv z;
  z.m = 0;
  z.exp = ZERO_EXP;
  z.sign = 1;
  return z;
}

/* RNDN(a + b) at precision PB with cu_fadd's semantics; every lane of the
 * wavefront must call it */
template <int PB, class L>
inline __host__ __device__ cv add (const L &w, const cv &a, const cv &b)
{
  constexpr int NL = (PB + 63) / 64, G = lanes_for (PB), SB = 64 * NL - PB;
  static_assert (G >= 4 && G <= 64 && NL + 2 <= G, "frame does not fit a lane group");
  const int lane = w.lane (), j = lane & (G - 1), base = lane & ~(G - 1);
  const bool za = a.exp == ZERO_EXP, zb = b.exp == ZERO_EXP;

  /* ---- order by magnitude (zero is the smallest) ---- */
  bool age = zb || (!za && a.exp > b.exp);
  if (w.ballot (!za && !zb && a.exp == b.exp) != 0)
    {
      const u64 neq = gbits<G> (w.ballot (a.m != b.m), base), gt = gbits<G> (w.ballot (a.m > b.m), base);
      const int top_d = 63 - clz64 (neq | 1ull);
      const bool lge = (neq == 0) || ((gt >> top_d) & 1);
      age = zb || (!za && (a.exp > b.exp || (a.exp == b.exp && lge)));
    }
  const u64 A = age ? a.m : b.m, Bv = age ? b.m : a.m;
  const int sA0 = age ? a.sign : b.sign, sB = age ? b.sign : a.sign;
  const i64 eA = age ? a.exp : b.exp, eB = age ? b.exp : a.exp;
  const bool zB = age ? zb : za;

  /* ---- align: T = B >> d within the frame ---- */
  i64 d = zB ? 64 * G : eA - eB;
  d = d > 64 * G ? 64 * G : d;
  const int ls = (int) (d >> 6), bs = (int) (d & 63);
  u64 Bs = Bv;
  bool lost = false;                        /* material below the guard limb, this lane */
  if (w.ballot (ls != 0) != 0)
    {
      /* whole-limb shifts: lanes move down by ls (everything when ls = G) */
      const bool all_out = ls >= G;
      lost = all_out && (Bs != 0);
      Bs = all_out ? 0ull : Bs;
      const int lsx = all_out ? 0 : ls;
      if constexpr (G >= 32)
        {
          /* groups across rows: one ds_bpermute move by lsx lanes */
          for (int q = 0; q < lsx && q < G; q++) lost = lost || (j == q && Bs != 0);
          const int src = j + lsx;
          u64 v = bc64 (w, Bs, base + (src < G ? src : 0));
          Bs = (src < G) ? v : 0ull;
        }
      if constexpr (G > 1 && G <= 16)
        {
          const bool on = lsx & 1;
          lost = lost || (on && j < 1 && Bs != 0);
          u64 v = up64<1> (w, Bs);
          v = (j + 1 < G) ? v : 0ull;
          Bs = on ? v : Bs;
        }
      if constexpr (G > 2 && G <= 16)
        {
          const bool on = (lsx >> 1) & 1;
          lost = lost || (on && j < 2 && Bs != 0);
          u64 v = up64<2> (w, Bs);
          v = (j + 2 < G) ? v : 0ull;
          Bs = on ? v : Bs;
        }
      if constexpr (G > 4 && G <= 16)
        {
          const bool on = (lsx >> 2) & 1;
          lost = lost || (on && j < 4 && Bs != 0);
          u64 v = up64<4> (w, Bs);
          v = (j + 4 < G) ? v : 0ull;
          Bs = on ? v : Bs;
        }
      if constexpr (G > 8 && G <= 16)
        {
          const bool on = (lsx >> 3) & 1;
          lost = lost || (on && j < 8 && Bs != 0);
          u64 v = up64<8> (w, Bs);
          v = (j + 8 < G) ? v : 0ull;
          Bs = on ? v : Bs;
        }
    }
  u64 nb = nxt64<G> (w, Bs);
  nb = (j + 1 < G) ? nb : 0ull;
  const u64 T = bs ? ((Bs >> bs) | (nb << (64 - bs))) : Bs;
  lost = lost || (j == 0 && bs != 0 && (Bs << (64 - bs)) != 0);
  /* the sticky bit, valid in lane 0 of each group (lost bits beyond lane 0
   * only come from whole-limb shifts) */
  bool st0 = lost;
  if (w.ballot (lost && j != 0) != 0) st0 = gbits<G> (w.ballot (lost), base) != 0;

  /* ---- D = A + T, or A - T - st (the sticky borrow), over lanes 0..N ---- */
  const bool sub = (sA0 != sB) && !zB;
  const u64 X = sub ? (j <= NL ? ~T : 0ull) : T;
  const u64 cin = (j == 0 && sub && !st0) ? 1ull : 0ull;
  const u64 s1 = A + X, s = s1 + cin;
  /* A's guard limb is zero, so a lane cannot both generate and propagate */
  const u64 Gm = w.ballot (s1 < A || s < s1), Pm = w.ballot (s == ~0ull);
  const u64 C = ((Gm << 1) + Pm) ^ Pm;
  u64 res = s + ((C >> lane) & 1ull);
  res = (sub && j > NL) ? 0ull : res;
  const int sA = sA0;                       /* |A| >= |B|: the difference is not negative */

  /* ---- normalise: leading one to bit 63 of lane N ---- */
  const u64 rc = gbc64<NL + 1, G> (w, res), rt = gbc64<NL, G> (w, res);
  const bool carry = rc != 0;
  bool deep = !carry && rt == 0;
  u64 Rn;
  i64 eR;
  bool zres = false;
  if (w.ballot (deep) != 0)
    {
      /* general path: the top nonzero lane from a ballot, its leading zeros
       * and the two source limbs of every lane by ds_bpermute */
      const u64 nz = gbits<G> (w.ballot (res != 0), base);
      zres = nz == 0;
      const int top = 63 - clz64 (nz | 1ull);
      const int lzl = clz64 (res | 1ull);
      const int lz = (int) w.bperm (base + top, (u32) lzl);
      const int Lsh = (NL - top) * 64 + lz;                 /* left shift, >= -1 */
      const int off = -Lsh, q = off >> 6, r = off & 63;
      const int c0 = j + q, c1 = j + q + 1;
      u64 w0 = bc64 (w, res, base + ((c0 >= 0 && c0 < G) ? c0 : 0));
      u64 w1 = bc64 (w, res, base + ((c1 >= 0 && c1 < G) ? c1 : 0));
      w0 = (c0 >= 0 && c0 < G) ? w0 : 0ull;
      w1 = (c1 >= 0 && c1 < G) ? w1 : 0ull;
      Rn = r ? ((w0 >> r) | (w1 << (64 - r))) : w0;
      /* a right shift (carry-out) drops bit 0 of the guard limb into the sticky bit */
      const bool any_low = gbits<G> (w.ballot (j == 0 && (res & 1ull)), base) != 0;   /* every lane ballots */
      const bool drop = off > 0 && any_low;
      st0 = st0 || drop;
      eR = eA - Lsh;
    }
  else
    {
      const int lz = clz64 (rt);
      u64 upn = nxt64<G> (w, res), dnn = prv64<G> (w, res);
      upn = (j + 1 < G) ? upn : 0ull;
      dnn = (j >= 1) ? dnn : 0ull;
      Rn = carry ? ((res >> 1) | (upn << 63)) : (lz ? ((res << lz) | (dnn >> (64 - lz))) : res);
      st0 = st0 || (carry && j == 0 && (res & 1ull));
      eR = carry ? eA + 1 : eA - lz;
    }

  /* ---- RNDN on the kept bits: lanes 1..N minus the low SB bits of lane 1 ---- */
  const u64 g0 = gbc64<0, G> (w, Rn), g1 = gbc64<1, G> (w, Rn);
  const bool st = gbc32<0, G> (w, st0 ? 1u : 0u) != 0;
  bool rb, sbit, lb;
  if constexpr (SB == 0)
    {
      rb = (g0 >> 63) != 0;
      sbit = (g0 << 1) != 0;
      lb = (g1 & 1ull) != 0;
    }
  else
    {
      rb = ((g1 >> (SB - 1)) & 1ull) != 0;
      sbit = (g1 & ((1ull << ((SB - 1) & 63)) - 1)) != 0 || g0 != 0;
      lb = ((g1 >> (SB & 63)) & 1ull) != 0;
    }
  const bool inc = rb && (st || sbit || lb);
  const u64 low = (SB == 0) ? 0ull : ((1ull << (SB & 63)) - 1);
  const u64 kept = (j == 1) ? (Rn | low) : Rn;
  const u64 P2 = w.ballot (j >= 1 && j <= NL && kept == ~0ull);
  const u64 C2 = (w.ballot (j == 1 && inc) + P2) ^ P2;
  const bool cj = (C2 >> lane) & 1ull;
  u64 out = Rn + (cj ? ((j == 1) ? (1ull << (SB & 63)) : 1ull) : 0ull);
  out = (j == 1) ? (out & ~low) : out;
  const bool ovf = (C2 >> (base + NL + 1)) & 1ull;              /* rounded up to 2^PB */
  out = ovf ? ((j == NL) ? (1ull << 63) : 0ull) : out;
  eR += ovf ? 1 : 0;
  out = (j == 0 || j > NL) ? 0ull : out;

  /* x + 0 -> x, 0 + y -> y (operands returned unchanged) */
  cv v;
  v.m = za ? b.m : zb ? a.m : zres ? 0ull : out;
  v.exp = za ? b.exp : zb ? a.exp : zres ? ZERO_EXP : eR;
  v.sign = za ? b.sign : zb ? a.sign : zres ? 1 : sA;
  return v;
}

/* a record of N limbs then the sign/exponent word, as one lane loads it:
 * limb j-1 (lanes outside 1..N load a limb of the same record and drop it in
 * decode) and the sign/exponent word; no use of the values before decode, so
 * a prefetch stays in flight until the step that needs it */
struct raw
{
  u64 m;
  i64 w;
};

template <int PB, class L>
inline __host__ __device__ raw load_raw (const L &w, const u64 *__restrict__ rec)
{
  constexpr int NL = (PB + 63) / 64, G = lanes_for (PB);
  const int j = w.lane () & (G - 1);
  const int jj = j < 1 ? 0 : (j > NL ? NL - 1 : j - 1);
  raw v;
  v.m = rec[jj];
  v.w = (i64) rec[NL];
  return v;
}

template <int PB, class L>
inline __host__ __device__ cv decode (const L &w, const raw &r)
{
  constexpr int NL = (PB + 63) / 64, G = lanes_for (PB);
  const int j = w.lane () & (G - 1);
  cv v;
  v.m = (j >= 1 && j <= NL) ? r.m : 0ull;
  v.sign = (r.w & 1) ? -1 : 1;
  v.exp = r.w >> 1;
  return v;
}

/* acc = +0; acc = RNDN(acc + p_s) for s < len, p_s the record at rec0 + s*step
 * (records of NL+1 words); the wavefront runs max_len steps (every group takes
 * part in the cross-lane operations), a group past its own length keeps its
 * value.  Records are prefetched PD steps ahead; the records at rec0 + s*step
 * must exist for every s < max_len (max_len is wave-uniform), also for a
 * group with len < max_len, and a prefetch past max_len - 1 loads record
 * max_len - 1 again. */
constexpr int PD = 8;

template <int PB, class L>
inline __host__ __device__ cv fold (const L &w, const u64 *__restrict__ recs, i64 rec0, i64 step, i64 len, i64 max_len)
{
  constexpr int NL = (PB + 63) / 64;
  cv acc = zero ();
  if (max_len <= 0) return acc;
  raw ring[PD];
#pragma unroll
  for (int d = 0; d < PD; d++)
    ring[d] = load_raw<PB> (w, recs + (rec0 + step * (d < max_len ? d : max_len - 1)) * (NL + 1));
  for (i64 s0 = 0; s0 < max_len; s0 += PD)
    {
#pragma unroll
      for (int d = 0; d < PD; d++)
        {
          const i64 s = s0 + d;
          if (s < max_len)
            {
              const cv nv = add<PB> (w, acc, decode<PB> (w, ring[d]));
              const bool live = s < len;
              acc.m = live ? nv.m : acc.m;
              acc.exp = live ? nv.exp : acc.exp;
              acc.sign = live ? nv.sign : acc.sign;
              const i64 sn = s + PD < max_len ? s + PD : max_len - 1;
              ring[d] = load_raw<PB> (w, recs + (rec0 + step * sn) * (NL + 1));
            }
        }
    }
  return acc;
}

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
/* the device's lane operations: DPP row shifts (row_shl / row_shr with zero
 * fill), row broadcasts (row_newbcast), quad permutations, ballots and
 * ds_bpermute */
struct DevLanes
{
  __device__ int lane () const { return (int) (threadIdx.x & 63); }
  __device__ u64 ballot (bool c) const { return __builtin_amdgcn_ballot_w64 (c); }
  template <int S> __device__ u32 row_shl (u32 v) const
  {
    return (u32) __builtin_amdgcn_update_dpp (0, (int) v, 0x100 + S, 0xf, 0xf, true);
  }
  template <int S> __device__ u32 row_shr (u32 v) const
  {
    return (u32) __builtin_amdgcn_update_dpp (0, (int) v, 0x110 + S, 0xf, 0xf, true);
  }
  template <int S> __device__ u32 row_bcast (u32 v) const
  {
    return (u32) __builtin_amdgcn_update_dpp (0, (int) v, 0x150 + S, 0xf, 0xf, true);
  }
  template <int P> __device__ u32 quad_perm (u32 v) const
  {
    return (u32) __builtin_amdgcn_update_dpp (0, (int) v, P, 0xf, 0xf, true);
  }
  __device__ u32 bperm (int src, u32 v) const { return (u32) __builtin_amdgcn_ds_bpermute (src << 2, (int) v); }
  /* lane i reads lane i+1 / i-1 of the wavefront (0 at the ends) */
  __device__ u32 wave_shl1 (u32 v) const { return (u32) __builtin_amdgcn_update_dpp (0, (int) v, 0x130, 0xf, 0xf, true); }
  __device__ u32 wave_shr1 (u32 v) const { return (u32) __builtin_amdgcn_update_dpp (0, (int) v, 0x138, 0xf, 0xf, true); }
  /* the value of lane `src` (a constant) in every lane */
  __device__ u32 readlane (u32 v, int src) const { return (u32) __builtin_amdgcn_readlane ((int) v, src); }
};
#endif

} // namespace coopd
} // namespace mpgfx
#endif
