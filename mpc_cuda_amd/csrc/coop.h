/* coop.h -- lanes-per-number RNDN addition for the sequential per-op folds
 * (GEMV rows, dot products): acc = RNDN(acc + p), bit-exact with cu_fadd
 * (reference cu_freal.cuh:279-292: sign dispatch, |a|+|b| / |a|-|b| with
 * round/sticky, exact cancellation gives +0, x + 0 and 0 + x return x).
 *
 * A fold is a chain of dependent additions, so one lane per chain leaves the
 * chip idle when chains are few (4096 GEMV rows are 64 wavefronts).  Here a
 * number occupies a group of G lanes (G = the power of two >= N + 2), one
 * 64-bit limb per lane, in a "frame":
 *
 *   lane j = 0        guard limb (bits below the kept precision)
 *   lane j = 1..N     limbs 0..N-1 of the significand
 *   lane j = N+1..    carry limb, then zeros
 *
 * with sign and exponent replicated in every lane of the group; a wavefront
 * holds R = 64 / G numbers.  Each step of the addition is a few VALU
 * instructions per lane plus cross-lane work done with ballots (wave-uniform
 * 64-bit masks) and ds_bpermute:
 *   - magnitude order: the highest differing limb from ballot(a != b), its
 *     direction from ballot(a > b);
 *   - alignment: B >> d with d = 64 ls + bs: lane j reads lanes j+ls, j+ls+1;
 *     the bits shifted out below the guard limb become a sticky bit (ballot);
 *   - addition / subtraction with sticky borrow: per-lane sums, then the carry
 *     into every lane at once from the generate / propagate masks,
 *     carries = ((G << 1 | cin) + P) ^ P  (a group's top lane never generates
 *     or propagates, so carries stay inside their group);
 *   - normalisation: the top nonzero lane (ballot) and its leading zeros give
 *     a signed shift, applied as a two-lane funnel shift;
 *   - RNDN on the kept PB bits (round bit, sticky, ties to even) and the
 *     increment, again with a propagate mask.
 * Everything is templated on the precision; the header is included inside
 * an anonymous namespace of each translation unit that folds (the
 * per-precision kernels, and the batch's all-precision fold kernel). */

/* one lane's part of a number in the frame */
struct cval
{
  u64 m;
  i64 exp;
  int sign;
};

__device__ __forceinline__ u64 bperm64 (u64 v, int src)
{
  const int a = src << 2;
  const u32 lo = (u32) __builtin_amdgcn_ds_bpermute (a, (int) (u32) v);
  const u32 hi = (u32) __builtin_amdgcn_ds_bpermute (a, (int) (u32) (v >> 32));
  return ((u64) hi << 32) | lo;
}

/* the group's bits of a wave mask, shifted down to bit 0 (groups of G lanes) */
template <int G>
__device__ __forceinline__ u64 gbits (u64 mask, int base)
{
  const u64 gm = (G == 64) ? ~0ull : ((1ull << (G & 63)) - 1);
  return (mask >> base) & gm;
}

__device__ __forceinline__ cval coop_zero ()
{
  cval z;
  z.m = 0;
  z.exp = CU_FREAL_ZERO_EXP;
  z.sign = 1;
  return z;
}

/* RNDN(a + b) with cu_fadd's semantics at precision PB; every lane of the
 * wavefront must call it */
template <int PB>
__device__ __forceinline__ cval coop_add (const cval &a, const cval &b)
{
  constexpr int NL = (PB + 63) / 64, COOP_G = coop_lanes (PB), COOP_SB = 64 * NL - PB;
  static_assert (NL + 2 <= COOP_G && COOP_G <= 64, "frame does not fit the lane group");
  const int lane = (int) (threadIdx.x & 63), j = lane & (COOP_G - 1), base = lane & ~(COOP_G - 1);
  const bool za = a.exp == CU_FREAL_ZERO_EXP, zb = b.exp == CU_FREAL_ZERO_EXP;

  /* order by magnitude (zero is the smallest) */
  const u64 neq = gbits<COOP_G> (__ballot (a.m != b.m), base), gt = gbits<COOP_G> (__ballot (a.m > b.m), base);
  const int top_d = 63 - __clzll (neq | 1ull);
  const bool lge = (neq == 0) || ((gt >> top_d) & 1);
  const bool age = zb || (!za && (a.exp > b.exp || (a.exp == b.exp && lge)));
  const u64 A = age ? a.m : b.m, B = age ? b.m : a.m;
  const int sA = age ? a.sign : b.sign, sB = age ? b.sign : a.sign;
  const i64 eA = age ? a.exp : b.exp, eB = age ? b.exp : a.exp;
  const bool zB = age ? zb : za;

  /* T = B >> d within the frame; bits below the guard limb -> sticky */
  i64 d = zB ? 64 * COOP_G : eA - eB;
  d = d > 64 * COOP_G ? 64 * COOP_G : d;
  const int ls = (int) (d >> 6), bs = (int) (d & 63);
  const int s0 = j + ls;
  u64 b0 = bperm64 (B, base + (s0 < COOP_G ? s0 : 0));
  u64 b1 = bperm64 (B, base + (s0 + 1 < COOP_G ? s0 + 1 : 0));
  b0 = s0 < COOP_G ? b0 : 0;
  b1 = s0 + 1 < COOP_G ? b1 : 0;
  const u64 T = bs ? ((b0 >> bs) | (b1 << (64 - bs))) : b0;
  const bool lost = (j < ls) ? (B != 0) : (j == ls && bs != 0 && (B << (64 - bs)) != 0);
  bool st = gbits<COOP_G> (__ballot (lost), base) != 0;

  /* D = A + T, or A - T - st (sticky borrow), over lanes 0..N */
  const bool sub = (sA != sB) && !zB;
  const u64 X = sub ? (j <= NL ? ~T : 0ull) : T;
  const u64 s = A + X;
  const u64 Gm = __ballot (s < A), Pm = __ballot (s == ~0ull);
  const u64 Cm = __ballot (j == 0 && sub && !st);
  const u64 C = (((Gm << 1) | Cm) + Pm) ^ Pm;
  u64 res = s + ((C >> lane) & 1ull);
  res = (sub && j > NL) ? 0ull : res;

  /* normalise: leading one to bit 63 of lane N */
  const u64 nz = gbits<COOP_G> (__ballot (res != 0), base);
  const bool zres = nz == 0;
  const int top = 63 - __clzll (nz | 1ull);
  const int lzl = __clzll (res | 1ull);                       /* this lane's leading zeros */
  const int lz = __builtin_amdgcn_ds_bpermute ((base + top) << 2, lzl);
  const int L = (NL - top) * 64 + lz;                         /* left shift, >= -1 */
  const int off = -L, q = off >> 6, r = off & 63;
  const int c0 = j + q, c1 = j + q + 1;
  u64 w0 = bperm64 (res, base + ((c0 >= 0 && c0 < COOP_G) ? c0 : 0));
  u64 w1 = bperm64 (res, base + ((c1 >= 0 && c1 < COOP_G) ? c1 : 0));
  w0 = (c0 >= 0 && c0 < COOP_G) ? w0 : 0ull;
  w1 = (c1 >= 0 && c1 < COOP_G) ? w1 : 0ull;
  u64 Rn = r ? ((w0 >> r) | (w1 << (64 - r))) : w0;
  /* a right shift (carry-out) drops bit 0 of the guard limb into the sticky bit */
  st = st || (off > 0 && gbits<COOP_G> (__ballot (j == 0 && (res & 1ull)), base) != 0);
  i64 eR = eA - L;

  /* RNDN on the kept bits: lanes 1..N minus the low SB bits of lane 1 */
  bool rf, sf, lf;
  if (COOP_SB == 0)
    {
      rf = (j == 0) && (Rn >> 63);
      sf = (j == 0) && (Rn << 1) != 0;
      lf = (j == 1) && (Rn & 1ull);
    }
  else
    {
      rf = (j == 1) && ((Rn >> (COOP_SB - 1)) & 1ull);
      sf = ((j == 1) && (Rn & ((1ull << ((COOP_SB - 1) & 63)) - 1)) != 0) || ((j == 0) && Rn != 0);
      lf = (j == 1) && ((Rn >> (COOP_SB & 63)) & 1ull);
    }
  const bool rb = gbits<COOP_G> (__ballot (rf), base) != 0;
  const bool sb = gbits<COOP_G> (__ballot (sf), base) != 0;
  const bool lb = gbits<COOP_G> (__ballot (lf), base) != 0;
  const bool inc = rb && (st || sb || lb);
  const u64 low = (COOP_SB == 0) ? 0ull : ((1ull << (COOP_SB & 63)) - 1);
  const u64 kept = (j == 1) ? (Rn | low) : Rn;
  const u64 P2 = __ballot (j >= 1 && j <= NL && kept == ~0ull);
  const u64 C2 = (__ballot (j == 1 && inc) + P2) ^ P2;
  const bool cj = (C2 >> lane) & 1ull;
  u64 out = Rn + (cj ? ((j == 1) ? (1ull << (COOP_SB & 63)) : 1ull) : 0ull);
  out = (j == 1) ? (out & ~low) : out;
  const bool ovf = (C2 >> (base + NL + 1)) & 1ull;              /* rounded up to 2^PB */
  out = ovf ? ((j == NL) ? (1ull << 63) : 0ull) : out;
  eR += ovf ? 1 : 0;
  out = (j == 0 || j > NL) ? 0ull : out;

  /* x + 0 -> x, 0 + y -> y (operands returned unchanged); field-wise selects */
  cval v;
  v.m = za ? b.m : zb ? a.m : zres ? 0ull : out;
  v.exp = za ? b.exp : zb ? a.exp : zres ? (i64) CU_FREAL_ZERO_EXP : eR;
  v.sign = za ? b.sign : zb ? a.sign : zres ? 1 : sA;
  return v;
}

/* a record of N limbs then the sign/exponent word (NL + 1 words), as one
 * lane loads it: limb j-1 (lanes outside 1..N load a limb of the same record
 * and drop it in coop_decode) and the sign/exponent word.  Every lane loads,
 * with no branch and no use of the values until coop_decode, so a prefetch
 * stays in flight until the step that needs it. */
struct craw
{
  u64 m;
  i64 w;
};

template <int PB>
__device__ __forceinline__ craw coop_load_raw (const u64 *__restrict__ rec)
{
  constexpr int NL = (PB + 63) / 64, COOP_G = coop_lanes (PB);
  const int j = (int) (threadIdx.x & (COOP_G - 1));
  const int jj = j < 1 ? 0 : (j > NL ? NL - 1 : j - 1);
  craw v;
  v.m = rec[jj];
  v.w = (i64) rec[NL];
  return v;
}

/* frame part of a loaded record: lanes 1..N hold limbs 0..N-1 */
template <int PB>
__device__ __forceinline__ cval coop_decode (const craw &r)
{
  constexpr int NL = (PB + 63) / 64, COOP_G = coop_lanes (PB);
  const int j = (int) (threadIdx.x & (COOP_G - 1));
  cval v;
  v.m = (j >= 1 && j <= NL) ? r.m : 0ull;
  v.sign = (r.w & 1) ? -1 : 1;
  v.exp = r.w >> 1;
  return v;
}

/* store the group's number to an MP array (limb planes + sign/exponent word) */
template <int PB>
__device__ __forceinline__ void coop_store (const cval &v, u64 *__restrict__ l, i64 *__restrict__ se, i64 n, i64 i)
{
  constexpr int NL = (PB + 63) / 64, COOP_G = coop_lanes (PB);
  const int j = (int) (threadIdx.x & (COOP_G - 1));
  if (j >= 1 && j <= NL) l[(i64) (j - 1) * n + i] = v.m;
  if (j == 0) se[i] = (i64) ((u64) v.exp << 1) | (v.sign < 0 ? 1 : 0);
}

/* record store of a whole value (one lane) */
template <int PB>
__device__ __forceinline__ void st_rec (const cu_freal<PB> &v, u64 *__restrict__ rec)
{
  constexpr int NL = (PB + 63) / 64;
#pragma unroll
  for (int k = 0; k < NL; k++) rec[k] = v.m[k];
  rec[NL] = (u64) (((i64) v.exp << 1) | (v.sign < 0 ? 1 : 0));
}

/* acc = +0; acc = RNDN(acc + p_s) for s < len, p_s the record at rec0 + s*step
 * (in records of NL+1 words); the wavefront runs max_len steps (every group
 * must take part in the cross-lane operations), a group past its own length
 * keeps its value.  Records are prefetched COOP_PD steps ahead.  The caller
 * guarantees that the records at rec0 + s*step exist for every s < max_len
 * (max_len is wave-uniform), also for a group with len < max_len: a group
 * loads, and then ignores, the records past its own length, and a prefetch
 * past max_len - 1 loads record max_len - 1 again. */
constexpr int COOP_PD = 8;

template <int PB>
__device__ __forceinline__ cval coop_fold (const u64 *__restrict__ recs, i64 rec0, i64 step, i64 len, i64 max_len)
{
  constexpr int NL = (PB + 63) / 64;
  cval acc = coop_zero ();
  if (max_len <= 0) return acc;
  craw ring[COOP_PD];
#pragma unroll
  for (int d = 0; d < COOP_PD; d++)
    ring[d] = coop_load_raw<PB> (recs + (rec0 + step * (d < max_len ? d : max_len - 1)) * (NL + 1));
  for (i64 s0 = 0; s0 < max_len; s0 += COOP_PD)
    {
#pragma unroll
      for (int d = 0; d < COOP_PD; d++)
        {
          const i64 s = s0 + d;
          if (s < max_len)
            {
              const cval nv = coop_add<PB> (acc, coop_decode<PB> (ring[d]));
              const bool live = s < len;
              acc.m = live ? nv.m : acc.m;
              acc.exp = live ? nv.exp : acc.exp;
              acc.sign = live ? nv.sign : acc.sign;
              const i64 sn = s + COOP_PD < max_len ? s + COOP_PD : max_len - 1;
              ring[d] = coop_load_raw<PB> (recs + (rec0 + step * sn) * (NL + 1));
            }
        }
    }
  return acc;
}
