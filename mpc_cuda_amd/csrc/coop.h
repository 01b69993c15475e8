/* coop.h -- the lanes-per-number frame shared by the per-op folds (GEMV rows,
 * dot products; the addition itself is in coop_dpp.h): a number occupies a
 * group of G lanes (G = the power of two >= N + 2), one 64-bit limb per lane,
 *
 *   lane j = 0        guard limb (bits below the kept precision)
 *   lane j = 1..N     limbs 0..N-1 of the significand
 *   lane j = N+1..    carry limb, then zeros
 *
 * with sign and exponent replicated in every lane of the group; a wavefront
 * holds R = 64 / G numbers.  Here: the lane's part of a number, the store of
 * a group's number into an MP array, and the record format of the product
 * scratch (N limbs then the sign/exponent word).  Included inside an
 * anonymous namespace of each translation unit that folds. */

/* one lane's part of a number in the frame */
struct cval
{
  u64 m;
  i64 exp;
  int sign;
};

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
