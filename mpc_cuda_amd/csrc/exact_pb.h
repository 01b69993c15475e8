/* exact_pb.h -- per-precision kernels of the exact-dot GEMM (see exact.h for
 * the scheme).  Included by kernels_pb.hip inside its anonymous namespace,
 * where PB (bits) and NL (u64 limbs) are defined. */

constexpr int KWD = PB / 32;                   /* significand words (pb bits) */
constexpr int SBW = cu_freal<PB>::SB / 32;     /* zero low words of storage   */

typedef int v4i __attribute__ ((ext_vector_type (4)));

/* load the pb-bit significand M' (storage >> SB) as KWD words */
__device__ __forceinline__ void load_sig (const u64 *__restrict__ l, i64 stride, i64 i, u32 (&w)[KWD])
{
  u32 all[2 * NL];
#pragma unroll
  for (int q = 0; q < NL; q++)
    {
      u64 v = l[(i64) q * stride + i];
      all[2 * q] = (u32) v;
      all[2 * q + 1] = (u32) (v >> 32);
    }
#pragma unroll
  for (int q = 0; q < KWD; q++) w[q] = all[q + SBW];
}

/* The S signed digits of one value, put(t, d) for every plane t < Spad
 * (planes past S, and every plane of a value that is not live, get 0): the
 * value as an integer aligned to its line's largest exponent em (M' * 2^sh,
 * sh = spread - (em - e)), negated for a negative sign, in base-256 digits
 * recoded to [-128, 127] with a carry (the top digit keeps the sign). */
template <typename Put>
__device__ __forceinline__ void exact_value_digits (const u64 *__restrict__ l, const i64 *__restrict__ se, i64 nv,
                                                    i64 idx, bool live, i64 em, i64 spread, int S, int Spad, Put &&put)
{
  const i64 w = live ? se[idx] : 0;
  const i64 e = w >> 1;
  live = live && (e != MPGFX_ZERO_EXP) && (em != EXACT_EMPTY);
  if (!live)
    {
      for (int t = 0; t < Spad; t++) put (t, 0);
      return;
    }
  u32 M[KWD];
  load_sig (l, nv, idx, M);
  const i64 sh = spread - (em - e);              /* 0 <= sh <= spread */
  const int q = (int) (sh >> 3), r = (int) (sh & 7);
  u32 X[KWD + 1];                                 /* M' << r */
#pragma unroll
  for (int j = KWD; j >= 0; j--)
    {
      u32 hi = (j < KWD) ? M[j] : 0u, lo = (j >= 1) ? M[j - 1] : 0u;
      X[j] = cu_fp::core::fshl (hi, lo, (u32) r);
    }
  const u32 neg = (u32) (w & 1);
  u32 cn = neg, c = 0;                            /* negation carry, recode carry */
  for (int t = 0; t < q && t < Spad; t++) put (t, 0);
#pragma unroll
  for (int b = 0; b < 4 * (KWD + 1); b++)
    {
      const int t = q + b;
      u32 y = (X[b >> 2] >> (8 * (b & 3))) & 255u;
      if (neg) { y = (~y & 255u) + cn; cn = y >> 8; y &= 255u; }
      int d;
      if (t == S - 1)
        { d = (int) (y >= 128u ? (int) y - 256 : (int) y) + (int) c; c = 0; }
      else
        { u32 v = y + c; c = v >= 128u; d = (int) v - (c ? 256 : 0); }
      if (t < S) put (t, d);
    }
  for (int t = q + 4 * (KWD + 1); t < S; t++)
    {
      u32 y = 0;
      if (neg) { y = 255u + cn; cn = y >> 8; y &= 255u; }
      int d;
      if (t == S - 1)
        { d = (int) (y >= 128u ? (int) y - 256 : (int) y) + (int) c; c = 0; }
      else
        { u32 v = y + c; c = v >= 128u; d = (int) v - (c ? 256 : 0); }
      put (t, d);
    }
  for (int t = (S > q ? S : q); t < Spad; t++) put (t, 0);
}

/* Digit planes of one operand, D[t][L][p] for t < Spad, L < Lpad, p < kp
 * (padding and planes >= S are 0).  Lines are the rows of A (value L*len +
 * p, positions contiguous; COLS false) or the columns of B (value p*lines +
 * L, lines contiguous; COLS true).  A workgroup of 1024 lanes takes LT lines
 * x PT positions, each lane one value, read so that consecutive lanes read
 * consecutive values (positions for rows, lines for columns); its digits go
 * to an LDS tile [t][line][position], which is written out as PT-byte runs
 * of positions per line and plane (four bytes per lane), so neither side is
 * strided in HBM.  The tile holds TP = 64 planes (64 KB, so two workgroups
 * share a CU and one's loads overlap the other's stores); a value with more
 * digit planes is recoded once per 64 of them.  With a device plan `pl` the
 * spread (sA for which = 0, sB for 1), S and Spad come from it, and nothing is
 * written when it takes the wide path. */
template <bool COLS>
struct exd_tile
{
  static constexpr int LT = COLS ? 16 : 4, PT = COLS ? 64 : 256, TP = 64;
};
template <bool COLS>
__global__ __launch_bounds__ (1024) void k_exact_digits (i64 lines, i64 len, const u64 *__restrict__ l,
                                                         const i64 *__restrict__ se, i64 nv,
                                                         const i64 *__restrict__ emax_line, i64 spread, int S,
                                                         int Spad, i64 Lpad, i64 kp, int8_t *__restrict__ D,
                                                         const exact_dplan *__restrict__ pl, int which)
{
  constexpr int LT = exd_tile<COLS>::LT, PT = exd_tile<COLS>::PT, TP = exd_tile<COLS>::TP, TS = LT * PT;
  static_assert (TS == 1024, "one value per lane");
  __shared__ __attribute__ ((aligned (16))) int8_t tile[TP * TS];
  if (pl)
    {
      if (pl->fallback) return;
      spread = which ? pl->sB : pl->sA;
      S = pl->S;
      Spad = pl->G * pl->nG;
    }
  const int jj = COLS ? threadIdx.x % LT : threadIdx.x / PT, pp = COLS ? threadIdx.x / LT : threadIdx.x % PT;
  const i64 L0 = (i64) blockIdx.x * LT, p0 = (i64) blockIdx.y * PT;
  const i64 L = L0 + jj, p = p0 + pp;
  const bool live = (L < lines) && (p < len);
  const i64 idx = live ? (COLS ? p * lines + L : L * len + p) : 0;
  const i64 em = live ? emax_line[L] : EXACT_EMPTY;
  int8_t *const mine = tile + jj * PT + pp;
  const i64 plane = Lpad * kp;
  constexpr int WL = PT / 4;                      /* 32-bit words per line and plane */
  for (int t0 = 0; t0 < Spad; t0 += TP)
    {
      /* planes t0 .. t0 + TP - 1 of this tile */
      const int np = Spad - t0 < TP ? Spad - t0 : TP;
      if (t0 > 0) __syncthreads ();               /* the previous pass's stores have read the tile */
      exact_value_digits (l, se, nv, idx, live, em, spread, S, Spad, [&] (int t, int d) {
        const int u = t - t0;
        if ((unsigned) u < (unsigned) TP) mine[u * TS] = (int8_t) d;
      });
      __syncthreads ();
      /* plane t0 + u, line r: PT bytes of positions, four per lane */
      for (int x = threadIdx.x; x < np * LT * WL; x += 1024)
        {
          const int u = x / (LT * WL), r = (x / WL) % LT, qd = x % WL;
          const i64 Lr = L0 + r, pr = p0 + 4 * qd;
          if (Lr < Lpad && pr < kp)
            *(u32 *) (D + (i64) (t0 + u) * plane + Lr * kp + pr) = *(const u32 *) (tile + u * TS + r * PT + 4 * qd);
        }
    }
}

/* Whether an error below 2^eb (integer units) could change the rounding of
 * the magnitude held in wbuf[0..nw) (top nonzero word h, -1 for zero) to pb
 * bits: the magnitude is below 2^(eb+2), the error is not below an eighth of
 * the last place (round bit rb < eb + 2), or bits eb .. rb read 2^(rb-eb) or
 * 2^(rb-eb) - 1, i.e. the value lies within 2^eb of a rounding midpoint
 * (tested on the top 64 of those bits when there are more). */
template <typename W>
__device__ bool exact_undecided (const W *__restrict__ wbuf, int nw, int h, int eb)
{
  if (h < 0) return true;
  const i64 tb = 32 * (i64) h + 31 - (i64) cu_fp::core::clz32 ((u32) wbuf[h]);
  const i64 rb = tb - PB;
  if (rb - eb < 2) return true;
  auto bits64 = [&] (i64 pos) -> u64 {           /* magnitude bits pos .. pos + 63, pos >= 0 */
    const i64 w = pos >> 5;
    const u32 b = (u32) (pos & 31);
    const u64 x0 = (w < nw) ? (u32) wbuf[w] : 0u, x1 = (w + 1 < nw) ? (u32) wbuf[w + 1] : 0u,
              x2 = (w + 2 < nw) ? (u32) wbuf[w + 2] : 0u;
    const u64 lo = x0 | (x1 << 32);
    return b ? ((lo >> b) | (x2 << (64 - b))) : lo;
  };
  const i64 sft = rb - eb;
  if (sft <= 63)
    {
      const u64 mask = (sft == 63) ? ~0ull : ((1ull << (sft + 1)) - 1ull);
      const u64 q = bits64 (eb) & mask, mid = 1ull << sft;
      return q == mid || q + 1 == mid;
    }
  const u64 q = bits64 (rb - 63), mid = 1ull << 63;
  return q == mid || q + 1 == mid;
}

/* Round the two's complement integer held in the first `nw` 32-bit words of
 * wbuf (stored one per 64-bit slot) times 2^scale to pb bits and store it as
 * value i of C.  The words are read at runtime positions from memory, so no
 * register array is indexed dynamically.  eb >= 0: the integer is within 2^eb
 * of the exact one, and an element whose rounding that could change gets the
 * word EXACT_FIX_MARK in Cse instead (the exact fix-up recomputes it). */
template <typename W>
__device__ void round_store_int (W *__restrict__ wbuf, int nw, i64 scale, u64 *__restrict__ Cl,
                                 i64 *__restrict__ Cse, i64 nC, i64 i, int eb = -1)
{
  /* sign from the top word, then magnitude in place */
  const bool neg = ((u32) wbuf[nw - 1] >> 31) != 0;
  if (neg)
    {
      u32 c = 1;
      for (int s = 0; s < nw; s++)
        {
          u32 v = ~(u32) wbuf[s];
          wbuf[s] = (W) cu_fp::core::addc (v, 0u, c);
        }
    }
  int h = -1;
  for (int s = nw - 1; s >= 0; s--)
    if ((u32) wbuf[s] != 0u) { h = s; break; }
  if (eb >= 0 && exact_undecided (wbuf, nw, h, eb))
    {
      Cse[i] = EXACT_FIX_MARK;
      return;
    }
  cu_freal<PB> r;
  if (h < 0)
    {
      r.set_zero ();
      st<PB> (r, Cl, Cse, nC, i);
      return;
    }
  constexpr int WL = KWD + 2;
  u32 win[WL];
  u32 sticky = 0;
#pragma unroll
  for (int q = 0; q < WL; q++)
    {
      int s = h - (WL - 1) + q;
      win[q] = (s >= 0) ? (u32) wbuf[s] : 0u;
    }
  for (int s = 0; s < h - (WL - 1); s++) sticky |= (u32) wbuf[s];
  const u32 lz = (u32) cu_fp::core::clz32 (win[WL - 1]);
#pragma unroll
  for (int q = WL - 1; q >= 0; q--)
    {
      u32 lo = (q >= 1) ? win[q - 1] : 0u;
      win[q] = cu_fp::core::fshl (win[q], lo, lz);
    }
  u32 K[KWD];
  u32 co = cu_fp::core::round_top<KWD, WL> (K, win, 0u, sticky);
  r.set_kept (K);
  r.sign = neg ? -1 : 1;
  r.exp = (long) (scale + 32 * (i64) (h + 1) - (i64) lz + (i64) co);
  st<PB> (r, Cl, Cse, nC, i);
}

/* Rounding stage: one lane per output element.  The MFMA kernel leaves, for
 * element (i, j), `slots` carry-save int64 slots (slot q weighs 2^(32q), slot-
 * major: S64[q * nE + i*np + j]); they are carry-normalised into 32-bit
 * two's complement words, staged in LDS and rounded once to pb bits. */
__global__ __launch_bounds__ (256) void k_exact_round (i64 m, i64 n, i64 np, i64 nE, int slots,
                                                       const i64 *__restrict__ S64, const i64 *__restrict__ emaxA,
                                                       const i64 *__restrict__ emaxB, i64 sA, i64 sB,
                                                       u64 *__restrict__ Cl, i64 *__restrict__ Cse,
                                                       const exact_dplan *__restrict__ pl)
{
  __shared__ u32 words[256 * EXACT_SLMAX];
  if (pl)
    {
      if (pl->fallback) return;
      slots = pl->slots;
      sA = pl->sA;
      sB = pl->sB;
    }
  const i64 e = (i64) blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m * n) return;
  const i64 i = e / n, j = e - i * n, el = i * np + j;
  /* slots below cut/4 hold only diagonals below the cut: not stored by the
   * MFMA stage, read as 0 (within the cut's error bound) */
  const int q0 = pl ? pl->cut / 4 : 0;
  i64 x[EXACT_SLMAX];
#pragma unroll
  for (int q = 0; q < EXACT_SLMAX; q++) x[q] = (q >= q0 && q < slots) ? S64[(i64) q * nE + el] : 0;
  u32 *wb = words + threadIdx.x * EXACT_SLMAX;
  i64 carry = 0;
#pragma unroll
  for (int q = 0; q < EXACT_SLMAX; q++)
    {
      const i64 y = x[q] + carry;
      if (q < slots) wb[q] = (u32) y;
      carry = y >> 32;
    }
  const i64 ea = emaxA[i], eb = emaxB[j];
  const i64 scale = (ea == EXACT_EMPTY || eb == EXACT_EMPTY) ? 0 : ea + eb - sA - sB - 2 * (i64) PB;
  round_store_int (wb, slots, scale, Cl, Cse, m * n, e, pl ? pl->eb : -1);
}

/* VALU fallback for one output (i, j): its exact sum accumulated in `nw`
 * two's complement 32-bit words (one per 64-bit slot of acc) whose LSB weighs
 * 2^(eminA_i + eminB_j - 2pb), then rounded once. */
__device__ void exact_fb_output (i64 m, i64 n, i64 k, i64 i, i64 j, const u64 *__restrict__ Al,
                                 const i64 *__restrict__ Ase, const u64 *__restrict__ Bl, const i64 *__restrict__ Bse,
                                 const i64 *__restrict__ eminA, const i64 *__restrict__ eminB, int nw,
                                 i64 *__restrict__ acc, u64 *__restrict__ Cl, i64 *__restrict__ Cse)
{
  for (int s = 0; s < nw; s++) acc[s] = 0;
  const i64 base = (eminA[i] == EXACT_EMPTY || eminB[j] == EXACT_EMPTY) ? 0 : eminA[i] + eminB[j];
  bool any = false;
  for (i64 kk = 0; kk < k; kk++)
    {
      const i64 ia = i * k + kk, ib = kk * n + j;
      const i64 wa = Ase[ia], wb = Bse[ib];
      if ((wa >> 1) == MPGFX_ZERO_EXP || (wb >> 1) == MPGFX_ZERO_EXP) continue;
      any = true;
      u32 MA[KWD], MB[KWD], P[2 * KWD];
      load_sig (Al, m * k, ia, MA);
      load_sig (Bl, k * n, ib, MB);
      cu_fp::core::mul_words<KWD> (P, MA, MB);
      const i64 off = (wa >> 1) + (wb >> 1) - base;      /* >= 0 */
      const int ws = (int) (off >> 5);
      const u32 bs = (u32) (off & 31);
      const bool sub = ((wa ^ wb) & 1) != 0;
      u32 c = sub ? 1u : 0u;                             /* two's complement add of +-P<<off */
      const u32 xm = sub ? 0xffffffffu : 0u;
      int s = ws;
#pragma unroll
      for (int q = 0; q <= 2 * KWD; q++, s++)
        {
          u32 hi = (q < 2 * KWD) ? P[q] : 0u, lo = (q >= 1) ? P[q - 1] : 0u;
          u32 word = cu_fp::core::fshl (hi, lo, bs) ^ xm;
          if (s < nw) acc[s] = (i64) cu_fp::core::addc ((u32) acc[s], word, c);
        }
      /* words above the shifted product: 0 (add) or all ones (subtract) */
      for (; s < nw; s++)
        {
          if (!sub && !c) break;
          acc[s] = (i64) cu_fp::core::addc ((u32) acc[s], xm, c);
        }
      /* below the product a subtraction adds ~0 words with the +1 carry:
       * ~0 + 1 = 0 with carry 1 -- already folded into the carry-in above */
    }
  if (!any)
    {
      cu_freal<PB> z;
      z.set_zero ();
      st<PB> (z, Cl, Cse, m * n, i * n + j);
      return;
    }
  round_store_int (acc, nw, base - 2 * (i64) PB, Cl, Cse, m * n, i * n + j);
}

/* VALU fallback: one lane per output, rows r0 .. r0+rows-1 of C, lane e's
 * accumulator at wsp + e * nw. */
__global__ __launch_bounds__ (64) void k_exact_fallback (i64 m, i64 n, i64 k, i64 r0, i64 rows,
                                                         const u64 *__restrict__ Al, const i64 *__restrict__ Ase,
                                                         const u64 *__restrict__ Bl, const i64 *__restrict__ Bse,
                                                         const i64 *__restrict__ eminA, const i64 *__restrict__ eminB,
                                                         int nw, i64 *__restrict__ wsp, u64 *__restrict__ Cl,
                                                         i64 *__restrict__ Cse)
{
  const i64 e = (i64) blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * n) return;
  exact_fb_output (m, n, k, r0 + e / n, e % n, Al, Ase, Bl, Bse, eminA, eminB, nw, wsp + e * nw, Cl, Cse);
}

/* The same on a device plan: on the wide path the plan's `lanes` lanes (at
 * most the grid's) each own an accumulator at wsp + lane * nw and take every
 * lanes-th output; a refused plan (status set) writes +0 everywhere.  On the
 * MFMA path with a cut (the exact fix-up) the same lanes recompute, exactly,
 * the outputs the rounding stage marked with EXACT_FIX_MARK; without a cut
 * the kernel returns at once. */
__global__ __launch_bounds__ (64) void k_exact_fallback_dev (i64 m, i64 n, i64 k, const u64 *__restrict__ Al,
                                                             const i64 *__restrict__ Ase, const u64 *__restrict__ Bl,
                                                             const i64 *__restrict__ Bse,
                                                             const i64 *__restrict__ eminA,
                                                             const i64 *__restrict__ eminB, i64 *__restrict__ wsp,
                                                             u64 *__restrict__ Cl, i64 *__restrict__ Cse,
                                                             const exact_dplan *__restrict__ pl)
{
  if (!pl->fallback && pl->cut == 0) return;
  const i64 tid = (i64) blockIdx.x * blockDim.x + threadIdx.x, nt = (i64) gridDim.x * blockDim.x;
  if (!pl->fallback)
    {
      const i64 lanes = pl->lanes < nt ? pl->lanes : nt;
      if (tid >= lanes) return;
      const int nw = (int) pl->nw;
      for (i64 e = tid; e < m * n; e += lanes)
        if (Cse[e] == EXACT_FIX_MARK)
          exact_fb_output (m, n, k, e / n, e % n, Al, Ase, Bl, Bse, eminA, eminB, nw, wsp + tid * nw, Cl, Cse);
      return;
    }
  if (pl->status != MPGFX_OK)
    {
      cu_freal<PB> z;
      z.set_zero ();
      for (i64 e = tid; e < m * n; e += nt) st<PB> (z, Cl, Cse, m * n, e);
      return;
    }
  const i64 lanes = pl->lanes < nt ? pl->lanes : nt;
  if (tid >= lanes) return;
  const int nw = (int) pl->nw;
  for (i64 e = tid; e < m * n; e += lanes)
    exact_fb_output (m, n, k, e / n, e % n, Al, Ase, Bl, Bse, eminA, eminB, nw, wsp + tid * nw, Cl, Cse);
}

int exact_digits_launch (i64 lines, i64 len, i64 ls, i64 ps, const u64 *l, const i64 *se, i64 nv, const i64 *emax,
                         i64 spread, int S, int Spad, i64 Lpad, i64 kp, int8_t *D, const exact_dplan *pl, int which,
                         hipStream_t s)
{
  if (ps == 1 && ls == len)
    {
      /* lines are rows (the A operand) */
      typedef exd_tile<false> T;
      dim3 grid ((unsigned) ((Lpad + T::LT - 1) / T::LT), (unsigned) ((kp + T::PT - 1) / T::PT));
      hipLaunchKernelGGL ((k_exact_digits<false>), grid, dim3 (1024), 0, s, lines, len, l, se, nv, emax, spread, S,
                          Spad, Lpad, kp, D, pl, which);
      return launch_status ("gemm_exact digits");
    }
  if (ls == 1 && ps == lines)
    {
      /* lines are columns (the B operand) */
      typedef exd_tile<true> T;
      dim3 grid ((unsigned) ((Lpad + T::LT - 1) / T::LT), (unsigned) ((kp + T::PT - 1) / T::PT));
      hipLaunchKernelGGL ((k_exact_digits<true>), grid, dim3 (1024), 0, s, lines, len, l, se, nv, emax, spread, S,
                          Spad, Lpad, kp, D, pl, which);
      return launch_status ("gemm_exact digits");
    }
  set_error_text ("gemm_exact digits: operand layout neither row- nor column-major");
  return MPGFX_EINVAL;
}

int exact_round_launch (i64 m, i64 n, i64 np, i64 nE, int slots, const i64 *S64, const i64 *emaxA, const i64 *emaxB,
                        i64 sA, i64 sB, u64 *Cl, i64 *Cse, const exact_dplan *pl, hipStream_t s)
{
  const i64 tot = m * n;
  hipLaunchKernelGGL (k_exact_round, dim3 ((unsigned) ((tot + 255) / 256)), dim3 (256), 0, s, m, n, np, nE, slots, S64,
                      emaxA, emaxB, sA, sB, Cl, Cse, pl);
  return launch_status ("gemm_exact round");
}

int exact_fallback_dev_launch (i64 m, i64 n, i64 k, const u64 *Al, const i64 *Ase, const u64 *Bl, const i64 *Bse,
                               const i64 *eminA, const i64 *eminB, i64 *wsp, u64 *Cl, i64 *Cse, const exact_dplan *pl,
                               hipStream_t s)
{
  const i64 tot = m * n, nb = (tot + 63) / 64;
  hipLaunchKernelGGL (k_exact_fallback_dev, dim3 ((unsigned) (nb < 4096 ? nb : 4096)), dim3 (64), 0, s, m, n, k, Al,
                      Ase, Bl, Bse, eminA, eminB, wsp, Cl, Cse, pl);
  return launch_status ("gemm_exact fallback");
}

int exact_fallback_launch (i64 m, i64 n, i64 k, i64 r0, i64 rows, const u64 *Al, const i64 *Ase, const u64 *Bl,
                           const i64 *Bse, const i64 *eminA, const i64 *eminB, int nw, i64 *wsp, u64 *Cl, i64 *Cse,
                           hipStream_t s)
{
  i64 tot = rows * n;
  hipLaunchKernelGGL (k_exact_fallback, dim3 ((unsigned) ((tot + 63) / 64)), dim3 (64), 0, s, m, n, k, r0, rows, Al,
                      Ase, Bl, Bse, eminA, eminB, nw, wsp, Cl, Cse);
  return launch_status ("gemm_exact fallback");
}
