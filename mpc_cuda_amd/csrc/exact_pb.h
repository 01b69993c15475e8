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

/* Digit planes of one operand.  Line L (row of A / column of B), position p
 * (k index); the value sits at index L*ls + p*ps of the MP array (nv values).
 * D[t][L][p] for t < Spad, L < Lpad, p < kp; padding and planes >= S are 0. */
__global__ __launch_bounds__ (256) void k_exact_digits (i64 lines, i64 len, i64 ls, i64 ps, const u64 *__restrict__ l,
                                                        const i64 *__restrict__ se, i64 nv,
                                                        const i64 *__restrict__ emax_line, i64 spread, int S,
                                                        int Spad, i64 Lpad, i64 kp, int8_t *__restrict__ D)
{
  const i64 p = (i64) blockIdx.x * blockDim.x + threadIdx.x;
  const i64 L = blockIdx.y;
  if (p >= kp || L >= Lpad) return;
  const i64 plane = Lpad * kp;
  int8_t *out = D + L * kp + p;
  bool live = (L < lines) && (p < len);
  i64 idx = live ? L * ls + p * ps : 0;
  i64 w = live ? se[idx] : 0;
  i64 e = w >> 1;
  i64 em = live ? emax_line[L] : EXACT_EMPTY;
  live = live && (e != MPGFX_ZERO_EXP) && (em != EXACT_EMPTY);
  if (!live)
    {
      for (int t = 0; t < Spad; t++) out[(i64) t * plane] = 0;
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
  for (int t = 0; t < q && t < Spad; t++) out[(i64) t * plane] = 0;
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
      if (t < S) out[(i64) t * plane] = (int8_t) d;
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
      out[(i64) t * plane] = (int8_t) d;
    }
  for (int t = (S > q ? S : q); t < Spad; t++) out[(i64) t * plane] = 0;
}

/* Round the two's complement integer held in the first `nw` 32-bit words of
 * wbuf (stored one per 64-bit slot) times 2^scale to pb bits and store it as
 * value i of C.  The words are read at runtime positions from memory, so no
 * register array is indexed dynamically. */
template <typename W>
__device__ void round_store_int (W *__restrict__ wbuf, int nw, i64 scale, u64 *__restrict__ Cl,
                                 i64 *__restrict__ Cse, i64 nC, i64 i)
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

/* The MFMA kernel.  A workgroup of 4 wavefronts computes a 32x32 output tile
 * (wave w: rows 16*(w>>1).., columns 16*(w&1)..); lane l of a wave owns the
 * outputs (row 4*(l>>4)+v, column l&15), v < 4 (the 16x16 MFMA accumulator
 * layout), with its carry-save slots in `slotbuf`.
 *
 * A pass pairs digit group ga of A with digit group gb of B (G digits each).
 * Per k step of 64 the workgroup stages the G planes of both groups for its 32
 * rows and 32 columns in LDS (64 KB, each byte fetched from memory once per
 * workgroup and read by two waves); the next step's planes are loaded into
 * registers while the MFMAs of the current one run.  LDS rows are padded to
 * 80 bytes so the 16-byte fragment reads of 8 consecutive rows hit distinct
 * banks.  A wave reads the G fragments of B into VGPRs and streams the G
 * fragments of A: acc[t+u] += A_t B_u on v_mfma_i32_16x16x64_i8 (2G-1
 * diagonal int32 accumulators, exact for G * EXACT_KCHUNK * 2^14 < 2^31).
 * Digits of a partial last group are zero; their loads and products are
 * skipped.  Each pass's diagonals start at bit 8*G*(ga+gb), a multiple of 32,
 * and are folded in registers into PS slot-aligned 64-bit contributions.
 *
 * Tiles are assigned XCD-aware: workgroup b runs on XCD b % 8, and the 32
 * workgroups of one XCD take 32 consecutive tiles (one row of tiles at 1024
 * columns), so the A rows they share stay in that XCD's L2. */
template <int G>
__global__ __launch_bounds__ (256, 1) void k_exact_lds (i64 m, i64 n, i64 mp, i64 np, i64 kp, int nG, int S, int slots,
                                                      const int8_t *__restrict__ DA, const int8_t *__restrict__ DB,
                                                      const i64 *__restrict__ emaxA, const i64 *__restrict__ emaxB,
                                                      i64 sA, i64 sB, i64 *__restrict__ slotbuf,
                                                      u64 *__restrict__ Cl, i64 *__restrict__ Cse)
{
  static_assert (G % 2 == 0, "staging assignment splits the planes over two thread halves");
  static_assert ((8 * G) % 32 == 0, "pass offsets must be slot aligned");
  constexpr int ND = 2 * G - 1;                    /* diagonals per pass */
  constexpr int PS = (8 * (ND - 1) + 31) / 32 + 1;  /* slots a pass touches */
  constexpr int KCH = EXACT_KCHUNK;
  static_assert ((long long) G * KCH * 16384 < (1ll << 31), "int32 accumulators could overflow");
  constexpr int RS = 80;                           /* LDS row stride (bytes) */
  constexpr int PL = 32 * RS;                      /* LDS plane: 32 rows */
  constexpr int NST = G;                           /* 16-byte chunks staged per thread per step */
  constexpr int HG = G / 2;
  __shared__ __attribute__ ((aligned (16))) int8_t lds[2 * G * PL];
  __shared__ u32 words[256 * EXACT_SLMAX];         /* epilogue: one lane's integer */

  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  const int r0 = l & 15, g = l >> 4;
  /* staging: chunk x of this thread is plane 2*(x % (G/2)) + tid/128 of A
   * (x < G/2) or of B (x >= G/2), row (tid/4) & 31, 16-byte column tid & 3 */
  const int srow = (tid >> 2) & 31, scol = tid & 3, shalf = tid >> 7;
  const i64 tn = np / 32, tiles = (mp / 32) * tn;
  const i64 pla = mp * kp, plb = np * kp;
  const i64 nb = gridDim.x;
  const i64 per_xcd = nb / 8, xslot = (i64) (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  i64 *my = slotbuf + ((i64) blockIdx.x * 256 + tid) * 4 * slots;
  for (i64 it = 0; it * nb < tiles; it++)
    {
      const i64 tile = it * nb + xslot;
      if (tile >= tiles) break;
      const i64 ti = tile / tn, tj = tile % tn;
      for (int q = 0; q < 4 * slots; q++) my[q] = 0;
      for (int ga = 0; ga < nG; ga++)
        for (int gb = 0; gb < nG; gb++)
          {
            const int tv = (S - ga * G < G) ? S - ga * G : G, uv = (S - gb * G < G) ? S - gb * G : G;
            const bool full = (tv == G) && (uv == G);
            const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc (
              (void *) (DA + (i64) (ga * G) * pla), (short) 0, (int) (G * pla), 0x00020000);
            const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc (
              (void *) (DB + (i64) (gb * G) * plb), (short) 0, (int) (G * plb), 0x00020000);
            const int goa = (int) ((ti * 32 + srow) * kp + 16 * scol);
            const int gob = (int) ((tj * 32 + srow) * kp + 16 * scol);
            for (i64 kc = 0; kc < kp; kc += KCH)
              {
                const i64 kend = (kc + KCH < kp) ? kc + KCH : kp;
                v4i stg[NST];
                auto fetch = [&] (i64 k0) {
#pragma unroll
                  for (int x = 0; x < NST; x++)
                    {
                      const int t = 2 * (x % HG) + shalf;
                      const bool isA = x < HG;
                      const bool need = isA ? (t < tv) : (t < uv);
                      stg[x] = need ? (isA ? __builtin_amdgcn_raw_buffer_load_b128 (ra, goa + (int) k0, (int) (t * pla), 0)
                                           : __builtin_amdgcn_raw_buffer_load_b128 (rb, gob + (int) k0, (int) (t * plb), 0))
                                    : (v4i) {0, 0, 0, 0};
                    }
                };
                auto stash = [&] () {
#pragma unroll
                  for (int x = 0; x < NST; x++)
                    {
                      const int t = 2 * (x % HG) + shalf;
                      *(v4i *) (lds + (x < HG ? 0 : G * PL) + t * PL + srow * RS + 16 * scol) = stg[x];
                    }
                };
                v4i acc[ND];
#pragma unroll
                for (int d = 0; d < ND; d++) acc[d] = (v4i) {0, 0, 0, 0};
                fetch (kc);
                stash ();
                __syncthreads ();
                for (i64 k0 = kc; k0 < kend; k0 += 64)
                  {
                    const bool more = k0 + 64 < kend;
                    if (more) fetch (k0 + 64);
                    v4i fb[G];
#pragma unroll
                    for (int u = 0; u < G; u++)
                      fb[u] = *(const v4i *) (lds + G * PL + u * PL + (wc * 16 + r0) * RS + 16 * g);
                    const int8_t *pa = lds + (wr * 16 + r0) * RS + 16 * g;
                    v4i fa = *(const v4i *) pa;
#pragma unroll
                    for (int t = 0; t < G; t++)
                      {
                        if (t < tv)
                          {
                            const v4i fn = (t + 1 < tv) ? *(const v4i *) (pa + (t + 1) * PL) : fa;
#pragma unroll
                            for (int u = 0; u < G; u++)
                              if (full || u < uv)
                                acc[t + u] = __builtin_amdgcn_mfma_i32_16x16x64_i8 (fa, fb[u], acc[t + u], 0, 0, 0);
                            fa = fn;
                          }
                      }
                    __syncthreads ();
                    if (more)
                      {
                        stash ();
                        __syncthreads ();
                      }
                  }
                /* fold the diagonals into PS slot-aligned contributions per element */
                const int s0 = (8 * G * (ga + gb)) >> 5;
#pragma unroll
                for (int v = 0; v < 4; v++)
                  {
                    i64 c[PS];
#pragma unroll
                    for (int q = 0; q < PS; q++) c[q] = 0;
#pragma unroll
                    for (int d = 0; d < ND; d++) c[d >> 2] += (i64) acc[d][v] << (8 * (d & 3));
                    i64 *sl = my + v * slots + s0;
                    i64 old[PS];
#pragma unroll
                    for (int q = 0; q < PS; q++) old[q] = sl[q];
#pragma unroll
                    for (int q = 0; q < PS; q++) sl[q] = old[q] + c[q];
                  }
              }
          }
      /* carry-normalise each element's slots into 32-bit words (all slot
       * loads in flight at once), stage them in LDS and round from there */
#pragma unroll 1
      for (int v = 0; v < 4; v++)
        {
          const i64 i = ti * 32 + wr * 16 + 4 * g + v, j = tj * 32 + wc * 16 + r0;
          const i64 *sl = my + v * slots;
          i64 x[EXACT_SLMAX];
#pragma unroll
          for (int q = 0; q < EXACT_SLMAX; q++) x[q] = (q < slots) ? sl[q] : 0;
          u32 *wb = words + tid * EXACT_SLMAX;
          i64 carry = 0;
#pragma unroll
          for (int q = 0; q < EXACT_SLMAX; q++)
            {
              const i64 y = x[q] + carry;
              if (q < slots) wb[q] = (u32) y;
              carry = y >> 32;
            }
          if (i < m && j < n)
            {
              const i64 ea = emaxA[i], eb = emaxB[j];
              const i64 scale = (ea == EXACT_EMPTY || eb == EXACT_EMPTY) ? 0 : ea + eb - sA - sB - 2 * (i64) PB;
              round_store_int (wb, slots, scale, Cl, Cse, m * n, i * n + j);
            }
        }
    }
}

/* VALU fallback: one lane per output, the exact sum accumulated in `nw`
 * two's complement 32-bit words (one per 64-bit workspace slot) whose LSB
 * weighs 2^(eminA_i + eminB_j - 2pb).  Rows r0 .. r0+rows-1 of C. */
__global__ __launch_bounds__ (64) void k_exact_fallback (i64 m, i64 n, i64 k, i64 r0, i64 rows,
                                                         const u64 *__restrict__ Al, const i64 *__restrict__ Ase,
                                                         const u64 *__restrict__ Bl, const i64 *__restrict__ Bse,
                                                         const i64 *__restrict__ eminA, const i64 *__restrict__ eminB,
                                                         int nw, i64 *__restrict__ wsp, u64 *__restrict__ Cl,
                                                         i64 *__restrict__ Cse)
{
  const i64 e = (i64) blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * n) return;
  const i64 i = r0 + e / n, j = e % n;
  i64 *acc = wsp + e * nw;
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

int exact_digits_launch (i64 lines, i64 len, i64 ls, i64 ps, const u64 *l, const i64 *se, i64 nv, const i64 *emax,
                         i64 spread, int S, int Spad, i64 Lpad, i64 kp, int8_t *D, hipStream_t s)
{
  dim3 grid ((unsigned) ((kp + 255) / 256), (unsigned) Lpad);
  hipLaunchKernelGGL (k_exact_digits, grid, dim3 (256), 0, s, lines, len, ls, ps, l, se, nv, emax, spread, S, Spad,
                      Lpad, kp, D);
  return launch_status ("gemm_exact digits");
}

int exact_mfma_launch (int G, const exact_plan &p, const int8_t *DA, const int8_t *DB, const i64 *emaxA,
                       const i64 *emaxB, i64 *slotbuf, int grid, u64 *Cl, i64 *Cse, hipStream_t s)
{
#define EXACT_G(GG)                                                                                        \
  if (G == GG)                                                                                             \
    {                                                                                                      \
      hipLaunchKernelGGL ((k_exact_lds<GG>), dim3 (grid), dim3 (256), 0, s, p.m, p.n, p.mp, p.np, p.kp,    \
                          p.nG, p.S, p.slots, DA, DB, emaxA, emaxB, p.sA, p.sB, slotbuf, Cl, Cse);         \
      return launch_status ("gemm_exact mfma");                                                            \
    }
  EXACT_G (16)
  EXACT_G (20)
#undef EXACT_G
  return MPGFX_EINVAL;
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
