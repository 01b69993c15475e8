/* exact_gemm.hip -- host orchestration and precision-independent kernels of
 * the exact-dot GEMM (scheme in exact.h; per-precision kernels in
 * exact_pb.h).  Two modes:
 *  - asynchronous (a caller workspace of at least exact_workspace_bound()
 *    bytes): the digit count, group size and path are decided on the device
 *    from the exponent scan (k_exact_plan) and every later stage is launched
 *    for the worst case, reading that plan; no synchronisation and no
 *    allocation, so the call can be captured in a graph;
 *  - synchronous (smaller or no workspace): the scan's spreads are copied to
 *    the host, which sizes the workspace for this call (allocating it
 *    stream-ordered, or refusing with noalloc) and launches exactly the
 *    stages it needs.
 * Also here: the static scratch of the no-allocation entry point
 * (mpgfx_gemm_exact_noalloc). */
#include "common.h"
#include "dispatch.h"
#include "exact.h"

#include <algorithm>
#include <mutex>

namespace mpgfx {

/* The exponent scan, in three launches that any number of workgroups can
 * share: k_exact_init sets every line's largest exponent to EXACT_EMPTY and
 * smallest to INT64_MAX; k_exact_scan folds 256-position pieces of the rows
 * of A and 64 x 64 tiles of B (64 consecutive columns per read) into them with
 * atomicMax / atomicMin; exact_scan_finish (one workgroup) marks the lines
 * without a nonzero value (both EXACT_EMPTY) and writes each operand's
 * largest spread, spreads[0] for A, spreads[1] for B. */
__global__ __launch_bounds__ (256) void k_exact_init (i64 m, i64 n, i64 *__restrict__ emaxA, i64 *__restrict__ eminA,
                                                      i64 *__restrict__ emaxB, i64 *__restrict__ eminB)
{
  const i64 t = (i64) blockIdx.x * 256 + threadIdx.x;
  if (t < m) { emaxA[t] = EXACT_EMPTY; eminA[t] = INT64_MAX; }
  if (t < n) { emaxB[t] = EXACT_EMPTY; eminB[t] = INT64_MAX; }
}

static unsigned exact_init_grid (i64 m, i64 n) { return (unsigned) (((m > n ? m : n) + 255) / 256); }

__global__ __launch_bounds__ (256) void k_exact_scan (i64 m, i64 n, i64 k, const i64 *__restrict__ Ase,
                                                      const i64 *__restrict__ Bse, i64 *__restrict__ emaxA,
                                                      i64 *__restrict__ eminA, i64 *__restrict__ emaxB,
                                                      i64 *__restrict__ eminB)
{
  __shared__ i64 smax[256], smin[256];
  const i64 kA = (k + 255) / 256, nbA = m * kA;
  i64 hi = INT64_MIN, lo = INT64_MAX;
  auto take = [&] (i64 w) {
    const i64 e = w >> 1;
    if (e != MPGFX_ZERO_EXP)
      {
        hi = e > hi ? e : hi;
        lo = e < lo ? e : lo;
      }
  };
  auto fold = [&] (i64 *emax, i64 *emin, i64 L) {
    if (hi == INT64_MIN) return;
    atomicMax ((long long *) (emax + L), (long long) hi);
    atomicMin ((long long *) (emin + L), (long long) lo);
  };
  if ((i64) blockIdx.x < nbA)
    {
      /* positions [c*256, c*256 + 256) of row L */
      const i64 L = blockIdx.x / kA, p = (blockIdx.x % kA) * 256 + threadIdx.x;
      if (p < k) take (Ase[L * k + p]);
      for (int o = 32; o >= 1; o >>= 1)
        {
          const i64 h2 = __shfl_xor (hi, o), l2 = __shfl_xor (lo, o);
          hi = h2 > hi ? h2 : hi;
          lo = l2 < lo ? l2 : lo;
        }
      if ((threadIdx.x & 63) == 0)
        {
          smax[threadIdx.x >> 6] = hi;
          smin[threadIdx.x >> 6] = lo;
        }
      __syncthreads ();
      if (threadIdx.x == 0)
        {
          for (int q = 1; q < 4; q++)
            {
              hi = smax[q] > hi ? smax[q] : hi;
              lo = smin[q] < lo ? smin[q] : lo;
            }
          fold (emaxA, eminA, L);
        }
    }
  else
    {
      /* columns [g*64, g*64 + 64) over positions [c*64, c*64 + 64): lane
       * (jj, ph) takes every fourth position */
      const i64 kB = (k + 63) / 64, b = (i64) blockIdx.x - nbA, g = b / kB, c = b % kB;
      const i64 L = g * 64 + (threadIdx.x & 63);
      const int ph = threadIdx.x >> 6;
      if (L < n)
        for (int t = 0; t < 16; t++)
          {
            const i64 p = c * 64 + ph + 4 * t;
            if (p < k) take (Bse[p * n + L]);
          }
      smax[threadIdx.x] = hi;
      smin[threadIdx.x] = lo;
      __syncthreads ();
      if (threadIdx.x < 64 && L < n)
        {
          for (int q = 1; q < 4; q++)
            {
              const i64 h2 = smax[threadIdx.x + 64 * q], l2 = smin[threadIdx.x + 64 * q];
              hi = h2 > hi ? h2 : hi;
              lo = l2 < lo ? l2 : lo;
            }
          fold (emaxB, eminB, L);
        }
    }
}

static unsigned exact_scan_grid (i64 m, i64 n, i64 k)
{
  return (unsigned) (m * ((k + 255) / 256) + ((n + 63) / 64) * ((k + 63) / 64));
}

/* one workgroup of 256 lanes; `red` is 512 words of LDS */
__device__ void exact_scan_finish (i64 m, i64 n, i64 *__restrict__ emaxA, i64 *__restrict__ eminA,
                                   i64 *__restrict__ emaxB, i64 *__restrict__ eminB,
                                   unsigned long long *__restrict__ spreads, unsigned long long *red)
{
  unsigned long long sa = 0, sb = 0;
  for (i64 L = threadIdx.x; L < m; L += 256)
    {
      const i64 hi = emaxA[L];
      if (hi == EXACT_EMPTY) eminA[L] = EXACT_EMPTY;
      else sa = (unsigned long long) (hi - eminA[L]) > sa ? (unsigned long long) (hi - eminA[L]) : sa;
    }
  for (i64 L = threadIdx.x; L < n; L += 256)
    {
      const i64 hi = emaxB[L];
      if (hi == EXACT_EMPTY) eminB[L] = EXACT_EMPTY;
      else sb = (unsigned long long) (hi - eminB[L]) > sb ? (unsigned long long) (hi - eminB[L]) : sb;
    }
  red[threadIdx.x] = sa;
  red[256 + threadIdx.x] = sb;
  __syncthreads ();
  for (int st = 128; st > 0; st >>= 1)
    {
      if ((int) threadIdx.x < st)
        {
          red[threadIdx.x] = red[threadIdx.x + st] > red[threadIdx.x] ? red[threadIdx.x + st] : red[threadIdx.x];
          red[256 + threadIdx.x] = red[256 + threadIdx.x + st] > red[256 + threadIdx.x] ? red[256 + threadIdx.x + st]
                                                                                        : red[256 + threadIdx.x];
        }
      __syncthreads ();
    }
  if (threadIdx.x == 0)
    {
      spreads[0] = red[0];
      spreads[1] = red[256];
    }
}

/* the synchronous mode's last scan step (the host reads the spreads) */
__global__ __launch_bounds__ (256) void k_exact_scan_finish (i64 m, i64 n, i64 *__restrict__ emaxA,
                                                             i64 *__restrict__ eminA, i64 *__restrict__ emaxB,
                                                             i64 *__restrict__ eminB,
                                                             unsigned long long *__restrict__ spreads)
{
  __shared__ unsigned long long red[512];
  exact_scan_finish (m, n, emaxA, eminA, emaxB, eminB, spreads, red);
}


typedef int v4i __attribute__ ((ext_vector_type (4)));
typedef __attribute__ ((address_space (3))) void *lds_ptr;

/* ---------------------------------------------------------------------------
 * The MFMA kernel (precision-independent: it sees only digit planes).
 *
 * A workgroup of 4 wavefronts computes a 32x32 output tile (wave w: rows
 * 16*(w>>1).., columns 16*(w&1)..); lane l of a wave owns the outputs (row
 * 4*(l>>4)+v, column l&15), v < 4 (the 16x16 MFMA accumulator layout).
 *
 * A pass pairs digit group ga of A with group gb of B (G digits each; planes
 * past S hold zero digits, so every pass is a full G x G pass and the MFMA
 * stream has no branches).  Per k step of 64 the G planes of both groups for
 * the tile's 32 rows and 32 columns are copied into LDS with direct
 * global->LDS loads (64 lanes x 16 bytes land at consecutive LDS addresses; the
 * lane->global mapping applies an XOR swizzle, chunk c of row r at position
 * c ^ ((r>>1)&3), so the 16-byte fragment reads of 8 consecutive rows hit
 * distinct banks).  Two workgroups share a CU: one computes while the other
 * waits for its loads.  A wave reads the G fragments of B into VGPRs and
 * streams the G fragments of A: acc[t+u] += A_t B_u on v_mfma_i32_16x16x64_i8
 * (2G-1 diagonal int32 accumulators, exact for G * EXACT_KCHUNK * 2^14 < 2^31).
 * Diagonal d of a pass sits at digit G*(ga+gb)+d, i.e. slot (G*(ga+gb)+d)/4
 * shifted by 8*((G*(ga+gb)+d) mod 4) bits; each pass adds its diagonals into
 * the element's carry-save int64 slots, slot-major S64[q*nE + i*np + j] (the
 * first pass of a tile stores them and zeroes the rest).  With a cut (the
 * asynchronous mode, exact_choose_cut) the pairs on global diagonals below it
 * are left out: passes entirely below it are skipped, the others skip the rows
 * of A fragments (G MFMAs each) entirely below it.
 *
 * Tiles are assigned XCD-aware: workgroup b runs on XCD b % 8, and the
 * workgroups of one XCD take consecutive tiles (rows of tiles), so the A rows
 * they share stay in that XCD's L2. */
template <int G>
struct exact_cfg
{
  static constexpr int ND = 2 * G - 1;                 /* diagonals per pass */
  static constexpr int PS = ((3 + ND - 1) >> 2) + 1;   /* slots a pass touches (digit offset up to 3) */
  static constexpr int PLANE = 32 * 64;                /* LDS bytes of one plane of the tile */
  static constexpr int LDS = 2 * G * PLANE;            /* A planes, then B planes */
};

/* add the diagonals of one pass into the element's slots, digit offset O */
template <int G, int O>
__device__ __forceinline__ void exact_fold (const v4i (&acc)[exact_cfg<G>::ND], int v, i64 (&c)[exact_cfg<G>::PS])
{
#pragma unroll
  for (int q = 0; q < exact_cfg<G>::PS; q++) c[q] = 0;
#pragma unroll
  for (int d = 0; d < exact_cfg<G>::ND; d++) c[(O + d) >> 2] += (i64) acc[d][v] << (8 * ((O + d) & 3));
}

/* the kernel body for group size G; `lds` holds at least exact_cfg<G>::LDS
 * bytes */
template <int G>
__device__ __forceinline__ void exact_mfma_body (int8_t *lds, i64 mp, i64 np, i64 kp, int nG, int slots, int cut,
                                                 const int8_t *__restrict__ DA, const int8_t *__restrict__ DB,
                                                 i64 *__restrict__ S64)
{
  typedef exact_cfg<G> C;
  constexpr int ND = C::ND, PS = C::PS, PLANE = C::PLANE;
  constexpr int KCH = EXACT_KCHUNK;
  static_assert ((long long) G * KCH * 16384 < (1ll << 31), "int32 accumulators could overflow");
  static_assert (4 * G % 4 == 0 && G <= 20, "group size");

  const int tid = threadIdx.x, l = tid & 63, w = __builtin_amdgcn_readfirstlane (tid >> 6), wr = w >> 1, wc = w & 1;
  const int r0 = l & 15, g = l >> 4;
  const i64 tn = np / 32, tiles = (mp / 32) * tn, nE = mp * np;
  /* slots below q0 hold only diagonals below the cut: never stored (the
   * rounding stage reads them as 0; the cut's error bound covers them) */
  const int q0 = cut / 4;
  const i64 pla = mp * kp, plb = np * kp;
  const i64 nb = gridDim.x;
  const i64 per_xcd = nb / 8, xslot = (i64) (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  /* staging: the wave copies G blocks of 1 KB (16 rows x 64 bytes) per step;
   * block b = w*G + x is half b&1 of plane b/2 of A (b < 2G), or of B.  Lane l
   * lands at row 16*half + l/4, position l&3, and fetches chunk
   * (l&3) ^ ((l/4 >> 1) & 3) of that row: one per-lane offset for all blocks */
  const u32 lane_off = (u32) ((l >> 2) * kp + 16 * ((l & 3) ^ (((l >> 2) >> 1) & 3)));
  /* fragment reads: row r of the tile plane, 16-byte chunk g, swizzled */
  const int ra = wr * 16 + r0, rb = wc * 16 + r0;
  const int offa = ra * 64 + 16 * (g ^ ((ra >> 1) & 3));
  const int offb = G * PLANE + rb * 64 + 16 * (g ^ ((rb >> 1) & 3));

  for (i64 it = 0; it * nb < tiles; it++)
    {
      const i64 tile = it * nb + xslot;
      if (tile >= tiles) break;
      const i64 ti = tile / tn, tj = tile % tn;
      bool inited = false;
      for (int ga = 0; ga < nG; ga++)
        for (int gb = 0; gb < nG; gb++)
          {
            /* digit pairs with t + u < lim (local digits) are left out; a
             * pass with none left is skipped, and the first pass run stores
             * every slot of the element */
            const int lim = cut - G * (ga + gb);
            if (lim > 2 * G - 2) continue;
            const bool first = !inited;
            inited = true;
            /* buffer resources over this pass's digit groups (G planes each);
             * 32-bit offsets: G * max(mp, np) * kp < 2^31 is checked on the host */
            const __amdgpu_buffer_rsrc_t rsa = __builtin_amdgcn_make_buffer_rsrc (
              (void *) (DA + (i64) (ga * G) * pla), (short) 0, (int) (G * pla), 0x00020000);
            const __amdgpu_buffer_rsrc_t rsb = __builtin_amdgcn_make_buffer_rsrc (
              (void *) (DB + (i64) (gb * G) * plb), (short) 0, (int) (G * plb), 0x00020000);
            for (i64 kc = 0; kc < kp; kc += KCH)
              {
                const i64 kend = (kc + KCH < kp) ? kc + KCH : kp;
                v4i acc[ND];
#pragma unroll
                for (int d = 0; d < ND; d++) acc[d] = (v4i) {0, 0, 0, 0};
                for (i64 k0 = kc; k0 < kend; k0 += 64)
                  {
                    /* stage the planes of this step */
#pragma unroll
                    for (int x = 0; x < G; x++)
                      {
                        const int b = w * G + x;                  /* wave-uniform */
                        const bool isB = b >= 2 * G;
                        const int pb = isB ? b - 2 * G : b, t = pb >> 1, half = pb & 1;
                        const u32 so = (u32) (isB ? t * plb + (tj * 32 + 16 * half) * kp + k0
                                                  : t * pla + (ti * 32 + 16 * half) * kp + k0);
                        __builtin_amdgcn_raw_ptr_buffer_load_lds (
                          isB ? rsb : rsa, (lds_ptr) (lds + (isB ? G * PLANE : 0) + t * PLANE + half * 1024), 16,
                          lane_off, so, 0, 0);
                      }
                    __builtin_amdgcn_s_waitcnt (0);
                    __syncthreads ();
                    v4i fb[G];
#pragma unroll
                    for (int u = 0; u < G; u++) fb[u] = *(const v4i *) (lds + offb + u * PLANE);
                    v4i fa = *(const v4i *) (lds + offa);
#pragma unroll
                    for (int t = 0; t < G; t++)
                      {
                        const v4i fn = (t + 1 < G) ? *(const v4i *) (lds + offa + (t + 1) * PLANE) : fa;
                        /* a row of A fragments whose diagonals all lie
                         * below the limit is skipped (one wave-uniform
                         * branch per row; the pairs of the other rows below
                         * it are computed, which only makes the sum closer) */
                        if (t + G - 1 >= lim)
                          {
#pragma unroll
                            for (int u = 0; u < G; u++)
                              acc[t + u] = __builtin_amdgcn_mfma_i32_16x16x64_i8 (fa, fb[u], acc[t + u], 0, 0, 0);
                          }
                        fa = fn;
                      }
                    __syncthreads ();
                  }
                /* add the diagonals into the slots of the 4 elements; the
                 * opaque copy of nE keeps the slot addresses from being
                 * precomputed (and held in registers) outside the k loop */
                const int base = G * (ga + gb);
                const int s0 = base >> 2, o = base & 3;
                i64 nEo = nE;
                asm volatile ("" : "+s"(nEo));
#pragma unroll
                for (int v = 0; v < 4; v++)
                  {
                    const i64 el = (ti * 32 + wr * 16 + 4 * g + v) * np + tj * 32 + wc * 16 + r0;
                    i64 cs[PS];
                    if (G % 4 == 0 || o == 0) exact_fold<G, 0> (acc, v, cs);
                    else if (o == 1) exact_fold<G, (G % 4 == 0) ? 0 : 1> (acc, v, cs);
                    else if (o == 2) exact_fold<G, (G % 4 == 0) ? 0 : 2> (acc, v, cs);
                    else exact_fold<G, (G % 4 == 0) ? 0 : 3> (acc, v, cs);
                    i64 *sl = S64 + (i64) s0 * nEo + el;
                    if (first && kc == 0)
                      {
#pragma unroll
                        for (int q = 0; q < PS; q++)
                          if (s0 + q >= q0 && s0 + q < slots) sl[(i64) q * nEo] = cs[q];
                        for (int q = q0; q < s0; q++) S64[(i64) q * nEo + el] = 0;
                        for (int q = s0 + PS; q < slots; q++) S64[(i64) q * nEo + el] = 0;
                      }
                    else
                      {
                        i64 old[PS];
#pragma unroll
                        for (int q = 0; q < PS; q++)
                          old[q] = (s0 + q >= q0 && s0 + q < slots) ? sl[(i64) q * nEo] : 0;
#pragma unroll
                        for (int q = 0; q < PS; q++)
                          if (s0 + q >= q0 && s0 + q < slots) sl[(i64) q * nEo] = old[q] + cs[q];
                      }
                  }
              }
          }
    }
}

/* the kernel for one group size (the synchronous mode) */
template <int G>
__global__ __launch_bounds__ (256, 2) void k_exact_mfma (i64 mp, i64 np, i64 kp, int nG, int slots,
                                                       const int8_t *__restrict__ DA, const int8_t *__restrict__ DB,
                                                       i64 *__restrict__ S64)
{
  __shared__ __attribute__ ((aligned (16))) int8_t lds[exact_cfg<G>::LDS];
  exact_mfma_body<G> (lds, mp, np, kp, nG, slots, 0, DA, DB, S64);
}

/* the kernel of the asynchronous mode: group size, group count and slots
 * from the device plan (a wave-uniform switch), nothing on the wide path */
__global__ __launch_bounds__ (256, 2) void k_exact_mfma_plan (i64 mp, i64 np, i64 kp, const int8_t *__restrict__ DA,
                                                            const int8_t *__restrict__ DB, i64 *__restrict__ S64,
                                                            const exact_dplan *__restrict__ pl)
{
  __shared__ __attribute__ ((aligned (16))) int8_t lds[exact_cfg<17>::LDS];
  if (pl->fallback) return;
  const int G = pl->G, nG = pl->nG, slots = pl->slots, cut = pl->cut;
  switch (G)
    {
    case 9: exact_mfma_body<9> (lds, mp, np, kp, nG, slots, cut, DA, DB, S64); break;
    case 11: exact_mfma_body<11> (lds, mp, np, kp, nG, slots, cut, DA, DB, S64); break;
    case 13: exact_mfma_body<13> (lds, mp, np, kp, nG, slots, cut, DA, DB, S64); break;
    case 15: exact_mfma_body<15> (lds, mp, np, kp, nG, slots, cut, DA, DB, S64); break;
    default: exact_mfma_body<17> (lds, mp, np, kp, nG, slots, cut, DA, DB, S64); break;
    }
}

int exact_mfma_run (const exact_plan &p, const int8_t *DA, const int8_t *DB, i64 *S64, int grid, hipStream_t s)
{
#define EXACT_G(GG)                                                                                         \
  if (p.G == GG)                                                                                            \
    {                                                                                                       \
      hipLaunchKernelGGL ((k_exact_mfma<GG>), dim3 (grid), dim3 (256), 0, s, p.mp, p.np, p.kp, p.nG, p.slots, \
                          DA, DB, S64);                                                                     \
      return launch_status ("gemm_exact mfma");                                                             \
    }
  EXACT_G (9)
  EXACT_G (11)
  EXACT_G (13)
  EXACT_G (15)
  EXACT_G (17)
#undef EXACT_G
  return MPGFX_EINVAL;
}

/* The last scan step, then the plan of the asynchronous mode (one lane): the
 * digit count from the spreads, the group size and count, the slots, or the
 * wide-accumulator path with its words per output and how many outputs the
 * workspace after the scan (`room` bytes) can accumulate at once; refused
 * (status) when the span exceeds EXACT_SPAN_MAX or not one accumulator fits. */
__global__ __launch_bounds__ (256) void k_exact_plan (int pb, i64 m, i64 n, i64 kp, i64 *__restrict__ emaxA,
                                                      i64 *__restrict__ eminA, i64 *__restrict__ emaxB,
                                                      i64 *__restrict__ eminB,
                                                      unsigned long long *__restrict__ spreads,
                                                      exact_dplan *__restrict__ pl, i64 room)
{
  __shared__ unsigned long long red[512];
  exact_scan_finish (m, n, emaxA, eminA, emaxB, eminB, spreads, red);
  if (threadIdx.x != 0) return;
  exact_dplan p;
  p.status = MPGFX_OK;
  p.sA = (i64) spreads[0];
  p.sB = (i64) spreads[1];
  p.nw = 0;
  p.lanes = 0;
  p.G = 17;
  p.nG = 1;
  p.slots = 0;
  p.cut = 0;
  p.eb = -1;
  const i64 smax = p.sA > p.sB ? p.sA : p.sB;
  p.S = (smax > 100000) ? EXACT_SMAX + 1 : exact_digits (pb, smax);
  bool fb = p.S > EXACT_SMAX;
  if (!fb)
    {
      exact_choose_groups (p.S, &p.G, &p.nG);
      fb = !exact_slots_hold (p.G * p.nG, kp);
    }
  /* accumulator words and lanes of the wide path, also those of the MFMA
   * path's exact fix-up (which takes the same memory once the slots are read) */
  const i64 span = exact_span_bits (pb, p.sA, p.sB);
  p.nw = (span + 31) / 32 + 1;
  p.lanes = (span > EXACT_SPAN_MAX) ? 0 : room / (p.nw * 8);
  if (!fb)
    {
      p.slots = exact_slots (p.G * p.nG);
      if (p.lanes >= 1) exact_choose_cut (pb, p.sA, p.sB, kp, p.S, &p.cut, &p.eb);
    }
  else if (p.lanes < 1)
    p.status = MPGFX_EUNSUPPORTED;
  p.fallback = fb ? 1 : 0;
  *pl = p;
}

static int hip_ok (const char *where, hipError_t e)
{
  if (e != hipSuccess) { set_error (where, e); return MPGFX_EHIP; }
  return MPGFX_OK;
}

/* workgroups of the MFMA kernel: two per CU (persistent over 32x32 tiles), a
 * multiple of 8 so the XCD-aware tile assignment covers every tile */
static i64 exact_grid (i64 tiles)
{
  const i64 g = tiles < 512 ? tiles : 512;
  return (g + 7) / 8 * 8;
}

/* bytes of workspace for the MFMA path with digit count up to EXACT_SMAX,
 * the device plan included: a workspace of this size runs the call in the
 * asynchronous mode */
size_t exact_workspace_bound (int pb, i64 m, i64 n, i64 k)
{
  (void) pb;
  const i64 mp = pad_to (m, 32), np = pad_to (n, 32), kp = pad_to (k > 0 ? k : 1, 64);
  const i64 spad = EXACT_SPAD_MAX;                       /* G*nG */
  size_t b = EXACT_DPLAN_BYTES;
  b += (size_t) (2 * m + 2 * n + 8) * 8;                 /* scan outputs + spread words */
  b = (size_t) pad_to ((i64) b, 256);
  b += (size_t) pad_to (spad * mp * kp, 256);            /* digit planes of A */
  b += (size_t) pad_to (spad * np * kp, 256);            /* digit planes of B */
  b += (size_t) mp * np * EXACT_SLMAX * 8;               /* carry-save slots */
  return b;
}

/* bytes the MFMA path needs after the scan buffer: digit planes of A and B,
 * then the carry-save slots */
static size_t exact_mfma_bytes (i64 mp, i64 np, i64 kp, int spad)
{
  size_t b = (size_t) pad_to ((i64) spad * mp * kp, 256);
  b += (size_t) pad_to ((i64) spad * np * kp, 256);
  b += (size_t) mp * np * exact_slots (spad) * 8;
  return b;
}

/* The asynchronous mode on a workspace of exact_workspace_bound() bytes: the
 * device plan at its start, then the scan outputs, the digit planes of A and
 * B for EXACT_SPAD_MAX digits, the carry-save slots; the wide-accumulator path
 * reuses everything after the scan outputs.  Every stage is launched; those
 * the plan does not take return at once. */
static int gemm_exact_async (const pb_ops *o, int pb, i64 m, i64 n, i64 k, const u64 *Al, const i64 *Ase,
                             const u64 *Bl, const i64 *Bse, u64 *Cl, i64 *Cse, void *ws, size_t ws_bytes,
                             hipStream_t s)
{
  const i64 mp = pad_to (m, 32), np = pad_to (n, 32), kp = pad_to (k, 64);
  const size_t scan_bytes = (size_t) pad_to ((i64) ((2 * m + 2 * n + 8) * 8), 256);
  char *const base = (char *) ws;
  exact_dplan *const pl = (exact_dplan *) base;
  i64 *emaxA = (i64 *) (base + EXACT_DPLAN_BYTES), *eminA = emaxA + m, *emaxB = eminA + m, *eminB = emaxB + n;
  unsigned long long *spreads = (unsigned long long *) (eminB + n);
  char *q = base + EXACT_DPLAN_BYTES + scan_bytes;
  const i64 room = (i64) (ws_bytes - EXACT_DPLAN_BYTES - scan_bytes);
  int8_t *DA = (int8_t *) q;
  int8_t *DB = DA + pad_to ((i64) EXACT_SPAD_MAX * mp * kp, 256);
  i64 *S64 = (i64 *) (DB + pad_to ((i64) EXACT_SPAD_MAX * np * kp, 256));
  int rc;
  hipLaunchKernelGGL (k_exact_init, dim3 (exact_init_grid (m, n)), dim3 (256), 0, s, m, n, emaxA, eminA, emaxB, eminB);
  hipLaunchKernelGGL (k_exact_scan, dim3 (exact_scan_grid (m, n, k)), dim3 (256), 0, s, m, n, k, Ase, Bse, emaxA,
                      eminA, emaxB, eminB);
  hipLaunchKernelGGL (k_exact_plan, dim3 (1), dim3 (256), 0, s, pb, m, n, kp, emaxA, eminA, emaxB, eminB, spreads, pl,
                      room);
  if ((rc = launch_status ("gemm_exact scan"))) return rc;
  if ((rc = o->exact_digits (m, k, k, 1, Al, Ase, m * k, emaxA, 0, 0, 0, mp, kp, DA, pl, 0, s))) return rc;
  if ((rc = o->exact_digits (n, k, 1, n, Bl, Bse, k * n, emaxB, 0, 0, 0, np, kp, DB, pl, 1, s))) return rc;
  hipLaunchKernelGGL (k_exact_mfma_plan, dim3 ((unsigned) exact_grid ((mp / 32) * (np / 32))), dim3 (256), 0, s, mp,
                      np, kp, DA, DB, S64, pl);
  if ((rc = launch_status ("gemm_exact mfma"))) return rc;
  if ((rc = o->exact_round (m, n, np, mp * np, 0, S64, emaxA, emaxB, 0, 0, Cl, Cse, pl, s))) return rc;
  return o->exact_fallback_dev (m, n, k, Al, Ase, Bl, Bse, eminA, eminB, (i64 *) q, Cl, Cse, pl, s);
}

/* whether a call with this workspace runs in the asynchronous mode: the
 * worst-case workspace, and the MFMA kernel's 32-bit plane offsets for the
 * largest group */
static bool exact_takes_async (int pb, i64 m, i64 n, i64 k, size_t ws_bytes)
{
  const i64 mp = pad_to (m, 32), np = pad_to (n, 32), kp = pad_to (k, 64);
  return ws_bytes >= exact_workspace_bound (pb, m, n, k) && (i64) 17 * (mp > np ? mp : np) * kp < ((i64) 1 << 31);
}

/* A caller workspace of exact_workspace_bound() bytes takes the asynchronous
 * mode (the MFMA kernel's 32-bit plane offsets checked for the largest
 * group).  Otherwise, and always for the no-allocation entry point (whose
 * caller has no pointer to a status word), the synchronous one: the workspace
 * is the caller's `ws` when it is large enough for this call; what it lacks
 * is allocated stream-ordered, or, with noalloc, refused (MPGFX_ENOMEM,
 * before any launch for the scan buffer, after the scan for the rest: the
 * digit count depends on the exponents). */
int gemm_exact (const pb_ops *o, int pb, i64 m, i64 n, i64 k, const u64 *Al, const i64 *Ase, const u64 *Bl,
                const i64 *Bse, u64 *Cl, i64 *Cse, void *ws, size_t ws_bytes, hipStream_t s, bool noalloc)
{
  if (k == 0)                      /* empty sums: +0 everywhere, as the per-op path */
    return o->gemm_perop (m, n, k, Al, Ase, Bl, Bse, Cl, Cse, s);
  const i64 mp = pad_to (m, 32), np = pad_to (n, 32), kp = pad_to (k, 64);
  if (!noalloc && ws && exact_takes_async (pb, m, n, k, ws_bytes))
    return gemm_exact_async (o, pb, m, n, k, Al, Ase, Bl, Bse, Cl, Cse, ws, ws_bytes, s);
  const size_t scan_bytes = (size_t) pad_to ((i64) ((2 * m + 2 * n + 8) * 8), 256);
  char *base = ws ? (char *) ws : nullptr;
  size_t avail = ws ? ws_bytes : 0;
  char *owned = nullptr, *owned2 = nullptr;
  auto refuse = [&] (const char *what, size_t need) {
    char msg[192];
    snprintf (msg, sizeof msg, "gemm_exact: the %s needs %zu bytes of workspace, %zu available", what, need, avail);
    set_error_text (msg);
    return (int) MPGFX_ENOMEM;
  };
  if (avail < scan_bytes)
    {
      if (noalloc) return refuse ("exponent scan", scan_bytes);
      if (hip_ok ("gemm_exact workspace", hipMallocAsync ((void **) &owned, scan_bytes, s))) return MPGFX_ENOMEM;
      base = owned;
      avail = scan_bytes;
    }
  i64 *emaxA = (i64 *) base, *eminA = emaxA + m, *emaxB = eminA + m, *eminB = emaxB + n;
  unsigned long long *spreads = (unsigned long long *) (eminB + n);
  int rc = MPGFX_OK;
  auto done = [&] (int r) {
    if (owned2) (void) hipFreeAsync (owned2, s);
    if (owned) (void) hipFreeAsync (owned, s);
    return r;
  };
  hipLaunchKernelGGL (k_exact_init, dim3 (exact_init_grid (m, n)), dim3 (256), 0, s, m, n, emaxA, eminA, emaxB, eminB);
  hipLaunchKernelGGL (k_exact_scan, dim3 (exact_scan_grid (m, n, k)), dim3 (256), 0, s, m, n, k, Ase, Bse, emaxA,
                      eminA, emaxB, eminB);
  hipLaunchKernelGGL (k_exact_scan_finish, dim3 (1), dim3 (256), 0, s, m, n, emaxA, eminA, emaxB, eminB, spreads);
  if ((rc = launch_status ("gemm_exact scan"))) return done (rc);
  unsigned long long hs[2] = {0, 0};
  if ((rc = hip_ok ("gemm_exact spread copy", hipMemcpyAsync (hs, spreads, sizeof hs, hipMemcpyDeviceToHost, s))))
    return done (rc);
  if ((rc = hip_ok ("gemm_exact sync", hipStreamSynchronize (s)))) return done (rc);
  exact_plan p;
  p.pb = pb; p.m = m; p.n = n; p.k = k; p.mp = mp; p.np = np; p.kp = kp;
  p.sA = (i64) hs[0]; p.sB = (i64) hs[1];
  const i64 smax = p.sA > p.sB ? p.sA : p.sB;
  p.S = (smax > 100000) ? EXACT_SMAX + 1 : exact_digits (pb, smax);
  p.fallback = p.S > EXACT_SMAX;
  if (!p.fallback)
    {
      exact_choose_groups (p.S, &p.G, &p.nG);
      p.fallback = !exact_slots_hold (p.G * p.nG, kp);
    }
  if (!p.fallback)
    {
      /* the MFMA kernel addresses one digit group with 32-bit offsets */
      if ((i64) p.G * (mp > np ? mp : np) * kp >= ((i64) 1 << 31)) return done (MPGFX_EUNSUPPORTED);
      const int spad = p.G * p.nG;
      p.slots = exact_slots (spad);
      const size_t need = exact_mfma_bytes (mp, np, kp, spad);
      char *q = base + scan_bytes;
      if (avail < scan_bytes + need)
        {
          if (noalloc) return done (refuse ("digit planes and slots", scan_bytes + need));
          if (hip_ok ("gemm_exact workspace", hipMallocAsync ((void **) &owned2, need, s))) return done (MPGFX_ENOMEM);
          q = owned2;
        }
      const int grid = (int) exact_grid ((mp / 32) * (np / 32));
      int8_t *DA = (int8_t *) q;
      q += pad_to ((i64) spad * mp * kp, 256);
      int8_t *DB = (int8_t *) q;
      q += pad_to ((i64) spad * np * kp, 256);
      i64 *S64 = (i64 *) q;
      if ((rc = o->exact_digits (m, k, k, 1, Al, Ase, m * k, emaxA, p.sA, p.S, spad, mp, kp, DA, nullptr, 0, s)))
        return done (rc);
      if ((rc = o->exact_digits (n, k, 1, n, Bl, Bse, k * n, emaxB, p.sB, p.S, spad, np, kp, DB, nullptr, 1, s)))
        return done (rc);
      if ((rc = exact_mfma_run (p, DA, DB, S64, grid, s))) return done (rc);
      rc = o->exact_round (m, n, np, mp * np, p.slots, S64, emaxA, emaxB, p.sA, p.sB, Cl, Cse, nullptr, s);
      return done (rc);
    }
  /* ---- VALU wide-accumulator fallback, in row chunks ---- */
  const i64 spanbits = exact_span_bits (pb, p.sA, p.sB);
  if (spanbits > EXACT_SPAN_MAX) return done (MPGFX_EUNSUPPORTED);
  const int nw = (int) ((spanbits + 31) / 32) + 1;
  const size_t row_bytes = (size_t) n * nw * 8;
  const size_t budget = (size_t) 256 << 20;
  i64 *wsp = nullptr;
  i64 rows;
  if (avail >= scan_bytes + row_bytes)
    {
      const size_t room = std::min (avail - scan_bytes, budget > row_bytes ? budget : row_bytes);
      rows = (i64) (room / row_bytes);
      wsp = (i64 *) (base + scan_bytes);
    }
  else
    {
      if (noalloc) return done (refuse ("wide accumulator", scan_bytes + row_bytes));
      rows = (i64) (budget / row_bytes);
      if (rows < 1) rows = 1;
      if (rows > m) rows = m;
      if ((rc = hip_ok ("gemm_exact fallback workspace", hipMallocAsync ((void **) &owned2, (size_t) rows * row_bytes, s))))
        return done (MPGFX_ENOMEM);
      wsp = (i64 *) owned2;
    }
  if (rows > m) rows = m;
  for (i64 r0 = 0; r0 < m && rc == MPGFX_OK; r0 += rows)
    {
      i64 nr = (r0 + rows <= m) ? rows : m - r0;
      rc = o->exact_fallback (m, n, k, r0, nr, Al, Ase, Bl, Bse, eminA, eminB, nw, wsp, Cl, Cse, s);
    }
  return done (rc);
}

/* ---- the entry point that allocates nothing: its workspace is a static
 * device array, one per GPU for the life of the process ---- */
static __device__ __align__ (256) unsigned char scratch[EXACT_NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[256];

/* device address of this GPU's copy of `scratch`, looked up once per device
 * (the host-side name of the array is not a device address) */
int exact_noalloc_scratch (void **out)
{
  int dev = 0;
  if (hip_ok ("gemm_exact_noalloc device", hipGetDevice (&dev))) return MPGFX_EHIP;
  if (dev < 0 || dev >= 256) return MPGFX_EINVAL;
  std::lock_guard<std::mutex> lock (g_scratch_mu);
  if (!g_scratch_ptr[dev])
    {
      void *p = nullptr;
      if (hip_ok ("gemm_exact_noalloc scratch", hipGetSymbolAddress (&p, HIP_SYMBOL (scratch)))) return MPGFX_EHIP;
      if (!p) return MPGFX_EHIP;
      g_scratch_ptr[dev] = p;
    }
  *out = g_scratch_ptr[dev];
  return MPGFX_OK;
}

/* The operands of the complex exact GEMM as real MP arrays over 2k terms, in
 * one pass: A2 = [Ar | Ai] (m rows of 2k values), Bre = [Br ; -Bi] and Bim =
 * [Bi ; Br] (2k rows of n values).  NL limb planes, then the sign/exponent
 * words; the sign flip of -Bi is se ^= 1 (zeros are skipped by the exact path
 * whatever their sign field). */
__global__ __launch_bounds__ (256) void k_exact_cplx_operands (int NL, i64 m, i64 n, i64 k,
                                                               const u64 *__restrict__ Al,
                                                               const i64 *__restrict__ Ase,
                                                               const u64 *__restrict__ Bl,
                                                               const i64 *__restrict__ Bse, u64 *__restrict__ A2l,
                                                               i64 *__restrict__ A2s, u64 *__restrict__ Brl,
                                                               i64 *__restrict__ Brs, u64 *__restrict__ Bil,
                                                               i64 *__restrict__ Bis)
{
  const i64 nA = m * k, nB = k * n, nA2 = 2 * nA, nB2 = 2 * nB;
  const i64 stride = (i64) gridDim.x * blockDim.x;
  for (i64 t = (i64) blockIdx.x * blockDim.x + threadIdx.x; t < nA2; t += stride)
    {
      /* A2[i, c]: c < k from Ar[i, c], else Ai[i, c - k] */
      const i64 i = t / (2 * k), c = t - i * 2 * k, part = c >= k ? 1 : 0, src = i * k + c - part * k;
      for (int q = 0; q < NL; q++) A2l[(i64) q * nA2 + t] = Al[((i64) part * NL + q) * nA + src];
      A2s[t] = Ase[part * nA + src];
    }
  for (i64 t = (i64) blockIdx.x * blockDim.x + threadIdx.x; t < nB2; t += stride)
    {
      /* row r of the stacked B: r < k from Br (Bre) and Bi (Bim), else Bi
       * negated (Bre) and Br (Bim) */
      const bool lo = t < nB;
      const i64 u = lo ? t : t - nB;
      for (int q = 0; q < NL; q++)
        {
          const u64 br = Bl[(i64) q * nB + u], bi = Bl[((i64) NL + q) * nB + u];
          Brl[(i64) q * nB2 + t] = lo ? br : bi;
          Bil[(i64) q * nB2 + t] = lo ? bi : br;
        }
      const i64 sr = Bse[u], si = Bse[nB + u];
      Brs[t] = lo ? sr : (si ^ 1);
      Bis[t] = lo ? si : sr;
    }
}

/* Complex exact GEMM: Re C = RNDN(sum_k ar br - ai bi), Im C = RNDN(sum_k
 * ar bi + ai br), each component one exact sum rounded once.  Both are real
 * exact GEMMs over 2k terms, Re = [Ar | Ai] [Br ; -Bi] and Im = [Ar | Ai]
 * [Bi ; Br], so they run on the MFMA path above, in the asynchronous mode on
 * one stream-ordered allocation (the assembled operands, then a full-size
 * workspace); the first GEMM's status word is copied aside before the second
 * reuses the workspace, and the call reads both once at its end. */
int gemm_exact_complex (const pb_ops *o, int pb, i64 m, i64 n, i64 k, const u64 *Al, const i64 *Ase, const u64 *Bl,
                        const i64 *Bse, u64 *Cl, i64 *Cse, hipStream_t s)
{
  const int NL = nlimbs (pb);
  u64 *Cil = Cl + (size_t) NL * m * n;
  i64 *Cis = Cse + m * n;
  if (k == 0)
    {
      int rc = o->gemm_perop (m, n, 0, Al, Ase, Bl, Bse, Cl, Cse, s);
      return rc ? rc : o->gemm_perop (m, n, 0, Al, Ase, Bl, Bse, Cil, Cis, s);
    }
  const i64 nA2 = 2 * m * k, nB2 = 2 * k * n;
  const size_t opbytes = (size_t) pad_to ((i64) ((nA2 + 2 * nB2) * (NL + 1) * 8 + 16), 256);
  const size_t wsb = exact_workspace_bound (pb, m, n, 2 * k);
  char *buf = nullptr;
  if (hip_ok ("gemm_exact complex workspace", hipMallocAsync ((void **) &buf, opbytes + wsb, s))) return MPGFX_ENOMEM;
  u64 *A2l = (u64 *) buf;
  i64 *A2s = (i64 *) (A2l + (size_t) NL * nA2);
  u64 *Brl = (u64 *) (A2s + nA2);
  i64 *Brs = (i64 *) (Brl + (size_t) NL * nB2);
  u64 *Bil = (u64 *) (Brs + nB2);
  i64 *Bis = (i64 *) (Bil + (size_t) NL * nB2);
  i64 *st0 = Bis + nB2;                             /* the first GEMM's status word */
  char *ws = buf + opbytes;
  const i64 tot = nA2 > nB2 ? nA2 : nB2, g = (tot + 255) / 256;
  hipLaunchKernelGGL (k_exact_cplx_operands, dim3 ((unsigned) (g < 65536 ? g : 65536)), dim3 (256), 0, s, NL, m, n,
                      k, Al, Ase, Bl, Bse, A2l, A2s, Brl, Brs, Bil, Bis);
  int rc = launch_status ("gemm_exact complex operands");
  if (!rc) rc = gemm_exact (o, pb, m, n, 2 * k, A2l, A2s, Brl, Brs, Cl, Cse, ws, wsb, s, false);
  if (!rc) rc = hip_ok ("gemm_exact complex status", hipMemcpyAsync (st0, ws, 8, hipMemcpyDeviceToDevice, s));
  if (!rc) rc = gemm_exact (o, pb, m, n, 2 * k, A2l, A2s, Bil, Bis, Cil, Cis, ws, wsb, s, false);
  i64 hst[2] = {MPGFX_OK, MPGFX_OK};
  const bool async = exact_takes_async (pb, m, n, 2 * k, wsb);   /* else the return codes carry the outcome */
  if (!rc && async)
    rc = hip_ok ("gemm_exact complex status", hipMemcpyAsync (&hst[0], st0, 8, hipMemcpyDeviceToHost, s));
  if (!rc && async)
    rc = hip_ok ("gemm_exact complex status", hipMemcpyAsync (&hst[1], ws, 8, hipMemcpyDeviceToHost, s));
  (void) hipFreeAsync (buf, s);
  if (!rc) rc = hip_ok ("gemm_exact complex sync", hipStreamSynchronize (s));
  if (!rc && (hst[0] != MPGFX_OK || hst[1] != MPGFX_OK))
    {
      set_error_text ("gemm_exact complex: an exponent spread needs more than 2^22 accumulator bits");
      rc = (int) (hst[0] != MPGFX_OK ? hst[0] : hst[1]);
    }
  return rc;
}

} // namespace mpgfx
