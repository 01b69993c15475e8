/* kernels_pb.hip -- the per-precision kernels of the mpgfx library, compiled
 * once per precision with -DMPGFX_PB=<bits> (one translation unit per PB, so
 * the precisions build in parallel).
 *
 *  elementwise real / complex   cu_fmul, cu_fadd, cu_fsub / cu_cmul, cu_cadd,
 *                               cu_csub per element  (tools/test_freal.cu:41-48,
 *                               tools/test_fcomplex.cu:147-153)
 *  AXPY real / complex          demos/bench_3way.cu:44-59
 *  GEMM per-op, real / complex  tools/linalg_fast.cu:53-68 (k-ordered fold)
 *  GEMV per-op                  tools/linalg_fast.cu:37-51
 *  AoS <-> SoA                  reference struct layout cu_freal.cuh:75-77
 *
 * Mapping: one value per lane.  Each lane keeps its significands as 32-bit
 * words in VGPRs and runs the cu_fp word engine; scalars broadcast to the whole
 * grid (the AXPY coefficient) are read at wave-uniform addresses so they live
 * in SGPRs.  Streamed arrays use non-temporal loads/stores.
 */
#include "common.h"
#include "dispatch.h"
#include "coop_dpp.h"
#include "axpy_mfma.h"

#ifndef MPGFX_PB
#error "compile with -DMPGFX_PB=<precision in bits>"
#endif

namespace mpgfx {
namespace {

constexpr int PB = MPGFX_PB;
constexpr int NL = cu_freal<PB>::N;
/* workgroup size: 256 lanes (one wave per SIMD of a CU) lets the compiler use
 * up to 512 registers per lane for the widest precisions */
constexpr int TPB = 256;

template <int OP>
__global__ __launch_bounds__ (TPB) void k_ew_real (i64 n, const u64 *__restrict__ al, const i64 *__restrict__ ase,
                                                   const u64 *__restrict__ bl, const i64 *__restrict__ bse,
                                                   u64 *__restrict__ ol, i64 *__restrict__ ose)
{
  i64 stride = (i64) gridDim.x * blockDim.x;
  for (i64 i = (i64) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    {
      cu_freal<PB> a = ld<PB> (al, ase, n, i), b = ld<PB> (bl, bse, n, i), r;
      if (OP == MPGFX_MUL) r = cu_fp::cu_fmul<PB> (a, b);
      else if (OP == MPGFX_ADD) r = cu_fp::cu_fadd<PB> (a, b);
      else if (OP == MPGFX_SUB) r = cu_fp::cu_fsub<PB> (a, b);
      else r = cu_fp::cu_fdiv<PB> (a, b);
      st<PB> (r, ol, ose, n, i);
    }
}

/* o = sqrt(x) per element (cu_sqrt: Newton at PB + 128 bits) */
__global__ __launch_bounds__ (TPB) void k_sqrt_real (i64 n, const u64 *__restrict__ xl, const i64 *__restrict__ xse,
                                                     u64 *__restrict__ ol, i64 *__restrict__ ose)
{
  i64 stride = (i64) gridDim.x * blockDim.x;
  for (i64 i = (i64) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    st<PB> (cu_fp::cu_sqrt<PB> (ld<PB> (xl, xse, n, i)), ol, ose, n, i);
}

template <int OP>
__global__ __launch_bounds__ (TPB) void k_ew_cplx (i64 n, const u64 *__restrict__ al, const i64 *__restrict__ ase,
                                                   const u64 *__restrict__ bl, const i64 *__restrict__ bse,
                                                   u64 *__restrict__ ol, i64 *__restrict__ ose)
{
  const i64 o = (i64) NL * n;
  i64 stride = (i64) gridDim.x * blockDim.x;
  for (i64 i = (i64) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    {
      cu_fcomplex<PB> a (ld<PB> (al, ase, n, i), ld<PB> (al + o, ase + n, n, i));
      cu_fcomplex<PB> b (ld<PB> (bl, bse, n, i), ld<PB> (bl + o, bse + n, n, i));
      cu_fcomplex<PB> r;
      if (OP == MPGFX_MUL) r = cu_fp::cu_cmul<PB> (a, b);
      else if (OP == MPGFX_ADD) r = cu_fp::cu_cadd<PB> (a, b);
      else r = cu_fp::cu_csub<PB> (a, b);
      st<PB> (r.re, ol, ose, n, i);
      st<PB> (r.im, ol + o, ose + n, n, i);
    }
}

/* out = RNDN(RNDN(a*x) + y) */
__global__ __launch_bounds__ (TPB) void k_axpy_real (i64 n, const u64 *__restrict__ al, const i64 *__restrict__ ase,
                                                     const u64 *__restrict__ xl, const i64 *__restrict__ xse,
                                                     const u64 *__restrict__ yl, const i64 *__restrict__ yse,
                                                     u64 *__restrict__ ol, i64 *__restrict__ ose, i64 i0)
{
  const cu_freal<PB> a = ldc<PB> (al, ase, 1, 0);
  i64 stride = (i64) gridDim.x * blockDim.x;
  for (i64 i = i0 + (i64) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    {
      cu_freal<PB> t = cu_fp::cu_fmul<PB> (a, ld<PB> (xl, xse, n, i));
      st<PB> (cu_fp::cu_fadd<PB> (t, ld<PB> (yl, yse, n, i)), ol, ose, n, i);
    }
}

/* z = a*x + y, MPC_RNDNN per operation.  Two lanes per element: lane 2i+p
 * computes part p (0 = Re, 1 = Im) of element i,
 *   p = 0: RNDN(RNDN(ar*xr - ai*xi) + yr),   p = 1: RNDN(RNDN(ar*xi + ai*xr) + yi),
 * with the operands selected so both lanes run the same instructions; each
 * lane holds half the complex product's exact terms, which keeps the wide
 * precisions in registers. */
__global__ __launch_bounds__ (TPB) void k_axpy_cplx (i64 n, const u64 *__restrict__ al, const i64 *__restrict__ ase,
                                                     const u64 *__restrict__ xl, const i64 *__restrict__ xse,
                                                     const u64 *__restrict__ yl, const i64 *__restrict__ yse,
                                                     u64 *__restrict__ zl, i64 *__restrict__ zse, i64 i0)
{
  const i64 o = (i64) NL * n;
  const cu_freal<PB> ar = ldc<PB> (al, ase, 1, 0), ai = ldc<PB> (al + NL, ase + 1, 1, 0);
  const i64 stride = (i64) gridDim.x * blockDim.x;
  for (i64 t = 2 * i0 + (i64) blockIdx.x * blockDim.x + threadIdx.x; t < 2 * n; t += stride)
    {
      const i64 i = t >> 1;
      const int p = (int) (t & 1);
      const cu_freal<PB> xr = ldc<PB> (xl, xse, n, i), xi = ldc<PB> (xl + o, xse + n, n, i);
      const cu_freal<PB> u = p ? xi : xr, v = p ? xr : xi;
      cu_freal<PB> prod = cu_fp::core::fmm<PB> (ar, u, ai, v, p == 0);
      cu_freal<PB> y = ld<PB> (yl + p * o, yse + p * n, n, i);
      st<PB> (cu_fp::cu_fadd<PB> (prod, y), zl + p * o, zse + p * n, n, i);
    }
}

/* The same complex AXPY with its operand traffic overlapped with arithmetic,
 * for the precisions whose register footprint leaves one wavefront per SIMD
 * (so no other wavefront hides a load).  Each wavefront loops over chunks of
 * 32 elements (64 lanes, two per element as above) and, while it computes
 * chunk c, copies chunk c + stride of x and y from HBM straight into LDS
 * (global_load_lds_dwordx4, no registers held).  LDS holds two buffers per
 * wavefront; a buffer is 4 NL + 4 rows of 256 bytes (32 values x 8 bytes):
 * rows 0..2NL-1 the limbs of x (real then imaginary), 2NL..4NL-1 those of y,
 * then the sign/exponent words of xr, xi, yr, yi.  Copy instruction q moves
 * rows 4q..4q+3 (16 lanes per row, 16 bytes per lane).  Only whole chunks go
 * through this kernel; the n mod 32 tail goes through k_axpy_cplx. */
constexpr int AXC_WAVES = TPB / 64;
constexpr int AXC_ROWS = 4 * NL + 4;
constexpr int AXC_BUF = AXC_ROWS * 256;
constexpr bool AXC_PIPE = PB >= 512 && 2 * AXC_WAVES * AXC_BUF <= 160 * 1024;
typedef __attribute__ ((address_space (3))) void *lds_ptr;

__device__ __forceinline__ cu_freal<PB> lds_val (const unsigned char *buf, int row, int srow, int e)
{
  cu_freal<PB> r;
#pragma unroll
  for (int k = 0; k < NL; k++) r.m[k] = *(const u64 *) (buf + (row + k) * 256 + e * 8);
  const i64 w = *(const i64 *) (buf + srow * 256 + e * 8);
  r.sign = (w & 1) ? -1 : 1;
  r.exp = (long) (w >> 1);
  return r;
}

__global__ __launch_bounds__ (TPB) void k_axpy_cplx_pipe (i64 nchunks, i64 n, const u64 *__restrict__ al,
                                                          const i64 *__restrict__ ase, const u64 *__restrict__ xl,
                                                          const i64 *__restrict__ xse, const u64 *__restrict__ yl,
                                                          const i64 *__restrict__ yse, u64 *__restrict__ zl,
                                                          i64 *__restrict__ zse)
{
  __shared__ __attribute__ ((aligned (16))) unsigned char lds[AXC_PIPE ? 2 * AXC_WAVES * AXC_BUF : 16];
  if constexpr (AXC_PIPE)
    {
      const int L = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane (threadIdx.x >> 6);
      const int e = L >> 1, p = L & 1, quarter = L >> 4, lane_e = 2 * (L & 15);
      const i64 o = (i64) NL * n, wstride = (i64) gridDim.x * AXC_WAVES;
      const cu_freal<PB> ar = ldc<PB> (al, ase, 1, 0), ai = ldc<PB> (al + NL, ase + 1, 1, 0);
      unsigned char *const mine = lds + w * 2 * AXC_BUF;
      i64 c = (i64) blockIdx.x * AXC_WAVES + w;
      if (c >= nchunks) return;
      auto copy_in = [&] (i64 cc, unsigned char *dst) {
        const i64 eoff = cc * 32 + lane_e;
#pragma unroll
        for (int q = 0; q <= NL; q++)
          {
            const int r = 4 * q + quarter;
            const void *src;
            if (r < 2 * NL) src = xl + (i64) r * n + eoff;
            else if (r < 4 * NL) src = yl + (i64) (r - 2 * NL) * n + eoff;
            else
              {
                const int sr = r - 4 * NL;
                src = (sr < 2 ? xse : yse) + (sr & 1) * n + eoff;
              }
            __builtin_amdgcn_global_load_lds (src, (lds_ptr) (dst + q * 1024), 16, 0, 0);
          }
      };
      copy_in (c, mine);
      __builtin_amdgcn_s_waitcnt (0);     /* LDS copies are not tracked by the compiler */
      for (int j = 0;; j++)
        {
          const unsigned char *cur = mine + (j & 1) * AXC_BUF;
          const i64 cn = c + wstride;
          const cu_freal<PB> u = lds_val (cur, p ? NL : 0, 4 * NL + p, e);
          const cu_freal<PB> v = lds_val (cur, p ? 0 : NL, 4 * NL + 1 - p, e);
          if (cn < nchunks) copy_in (cn, mine + ((j + 1) & 1) * AXC_BUF);
          bool amb;
          cu_freal<PB> prod = cu_fp::core::fmm_short<PB> (ar, u, ai, v, p == 0, amb);
          if (__builtin_amdgcn_ballot_w64 (amb) != 0)
            {
              /* exact products for the lanes the short ones leave open; the
               * operands are read again from LDS rather than held */
              const cu_freal<PB> u2 = lds_val (cur, p ? NL : 0, 4 * NL + p, e);
              const cu_freal<PB> v2 = lds_val (cur, p ? 0 : NL, 4 * NL + 1 - p, e);
              const cu_freal<PB> ex = cu_fp::core::fmm_exact<PB> (ar, u2, ai, v2, p == 0);
              if (amb) prod = ex;
            }
          const cu_freal<PB> y = lds_val (cur, 2 * NL + p * NL, 4 * NL + 2 + p, e);
          const cu_freal<PB> z = cu_fp::cu_fadd<PB> (prod, y);
          /* the next chunk has landed long since; waiting here, before this
           * chunk's stores, keeps the next iteration's LDS reads from waiting
           * on them */
          __builtin_amdgcn_s_waitcnt (0);
          st<PB> (z, zl + p * o, zse + p * n, n, c * 32 + e);
          if (cn >= nchunks) break;
          c = cn;
        }
    }
}

/* The complex AXPY with its products on INT8 MFMA (axpy_mfma.h): whole
 * 32-element chunks, one chunk per wavefront at a time, AXM_WAVES wavefronts
 * per workgroup sharing the A tiles and the words of a_r and a_i in LDS, two
 * wavefronts per SIMD.  Per wavefront, LDS holds the x staging rows of a
 * chunk (limb-major, as in HBM: the digits are recoded in place and the B
 * tiles read from there, after which the next chunk's x is copied in with
 * global_load_lds while this one computes) and a word-value buffer through
 * which a product's words pass in groups of TG output tiles. */
constexpr bool AXM = PB == 1024;
constexpr int AXM_WAVES = 8;
constexpr int AXM_TPB = 64 * AXM_WAVES;
namespace axm {
typedef axmf::cfg<AXM ? PB : 1024> C;
constexpr int TG = 3;                                  /* output tiles per pass through LDS */
constexpr int NG = (C::NT + TG - 1) / TG;
constexpr int PBS = 64 + 2;                            /* word-value row: 64 lanes, padded */
constexpr int PBUF = 4 * TG * PBS * 8;
constexpr int STG_ROWS = 2 * NL + 4;                   /* x limbs (Re, Im), both sign/exponent rows, padding */
constexpr int STG = STG_ROWS * 256;
constexpr int WAVE = STG + PBUF;
constexpr int ATAB = 2 * C::NF * 64 * 16;              /* A tiles [q][tile][lane], 16 bytes a lane */
constexpr int TPS = 2 * C::TPW * 4;                    /* padded digit strings of a_r, a_i */
constexpr int AWS = 2 * C::KW * 4;                     /* words of a_r, a_i */
constexpr int LDSB = AXM ? ATAB + AXM_WAVES * WAVE + TPS + AWS : 16;
static_assert (!AXM || (C::KW == cu_freal<PB>::KW && cu_freal<PB>::SB == 0), "layout");
static_assert (LDSB <= 160 * 1024, "LDS");
}

/* LDS written by some lanes of a wavefront and read by others: a wavefront's
 * LDS instructions execute in order, so only the compiler must keep them in
 * program order */
__device__ __forceinline__ void wave_lds_sync ()
{
  __builtin_amdgcn_fence (__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier ();
  __builtin_amdgcn_fence (__ATOMIC_ACQUIRE, "wavefront");
}

typedef int v4i __attribute__ ((ext_vector_type (4)));

/* MFMA tiles [TA, TB) of pass q for every x part and column tile; lane l
 * writes the word values of its rows for the lane that carries the product */
template <int Q, int TA, int TB>
__device__ __forceinline__ void axm_tiles (const v4i *atab, const v4i (&bf)[2][2][axm::C::KB], i64 *pw, int L)
{
  using cu_fp::core::sfor;
  typedef axm::C C;
  const int col = L & 15, g = L >> 4;
  /* the A tiles of this group */
  v4i af[(TB - TA) * C::KB];
  sfor<TA, TB> ([&] (auto Tc) {
    constexpr int T = decltype (Tc)::value;
    sfor<0, C::KB> ([&] (auto Bc) {
      constexpr int b = decltype (Bc)::value;
      if constexpr (C::live (T, b)) af[(T - TA) * C::KB + b] = atab[(Q * C::NF + C::fidx (T, b)) * 64 + L];
    });
  });
#pragma unroll
  for (int pp = 0; pp < 2; pp++)
#pragma unroll
    for (int nt = 0; nt < 2; nt++)
      {
        v4i acc[TB - TA];
#pragma unroll
        for (int T = 0; T < TB - TA; T++) acc[T] = (v4i) {0, 0, 0, 0};
        /* K block outer: the two products into one accumulator are far apart */
        sfor<0, C::KB> ([&] (auto Bc) {
          constexpr int b = decltype (Bc)::value;
          sfor<TA, TB> ([&] (auto Tc) {
            constexpr int T = decltype (Tc)::value;
            if constexpr (C::live (T, b))
              acc[T - TA] = __builtin_amdgcn_mfma_i32_16x16x64_i8 (af[(T - TA) * C::KB + b], bf[pp][nt][b],
                                                                   acc[T - TA], 0, 0, 0);
          });
        });
        const int tl = 2 * (nt * 16 + col) + (pp ^ Q);
#pragma unroll
        for (int T = 0; T < TB - TA; T++)
          pw[(4 * T + g) * axm::PBS + tl] = axmf::assemble (acc[T][0], acc[T][1], acc[T][2], acc[T][3]);
      }
}

/* RNDN(a_r x_r - a_i x_i) (p = 0) or RNDN(a_r x_i + a_i x_r) (p = 1) of
 * element i from the exact products (rare: called, so that its registers do
 * not count against the loop's) */
__device__ __attribute__ ((noinline)) cu_freal<PB> axm_exact (const u64 *al, const i64 *ase, const u64 *xl,
                                                              const i64 *xse, i64 n, i64 i, int p)
{
  const cu_freal<PB> ar = ldc<PB> (al, ase, 1, 0), ai = ldc<PB> (al + NL, ase + 1, 1, 0);
  const cu_freal<PB> xr = ldc<PB> (xl, xse, n, i), xi = ldc<PB> (xl + (i64) NL * n, xse + n, n, i);
  return cu_fp::core::fmm_exact<PB> (ar, p ? xi : xr, ai, p ? xr : xi, p == 0);
}

template <int P>
__global__ __launch_bounds__ (AXM_TPB) void k_axpy_cplx_mfma (i64 nchunks, i64 n, const u64 *__restrict__ al,
                                                              const i64 *__restrict__ ase, const u64 *__restrict__ xl,
                                                              const i64 *__restrict__ xse, const u64 *__restrict__ yl,
                                                              const i64 *__restrict__ yse, u64 *__restrict__ zl,
                                                              i64 *__restrict__ zse)
{
  static_assert (P == PB, "one instantiation, at the library's precision");
  __shared__ __attribute__ ((aligned (16))) unsigned char lds[axm::LDSB];
  if constexpr (P == 1024)
    {
      using namespace axm;
      typedef axmf::cfg<P> C;
      using cu_fp::core::sfor;
      constexpr int KW = C::KW, NF = C::NF, NT = C::NT, KB = C::KB, H = C::H, TPW = C::TPW;
      const int tid = threadIdx.x, L = tid & 63, w = __builtin_amdgcn_readfirstlane (tid >> 6);
      const i64 o = (i64) NL * n;
      v4i *const atab = (v4i *) lds;
      u32 *const tpw = (u32 *) (lds + ATAB + AXM_WAVES * WAVE);
      u32 *const awl = (u32 *) (lds + ATAB + AXM_WAVES * WAVE + TPS);

      /* a_r, a_i: words, padded signed-digit strings, then their A tiles */
      if (tid < 2 * KW) awl[tid] = ((const u32 *) al)[tid];
      if (tid < 2 * TPW)
        {
          const int q = tid >= TPW ? 1 : 0;
          u32 aq[KW], d[KW];
          ldc<P> (al + q * NL, ase + q, 1, 0).words (aq);
          axmf::recode<KW> (d, aq);
          tpw[tid] = axmf::tp_dword<P> (tid - q * TPW, d);
        }
      __syncthreads ();
      sfor<0, NT> ([&] (auto Tc) {
        constexpr int T = decltype (Tc)::value;
        sfor<0, KB> ([&] (auto Bc) {
          constexpr int b = decltype (Bc)::value;
          if constexpr (C::live (T, b))
            {
#pragma unroll
              for (int q = 0; q < 2; q++)
                {
                  const int f = q * NF + C::fidx (T, b);
                  if (f % AXM_WAVES == w)
                    {
                      const int k0 = axmf::frag_k0<P> (T, b, L);
                      v4i fr;
#pragma unroll
                      for (int i = 0; i < 4; i++) fr[i] = (int) axmf::frag_dword (tpw + q * TPW, k0, i);
                      atab[f * 64 + L] = fr;
                    }
                }
            }
        });
      });
      __syncthreads ();

      const int e = L >> 1, p = L & 1, col = L & 15, g = L >> 4, quarter = L >> 4, lane_e = 2 * (L & 15);
      /* a's sign/exponent words */
      const i64 sea[2] = {ase[0], ase[1]};
      unsigned char *const stg = lds + ATAB + w * WAVE;
      i64 *const pw = (i64 *) (stg + STG);
      const i64 wstride = (i64) gridDim.x * AXM_WAVES;
      auto copy_in = [&] (i64 cc) {
        const u64 *src = xl + (i64) quarter * n + cc * 32 + lane_e;
#pragma unroll
        for (int qq = 0; qq < 2 * NL / 4; qq++)
          {
            __builtin_amdgcn_global_load_lds (src, (lds_ptr) (stg + qq * 1024), 16, 0, 0);
            src += 4 * n;
            asm volatile ("" : "+v"(src));
          }
        const i64 *ssrc = xse + (i64) (quarter & 1) * n + cc * 32 + lane_e;
        __builtin_amdgcn_global_load_lds (ssrc, (lds_ptr) (stg + (2 * NL / 4) * 1024), 16, 0, 0);
      };
      if ((i64) blockIdx.x * AXM_WAVES + w < nchunks) copy_in ((i64) blockIdx.x * AXM_WAVES + w);
      __builtin_amdgcn_s_waitcnt (0);
      for (i64 c = (i64) blockIdx.x * AXM_WAVES + w; c < nchunks; c += wstride)
        {
          const i64 i = c * 32 + e, cn = c + wstride;
          /* this lane's x part (one per-lane pointer stepped by a plane, kept
           * opaque so that the compiler holds no address per plane), both
           * parts' sign/exponent words */
          const i64 se0 = *(const i64 *) (stg + (2 * NL) * 256 + e * 8);
          const i64 se1 = *(const i64 *) (stg + (2 * NL + 1) * 256 + e * 8);
          {
            u32 xw[KW], dgw[KW];
#pragma unroll
            for (int k = 0; k < NL; k++)
              {
                const u64 v = *(const u64 *) (stg + (p * NL + k) * 256 + e * 8);
                xw[2 * k] = (u32) v;
                xw[2 * k + 1] = (u32) (v >> 32);
              }
            axmf::recode<KW> (dgw, xw);
#pragma unroll
            for (int k = 0; k < NL; k++)
              *(u64 *) (stg + (p * NL + k) * 256 + e * 8) = ((u64) dgw[2 * k + 1] << 32) | dgw[2 * k];
          }
          wave_lds_sync ();
          v4i bf[2][2][KB];
#pragma unroll
          for (int pp = 0; pp < 2; pp++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
#pragma unroll
              for (int b = 0; b < KB; b++)
                {
                  const unsigned char *r0 = stg + (pp * NL + 8 * b + 2 * g) * 256 + (nt * 16 + col) * 8;
                  const u64 lo = *(const u64 *) r0, hi = *(const u64 *) (r0 + 256);
                  bf[pp][nt][b] = (v4i) {(int) (u32) lo, (int) (u32) (lo >> 32), (int) (u32) hi, (int) (u32) (hi >> 32)};
                }
          wave_lds_sync ();
          __builtin_amdgcn_s_waitcnt (0xC07F);
          if (cn < nchunks) copy_in (cn);
          axmf::term<P> tm[2];
          cu_freal<P> yv;
          sfor<0, 2> ([&] (auto Qc) {
            constexpr int q = decltype (Qc)::value;
            if constexpr (q == 1)
              {
                const u64 *yp = yl + p * o + i;
#pragma unroll
                for (int k = 0; k < NL; k++)
                  {
                    yv.m[k] = __builtin_nontemporal_load (yp);
                    yp += n;
                    asm volatile ("" : "+v"(yp));
                  }
                const i64 yw = __builtin_nontemporal_load (yse + p * n + i);
                yv.sign = (yw & 1) ? -1 : 1;
                yv.exp = (long) (yw >> 1);
              }
            axmf::carry<P> cs;
            cs.c = 0;
            auto aw = [&] (int j) { return awl[q * KW + j]; };
            sfor<0, NG> ([&] (auto Gc) {
              constexpr int TA = decltype (Gc)::value * TG, TB = (TA + TG < NT) ? TA + TG : NT;
              axm_tiles<q, TA, TB> (atab, bf, pw, L);
              wave_lds_sync ();
              axmf::sweep_range<PB, 4 * TA, 4 * TB> (cs, [&] (int ww) { return pw[ww * PBS + L]; }, aw);
              wave_lds_sync ();
            });
            /* this lane's product: a_q times x part p ^ q of element e */
            const i64 sew = (p ^ q) ? se1 : se0;
            const i64 aqe = sea[q] >> 1;
            axmf::term<P> &t = tm[q];
            const bool zero = aqe == (i64) CU_FREAL_ZERO_EXP || (sew >> 1) == (i64) CU_FREAL_ZERO_EXP;
            t.exp = aqe + (sew >> 1);
            axmf::sweep_finish_raw<P> (t, cs);
            t.zero = zero;
            t.sign = zero ? 1 : (((sea[q] ^ sew) & 1) ? -1 : 1);
            t.exp = zero ? 0 : t.exp;
            t.odd = t.odd && !zero;
#pragma unroll
            for (int j = 0; j < H; j++) t.M[j] = zero ? 0u : t.M[j];
          });
          /* RNDN(p0 - p1) (real part, p = 0) or RNDN(p0 + p1), one rounding */
          tm[1].sign = (p == 0) ? -tm[1].sign : tm[1].sign;
          const cu_fp::core::rres<KW> r = cu_fp::core::add_round<KW, H, axmf::ERRB, true> (
            tm[0].sign, tm[0].exp, tm[0].zero, tm[0].M, tm[1].sign, tm[1].exp, tm[1].zero, tm[1].M);
          const bool amb = r.amb || tm[0].odd || tm[1].odd;
          cu_freal<P> prod;
          prod.set_kept (r.k);
          prod.sign = r.sign;
          prod.exp = r.exp;
          if (r.zero) prod.set_zero ();
          /* the exact products for the lanes the bound leaves open (rare) */
          if (__builtin_amdgcn_ballot_w64 (amb) != 0 && amb) prod = axm_exact (al, ase, xl, xse, n, i, p);
          const cu_freal<P> z = cu_fp::cu_fadd<P> (prod, yv);
          __builtin_amdgcn_s_waitcnt (0);
          {
            u64 *zp = zl + p * o + i;
#pragma unroll
            for (int k = 0; k < NL; k++)
              {
                __builtin_nontemporal_store (z.m[k], zp);
                zp += n;
                asm volatile ("" : "+v"(zp));
              }
            __builtin_nontemporal_store ((i64) ((u64) (i64) z.exp << 1) | (z.sign < 0 ? 1 : 0), zse + p * n + i);
          }
        }
    }
}

/* The real AXPY out = RNDN(RNDN(a x) + y) with the product on INT8 MFMA
 * (axpy_mfma.h), pb 1024 and 2048: whole 64-element chunks, one lane per
 * element.  Per chunk: lane e loads x[e], recodes it and writes the digits to
 * LDS; the B tiles of the four 16-element column tiles are read into
 * registers; the output tiles are formed in groups of AXR_TG, each lane
 * writing the word values of its rows for the lane of that element, which
 * carries them (adding a's words at offset KW); the window is rounded once
 * (round_term), lanes whose rounding it leaves open take the exact product;
 * then y is added.  Workgroup: AXR_WAVES wavefronts sharing a's A tiles. */
constexpr bool AXR = PB == 1024 || PB == 2048;
constexpr int AXR_WAVES = PB == 2048 ? 4 : 8;
constexpr int AXR_TPB = 64 * AXR_WAVES;
namespace axr {
typedef axmf::cfg<AXR ? PB : 1024> C;
constexpr int TG = (C::NT + 1) / 2;                    /* output tiles per pass through LDS */
constexpr int NG = (C::NT + TG - 1) / TG;
constexpr int PBS = 64 + 2;                            /* word-value row: 64 lanes, padded */
constexpr int DGS = C::KW + 4;                         /* digit row stride (words), 16-byte rows */
constexpr int PBUF = (4 * TG * PBS * 8 > 64 * DGS * 4) ? 4 * TG * PBS * 8 : 64 * DGS * 4;
constexpr int ATAB = C::NF * 64 * 16;                  /* A tiles [tile][lane], 16 bytes a lane */
constexpr int TPS = C::TPW * 4;                        /* padded digit string of a */
constexpr int AWS = C::KW * 4;                         /* words of a */
constexpr int LDSB = AXR ? ATAB + AXR_WAVES * PBUF + TPS + AWS : 16;
static_assert (!AXR || (C::KW == cu_freal<PB>::KW && cu_freal<PB>::SB == 0), "layout");
static_assert (LDSB <= 160 * 1024, "LDS");
}

/* MFMA tiles [TA, TB) of the real product for the four column tiles */
template <int TA, int TB>
__device__ __forceinline__ void axr_tiles (const v4i *atab, const v4i (&bf)[4][axr::C::KB], i64 *pw, int L)
{
  using cu_fp::core::sfor;
  typedef axr::C C;
  const int col = L & 15, g = L >> 4;
  v4i af[(TB - TA) * C::KB];
  sfor<TA, TB> ([&] (auto Tc) {
    constexpr int T = decltype (Tc)::value;
    sfor<0, C::KB> ([&] (auto Bc) {
      constexpr int b = decltype (Bc)::value;
      if constexpr (C::live (T, b)) af[(T - TA) * C::KB + b] = atab[C::fidx (T, b) * 64 + L];
    });
  });
#pragma unroll
  for (int nt = 0; nt < 4; nt++)
    {
      v4i acc[TB - TA];
#pragma unroll
      for (int T = 0; T < TB - TA; T++) acc[T] = (v4i) {0, 0, 0, 0};
      sfor<0, C::KB> ([&] (auto Bc) {
        constexpr int b = decltype (Bc)::value;
        sfor<TA, TB> ([&] (auto Tc) {
          constexpr int T = decltype (Tc)::value;
          if constexpr (C::live (T, b))
            acc[T - TA] = __builtin_amdgcn_mfma_i32_16x16x64_i8 (af[(T - TA) * C::KB + b], bf[nt][b], acc[T - TA], 0,
                                                                 0, 0);
        });
      });
#pragma unroll
      for (int T = 0; T < TB - TA; T++)
        pw[(4 * T + g) * axr::PBS + nt * 16 + col] = axmf::assemble (acc[T][0], acc[T][1], acc[T][2], acc[T][3]);
    }
}

/* RNDN(a x[i]) from the exact product (rare: called, so that its registers do
 * not count against the loop's) */
__device__ __attribute__ ((noinline)) cu_freal<PB> axr_exact (const u64 *al, const i64 *ase, const u64 *xl,
                                                              const i64 *xse, i64 n, i64 i)
{
  return cu_fp::core::fmul_exact<PB> (ldc<PB> (al, ase, 1, 0), ldc<PB> (xl, xse, n, i));
}

template <int P>
__global__ __launch_bounds__ (AXR_TPB) void k_axpy_real_mfma (i64 nchunks, i64 n, const u64 *__restrict__ al,
                                                              const i64 *__restrict__ ase, const u64 *__restrict__ xl,
                                                              const i64 *__restrict__ xse, const u64 *__restrict__ yl,
                                                              const i64 *__restrict__ yse, u64 *__restrict__ zl,
                                                              i64 *__restrict__ zse)
{
  static_assert (P == PB, "one instantiation, at the library's precision");
  __shared__ __attribute__ ((aligned (16))) unsigned char lds[axr::LDSB];
  if constexpr (P == 1024 || P == 2048)
    {
      using namespace axr;
      using cu_fp::core::sfor;
      typedef axmf::cfg<P> C;
      constexpr int KW = C::KW, NT = C::NT, KB = C::KB, H = C::H, TPW = C::TPW;
      const int tid = threadIdx.x, L = tid & 63, w = __builtin_amdgcn_readfirstlane (tid >> 6);
      v4i *const atab = (v4i *) lds;
      u32 *const tpw = (u32 *) (lds + ATAB + AXR_WAVES * PBUF);
      u32 *const awl = (u32 *) (lds + ATAB + AXR_WAVES * PBUF + TPS);

      /* a: its words, its padded signed-digit string, then its A tiles */
      if (tid < KW) awl[tid] = ((const u32 *) al)[tid];
      if (tid < TPW)
        {
          u32 aw[KW], d[KW];
          ldc<P> (al, ase, 1, 0).words (aw);
          axmf::recode<KW> (d, aw);
          tpw[tid] = axmf::tp_dword<P> (tid, d);
        }
      __syncthreads ();
      sfor<0, NT> ([&] (auto Tc) {
        constexpr int T = decltype (Tc)::value;
        sfor<0, KB> ([&] (auto Bc) {
          constexpr int b = decltype (Bc)::value;
          if constexpr (C::live (T, b))
            {
              const int f = C::fidx (T, b);
              if (f % AXR_WAVES == w)
                {
                  const int k0 = axmf::frag_k0<P> (T, b, L);
                  v4i fr;
#pragma unroll
                  for (int i = 0; i < 4; i++) fr[i] = (int) axmf::frag_dword (tpw, k0, i);
                  atab[f * 64 + L] = fr;
                }
            }
        });
      });
      __syncthreads ();

      const int col = L & 15, g = L >> 4;
      const i64 sea = ase[0], aqe = sea >> 1;
      i64 *const pw = (i64 *) (lds + ATAB + w * PBUF);
      u32 *const dg = (u32 *) pw;
      const i64 wstride = (i64) gridDim.x * AXR_WAVES;
      for (i64 c = (i64) blockIdx.x * AXR_WAVES + w; c < nchunks; c += wstride)
        {
          const i64 i = c * 64 + L;
          const i64 sex = xse[i];
          {
            u32 xw[KW], dgw[KW];
            const u64 *xp = xl + i;
#pragma unroll
            for (int k = 0; k < NL; k++)
              {
                const u64 v = __builtin_nontemporal_load (xp);
                xp += n;
                asm volatile ("" : "+v"(xp));
                xw[2 * k] = (u32) v;
                xw[2 * k + 1] = (u32) (v >> 32);
              }
            axmf::recode<KW> (dgw, xw);
            u32 *row = dg + L * DGS;
#pragma unroll
            for (int j = 0; j < KW / 4; j++)
              *(v4i *) (row + 4 * j) = (v4i) {(int) dgw[4 * j], (int) dgw[4 * j + 1], (int) dgw[4 * j + 2],
                                              (int) dgw[4 * j + 3]};
          }
          wave_lds_sync ();
          /* B tiles: digit words 16b + 4g .. +3 of element 16 nt + col */
          v4i bf[4][KB];
#pragma unroll
          for (int nt = 0; nt < 4; nt++)
#pragma unroll
            for (int b = 0; b < KB; b++) bf[nt][b] = *(const v4i *) (dg + (nt * 16 + col) * DGS + 16 * b + 4 * g);
          wave_lds_sync ();
          axmf::carry<P> cs;
          cs.c = 0;
          auto aw = [&] (int j) { return awl[j]; };
          sfor<0, NG> ([&] (auto Gc) {
            constexpr int TA = decltype (Gc)::value * TG, TB = (TA + TG < NT) ? TA + TG : NT;
            axr_tiles<TA, TB> (atab, bf, pw, L);
            wave_lds_sync ();
            axmf::sweep_range<P, 4 * TA, 4 * TB> (cs, [&] (int ww) { return pw[ww * PBS + L]; }, aw);
            wave_lds_sync ();
          });
          cu_freal<P> yv;
          {
            const u64 *yp = yl + i;
#pragma unroll
            for (int k = 0; k < NL; k++)
              {
                yv.m[k] = __builtin_nontemporal_load (yp);
                yp += n;
                asm volatile ("" : "+v"(yp));
              }
            const i64 yw = __builtin_nontemporal_load (yse + i);
            yv.sign = (yw & 1) ? -1 : 1;
            yv.exp = (long) (yw >> 1);
          }
          axmf::term<P> t;
          t.exp = aqe + (sex >> 1);
          axmf::sweep_finish<P> (t, cs);
          const bool zero = aqe == (i64) CU_FREAL_ZERO_EXP || (sex >> 1) == (i64) CU_FREAL_ZERO_EXP;
          u32 K[KW];
          bool amb;
          const u32 co = axmf::round_term<P> (K, t, amb);
          amb = amb && !zero;
          cu_freal<P> prod;
          prod.set_kept (K);
          prod.sign = ((sea ^ sex) & 1) ? -1 : 1;
          prod.exp = t.exp + (i64) co;
          if (zero) prod.set_zero ();
          /* the exact product for the lanes the window leaves open (rare) */
          if (__builtin_amdgcn_ballot_w64 (amb) != 0 && amb) prod = axr_exact (al, ase, xl, xse, n, i);
          const cu_freal<P> z = cu_fp::cu_fadd<P> (prod, yv);
          {
            u64 *zp = zl + i;
#pragma unroll
            for (int k = 0; k < NL; k++)
              {
                __builtin_nontemporal_store (z.m[k], zp);
                zp += n;
                asm volatile ("" : "+v"(zp));
              }
            __builtin_nontemporal_store ((i64) ((u64) (i64) z.exp << 1) | (z.sign < 0 ? 1 : 0), zse + i);
          }
        }
    }
}

/* C[i,j] = fold_k RNDN(acc + RNDN(A[i,k]*B[k,j])): one lane per C element; the
 * lanes of a workgroup share the row i, so A[i,k] is one broadcast load and
 * B[k, j..j+255] one coalesced row segment per step. */
__global__ __launch_bounds__ (TPB) void k_gemm_perop (i64 m, i64 n, i64 k, const u64 *__restrict__ Al,
                                                      const i64 *__restrict__ Ase, const u64 *__restrict__ Bl,
                                                      const i64 *__restrict__ Bse, u64 *__restrict__ Cl,
                                                      i64 *__restrict__ Cse)
{
  const i64 tiles = (n + TPB - 1) / TPB;
  const i64 nA = m * k, nB = k * n, nC = m * n;
  for (i64 t = blockIdx.x; t < m * tiles; t += gridDim.x)
    {
      const i64 i = t / tiles, j = (t % tiles) * TPB + threadIdx.x;
      const bool on = j < n;
      const i64 jj = on ? j : 0;
      cu_freal<PB> acc;
      acc.set_zero ();
      for (i64 kk = 0; kk < k; kk++)
        {
          cu_freal<PB> a = ldc<PB> (Al, Ase, nA, i * k + kk);
          cu_freal<PB> b = ldc<PB> (Bl, Bse, nB, kk * n + jj);
          acc = cu_fp::cu_fadd<PB> (acc, cu_fp::cu_fmul<PB> (a, b));
        }
      if (on) st<PB> (acc, Cl, Cse, nC, i * n + j);
    }
}

/* The per-op GEMM on 16 x 16 output tiles, one lane per C element, the
 * operands staged through LDS: per k block of 16, the workgroup copies
 * A[i0..i0+15, k0..k0+15] and B[k0..k0+15, j0..j0+15] (coalesced 16-value row
 * segments of every limb plane) into LDS once, and each lane reads its A
 * value (shared by the 16 lanes of its row) and its B value (shared by the 16
 * lanes of its column) from there: the global reads per step fall from one
 * value per lane to one eighth.  CPLX: both parts of each value are staged,
 * and the lane folds acc = acc + a b with MPC_RNDNN products and sums
 * (demos/matmul_mpc.cu:33-52), both components of acc in its registers. */
constexpr int GT = 16;
/* one value of an LDS tile: limb planes, then the sign/exponent plane */
template <int NP>
__device__ __forceinline__ cu_freal<PB> tile_val (const u64 (&t)[NP][GT * GT], int x)
{
  cu_freal<PB> v;
#pragma unroll
  for (int q = 0; q < NP - 1; q++) v.m[q] = t[q][x];
  const i64 w = (i64) t[NP - 1][x];
  v.sign = (w & 1) ? -1 : 1;
  v.exp = (long) (w >> 1);
  return v;
}

template <int CPLX>
__global__ __launch_bounds__ (256) void k_gemm_perop_tiled (i64 m, i64 n, i64 k, const u64 *__restrict__ Al,
                                                           const i64 *__restrict__ Ase, const u64 *__restrict__ Bl,
                                                           const i64 *__restrict__ Bse, u64 *__restrict__ Cl,
                                                           i64 *__restrict__ Cse)
{
  constexpr int NP = cu_freal<PB>::N + 1;                /* limb planes and the sign/exponent plane */
  constexpr int NPART = CPLX ? 2 : 1;                    /* real part, imaginary part */
  /* launched up to 8 limbs (pb 512); wider precisions compile an empty body */
  constexpr bool ON = cu_freal<PB>::N <= 8;
  __shared__ u64 sa[NPART][ON ? NP : 1][GT * GT], sb[NPART][ON ? NP : 1][GT * GT];
  if constexpr (ON)
    {
      const int r = threadIdx.x >> 4, c = threadIdx.x & 15;
      const i64 tm = (m + GT - 1) / GT, tn = (n + GT - 1) / GT;
      const i64 nA = m * k, nB = k * n, nC = m * n;
      for (i64 t = blockIdx.x; t < tm * tn; t += gridDim.x)
        {
          const i64 i0 = (t / tn) * GT, j0 = (t % tn) * GT;
          const i64 i = i0 + r, j = j0 + c;
          cu_freal<PB> acc[NPART];
#pragma unroll
          for (int h = 0; h < NPART; h++) acc[h].set_zero ();
          for (i64 k0 = 0; k0 < k; k0 += GT)
            {
              /* stage: lane (r, c) copies A[i0 + r, k0 + c] and B[k0 + r, j0 + c] */
              {
                const bool va = (i < m) && (k0 + c < k), vb = (k0 + r < k) && (j < n);
                const i64 ia = va ? i * k + k0 + c : 0, ib = vb ? (k0 + r) * n + j : 0;
#pragma unroll
                for (int h = 0; h < NPART; h++)
                  {
#pragma unroll
                    for (int q = 0; q < NP - 1; q++)
                      {
                        sa[h][q][threadIdx.x] = va ? Al[((i64) h * (NP - 1) + q) * nA + ia] : 0ull;
                        sb[h][q][threadIdx.x] = vb ? Bl[((i64) h * (NP - 1) + q) * nB + ib] : 0ull;
                      }
                    sa[h][NP - 1][threadIdx.x] = va ? (u64) Ase[h * nA + ia] : 0ull;
                    sb[h][NP - 1][threadIdx.x] = vb ? (u64) Bse[h * nB + ib] : 0ull;
                  }
              }
              __syncthreads ();
              const int kb = (k - k0 < GT) ? (int) (k - k0) : GT;
              for (int kk = 0; kk < kb; kk++)
                {
                  if constexpr (CPLX)
                    {
                      const cu_fcomplex<PB> a (tile_val (sa[0], r * GT + kk), tile_val (sa[1], r * GT + kk));
                      const cu_fcomplex<PB> b (tile_val (sb[0], kk * GT + c), tile_val (sb[1], kk * GT + c));
                      cu_fcomplex<PB> s (acc[0], acc[NPART - 1]);
                      s = cu_fp::cu_cadd<PB> (s, cu_fp::cu_cmul<PB> (a, b));
                      acc[0] = s.re;
                      acc[NPART - 1] = s.im;
                    }
                  else
                    acc[0] = cu_fp::cu_fadd<PB> (acc[0], cu_fp::cu_fmul<PB> (tile_val (sa[0], r * GT + kk),
                                                                             tile_val (sb[0], kk * GT + c)));
                }
              __syncthreads ();
            }
          if (i < m && j < n)
#pragma unroll
            for (int h = 0; h < NPART; h++) st<PB> (acc[h], Cl + (i64) h * (NP - 1) * nC, Cse + h * nC, nC, i * n + j);
        }
    }
}

__global__ __launch_bounds__ (TPB) void k_gemm_cplx (i64 m, i64 n, i64 k, const u64 *__restrict__ Al,
                                                     const i64 *__restrict__ Ase, const u64 *__restrict__ Bl,
                                                     const i64 *__restrict__ Bse, u64 *__restrict__ Cl,
                                                     i64 *__restrict__ Cse)
{
  const i64 tiles = (n + TPB - 1) / TPB;
  const i64 nA = m * k, nB = k * n, nC = m * n;
  const i64 oA = (i64) NL * nA, oB = (i64) NL * nB, oC = (i64) NL * nC;
  for (i64 t = blockIdx.x; t < m * tiles; t += gridDim.x)
    {
      const i64 i = t / tiles, j = (t % tiles) * TPB + threadIdx.x;
      const bool on = j < n;
      const i64 jj = on ? j : 0;
      cu_fcomplex<PB> acc;
      acc.re.set_zero ();
      acc.im.set_zero ();
      for (i64 kk = 0; kk < k; kk++)
        {
          i64 ia = i * k + kk, ib = kk * n + jj;
          cu_fcomplex<PB> a (ldc<PB> (Al, Ase, nA, ia), ldc<PB> (Al + oA, Ase + nA, nA, ia));
          cu_fcomplex<PB> b (ldc<PB> (Bl, Bse, nB, ib), ldc<PB> (Bl + oB, Bse + nB, nB, ib));
          acc = cu_fp::cu_cadd<PB> (acc, cu_fp::cu_cmul<PB> (a, b));
        }
      if (on)
        {
          st<PB> (acc.re, Cl, Cse, nC, i * n + j);
          st<PB> (acc.im, Cl + oC, Cse + nC, nC, i * n + j);
        }
    }
}

#include "coop.h"
constexpr int COOP_G = coop_lanes (PB);
constexpr int COOP_R = 64 / COOP_G;

/* GEMV y_i = fold_k RNDN(acc + RNDN(A[i,k] x[k])), acc from +0, k ascending
 * (tools/linalg_fast.cu:37-51), with the products computed beside the
 * cooperative fold (coop_dpp.h).  A fold wavefront holds one chain per row
 * (one folder per SIMD at 4096 rows), so the products go to a second
 * wavefront that would otherwise leave the SIMD idle: a workgroup is two
 * wavefronts over R = 64/G rows: the producer (wavefront 1) computes, per block of G steps, one
 * product per lane, RNDN(A[i_r, s0+j] x[s0+j]) for row r and step s0 + j,
 * and writes it as a record (N limbs, sign/exponent word) to one of two LDS
 * buffers; the folder (wavefront 0) reads block b's records of its row from
 * the other buffer and folds its G dependent steps while the producer
 * computes block b + 1.  One barrier per block in each wavefront, the same
 * count in both (the producer's last one after its loop, the folder's first
 * one before its loop).  No scratch memory. */
constexpr int FG_G = mpgfx::coopd::lanes_for (PB);
constexpr int FG_R = 64 / FG_G;

__global__ __launch_bounds__ (128) void k_gemv_fused (i64 m, i64 k, const u64 *__restrict__ Al,
                                                      const i64 *__restrict__ Ase, const u64 *__restrict__ xl,
                                                      const i64 *__restrict__ xse, u64 *__restrict__ yl,
                                                      i64 *__restrict__ yse)
{
  static_assert (FG_G >= 4, "precision beyond the cooperative frame");
  namespace cd = mpgfx::coopd;
  constexpr int G = FG_G, R = FG_R, REC = NL + 1, BUF = R * G * REC;
  __shared__ u64 lds[2 * BUF];
  const int lane = (int) (threadIdx.x & 63), r = lane / G, j = lane % G;
  const i64 i = (i64) blockIdx.x * R + r, ii = i < m ? i : m - 1;
  const i64 nA = m * k;
  if (threadIdx.x >= 64)
    {
      /* producer: block b's products into buffer b & 1, then a barrier */
      i64 kk = j < k ? j : k - 1;
      cu_freal<PB> a = ld<PB> (Al, Ase, nA, ii * k + kk), xv = ldc<PB> (xl, xse, k, kk);
      int buf = 0;
      for (i64 s0 = 0; s0 < k; s0 += G)
        {
          const cu_freal<PB> p = cu_fp::cu_fmul<PB> (a, xv);
          /* the next block's operands, loaded while this one is stored */
          const i64 sn = s0 + G + j;
          kk = sn < k ? sn : k - 1;
          a = ld<PB> (Al, Ase, nA, ii * k + kk);
          xv = ldc<PB> (xl, xse, k, kk);
          u64 *rec = lds + buf * BUF + (r * G + j) * REC;
#pragma unroll
          for (int l = 0; l < NL; l++) rec[l] = p.m[l];
          rec[NL] = (u64) (((i64) p.exp << 1) | (p.sign < 0 ? 1 : 0));
          __syncthreads ();
          buf ^= 1;
        }
      __syncthreads ();
    }
  else
    {
      /* folder: lane j of a row holds limb j - 1 */
      const cd::DevLanes w;
      const int jj = j < 1 ? 0 : (j > NL ? NL - 1 : j - 1);
      cd::cv acc = cd::zero ();
      int buf = 0;
      __syncthreads ();
      for (i64 s0 = 0; s0 < k; s0 += G)
        {
          cd::raw q[G];
#pragma unroll
          for (int d = 0; d < G; d++)
            {
              const u64 *src = lds + buf * BUF + (r * G + d) * REC;
              q[d].m = src[jj];
              q[d].w = (i64) src[NL];
            }
          __syncthreads ();
          buf ^= 1;
          const i64 steps = k - s0;
#pragma unroll
          for (int d = 0; d < G; d++)
            {
              const cd::cv nv = cd::add<PB> (w, acc, cd::decode<PB> (w, q[d]));
              const bool live = d < steps;
              acc.m = live ? nv.m : acc.m;
              acc.exp = live ? nv.exp : acc.exp;
              acc.sign = live ? nv.sign : acc.sign;
            }
        }
      cval out;
      out.m = acc.m;
      out.exp = acc.exp;
      out.sign = acc.sign;
      if (i < m) coop_store<PB> (out, yl, yse, m, i);
    }
}

/* Complex GEMV y_i = fold_k (acc + RNDN_c(A[i,k] x[k])), per-op in k order
 * (each product component rounded once, cu_cmul; the complex addition
 * componentwise, cu_cadd), with the products inside the cooperative fold
 * (the fold wavefront computes them itself: it holds two chains per row, two
 * folders share a SIMD, and a separate producer wavefront, as in k_gemv_fused,
 * measured slower, 6.5 against 5.6 ms at pb 512 4096^2).  Complex addition is
 * componentwise, so the real and the
 * imaginary parts of a row are two independent real folds: a wavefront holds
 * 64/G chains, chain c the part c & 1 of row c >> 1.  Per block of G steps
 * every lane computes one product component, RNDN(ar u -/+ ai v) for its
 * row, part and step (core::fmm, the operands chosen per part), and writes it
 * as a record to LDS; the G steps of the block fold their chain's records.
 * Precisions whose group is the whole wavefront (G = 64) use k_gemm_cplx. */
constexpr bool FGC = FG_G <= 32;

__global__ __launch_bounds__ (64) void k_gemv_cplx_fused (i64 m, i64 k, const u64 *__restrict__ Al,
                                                          const i64 *__restrict__ Ase, const u64 *__restrict__ xl,
                                                          const i64 *__restrict__ xse, u64 *__restrict__ yl,
                                                          i64 *__restrict__ yse)
{
  if constexpr (FGC)
    {
      namespace cd = mpgfx::coopd;
      constexpr int G = FG_G, CH = 64 / G, RC = CH / 2, REC = NL + 1;
      __shared__ u64 lds[CH * G * REC];
      const cd::DevLanes w;
      const int lane = (int) threadIdx.x, ch = lane / G, r = ch >> 1, p = ch & 1, j = lane % G;
      const int jj = j < 1 ? 0 : (j > NL ? NL - 1 : j - 1);
      const i64 i = (i64) blockIdx.x * RC + r, ii = i < m ? i : m - 1;
      const i64 nA = m * k, oA = (i64) NL * nA, oX = (i64) NL * k;
      cd::cv acc = cd::zero ();
      /* operands of this lane's first product */
      i64 kk = j < k ? j : k - 1;
      cu_freal<PB> ar = ld<PB> (Al, Ase, nA, ii * k + kk), ai = ld<PB> (Al + oA, Ase + nA, nA, ii * k + kk);
      cu_freal<PB> xr = ldc<PB> (xl, xse, k, kk), xi = ldc<PB> (xl + oX, xse + k, k, kk);
      for (i64 s0 = 0; s0 < k; s0 += G)
        {
          const cu_freal<PB> pr = cu_fp::core::fmm<PB> (ar, p ? xi : xr, ai, p ? xr : xi, p == 0);
          u64 *rec = lds + (ch * G + j) * REC;
#pragma unroll
          for (int l = 0; l < NL; l++) rec[l] = pr.m[l];
          rec[NL] = (u64) (((i64) pr.exp << 1) | (pr.sign < 0 ? 1 : 0));
          __syncthreads ();
          /* this block's records of the chain (lane j' holds limb j'-1) */
          cd::raw q[G];
#pragma unroll
          for (int d = 0; d < G; d++)
            {
              const u64 *src = lds + (ch * G + d) * REC;
              q[d].m = src[jj];
              q[d].w = (i64) src[NL];
            }
          /* the next block's operands, loaded while this block folds */
          const i64 sn = s0 + G + j;
          kk = sn < k ? sn : k - 1;
          ar = ld<PB> (Al, Ase, nA, ii * k + kk);
          ai = ld<PB> (Al + oA, Ase + nA, nA, ii * k + kk);
          xr = ldc<PB> (xl, xse, k, kk);
          xi = ldc<PB> (xl + oX, xse + k, k, kk);
          const i64 steps = k - s0;
#pragma unroll
          for (int d = 0; d < G; d++)
            {
              const cd::cv nv = cd::add<PB> (w, acc, cd::decode<PB> (w, q[d]));
              const bool live = d < steps;
              acc.m = live ? nv.m : acc.m;
              acc.exp = live ? nv.exp : acc.exp;
              acc.sign = live ? nv.sign : acc.sign;
            }
          __syncthreads ();
        }
      cval out;
      out.m = acc.m;
      out.exp = acc.exp;
      out.sign = acc.sign;
      if (i < m) coop_store<PB> (out, yl + (i64) p * NL * m, yse + (i64) p * m, m, i);
    }
}

/* AoS (reference struct layout) <-> SoA, one lane per real value */
__global__ void k_pack (i64 n, int cplx, const cu_freal<PB> *__restrict__ aos, u64 *__restrict__ l,
                        i64 *__restrict__ se)
{
  i64 i = (i64) blockIdx.x * blockDim.x + threadIdx.x;
  i64 tot = cplx ? 2 * n : n;
  if (i >= tot) return;
  i64 part = (cplx && i >= n) ? 1 : 0, j = i - part * n;
  i64 src = cplx ? 2 * j + part : i;
  cu_freal<PB> v = aos[src];
  st<PB> (v, l + part * (i64) NL * n, se + part * n, n, j);
}
__global__ void k_unpack (i64 n, int cplx, const u64 *__restrict__ l, const i64 *__restrict__ se,
                          cu_freal<PB> *__restrict__ aos)
{
  i64 i = (i64) blockIdx.x * blockDim.x + threadIdx.x;
  i64 tot = cplx ? 2 * n : n;
  if (i >= tot) return;
  i64 part = (cplx && i >= n) ? 1 : 0, j = i - part * n;
  i64 dst = cplx ? 2 * j + part : i;
  aos[dst] = ldc<PB> (l + part * (i64) NL * n, se + part * n, n, j);
}

unsigned grid_for (i64 work, int threads, int per_cu)
{
  i64 b = (work + threads - 1) / threads, cap = (i64) 256 * per_cu;
  return (unsigned) (b > cap ? cap : (b < 1 ? 1 : b));
}

int ew_real (int op, i64 n, const u64 *al, const i64 *ase, const u64 *bl, const i64 *bse, u64 *ol, i64 *ose,
             hipStream_t s)
{
  unsigned g = grid_for (n, TPB, 32);
  if (op == MPGFX_MUL) hipLaunchKernelGGL ((k_ew_real<MPGFX_MUL>), dim3 (g), dim3 (TPB), 0, s, n, al, ase, bl, bse, ol, ose);
  else if (op == MPGFX_ADD) hipLaunchKernelGGL ((k_ew_real<MPGFX_ADD>), dim3 (g), dim3 (TPB), 0, s, n, al, ase, bl, bse, ol, ose);
  else if (op == MPGFX_SUB) hipLaunchKernelGGL ((k_ew_real<MPGFX_SUB>), dim3 (g), dim3 (TPB), 0, s, n, al, ase, bl, bse, ol, ose);
  else hipLaunchKernelGGL ((k_ew_real<MPGFX_DIV>), dim3 (g), dim3 (TPB), 0, s, n, al, ase, bl, bse, ol, ose);
  return launch_status ("elementwise_real");
}
int sqrt_real (i64 n, const u64 *xl, const i64 *xse, u64 *ol, i64 *ose, hipStream_t s)
{
  hipLaunchKernelGGL (k_sqrt_real, dim3 (grid_for (n, TPB, 32)), dim3 (TPB), 0, s, n, xl, xse, ol, ose);
  return launch_status ("sqrt_real");
}
int ew_cplx (int op, i64 n, const u64 *al, const i64 *ase, const u64 *bl, const i64 *bse, u64 *ol, i64 *ose,
             hipStream_t s)
{
  unsigned g = grid_for (n, TPB, 32);
  if (op == MPGFX_MUL) hipLaunchKernelGGL ((k_ew_cplx<MPGFX_MUL>), dim3 (g), dim3 (TPB), 0, s, n, al, ase, bl, bse, ol, ose);
  else if (op == MPGFX_ADD) hipLaunchKernelGGL ((k_ew_cplx<MPGFX_ADD>), dim3 (g), dim3 (TPB), 0, s, n, al, ase, bl, bse, ol, ose);
  else hipLaunchKernelGGL ((k_ew_cplx<MPGFX_SUB>), dim3 (g), dim3 (TPB), 0, s, n, al, ase, bl, bse, ol, ose);
  return launch_status ("elementwise_complex");
}
int axpy_real (i64 n, const u64 *al, const i64 *ase, const u64 *xl, const i64 *xse, const u64 *yl, const i64 *yse,
               u64 *ol, i64 *ose, hipStream_t s)
{
  i64 i0 = 0;
  if (AXR && n >= 64)
    {
      /* one workgroup per CU (its LDS) */
      const i64 nchunks = n / 64, wg = (nchunks + AXR_WAVES - 1) / AXR_WAVES;
      hipLaunchKernelGGL ((k_axpy_real_mfma<PB>), dim3 ((unsigned) (wg < 256 ? wg : 256)), dim3 (AXR_TPB), 0, s, nchunks,
                          n, al, ase, xl, xse, yl, yse, ol, ose);
      i0 = nchunks * 64;
    }
  if (i0 < n)
    hipLaunchKernelGGL (k_axpy_real, dim3 (grid_for (n - i0, TPB, 32)), dim3 (TPB), 0, s, n, al, ase, xl, xse, yl, yse,
                        ol, ose, i0);
  return launch_status ("axpy_real");
}
int axpy_cplx (i64 n, const u64 *al, const i64 *ase, const u64 *xl, const i64 *xse, const u64 *yl, const i64 *yse,
               u64 *ol, i64 *ose, hipStream_t s)
{
  i64 i0 = 0;
  if (AXM && n >= 32)
    {
      /* one workgroup per CU (its LDS) */
      const i64 nchunks = n / 32, wg = (nchunks + AXM_WAVES - 1) / AXM_WAVES;
      const i64 grid = wg < 256 ? wg : 256;
      hipLaunchKernelGGL ((k_axpy_cplx_mfma<PB>), dim3 ((unsigned) grid), dim3 (AXM_TPB), 0, s, nchunks, n, al, ase, xl, xse,
                          yl, yse, ol, ose);
      i0 = nchunks * 32;
    }
  else if (AXC_PIPE && n >= 32)
    {
      const i64 nchunks = n / 32, wg = (nchunks + AXC_WAVES - 1) / AXC_WAVES;
      hipLaunchKernelGGL (k_axpy_cplx_pipe, dim3 ((unsigned) (wg < 256 ? wg : 256)), dim3 (TPB), 0, s, nchunks, n, al,
                          ase, xl, xse, yl, yse, ol, ose);
      i0 = nchunks * 32;
    }
  if (i0 < n)
    hipLaunchKernelGGL (k_axpy_cplx, dim3 (grid_for (2 * (n - i0), TPB, 32)), dim3 (TPB), 0, s, n, al, ase, xl, xse,
                        yl, yse, ol, ose, i0);
  return launch_status ("axpy_complex");
}
int gemm_perop (i64 m, i64 n, i64 k, const u64 *Al, const i64 *Ase, const u64 *Bl, const i64 *Bse, u64 *Cl,
                i64 *Cse, hipStream_t s)
{
  if (NL <= 8 && k > 0)
    {
      const i64 tiles = ((m + GT - 1) / GT) * ((n + GT - 1) / GT);
      const unsigned g = (unsigned) (tiles > 65536 ? 65536 : tiles);
      hipLaunchKernelGGL ((k_gemm_perop_tiled<0>), dim3 (g), dim3 (256), 0, s, m, n, k, Al, Ase, Bl, Bse, Cl, Cse);
      return launch_status ("gemm_real");
    }
  i64 tiles = m * ((n + TPB - 1) / TPB);
  unsigned g = (unsigned) (tiles > 65536 ? 65536 : tiles);
  hipLaunchKernelGGL (k_gemm_perop, dim3 (g), dim3 (TPB), 0, s, m, n, k, Al, Ase, Bl, Bse, Cl, Cse);
  return launch_status ("gemm_real");
}
int gemm_cplx (i64 m, i64 n, i64 k, const u64 *Al, const i64 *Ase, const u64 *Bl, const i64 *Bse, u64 *Cl,
               i64 *Cse, hipStream_t s)
{
  if (FGC && n == 1 && k > 0)
    {
      /* a GEMV: the products inside the cooperative fold */
      constexpr i64 RC = 32 / FG_G;
      hipLaunchKernelGGL (k_gemv_cplx_fused, dim3 ((unsigned) ((m + RC - 1) / RC)), dim3 (64), 0, s, m, k, Al, Ase, Bl,
                          Bse, Cl, Cse);
      return launch_status ("gemv_complex");
    }
  if (NL <= 8 && k > 0)
    {
      const i64 tiles = ((m + GT - 1) / GT) * ((n + GT - 1) / GT);
      const unsigned g = (unsigned) (tiles > 65536 ? 65536 : tiles);
      hipLaunchKernelGGL ((k_gemm_perop_tiled<1>), dim3 (g), dim3 (256), 0, s, m, n, k, Al, Ase, Bl, Bse, Cl, Cse);
      return launch_status ("gemm_complex");
    }
  i64 tiles = m * ((n + TPB - 1) / TPB);
  unsigned g = (unsigned) (tiles > 65536 ? 65536 : tiles);
  hipLaunchKernelGGL (k_gemm_cplx, dim3 (g), dim3 (TPB), 0, s, m, n, k, Al, Ase, Bl, Bse, Cl, Cse);
  return launch_status ("gemm_complex");
}
int gemv (i64 m, i64 k, const u64 *Al, const i64 *Ase, const u64 *xl, const i64 *xse, u64 *yl, i64 *yse,
          hipStream_t s)
{
  if (k == 0)
    {
      /* empty sums: +0 (the per-op GEMM kernel writes them) */
      hipLaunchKernelGGL (k_gemm_perop, dim3 ((unsigned) (m < 65536 ? m : 65536)), dim3 (TPB), 0, s, m, (i64) 1,
                          k, Al, Ase, xl, xse, yl, yse);
      return launch_status ("gemv_real");
    }
  /* products inside the fold: no scratch */
  hipLaunchKernelGGL (k_gemv_fused, dim3 ((unsigned) ((m + FG_R - 1) / FG_R)), dim3 (128), 0, s, m, k, Al, Ase, xl,
                      xse, yl, yse);
  return launch_status ("gemv_real");
}
int pack (int cplx, i64 n, const void *aos, u64 *l, i64 *se, hipStream_t s)
{
  i64 tot = cplx ? 2 * n : n;
  hipLaunchKernelGGL (k_pack, dim3 ((unsigned) ((tot + 255) / 256)), dim3 (256), 0, s, n, cplx,
                      (const cu_freal<PB> *) aos, l, se);
  return launch_status ("pack");
}
int unpack (int cplx, i64 n, const u64 *l, const i64 *se, void *aos, hipStream_t s)
{
  i64 tot = cplx ? 2 * n : n;
  hipLaunchKernelGGL (k_unpack, dim3 ((unsigned) ((tot + 255) / 256)), dim3 (256), 0, s, n, cplx, l, se,
                      (cu_freal<PB> *) aos);
  return launch_status ("unpack");
}

#include "exact_pb.h"
#include "batch_pb.h"

const pb_ops ops = { PB, ew_real, ew_cplx, sqrt_real, axpy_real, axpy_cplx, gemm_perop, gemm_cplx, gemv, pack, unpack,
                     exact_digits_launch, exact_round_launch, exact_fallback_launch, exact_fallback_dev_launch,
                     batch_launch };

} // namespace

/* registration hook, one per compiled precision */
#define MPGFX_CAT2(a, b) a##b
#define MPGFX_CAT(a, b) MPGFX_CAT2 (a, b)
const pb_ops *MPGFX_CAT (ops_pb_, MPGFX_PB) () { return &ops; }

} // namespace mpgfx
