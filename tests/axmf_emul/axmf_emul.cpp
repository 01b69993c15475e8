/* tests/axmf_emul/axmf_emul.cpp -- host emulation of the complex AXPY whose
 * products run on INT8 MFMA (mpc_cuda_amd/csrc/axpy_mfma.h, kernel
 * k_axpy_cplx_mfma): the same digit recoding, padded digit strings, A-tile
 * bytes, word assembly, carry sweep and rounding helpers, with the wavefront's
 * lanes, its LDS buffers and v_mfma_i32_16x16x64_i8 emulated on the host (lane
 * l: row / column l&15, K bytes 16(l>>4)..+15 of the block; result rows
 * 4(l>>4)+v of column l&15).  Whole 32-element chunks only.  Test harness
 * only. */
#include <cstdint>
#include <cstring>
#include <vector>
#include "mpc_cuda.cuh"
#include "axpy_mfma.h"

using namespace mpgfx::axmf;
using cu_fp::cu_freal;

namespace {

struct v4 { int32_t x[4]; };

/* acc[l] += A * B for one 16x16x64 block, in the lane layout of the kernel */
void mfma (v4 (&acc)[64], const v4 (&A)[64], const v4 (&B)[64])
{
  int8_t a[16][64], b[64][16];
  for (int l = 0; l < 64; l++)
    for (int j = 0; j < 16; j++)
      {
        a[l & 15][16 * (l >> 4) + j] = (int8_t) ((uint32_t) A[l].x[j >> 2] >> (8 * (j & 3)));
        b[16 * (l >> 4) + j][l & 15] = (int8_t) ((uint32_t) B[l].x[j >> 2] >> (8 * (j & 3)));
      }
  for (int l = 0; l < 64; l++)
    for (int v = 0; v < 4; v++)
      {
        const int r = 4 * (l >> 4) + v, c = l & 15;
        int32_t s = 0;
        for (int k = 0; k < 64; k++) s += (int32_t) a[r][k] * b[k][c];
        acc[l].x[v] += s;
      }
}

template <int PB>
cu_freal<PB> ldv (const uint64_t *l, const int64_t *se, int64_t stride, int64_t i)
{
  cu_freal<PB> r;
  for (int k = 0; k < cu_freal<PB>::N; k++) r.m[k] = l[k * stride + i];
  const int64_t w = se[i];
  r.sign = (w & 1) ? -1 : 1;
  r.exp = (long) (w >> 1);
  return r;
}
template <int PB>
void stv (const cu_freal<PB> &r, uint64_t *l, int64_t *se, int64_t stride, int64_t i)
{
  for (int k = 0; k < cu_freal<PB>::N; k++) l[k * stride + i] = r.m[k];
  se[i] = (int64_t) ((uint64_t) (int64_t) r.exp << 1) | (r.sign < 0 ? 1 : 0);
}

template <int PB>
long run (int64_t n, const uint64_t *al, const int64_t *ase, const uint64_t *xl, const int64_t *xse,
          const uint64_t *yl, const int64_t *yse, uint64_t *zl, int64_t *zse)
{
  typedef cfg<PB> C;
  constexpr int NL = cu_freal<PB>::N, KW = C::KW, NF = C::NF, NT = C::NT, KB = C::KB, H = C::H, TPW = C::TPW;
  const int64_t o = (int64_t) NL * n;
  const cu_freal<PB> ar = ldv<PB> (al, ase, 1, 0), ai = ldv<PB> (al + NL, ase + 1, 1, 0);
  uint32_t arw[KW], aiw[KW];
  ar.words (arw);
  ai.words (aiw);
  /* padded digit strings and A tiles (workgroup prologue) */
  std::vector<uint32_t> tpw (2 * TPW);
  for (int tid = 0; tid < 2 * TPW; tid++)
    {
      const int q = tid >= TPW ? 1 : 0;
      uint32_t aq[KW], d[KW];
      for (int j = 0; j < KW; j++) aq[j] = q ? aiw[j] : arw[j];
      recode<KW> (d, aq);
      tpw[tid] = tp_dword<PB> (tid - q * TPW, d);
    }
  std::vector<v4> atab (2 * NF * 64);
  for (int T = 0; T < NT; T++)
    for (int b = 0; b < KB; b++)
      if (C::live (T, b))
        for (int q = 0; q < 2; q++)
          for (int L = 0; L < 64; L++)
            {
              const int k0 = frag_k0<PB> (T, b, L);
              v4 fr;
              for (int i = 0; i < 4; i++) fr.x[i] = (int32_t) frag_dword (tpw.data () + q * TPW, k0, i);
              atab[(q * NF + C::fidx (T, b)) * 64 + L] = fr;
            }
  long nodd = 0;
  const int64_t nchunks = n / 32;
  std::vector<int64_t> pw (C::NWW * 64);
  std::vector<uint32_t> dg (64 * (KW + 4));
  constexpr int DGS = KW + 4;
  for (int64_t c = 0; c < nchunks; c++)
    {
      /* step 1: each lane recodes its x part */
      for (int L = 0; L < 64; L++)
        {
          const int e = L >> 1, p = L & 1;
          const cu_freal<PB> xv = ldv<PB> (xl + p * o, xse + p * n, n, c * 32 + e);
          uint32_t xw[KW], d[KW];
          xv.words (xw);
          recode<KW> (d, xw);
          for (int j = 0; j < KW; j++) dg[(p * 32 + e) * DGS + j] = d[j];
        }
      /* step 2: B tiles */
      static v4 bf[2][2][8][64];
      for (int pp = 0; pp < 2; pp++)
        for (int nt = 0; nt < 2; nt++)
          for (int b = 0; b < KB; b++)
            for (int L = 0; L < 64; L++)
              for (int i = 0; i < 4; i++)
                bf[pp][nt][b][L].x[i] = (int32_t) dg[(pp * 32 + nt * 16 + (L & 15)) * DGS + 16 * b + 4 * (L >> 4) + i];
      term<PB> tm[64][2];
      for (int q = 0; q < 2; q++)
        {
          for (int pp = 0; pp < 2; pp++)
            for (int nt = 0; nt < 2; nt++)
              {
                static v4 acc[32][64];
                for (int T = 0; T < NT; T++)
                  for (int L = 0; L < 64; L++) acc[T][L] = v4{{0, 0, 0, 0}};
                for (int T = 0; T < NT; T++)
                  for (int b = 0; b < KB; b++)
                    if (C::live (T, b))
                      {
                        v4 A[64];
                        for (int L = 0; L < 64; L++) A[L] = atab[(q * NF + C::fidx (T, b)) * 64 + L];
                        mfma (acc[T], A, bf[pp][nt][b]);
                      }
                for (int L = 0; L < 64; L++)
                  {
                    const int tl = 2 * (nt * 16 + (L & 15)) + (pp ^ q);
                    for (int T = 0; T < NT; T++)
                      pw[(4 * T + (L >> 4)) * 64 + tl] =
                        assemble (acc[T][L].x[0], acc[T][L].x[1], acc[T][L].x[2], acc[T][L].x[3]);
                  }
              }
          for (int L = 0; L < 64; L++)
            {
              const int e = L >> 1, p = L & 1;
              const int64_t i = c * 32 + e;
              const int64_t sew = ((p ^ q) ? xse + n : xse)[i];
              const cu_freal<PB> &aq = q ? ai : ar;
              term<PB> &t = tm[L][q];
              const bool zero = aq.is_zero () || (sew >> 1) == (int64_t) CU_FREAL_ZERO_EXP;
              t.exp = (int64_t) aq.exp + (sew >> 1);
              /* in groups of three output tiles, as the kernel carries them */
              carry<PB> cs;
              cs.c = 0;
              auto aw = [&] (int j) { return (q ? aiw : arw)[j]; };
              cu_fp::core::sfor<0, (NT + 2) / 3> ([&] (auto Gc) {
                constexpr int TA = 3 * decltype (Gc)::value, TB = TA + 3 < NT ? TA + 3 : NT;
                sweep_range<PB, 4 * TA, 4 * TB> (cs, [&] (int ww) { return pw[(4 * TA + ww) * 64 + L]; }, aw);
              });
              sweep_finish_raw<PB> (t, cs);
              t.zero = zero;
              t.sign = zero ? 1 : aq.sign * ((sew & 1) ? -1 : 1);
              t.exp = zero ? 0 : t.exp;
              t.odd = t.odd && !zero;
              for (int j = 0; j < H; j++) t.M[j] = zero ? 0u : t.M[j];
            }
        }
      for (int L = 0; L < 64; L++)
        {
          const int e = L >> 1, p = L & 1;
          const int64_t i = c * 32 + e;
          term<PB> t0 = tm[L][0], t1 = tm[L][1];
          t1.sign = (p == 0) ? -t1.sign : t1.sign;
          const cu_fp::core::rres<KW> r = cu_fp::core::add_round<KW, H, ERRB, true> (t0.sign, t0.exp, t0.zero, t0.M, t1.sign,
                                                                              t1.exp, t1.zero, t1.M);
          const bool amb = r.amb || t0.odd || t1.odd;
          cu_freal<PB> prod;
          prod.set_kept (r.k);
          prod.sign = r.sign;
          prod.exp = r.exp;
          if (r.zero) prod.set_zero ();
          if (amb)
            {
              nodd++;
              const cu_freal<PB> xr = ldv<PB> (xl, xse, n, i), xi = ldv<PB> (xl + o, xse + n, n, i);
              prod = cu_fp::core::fmm_exact<PB> (ar, p ? xi : xr, ai, p ? xr : xi, p == 0);
            }
          const cu_freal<PB> yv = ldv<PB> (yl + p * o, yse + p * n, n, i);
          stv<PB> (cu_fp::cu_fadd<PB> (prod, yv), zl + p * o, zse + p * n, n, i);
        }
    }
  return nodd;
}

/* the real AXPY out = RNDN(RNDN(a x) + y): 64-element chunks, one lane per
 * element, four column tiles, the output tiles in groups of (NT + 1) / 2 */
template <int PB>
long run_real (int64_t n, const uint64_t *al, const int64_t *ase, const uint64_t *xl, const int64_t *xse,
               const uint64_t *yl, const int64_t *yse, uint64_t *zl, int64_t *zse)
{
  typedef cfg<PB> C;
  constexpr int KW = C::KW, NF = C::NF, NT = C::NT, KB = C::KB, TPW = C::TPW, DGS = KW + 4;
  constexpr int TG = (NT + 1) / 2;
  const cu_freal<PB> a = ldv<PB> (al, ase, 1, 0);
  uint32_t aw[KW];
  a.words (aw);
  std::vector<uint32_t> tpw (TPW);
  {
    uint32_t d[KW];
    recode<KW> (d, aw);
    for (int i = 0; i < TPW; i++) tpw[i] = tp_dword<PB> (i, d);
  }
  std::vector<v4> atab (NF * 64);
  for (int T = 0; T < NT; T++)
    for (int b = 0; b < KB; b++)
      if (C::live (T, b))
        for (int L = 0; L < 64; L++)
          {
            const int k0 = frag_k0<PB> (T, b, L);
            v4 fr;
            for (int i = 0; i < 4; i++) fr.x[i] = (int32_t) frag_dword (tpw.data (), k0, i);
            atab[C::fidx (T, b) * 64 + L] = fr;
          }
  long nodd = 0;
  const int64_t nchunks = n / 64;
  std::vector<uint32_t> dg (64 * DGS);
  std::vector<int64_t> pw (C::NWW * 64);
  static v4 bf[4][8][64], acc[40][64];
  for (int64_t c = 0; c < nchunks; c++)
    {
      for (int L = 0; L < 64; L++)
        {
          const cu_freal<PB> xv = ldv<PB> (xl, xse, n, c * 64 + L);
          uint32_t xw[KW], d[KW];
          xv.words (xw);
          recode<KW> (d, xw);
          for (int j = 0; j < KW; j++) dg[L * DGS + j] = d[j];
        }
      for (int nt = 0; nt < 4; nt++)
        for (int b = 0; b < KB; b++)
          for (int L = 0; L < 64; L++)
            for (int i = 0; i < 4; i++)
              bf[nt][b][L].x[i] = (int32_t) dg[(nt * 16 + (L & 15)) * DGS + 16 * b + 4 * (L >> 4) + i];
      for (int nt = 0; nt < 4; nt++)
        {
          for (int T = 0; T < NT; T++)
            for (int L = 0; L < 64; L++) acc[T][L] = v4{{0, 0, 0, 0}};
          for (int T = 0; T < NT; T++)
            for (int b = 0; b < KB; b++)
              if (C::live (T, b))
                {
                  v4 A[64];
                  for (int L = 0; L < 64; L++) A[L] = atab[C::fidx (T, b) * 64 + L];
                  mfma (acc[T], A, bf[nt][b]);
                }
          for (int L = 0; L < 64; L++)
            for (int T = 0; T < NT; T++)
              pw[(4 * T + (L >> 4)) * 64 + nt * 16 + (L & 15)] =
                assemble (acc[T][L].x[0], acc[T][L].x[1], acc[T][L].x[2], acc[T][L].x[3]);
        }
      for (int L = 0; L < 64; L++)
        {
          const int64_t i = c * 64 + L;
          const int64_t sex = xse[i];
          carry<PB> cs;
          cs.c = 0;
          sweep_range<PB, 0, 4 * TG> (cs, [&] (int ww) { return pw[ww * 64 + L]; }, [&] (int j) { return aw[j]; });
          sweep_range<PB, 4 * TG, 4 * NT> (cs, [&] (int ww) { return pw[(4 * TG + ww) * 64 + L]; },
                                          [&] (int j) { return aw[j]; });
          term<PB> t;
          t.exp = (int64_t) a.exp + (sex >> 1);
          sweep_finish<PB> (t, cs);
          const bool zero = a.is_zero () || (sex >> 1) == (int64_t) CU_FREAL_ZERO_EXP;
          uint32_t K[KW];
          bool amb;
          const uint32_t co = round_term<PB> (K, t, amb);
          amb = amb && !zero;
          cu_freal<PB> prod;
          prod.set_kept (K);
          prod.sign = ((ase[0] ^ sex) & 1) ? -1 : 1;
          prod.exp = t.exp + (int64_t) co;
          if (zero) prod.set_zero ();
          if (amb)
            {
              nodd++;
              prod = cu_fp::core::fmul_exact<PB> (a, ldv<PB> (xl, xse, n, i));
            }
          stv<PB> (cu_fp::cu_fadd<PB> (prod, ldv<PB> (yl, yse, n, i)), zl, zse, n, i);
        }
    }
  return nodd;
}

} // namespace

extern "C" {
/* z = a x + y over the whole 32-element chunks of n; returns the number of
 * lanes that took the exact products, or -1 for an unsupported precision */
long axmf_emul_axpy (int pb, int64_t n, const uint64_t *al, const int64_t *ase, const uint64_t *xl,
                     const int64_t *xse, const uint64_t *yl, const int64_t *yse, uint64_t *zl, int64_t *zse)
{
  if (pb == 512) return run<512> (n, al, ase, xl, xse, yl, yse, zl, zse);
  if (pb == 1024) return run<1024> (n, al, ase, xl, xse, yl, yse, zl, zse);
  if (pb == 2048) return run<2048> (n, al, ase, xl, xse, yl, yse, zl, zse);
  return -1;
}
/* real: out = a x + y over the whole 64-element chunks of n */
long axmf_emul_axpy_real (int pb, int64_t n, const uint64_t *al, const int64_t *ase, const uint64_t *xl,
                          const int64_t *xse, const uint64_t *yl, const int64_t *yse, uint64_t *zl, int64_t *zse)
{
  if (pb == 1024) return run_real<1024> (n, al, ase, xl, xse, yl, yse, zl, zse);
  if (pb == 2048) return run_real<2048> (n, al, ase, xl, xse, yl, yse, zl, zse);
  return -1;
}
}
