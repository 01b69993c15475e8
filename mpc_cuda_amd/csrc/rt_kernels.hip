/* rt_kernels.hip -- run-time precision kernels of the mpgfx_rt_* entry points,
 * compiled once per container size with -DMPGFX_RT_CW=<32-bit words> (2, 4,
 * 8, 16, 32, 64: precisions up to 64 .. 2048 bits).
 *
 *  elementwise real   mpfr_mul / mpfr_add / mpfr_sub per element, RNDN at p
 *  AXPY real          y = RNDN(RNDN(a x) + y)          demos/axpy_mpfr.cu:36-50
 *  AXPY complex       z = MPC_RNDNN(MPC_RNDNN(a x) + y) demos/axpy_mpc.cu:40-53
 *  GEMV / dot         y_i = fold_k RNDN(acc + RNDN(A_ik x_k)), acc from +0, k
 *                     ascending (mpfr_mul + mpfr_add; complex: mpc_mul +
 *                     mpc_add)                          demos/matvec_mpfr.cu,
 *                                                       demos/matvec_mpc.cu
 *  GEMM               the same fold per element         demos/matmul_mpfr.cu,
 *                                                       demos/matmul_mpc.cu
 *
 * Mapping as in the fixed-precision kernels (kernels_pb.hip): one value per
 * lane, and two lanes per complex element (one per component: complex
 * addition is componentwise, so each component of a complex fold is a real
 * fold of that component of the products).  GEMV: a lane per row (per row
 * component), x[k] read at a wave-uniform address; GEMM: a lane per output
 * (component), the lanes of a workgroup sharing the row.
 */
#include "rt.h"

#ifndef MPGFX_RT_CW
#error "compile with -DMPGFX_RT_CW=<container words>"
#endif

namespace mpgfx {
namespace {

constexpr int CW = MPGFX_RT_CW;
constexpr int TPB = 256;

inline int limbs_of (int p) { return (p + 63) / 64; }

inline unsigned grid (i64 work)
{
  const i64 b = (work + TPB - 1) / TPB, cap = (i64) 256 * 32;
  return (unsigned) (b > cap ? cap : (b < 1 ? 1 : b));
}

template <int OP>
__global__ __launch_bounds__ (TPB) void k_rt_ew (int p, int nl, i64 n, const u64 *__restrict__ al,
                                                 const i64 *__restrict__ ase, const u64 *__restrict__ bl,
                                                 const i64 *__restrict__ bse, u64 *__restrict__ ol,
                                                 i64 *__restrict__ ose)
{
  const i64 stride = (i64) gridDim.x * blockDim.x;
  for (i64 i = (i64) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    {
      const rt_val<CW> a = rt_ld<CW> (al, ase, n, i, nl), b = rt_ld<CW> (bl, bse, n, i, nl);
      const rt_val<CW> r = (OP == MPGFX_MUL) ? rt_mul<CW> (a, b, p) : rt_add<CW> (a, b, p, OP == MPGFX_SUB);
      rt_st<CW> (r, ol, ose, n, i, nl);
    }
}

__global__ __launch_bounds__ (TPB) void k_rt_axpy_real (int p, int nl, i64 n, const u64 *__restrict__ al,
                                                        const i64 *__restrict__ ase, const u64 *__restrict__ xl,
                                                        const i64 *__restrict__ xse, const u64 *__restrict__ yl,
                                                        const i64 *__restrict__ yse, u64 *__restrict__ ol,
                                                        i64 *__restrict__ ose)
{
  const rt_val<CW> a = rt_ld<CW> (al, ase, 1, 0, nl);
  const i64 stride = (i64) gridDim.x * blockDim.x;
  for (i64 i = (i64) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    {
      const rt_val<CW> t = rt_mul<CW> (a, rt_ld<CW> (xl, xse, n, i, nl), p);
      rt_st<CW> (rt_add<CW> (t, rt_ld<CW> (yl, yse, n, i, nl), p, false), ol, ose, n, i, nl);
    }
}

/* lane 2i+q computes component q of element i (0 = Re: RNDN(ar xr - ai xi),
 * 1 = Im: RNDN(ar xi + ai xr)), then adds y's component */
__global__ __launch_bounds__ (TPB) void k_rt_axpy_cplx (int p, int nl, i64 n, const u64 *__restrict__ al,
                                                        const i64 *__restrict__ ase, const u64 *__restrict__ xl,
                                                        const i64 *__restrict__ xse, const u64 *__restrict__ yl,
                                                        const i64 *__restrict__ yse, u64 *__restrict__ zl,
                                                        i64 *__restrict__ zse)
{
  const i64 o = (i64) nl * n;
  const rt_val<CW> ar = rt_ld<CW> (al, ase, 1, 0, nl), ai = rt_ld<CW> (al + nl, ase + 1, 1, 0, nl);
  const i64 stride = (i64) gridDim.x * blockDim.x;
  for (i64 t = (i64) blockIdx.x * blockDim.x + threadIdx.x; t < 2 * n; t += stride)
    {
      const i64 i = t >> 1;
      const int q = (int) (t & 1);
      const rt_val<CW> u = rt_ld<CW> (xl + q * o, xse + q * n, n, i, nl);
      const rt_val<CW> v = rt_ld<CW> (xl + (1 - q) * o, xse + (1 - q) * n, n, i, nl);
      const rt_val<CW> prod = rt_fmm<CW> (ar, u, ai, v, p, q == 0);
      const rt_val<CW> y = rt_ld<CW> (yl + q * o, yse + q * n, n, i, nl);
      rt_st<CW> (rt_add<CW> (prod, y, p, false), zl + q * o, zse + q * n, n, i, nl);
    }
}

int ew (int op, int p, i64 n, const u64 *al, const i64 *ase, const u64 *bl, const i64 *bse, u64 *ol, i64 *ose,
        hipStream_t s)
{
  const int nl = limbs_of (p);
  if (op == MPGFX_MUL)
    hipLaunchKernelGGL ((k_rt_ew<MPGFX_MUL>), dim3 (grid (n)), dim3 (TPB), 0, s, p, nl, n, al, ase, bl, bse, ol, ose);
  else if (op == MPGFX_ADD)
    hipLaunchKernelGGL ((k_rt_ew<MPGFX_ADD>), dim3 (grid (n)), dim3 (TPB), 0, s, p, nl, n, al, ase, bl, bse, ol, ose);
  else
    hipLaunchKernelGGL ((k_rt_ew<MPGFX_SUB>), dim3 (grid (n)), dim3 (TPB), 0, s, p, nl, n, al, ase, bl, bse, ol, ose);
  return launch_status ("rt_elementwise_real");
}

int axpy_real (int p, i64 n, const u64 *al, const i64 *ase, const u64 *xl, const i64 *xse, const u64 *yl,
               const i64 *yse, u64 *ol, i64 *ose, hipStream_t s)
{
  hipLaunchKernelGGL (k_rt_axpy_real, dim3 (grid (n)), dim3 (TPB), 0, s, p, limbs_of (p), n, al, ase, xl, xse, yl, yse,
                      ol, ose);
  return launch_status ("rt_axpy_real");
}

int axpy_cplx (int p, i64 n, const u64 *al, const i64 *ase, const u64 *xl, const i64 *xse, const u64 *yl,
               const i64 *yse, u64 *zl, i64 *zse, hipStream_t s)
{
  hipLaunchKernelGGL (k_rt_axpy_cplx, dim3 (grid (2 * n)), dim3 (TPB), 0, s, p, limbs_of (p), n, al, ase, xl, xse,
                      yl, yse, zl, zse);
  return launch_status ("rt_axpy_complex");
}

/* y = A x, a lane per row (CPLX: two lanes per row, component q = t & 1) */
template <int CPLX>
__global__ __launch_bounds__ (TPB) void k_rt_gemv (int p, int nl, i64 m, i64 k, const u64 *__restrict__ Al,
                                                   const i64 *__restrict__ Ase, const u64 *__restrict__ xl,
                                                   const i64 *__restrict__ xse, u64 *__restrict__ yl,
                                                   i64 *__restrict__ yse)
{
  const i64 nA = m * k, oA = (i64) nl * nA, ox = (i64) nl * k, oy = (i64) nl * m;
  const i64 stride = (i64) gridDim.x * blockDim.x;
  for (i64 t = (i64) blockIdx.x * blockDim.x + threadIdx.x; t < (m << CPLX); t += stride)
    {
      const i64 i = t >> CPLX;
      const int q = CPLX ? (int) (t & 1) : 0;
      rt_val<CW> acc = rt_zero<CW> (1);
      for (i64 kk = 0; kk < k; kk++)
        {
          rt_val<CW> prod;
          if (!CPLX)
            prod = rt_mul<CW> (rt_ld<CW> (Al, Ase, nA, i * k + kk, nl), rt_ld<CW> (xl, xse, k, kk, nl), p);
          else
            {
              const rt_val<CW> ar = rt_ld<CW> (Al, Ase, nA, i * k + kk, nl);
              const rt_val<CW> ai = rt_ld<CW> (Al + oA, Ase + nA, nA, i * k + kk, nl);
              const rt_val<CW> u = rt_ld<CW> (xl + q * ox, xse + q * k, k, kk, nl);
              const rt_val<CW> v = rt_ld<CW> (xl + (1 - q) * ox, xse + (1 - q) * k, k, kk, nl);
              prod = rt_fmm<CW> (ar, u, ai, v, p, q == 0);
            }
          acc = rt_add<CW> (acc, prod, p, false);
        }
      rt_st<CW> (acc, yl + q * oy, yse + q * m, m, i, nl);
    }
}

/* C = A B, a lane per output (CPLX: per output component); a workgroup's
 * lanes share the row i, so A[i,k] is a broadcast load and B[k, j..] a
 * coalesced row segment */
template <int CPLX>
__global__ __launch_bounds__ (TPB) void k_rt_gemm (int p, int nl, i64 m, i64 n, i64 k, const u64 *__restrict__ Al,
                                                   const i64 *__restrict__ Ase, const u64 *__restrict__ Bl,
                                                   const i64 *__restrict__ Bse, u64 *__restrict__ Cl,
                                                   i64 *__restrict__ Cse)
{
  const i64 per = TPB >> CPLX, tiles = (n + per - 1) / per;
  const i64 nA = m * k, nB = k * n, nC = m * n;
  const i64 oA = (i64) nl * nA, oB = (i64) nl * nB, oC = (i64) nl * nC;
  for (i64 t = blockIdx.x; t < m * tiles; t += gridDim.x)
    {
      const i64 i = t / tiles, j = (t % tiles) * per + (threadIdx.x >> CPLX);
      const int q = CPLX ? (int) (threadIdx.x & 1) : 0;
      const bool on = j < n;
      const i64 jj = on ? j : 0;
      rt_val<CW> acc = rt_zero<CW> (1);
      for (i64 kk = 0; kk < k; kk++)
        {
          rt_val<CW> prod;
          if (!CPLX)
            prod = rt_mul<CW> (rt_ld<CW> (Al, Ase, nA, i * k + kk, nl), rt_ld<CW> (Bl, Bse, nB, kk * n + jj, nl), p);
          else
            {
              const rt_val<CW> ar = rt_ld<CW> (Al, Ase, nA, i * k + kk, nl);
              const rt_val<CW> ai = rt_ld<CW> (Al + oA, Ase + nA, nA, i * k + kk, nl);
              const rt_val<CW> u = rt_ld<CW> (Bl + q * oB, Bse + q * nB, nB, kk * n + jj, nl);
              const rt_val<CW> v = rt_ld<CW> (Bl + (1 - q) * oB, Bse + (1 - q) * nB, nB, kk * n + jj, nl);
              prod = rt_fmm<CW> (ar, u, ai, v, p, q == 0);
            }
          acc = rt_add<CW> (acc, prod, p, false);
        }
      if (on) rt_st<CW> (acc, Cl + q * oC, Cse + q * nC, nC, i * n + j, nl);
    }
}

int gemv (int cplx, int p, i64 m, i64 k, const u64 *Al, const i64 *Ase, const u64 *xl, const i64 *xse, u64 *yl,
          i64 *yse, hipStream_t s)
{
  if (cplx)
    hipLaunchKernelGGL ((k_rt_gemv<1>), dim3 (grid (2 * m)), dim3 (TPB), 0, s, p, limbs_of (p), m, k, Al, Ase, xl, xse,
                        yl, yse);
  else
    hipLaunchKernelGGL ((k_rt_gemv<0>), dim3 (grid (m)), dim3 (TPB), 0, s, p, limbs_of (p), m, k, Al, Ase, xl, xse, yl,
                        yse);
  return launch_status ("rt_gemv");
}

int gemm (int cplx, int p, i64 m, i64 n, i64 k, const u64 *Al, const i64 *Ase, const u64 *Bl, const i64 *Bse, u64 *Cl,
          i64 *Cse, hipStream_t s)
{
  const i64 per = TPB >> cplx, tiles = m * ((n + per - 1) / per);
  const unsigned g = (unsigned) (tiles > 65536 ? 65536 : tiles);
  if (cplx)
    hipLaunchKernelGGL ((k_rt_gemm<1>), dim3 (g), dim3 (TPB), 0, s, p, limbs_of (p), m, n, k, Al, Ase, Bl, Bse, Cl, Cse);
  else
    hipLaunchKernelGGL ((k_rt_gemm<0>), dim3 (g), dim3 (TPB), 0, s, p, limbs_of (p), m, n, k, Al, Ase, Bl, Bse, Cl, Cse);
  return launch_status ("rt_gemm");
}

const rt_ops ops = { CW, ew, axpy_real, axpy_cplx, gemv, gemm };

} // namespace

#define MPGFX_RT_CAT2(a, b) a##b
#define MPGFX_RT_CAT(a, b) MPGFX_RT_CAT2 (a, b)
const rt_ops *MPGFX_RT_CAT (rt_ops_cw_, MPGFX_RT_CW) () { return &ops; }

} // namespace mpgfx
