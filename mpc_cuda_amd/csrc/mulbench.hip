/* mulbench.hip -- microbenchmark of the core limb multiply on gfx950 (the role
 * of the reference's tools/bench_mpn_mul.cu: exact N x N word products,
 * XOR-reduced so nothing is optimised away, several formulations checked to
 * agree).
 *
 *   peak     : independent v_mad_u64_u32 chains per lane -> the VALU integer
 *              multiply-add peak of the chip (word products / s)
 *   register : fully unrolled 32x32-word (1024-bit) schoolbook in VGPRs, the
 *              engine the library uses (cu_fp::core::mul_words)
 *   generic  : the same product with a runtime word count (scratch arrays)
 *
 * Prints one JSON object per line.  Usage: mulbench [pairs] [reps]
 */
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include "mpc_cuda/cu_fp_core.cuh"

using namespace cu_fp::core;

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
  fprintf (stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString (e_)); exit (1); } } while (0)

/* 8 independent 64-bit multiply-add chains per lane, `iters` rounds */
__global__ __launch_bounds__ (256) void k_peak (int iters, uint64_t seed, uint64_t *out)
{
  u32 a = (u32) (seed + threadIdx.x), b = (u32) (seed >> 32) ^ blockIdx.x;
  u64 acc[8];
#pragma unroll
  for (int j = 0; j < 8; j++) acc[j] = seed * (j + 1) + threadIdx.x;
  for (int it = 0; it < iters; it++)
    {
#pragma unroll
      for (int j = 0; j < 8; j++) acc[j] = (u64) (a + j) * (b ^ (u32) it) + acc[j];
    }
  u64 x = 0;
#pragma unroll
  for (int j = 0; j < 8; j++) x ^= acc[j];
  out[(size_t) blockIdx.x * blockDim.x + threadIdx.x] = x;
}

/* the same 8 chains with LDSB bytes of LDS per 256-lane workgroup, which caps
 * the workgroups per CU and so the wavefronts per SIMD (96 KB: one, 64 KB:
 * two, 4 KB: eight); the LDS is allocated, never used in the loop */
template <int LDSB>
__global__ __launch_bounds__ (256) void k_mad_issue (int iters, uint64_t seed, uint64_t *out)
{
  __shared__ u32 hog[LDSB / 4];
  if (seed == 0) hog[threadIdx.x] = 1;              /* never taken: keeps the allocation */
  u32 a = (u32) (seed + threadIdx.x), b = (u32) (seed >> 32) ^ blockIdx.x;
  u64 acc[8];
#pragma unroll
  for (int j = 0; j < 8; j++) acc[j] = seed * (j + 1) + threadIdx.x;
  for (int it = 0; it < iters; it++)
    {
#pragma unroll
      for (int j = 0; j < 8; j++) acc[j] = (u64) (a + j) * (b ^ (u32) it) + acc[j];
    }
  u64 x = 0;
#pragma unroll
  for (int j = 0; j < 8; j++) x ^= acc[j];
  if (seed == 0) x += hog[threadIdx.x & 63];
  out[(size_t) blockIdx.x * blockDim.x + threadIdx.x] = x;
}

constexpr int W = 32;        /* 1024-bit operands */

__global__ __launch_bounds__ (256) void k_register (int pairs, int reps, const u32 *A, const u32 *B, uint64_t *out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pairs) return;
  u32 a[W], b[W], p[2 * W];
#pragma unroll
  for (int q = 0; q < W; q++) { a[q] = A[(size_t) q * pairs + i]; b[q] = B[(size_t) q * pairs + i]; }
  u64 x = 0;
  for (int r = 0; r < reps; r++)
    {
      mul_words<W> (p, a, b);
#pragma unroll
      for (int q = 0; q < 2 * W; q++) x ^= (u64) p[q] << (q & 1 ? 32 : 0);
      a[r & (W - 1)] ^= (u32) x;                    /* carry a dependency across reps */
    }
  out[i] = x;
}

__global__ __launch_bounds__ (256) void k_generic (int pairs, int reps, int w, const u32 *A, const u32 *B, uint64_t *out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pairs) return;
  u32 a[64], b[64], p[128];
  for (int q = 0; q < w; q++) { a[q] = A[(size_t) q * pairs + i]; b[q] = B[(size_t) q * pairs + i]; }
  u64 x = 0;
  for (int r = 0; r < reps; r++)
    {
      for (int q = 0; q < 2 * w; q++) p[q] = 0;
      for (int j = 0; j < w; j++)
        {
          u32 c = 0;
          for (int q = 0; q < w; q++)
            {
              u64 t = (u64) a[q] * b[j] + p[q + j] + c;
              p[q + j] = (u32) t; c = (u32) (t >> 32);
            }
          p[w + j] = c;
        }
      for (int q = 0; q < 2 * w; q++) x ^= (u64) p[q] << (q & 1 ? 32 : 0);
      a[r & (w - 1)] ^= (u32) x;
    }
  out[i] = x;
}

/* INT8 MFMA peak: 8 independent v_mfma_i32_16x16x64_i8 accumulator chains
 * per wavefront (16*16*64*2 = 32768 ops each), operands fixed in registers */
typedef int v4i __attribute__ ((ext_vector_type (4)));
__global__ __launch_bounds__ (256) void k_mfma_peak (int iters, int seed, int *out)
{
  v4i a = {seed + (int) threadIdx.x, seed ^ 0x1234567, seed * 3, seed - 7};
  v4i b = {seed ^ (int) blockIdx.x, seed + 11, seed * 5, seed ^ 0x7654321};
  v4i acc[8];
#pragma unroll
  for (int j = 0; j < 8; j++) acc[j] = (v4i) {j, 0, 0, 0};
  for (int it = 0; it < iters; it++)
    {
#pragma unroll
      for (int j = 0; j < 8; j++) acc[j] = __builtin_amdgcn_mfma_i32_16x16x64_i8 (a, b, acc[j], 0, 0, 0);
    }
  int x = 0;
#pragma unroll
  for (int j = 0; j < 8; j++) x ^= acc[j].x ^ acc[j].y ^ acc[j].z ^ acc[j].w;
  out[(size_t) blockIdx.x * blockDim.x + threadIdx.x] = x;
}

int main (int argc, char **argv)
{
  int pairs = argc > 1 ? atoi (argv[1]) : (1 << 16);
  int reps = argc > 2 ? atoi (argv[2]) : 64;
  hipDeviceProp_t prop;
  HIPCHK (hipGetDeviceProperties (&prop, 0));
  int cus = prop.multiProcessorCount;
  double clk_ghz = prop.clockRate / 1e6;
  hipEvent_t e0, e1;
  HIPCHK (hipEventCreate (&e0)); HIPCHK (hipEventCreate (&e1));
  float ms;

  /* ---- peak ---- */
  {
    int blocks = cus * 16, iters = 1 << 14;
    uint64_t *out;
    HIPCHK (hipMalloc (&out, (size_t) blocks * 256 * 8));
    hipLaunchKernelGGL (k_peak, dim3 (blocks), dim3 (256), 0, 0, 16, 12345ull, out);
    HIPCHK (hipEventRecord (e0));
    hipLaunchKernelGGL (k_peak, dim3 (blocks), dim3 (256), 0, 0, iters, 12345ull, out);
    HIPCHK (hipEventRecord (e1));
    HIPCHK (hipEventSynchronize (e1));
    HIPCHK (hipEventElapsedTime (&ms, e0, e1));
    double mads = (double) blocks * 256 * iters * 8;
    printf ("{\"bench\": \"mad_u64_u32_peak\", \"cus\": %d, \"clock_ghz\": %.3f, \"ms\": %.4f, "
            "\"word_products_per_s\": %.4e, \"per_cu_per_clk\": %.3f}\n",
            cus, clk_ghz, ms, mads / (ms * 1e-3), mads / (ms * 1e-3) / cus / (clk_ghz * 1e9));
    HIPCHK (hipFree (out));
  }
  /* ---- v_mad_u64_u32 issue rate at 1, 2 and 8 wavefronts per SIMD ---- */
  {
    const int iters = 1 << 14;
    uint64_t *out;
    HIPCHK (hipMalloc (&out, (size_t) cus * 8 * 256 * 8));
    auto one = [&] (int waves, int blocks, auto kern) {
      hipLaunchKernelGGL (kern, dim3 (blocks), dim3 (256), 0, 0, 16, 12345ull, out);
      HIPCHK (hipEventRecord (e0));
      hipLaunchKernelGGL (kern, dim3 (blocks), dim3 (256), 0, 0, iters, 12345ull, out);
      HIPCHK (hipEventRecord (e1));
      HIPCHK (hipEventSynchronize (e1));
      HIPCHK (hipEventElapsedTime (&ms, e0, e1));
      double per_simd = (double) blocks * 4 * iters * 8 / (cus * 4.0);     /* wave-instructions per SIMD */
      printf ("{\"bench\": \"mad_u64_u32_issue\", \"waves_per_simd\": %d, \"ms\": %.4f, "
              "\"cycles_per_mad_per_simd\": %.2f}\n", waves, ms, (ms * 1e-3) * clk_ghz * 1e9 / per_simd);
    };
    one (1, cus, k_mad_issue<98304>);
    one (2, cus * 2, k_mad_issue<65536>);
    one (8, cus * 8, k_mad_issue<4096>);
    HIPCHK (hipFree (out));
  }
  /* ---- INT8 MFMA peak ---- */
  {
    int blocks = cus * 8, iters = 1 << 12;
    int *out;
    HIPCHK (hipMalloc (&out, (size_t) blocks * 256 * 4));
    hipLaunchKernelGGL (k_mfma_peak, dim3 (blocks), dim3 (256), 0, 0, 16, 5, out);
    HIPCHK (hipEventRecord (e0));
    hipLaunchKernelGGL (k_mfma_peak, dim3 (blocks), dim3 (256), 0, 0, iters, 5, out);
    HIPCHK (hipEventRecord (e1));
    HIPCHK (hipEventSynchronize (e1));
    HIPCHK (hipEventElapsedTime (&ms, e0, e1));
    double ops = (double) blocks * 4 * iters * 8 * 32768.0;
    printf ("{\"bench\": \"mfma_i32_16x16x64_i8_peak\", \"ms\": %.4f, \"ops_per_s\": %.4e, "
            "\"cycles_per_mfma_per_simd\": %.2f}\n",
            ms, ops / (ms * 1e-3), (ms * 1e-3) * clk_ghz * 1e9 * cus * 4 / (ops / 32768.0));
    HIPCHK (hipFree (out));
  }
  /* ---- 1024-bit products: register-resident vs generic ---- */
  {
    size_t bytes = (size_t) W * pairs * 4;
    u32 *hA = (u32 *) malloc (bytes), *hB = (u32 *) malloc (bytes);
    uint64_t s = 0x9e3779b97f4a7c15ull;
    for (size_t q = 0; q < (size_t) W * pairs; q++)
      {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17; hA[q] = (u32) s;
        s ^= s << 13; s ^= s >> 7; s ^= s << 17; hB[q] = (u32) s;
      }
    u32 *A, *B; uint64_t *o1, *o2;
    HIPCHK (hipMalloc (&A, bytes)); HIPCHK (hipMalloc (&B, bytes));
    HIPCHK (hipMalloc (&o1, (size_t) pairs * 8)); HIPCHK (hipMalloc (&o2, (size_t) pairs * 8));
    HIPCHK (hipMemcpy (A, hA, bytes, hipMemcpyHostToDevice));
    HIPCHK (hipMemcpy (B, hB, bytes, hipMemcpyHostToDevice));
    int blocks = (pairs + 255) / 256;
    hipLaunchKernelGGL (k_register, dim3 (blocks), dim3 (256), 0, 0, pairs, 2, A, B, o1);
    HIPCHK (hipEventRecord (e0));
    hipLaunchKernelGGL (k_register, dim3 (blocks), dim3 (256), 0, 0, pairs, reps, A, B, o1);
    HIPCHK (hipEventRecord (e1));
    HIPCHK (hipEventSynchronize (e1));
    HIPCHK (hipEventElapsedTime (&ms, e0, e1));
    double prods = (double) pairs * reps;
    printf ("{\"bench\": \"mul1024_register\", \"pairs\": %d, \"reps\": %d, \"ms\": %.4f, \"products_per_s\": %.4e, "
            "\"word_products_per_s\": %.4e}\n", pairs, reps, ms, prods / (ms * 1e-3), prods * W * W / (ms * 1e-3));
    hipLaunchKernelGGL (k_generic, dim3 (blocks), dim3 (256), 0, 0, pairs, 2, W, A, B, o2);
    HIPCHK (hipEventRecord (e0));
    hipLaunchKernelGGL (k_generic, dim3 (blocks), dim3 (256), 0, 0, pairs, reps, W, A, B, o2);
    HIPCHK (hipEventRecord (e1));
    HIPCHK (hipEventSynchronize (e1));
    HIPCHK (hipEventElapsedTime (&ms, e0, e1));
    printf ("{\"bench\": \"mul1024_generic\", \"pairs\": %d, \"reps\": %d, \"ms\": %.4f, \"products_per_s\": %.4e, "
            "\"word_products_per_s\": %.4e}\n", pairs, reps, ms, prods / (ms * 1e-3), prods * W * W / (ms * 1e-3));
    uint64_t *h1 = (uint64_t *) malloc ((size_t) pairs * 8), *h2 = (uint64_t *) malloc ((size_t) pairs * 8);
    HIPCHK (hipMemcpy (h1, o1, (size_t) pairs * 8, hipMemcpyDeviceToHost));
    HIPCHK (hipMemcpy (h2, o2, (size_t) pairs * 8, hipMemcpyDeviceToHost));
    long bad = 0;
    for (int q = 0; q < pairs; q++) bad += h1[q] != h2[q];
    printf ("{\"bench\": \"mul1024_agree\", \"mismatches\": %ld}\n", bad);
    return bad ? 1 : 0;
  }
}
