/* batch_pb.h -- per-precision kernels of the batched execution (see batch.h):
 * the counter reset and the phase-1 products.  Included by kernels_pb.hip
 * inside its anonymous namespace (PB, NL, COOP_R defined there). */

/* record of step s of the chain at position p (records of NL+1 words) */
__device__ __forceinline__ i64 bt_scr_index (const bt_class &c, i64 p, i64 s)
{
  return c.gbase[p / COOP_R] + COOP_R * s + p % COOP_R;
}

/* one part (0 real, 1 imaginary) of a product: for CPLX, the two lanes of an
 * element run the same code, RNDN(ar*ur -/+ ai*vi) with the operands swapped
 * and the operation chosen per lane (MPC's correctly rounded components) */
template <int CPLX>
__device__ __forceinline__ cu_freal<PB> bt_prod (const u64 *al, const i64 *as, i64 an, i64 ai_, const u64 *bl,
                                                 const i64 *bs, i64 bn, i64 bi, int part)
{
  if (!CPLX) return cu_fp::cu_fmul<PB> (ldc<PB> (al, as, an, ai_), ldc<PB> (bl, bs, bn, bi));
  const cu_freal<PB> ar = ldc<PB> (al, as, an, ai_), ai = ldc<PB> (al + (i64) NL * an, as + an, an, ai_);
  const cu_freal<PB> br = ldc<PB> (bl, bs, bn, bi), bim = ldc<PB> (bl + (i64) NL * bn, bs + bn, bn, bi);
  const cu_freal<PB> u = part ? bim : br, v = part ? br : bim;
  return cu_fp::core::fmm<PB> (ar, u, ai, v, part == 0);
}

/* phase 1: workgroups pull items from a queue; lane per real value, two
 * lanes (real and imaginary part) per complex value */
template <int CPLX>
__global__ __launch_bounds__ (256) void k_batch_products (bt_class c)
{
  __shared__ unsigned slot;
  for (;;)
    {
      if (threadIdx.x == 0) slot = atomicAdd (&c.counters[0], 1u);
      __syncthreads ();
      const i64 it = slot;
      __syncthreads ();
      if (it >= c.nitems) break;
      const bt_item w = c.items[it];
      const bt_task &t = c.tasks[w.task];
      for (i64 e = threadIdx.x; e < (w.count << CPLX); e += blockDim.x)
        {
          const int part = CPLX ? (int) (e & 1) : 0;
          const i64 g = w.start + (e >> CPLX);
          if (t.kind == MPGFX_TASK_AXPY)
            {
              const cu_freal<PB> p = bt_prod<CPLX> (t.in0l, t.in0s, 1, 0, t.in1l, t.in1s, t.n, g, part);
              const cu_freal<PB> y = ldc<PB> (t.in2l + (i64) part * NL * t.n, t.in2s + (i64) part * t.n, t.n, g);
              st<PB> (cu_fp::cu_fadd<PB> (p, y), t.outl + (i64) part * NL * t.n, t.outs + (i64) part * t.n, t.n, g);
            }
          else
            {
              i64 q;
              cu_freal<PB> p;
              if (t.kind == MPGFX_TASK_DOT)
                {
                  p = bt_prod<CPLX> (t.in0l, t.in0s, t.n, g, t.in1l, t.in1s, t.n, g, part);
                  q = bt_scr_index (c, t.p0, g);
                }
              else
                {
                  const i64 i = g / t.n, k = g - i * t.n;
                  p = bt_prod<CPLX> (t.in0l, t.in0s, t.m * t.n, g, t.in1l, t.in1s, t.n, k, part);
                  q = bt_scr_index (c, t.p0 + i, k);
                }
              st_rec<PB> (p, c.scr + ((i64) part * c.scr_n + q) * (NL + 1));
            }
        }
    }
}

__global__ void k_batch_reset (unsigned *counters)
{
  if (threadIdx.x < 2) counters[threadIdx.x] = 0u;
}

int batch_launch (int phase, int cplx, const bt_class &c, int grid, int block, hipStream_t s)
{
  if (phase == 0)
    hipLaunchKernelGGL (k_batch_reset, dim3 (1), dim3 (64), 0, s, c.counters);
  else if (phase == 1)
    {
      if (cplx) hipLaunchKernelGGL ((k_batch_products<1>), dim3 (grid), dim3 (block), 0, s, c);
      else hipLaunchKernelGGL ((k_batch_products<0>), dim3 (grid), dim3 (block), 0, s, c);
    }
  else
    return MPGFX_EINVAL;            /* the folds of all classes run in one kernel (batch_folds.hip) */
  return launch_status ("batch");
}
