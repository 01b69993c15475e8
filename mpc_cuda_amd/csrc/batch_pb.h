/* batch_pb.h -- per-precision persistent kernels of the batched execution (see
 * batch.h).  Included by kernels_pb.hip inside its anonymous namespace (PB, NL
 * defined there). */

/* value holder for the real / complex instantiation */
template <int CPLX> struct bval;
template <> struct bval<0>
{
  cu_freal<PB> v;
  __device__ static bval load (const u64 *l, const i64 *s, i64 n, i64 i)
  { bval r; r.v = ldc<PB> (l, s, n, i); return r; }
  __device__ void store (u64 *l, i64 *s, i64 n, i64 i) const { st<PB> (v, l, s, n, i); }
  __device__ static bval mul (const bval &a, const bval &b) { bval r; r.v = cu_fp::cu_fmul<PB> (a.v, b.v); return r; }
  __device__ static bval add (const bval &a, const bval &b) { bval r; r.v = cu_fp::cu_fadd<PB> (a.v, b.v); return r; }
  __device__ void zero () { v.set_zero (); }
};
template <> struct bval<1>
{
  cu_fcomplex<PB> v;
  __device__ static bval load (const u64 *l, const i64 *s, i64 n, i64 i)
  {
    bval r;
    r.v = cu_fcomplex<PB> (ldc<PB> (l, s, n, i), ldc<PB> (l + (i64) NL * n, s + n, n, i));
    return r;
  }
  __device__ void store (u64 *l, i64 *s, i64 n, i64 i) const
  {
    st<PB> (v.re, l, s, n, i);
    st<PB> (v.im, l + (i64) NL * n, s + n, n, i);
  }
  __device__ static bval mul (const bval &a, const bval &b) { bval r; r.v = cu_fp::cu_cmul<PB> (a.v, b.v); return r; }
  __device__ static bval add (const bval &a, const bval &b) { bval r; r.v = cu_fp::cu_cadd<PB> (a.v, b.v); return r; }
  __device__ void zero () { v.re.set_zero (); v.im.set_zero (); }
};

__device__ __forceinline__ i64 bt_scr_index (const bt_class &c, i64 p, i64 s)
{
  return c.gbase[p >> 6] + 64 * s + (p & 63);
}

/* phase 1: workgroups pull items from a queue; lane per value */
template <int CPLX>
__global__ __launch_bounds__ (256) void k_batch_products (bt_class c)
{
  __shared__ unsigned slot;
  typedef bval<CPLX> V;
  for (;;)
    {
      if (threadIdx.x == 0) slot = atomicAdd (&c.counters[0], 1u);
      __syncthreads ();
      const i64 it = slot;
      __syncthreads ();
      if (it >= c.nitems) break;
      const bt_item w = c.items[it];
      const bt_task &t = c.tasks[w.task];
      for (i64 e = threadIdx.x; e < w.count; e += blockDim.x)
        {
          const i64 g = w.start + e;
          if (t.kind == MPGFX_TASK_AXPY)
            {
              V a = V::load (t.in0l, t.in0s, 1, 0);
              V r = V::add (V::mul (a, V::load (t.in1l, t.in1s, t.n, g)), V::load (t.in2l, t.in2s, t.n, g));
              r.store (t.outl, t.outs, t.n, g);
            }
          else if (t.kind == MPGFX_TASK_DOT)
            {
              V p = V::mul (V::load (t.in0l, t.in0s, t.n, g), V::load (t.in1l, t.in1s, t.n, g));
              p.store (c.scr_l, c.scr_s, c.scr_n, bt_scr_index (c, t.p0, g));
            }
          else
            {
              const i64 i = g / t.n, k = g - i * t.n;
              V p = V::mul (V::load (t.in0l, t.in0s, t.m * t.n, g), V::load (t.in1l, t.in1s, t.n, k));
              p.store (c.scr_l, c.scr_s, c.scr_n, bt_scr_index (c, t.p0 + i, k));
            }
        }
    }
}

/* phase 2: a wavefront pulls a group of 64 chains; lane per chain */
template <int CPLX>
__global__ __launch_bounds__ (256) void k_batch_folds (bt_class c)
{
  typedef bval<CPLX> V;
  const int lane = threadIdx.x & 63;
  for (;;)
    {
      unsigned g = 0;
      if (lane == 0) g = atomicAdd (&c.counters[1], 1u);
      g = __shfl (g, 0);
      if ((i64) g >= c.ngroups) break;
      const bt_chain ch = c.chains[(i64) g * 64 + lane];
      if (ch.task >= 0)
        {
          const bt_task &t = c.tasks[ch.task];
          V acc;
          acc.zero ();
          const i64 base = c.gbase[g] + lane;
          for (i64 s = 0; s < ch.len; s++)
            acc = V::add (acc, V::load (c.scr_l, c.scr_s, c.scr_n, base + 64 * s));
          if (t.kind == MPGFX_TASK_DOT) acc.store (t.outl, t.outs, 1, 0);
          else acc.store (t.outl, t.outs, t.m, ch.row);
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
    {
      if (cplx) hipLaunchKernelGGL ((k_batch_folds<1>), dim3 (grid), dim3 (block), 0, s, c);
      else hipLaunchKernelGGL ((k_batch_folds<0>), dim3 (grid), dim3 (block), 0, s, c);
    }
  return launch_status ("batch");
}
