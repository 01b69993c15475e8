/* batch_folds.hip -- phase 2 of the batched execution: the per-op folds of
 * every class in one persistent kernel.
 *
 * The planner lists the fold units of all classes (a group of R chains of one
 * class, and for complex classes its real or imaginary parts: complex
 * addition is componentwise, so the two parts fold independently) sorted by
 * their longest chain; a wavefront pulls units from one atomic queue and folds
 * each with the cooperative addition at the unit's precision (a switch over
 * the compiled precisions).  One kernel for all classes means the longest
 * chains of every class start at once instead of one class after another. */
#include "common.h"
#include "batch.h"
#include "coop_dpp.h"

namespace mpgfx {
namespace {

#include "coop.h"

template <int PB>
__device__ __forceinline__ void fold_unit (const bt_class &c, i64 fg, int part)
{
  constexpr int NL = (PB + 63) / 64, G = coop_lanes (PB), R = 64 / G;
  const int lane = (int) (threadIdx.x & 63), r = lane / G;
  const bt_chain ch = c.chains[fg * R + r];
  const i64 max_len = c.chains[fg * R].len;                     /* longest first */
  const i64 len = ch.task >= 0 ? ch.len : 0;
  const coopd::cv a = coopd::fold<PB> (coopd::DevLanes{}, c.scr + (i64) part * c.scr_n * (NL + 1), c.gbase[fg] + r,
                                       R, len, max_len);
  cval acc;
  acc.m = a.m;
  acc.exp = a.exp;
  acc.sign = a.sign;
  if (ch.task >= 0)
    {
      const bt_task &t = c.tasks[ch.task];
      const i64 nout = (t.kind == MPGFX_TASK_DOT) ? 1 : t.m;
      const i64 idx = (t.kind == MPGFX_TASK_DOT) ? 0 : ch.row;
      coop_store<PB> (acc, t.outl + (i64) part * NL * nout, t.outs + (i64) part * nout, nout, idx);
    }
}

__global__ __launch_bounds__ (256) void k_batch_folds_all (const bt_class *__restrict__ cls,
                                                           const bt_unit *__restrict__ units, i64 nunits,
                                                           unsigned *__restrict__ counter)
{
  const int lane = (int) (threadIdx.x & 63);
  for (;;)
    {
      unsigned w = 0;
      if (lane == 0) w = atomicAdd (counter, 1u);
      w = __shfl (w, 0);
      if ((i64) w >= nunits) break;
      const bt_unit u = units[w];
      const bt_class &c = cls[u.cls];
      switch (c.pb)
        {
#define FOLD_CASE(P) case P: fold_unit<P> (c, u.fg, u.part); break;
          MPGFX_PB_LIST (FOLD_CASE)
#undef FOLD_CASE
          default: break;
        }
    }
}

__global__ void k_zero_counter (unsigned *counter)
{
  if (threadIdx.x == 0) *counter = 0u;
}

} // namespace

int batch_folds_launch (const bt_class *cls, const bt_unit *units, i64 nunits, unsigned *counter, int grid, int block,
                        hipStream_t s)
{
  hipLaunchKernelGGL (k_zero_counter, dim3 (1), dim3 (64), 0, s, counter);
  if (nunits > 0)
    hipLaunchKernelGGL (k_batch_folds_all, dim3 (grid), dim3 (block), 0, s, cls, units, nunits, counter);
  return launch_status ("batch folds");
}

} // namespace mpgfx
