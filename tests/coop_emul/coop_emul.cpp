/* tests/coop_emul/coop_emul.cpp -- host emulation of a 64-lane wavefront
 * running the DPP cooperative fold of mpc_cuda_amd/csrc/coop_dpp.h: one host
 * thread per lane, every cross-lane operation (ballot, row shift, row
 * broadcast, quad permutation, bpermute) an exchange through shared slots
 * between two barriers, with the lane mappings of gfx950 (row_shl:n -- lane i
 * reads lane i+n of its 16-lane row, 0 past the row; row_shr:n -- lane i-n;
 * row_newbcast:n -- lane n of the row; quad_perm; wave_shl1 / wave_shr1 --
 * lane i+1 / i-1 of the wavefront; readlane; ds_bpermute -- the lane given by
 * the address).  Test harness only. */
#include <atomic>
#include <barrier>
#include <cstdint>
#include <cstdio>
#include <thread>
#include <vector>
#include "coop_dpp.h"

using namespace mpgfx::coopd;

namespace {
struct Wave
{
  std::barrier<> bar{64};
  u32 x[64];
  u64 bits[64];
  long ops[64] = {};
  int last[64][8] = {};
};
std::atomic<long> g_phases{0};
Wave *g_cur = nullptr;

struct EmuLanes
{
  Wave *W;
  int l;
  int lane () const { return l; }
  u64 ballot (bool c, int line = __builtin_LINE ()) const
  {
    W->bits[l] = c ? 1u : 0u;
    W->last[l][W->ops[l] & 7] = line;
    W->ops[l]++;
    if (l == 0) g_phases++;
    W->bar.arrive_and_wait ();
    u64 m = 0;
    for (int i = 0; i < 64; i++) m |= (W->bits[i] & 1ull) << i;
    W->bar.arrive_and_wait ();
    return m;
  }
  u32 xchg (u32 v, int src, bool valid, int line) const
  {
    W->x[l] = v;
    W->last[l][W->ops[l] & 7] = line;
    W->ops[l]++;
    if (l == 0) g_phases++;
    W->bar.arrive_and_wait ();
    const u32 r = valid ? W->x[src] : 0u;
    W->bar.arrive_and_wait ();
    return r;
  }
  template <int S> u32 row_shl (u32 v, int line = __builtin_LINE ()) const { return xchg (v, l + S, (l & 15) + S < 16, line); }
  template <int S> u32 row_shr (u32 v, int line = __builtin_LINE ()) const { return xchg (v, l - S, (l & 15) - S >= 0, line); }
  template <int S> u32 row_bcast (u32 v, int line = __builtin_LINE ()) const { return xchg (v, (l & ~15) + S, true, line); }
  template <int P> u32 quad_perm (u32 v, int line = __builtin_LINE ()) const { return xchg (v, (l & ~3) + ((P >> (2 * (l & 3))) & 3), true, line); }
  u32 bperm (int src, u32 v, int line = __builtin_LINE ()) const { return xchg (v, src & 63, true, line); }
  u32 wave_shl1 (u32 v, int line = __builtin_LINE ()) const { return xchg (v, l + 1, l + 1 < 64, line); }
  u32 wave_shr1 (u32 v, int line = __builtin_LINE ()) const { return xchg (v, l - 1, l >= 1, line); }
  u32 readlane (u32 v, int src, int line = __builtin_LINE ()) const { return xchg (v, src & 63, true, line); }
};

/* chains laid out as the GEMV's fold groups: R = 64 / G chains per wavefront,
 * record of chain r of wave q at step s = (q * maxlen_q-block) ... given by
 * rec0[c] and `step` = R; lens[c] <= maxlen[wave] */
template <int PB>
void run (int64_t nchains, const int64_t *lens, const int64_t *rec0, int64_t step, const int64_t *wave_max,
          const u64 *recs, u64 *ol, int64_t *ose)
{
  constexpr int NL = (PB + 63) / 64, G = lanes_for (PB), R = 64 / G;
  const int64_t nw = (nchains + R - 1) / R;
  for (int64_t q = 0; q < nw; q++)
    {
      Wave W;
      g_cur = &W;
      std::vector<std::thread> th;
      for (int l = 0; l < 64; l++)
        th.emplace_back ([&, l] {
          EmuLanes w{&W, l};
          const int r = l / G, j = l & (G - 1);
          const int64_t c = q * R + r;
          const bool real = c < nchains;
          const cv acc = fold<PB> (w, recs, real ? rec0[c] : rec0[q * R], step, real ? lens[c] : 0, wave_max[q]);
          if (real && j >= 1 && j <= NL) ol[(int64_t) (j - 1) * nchains + c] = acc.m;
          if (real && j == 0) ose[c] = (int64_t) ((u64) acc.exp << 1) | (acc.sign < 0 ? 1 : 0);
        });
      for (auto &t : th) t.join ();
    }
}
} // namespace

extern "C" int emul_lanes (int pb) { return lanes_for (pb); }
extern "C" long emul_phases (void) { return g_phases.load (); }
extern "C" void emul_dump (void)
{
  if (!g_cur) return;
  for (int l = 0; l < 64; l += 16)
    {
      fprintf (stderr, "lane %d ops %ld:", l, g_cur->ops[l]);
      for (int k = 8; k >= 1; k--) fprintf (stderr, " %d", g_cur->last[l][(g_cur->ops[l] - k) & 7]);
      fprintf (stderr, "\n");
    }
}

extern "C" int emul_fold (int pb, int64_t nchains, const int64_t *lens, const int64_t *rec0, int64_t step,
                          const int64_t *wave_max, const u64 *recs, u64 *ol, int64_t *ose)
{
#define C(P) if (pb == P) { run<P> (nchains, lens, rec0, step, wave_max, recs, ol, ose); return 0; }
  C (32) C (64) C (96) C (128) C (160) C (192) C (256) C (288) C (320) C (384) C (512) C (1024) C (1056) C (2048)
#undef C
  return -1;
}
