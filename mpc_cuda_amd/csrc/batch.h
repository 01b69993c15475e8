/* batch.h -- device-side structures of the batched persistent execution
 * (mpgfx_batch_*, include/mpgfx.h).
 *
 * One class = all tasks of one (precision, real/complex).  Its work:
 *  phase 1  `items`: contiguous element ranges of one task -- AXPY outputs,
 *           or products of a dot / GEMV, each written (rounded, as the per-op
 *           fold needs) into the product scratch;
 *  phase 2  `chains`: one per dot and per GEMV row, sorted by length
 *           (longest first) and grouped by R = 64 / coop_lanes(pb); a
 *           wavefront folds one group, coop_lanes(pb) lanes per chain.  The
 *           product scratch holds records of N limbs + the sign/exponent word;
 *           step s of the chain at position p is record gbase[p/R] + R*s + p%R,
 *           so a wavefront's loads at step s are contiguous.  The fold units of
 *           all classes (one per group, per part for complex classes) form one
 *           list sorted by length, folded by a single kernel so the longest
 *           chains of every class start first and run side by side.
 */
#ifndef MPGFX_BATCH_H
#define MPGFX_BATCH_H

#include "common.h"

namespace mpgfx {

struct bt_task
{
  int32_t kind, cplx;
  i64 m, n;
  const u64 *in0l; const i64 *in0s;
  const u64 *in1l; const i64 *in1s;
  const u64 *in2l; const i64 *in2s;
  u64 *outl; i64 *outs;
  i64 p0;                                   /* first chain position (dot / GEMV) */
};

struct bt_item
{
  int32_t task, count;
  i64 start;
};

struct bt_chain
{
  int32_t task, row;                        /* task < 0: padding lane */
  i64 len;
};

struct bt_class
{
  int32_t pb, cplx;
  const bt_task *tasks;
  const bt_item *items;
  i64 nitems;
  const bt_chain *chains;
  const i64 *gbase;
  i64 ngroups;
  u64 *scr;                                  /* product records (re part, then im part) */
  i64 scr_n;                                /* records per part */
  unsigned *counters;                       /* [0] phase-1 queue, [1] phase-2 queue */
};

/* one wavefront's work in phase 2: fold group fg (part 0 real / 1 imaginary)
 * of class cls */
struct bt_unit
{
  int32_t cls, part;
  i64 fg;
};

} // namespace mpgfx
#endif
