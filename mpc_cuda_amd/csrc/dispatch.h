/* dispatch.h -- per-precision operation table of the mpgfx library. */
#ifndef MPGFX_DISPATCH_H
#define MPGFX_DISPATCH_H

#include "common.h"
#include "exact.h"
#include "batch.h"

namespace mpgfx {

struct pb_ops
{
  int pb;
  int (*ew_real) (int op, i64 n, const u64 *, const i64 *, const u64 *, const i64 *, u64 *, i64 *, hipStream_t);
  int (*ew_cplx) (int op, i64 n, const u64 *, const i64 *, const u64 *, const i64 *, u64 *, i64 *, hipStream_t);
  int (*sqrt_real) (i64 n, const u64 *, const i64 *, u64 *, i64 *, hipStream_t);
  int (*axpy_real) (i64 n, const u64 *, const i64 *, const u64 *, const i64 *, const u64 *, const i64 *, u64 *,
                    i64 *, hipStream_t);
  int (*axpy_cplx) (i64 n, const u64 *, const i64 *, const u64 *, const i64 *, const u64 *, const i64 *, u64 *,
                    i64 *, hipStream_t);
  int (*gemm_perop) (i64 m, i64 n, i64 k, const u64 *, const i64 *, const u64 *, const i64 *, u64 *, i64 *,
                     hipStream_t);
  int (*gemm_cplx) (i64 m, i64 n, i64 k, const u64 *, const i64 *, const u64 *, const i64 *, u64 *, i64 *,
                    hipStream_t);
  int (*gemv) (i64 m, i64 k, const u64 *, const i64 *, const u64 *, const i64 *, u64 *, i64 *, hipStream_t);
  int (*pack) (int cplx, i64 n, const void *aos, u64 *, i64 *, hipStream_t);
  int (*unpack) (int cplx, i64 n, const u64 *, const i64 *, void *aos, hipStream_t);
  /* exact-dot GEMM stages (exact_pb.h); with a device plan `pl` the digit
   * count, spreads and slots are read from it on the device */
  int (*exact_digits) (i64 lines, i64 len, i64 ls, i64 ps, const u64 *l, const i64 *se, i64 nv, const i64 *emax,
                       i64 spread, int S, int Spad, i64 Lpad, i64 kp, int8_t *D, const exact_dplan *pl, int which,
                       hipStream_t);
  int (*exact_round) (i64 m, i64 n, i64 np, i64 nE, int slots, const i64 *S64, const i64 *emaxA, const i64 *emaxB,
                      i64 sA, i64 sB, u64 *Cl, i64 *Cse, const exact_dplan *pl, hipStream_t);
  int (*exact_fallback) (i64 m, i64 n, i64 k, i64 r0, i64 rows, const u64 *Al, const i64 *Ase, const u64 *Bl,
                         const i64 *Bse, const i64 *eminA, const i64 *eminB, int nw, i64 *wsp, u64 *Cl, i64 *Cse,
                         hipStream_t);
  int (*exact_fallback_dev) (i64 m, i64 n, i64 k, const u64 *Al, const i64 *Ase, const u64 *Bl, const i64 *Bse,
                             const i64 *eminA, const i64 *eminB, i64 *wsp, u64 *Cl, i64 *Cse, const exact_dplan *pl,
                             hipStream_t);
  /* batched persistent kernels (batch_pb.h): phase 0 reset, 1 products */
  int (*batch_launch) (int phase, int cplx, const bt_class &c, int grid, int block, hipStream_t);
};

#define MPGFX_DECL(P) const pb_ops *ops_pb_##P ();
MPGFX_PB_LIST (MPGFX_DECL)
#undef MPGFX_DECL

/* table lookup; NULL if pb is not compiled in */
const pb_ops *find_ops (int pb);

/* phase 2 of the batched execution: the folds of all classes (batch_folds.hip);
 * resets `counter` first */
int batch_folds_launch (const bt_class *cls, const bt_unit *units, i64 nunits, unsigned *counter, int grid, int block,
                        hipStream_t s);

/* exact-dot GEMM (exact_gemm.hip) */
size_t exact_workspace_bound (int pb, i64 m, i64 n, i64 k);
int gemm_exact (const pb_ops *o, int pb, i64 m, i64 n, i64 k, const u64 *Al, const i64 *Ase, const u64 *Bl,
                const i64 *Bse, u64 *Cl, i64 *Cse, void *ws, size_t ws_bytes, hipStream_t s, bool noalloc);
/* device address of this GPU's static scratch of the no-allocation entry point */
int exact_noalloc_scratch (void **out);
int gemm_exact_complex (const pb_ops *o, int pb, i64 m, i64 n, i64 k, const u64 *Al, const i64 *Ase, const u64 *Bl,
                        const i64 *Bse, u64 *Cl, i64 *Cse, hipStream_t s);

} // namespace mpgfx
#endif
