/* dispatch.h -- per-precision operation table of the mpgfx library. */
#ifndef MPGFX_DISPATCH_H
#define MPGFX_DISPATCH_H

#include "common.h"

namespace mpgfx {

struct pb_ops
{
  int pb;
  int (*ew_real) (int op, i64 n, const u64 *, const i64 *, const u64 *, const i64 *, u64 *, i64 *, hipStream_t);
  int (*ew_cplx) (int op, i64 n, const u64 *, const i64 *, const u64 *, const i64 *, u64 *, i64 *, hipStream_t);
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
};

#define MPGFX_DECL(P) const pb_ops *ops_pb_##P ();
MPGFX_PB_LIST (MPGFX_DECL)
#undef MPGFX_DECL

/* table lookup; NULL if pb is not compiled in */
const pb_ops *find_ops (int pb);

} // namespace mpgfx
#endif
