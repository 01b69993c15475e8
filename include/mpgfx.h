/* mpgfx.h -- C ABI of the MI355X (gfx950) fixed-precision multiple-precision
 * library: batched RNDN arithmetic, AXPY, GEMV and GEMM on cu_fp::cu_freal<PB>
 * / cu_fcomplex<PB> values held in device memory.
 *
 * The reference (tkouya/mpc_cuda) exposes this path only as a header-only C++
 * template API plus example __global__ kernels; each entry point below names
 * the reference kernel or function whose behaviour it provides.  The template
 * API itself is available unchanged in include/mpc_cuda/cu_freal.cuh and
 * include/mpc_cuda/cu_fcomplex.cuh.
 *
 * Conventions
 *  - All array pointers are DEVICE pointers owned by the caller.  `stream` is a
 *    hipStream_t passed as void* (NULL = the default stream).  Every call is
 *    stream-ordered and asynchronous; it returns after enqueueing work.
 *  - Precision pb is in bits, a multiple of 32; mpgfx_supported_pb() tells
 *    which precisions are compiled in.
 *  - Value arrays are structure-of-arrays ("MP arrays").  A real MP array of
 *    n values at precision pb is
 *        limbs : N = ceil(pb/64) planes of n uint64; limb l of value i is
 *                limbs[l*n + i] (little-endian limbs, MPFR-compatible,
 *                bit 63 of limb N-1 set, the low 64N-pb bits zero)
 *        se    : n int64 sign/exponent words, se = exp*2 + (sign<0 ? 1 : 0);
 *                exp is the MPFR exponent, value = sign*M*2^(exp-64N); a zero
 *                value has exp == MPGFX_ZERO_EXP and all limbs zero.
 *    A complex MP array of n values is the real MP array of the n real parts
 *    followed by the real MP array of the n imaginary parts: limbs has 2*N*n
 *    entries (re planes, then im planes), se has 2*n entries (re, then im).
 *  - Matrices are MP arrays in row-major element order: element (i, j) of an
 *    m x n matrix is value i*n + j of an MP array of m*n values.
 *  - Results are bit-exact with MPFR (RNDN) / MPC (MPC_RNDNN) for the same
 *    operation order; zeros follow the reference's conventions (exact
 *    cancellation and products with a zero factor give +0).
 *  - Return: MPGFX_OK, or an error code; nothing is launched on error.
 */
#ifndef MPGFX_H
#define MPGFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPGFX_ZERO_EXP (-0x7fffffffffffffLL)

enum mpgfx_status {
  MPGFX_OK = 0,
  MPGFX_EINVAL = 1,        /* bad size, null pointer, bad mode */
  MPGFX_EUNSUPPORTED = 2,  /* precision not compiled in */
  MPGFX_EHIP = 3,          /* HIP launch or runtime error */
  MPGFX_ENOMEM = 4         /* workspace allocation failed */
};

enum mpgfx_op { MPGFX_MUL = 0, MPGFX_ADD = 1, MPGFX_SUB = 2, MPGFX_DIV = 3 };
enum mpgfx_gemm_mode { MPGFX_GEMM_PEROP = 0, MPGFX_GEMM_EXACT = 1 };

/* Library version (major*10000 + minor*100 + patch). */
int mpgfx_version (void);

/* 1 if precision pb is compiled into the elementwise / AXPY / GEMV / GEMM
 * kernels, else 0. */
int mpgfx_supported_pb (int pb);

/* Text of the last HIP error seen by this thread (empty if none). */
const char *mpgfx_last_error (void);

/* o[i] = a[i] OP b[i] for real MP arrays of n values, OP in {mul, add, sub},
 * each RNDN.  Replaces the reference's per-element use of cu_fmul / cu_fadd /
 * cu_fsub: tools/test_freal.cu:41-48 (kern<PB>), cu_freal.cuh:203-213, 280-292.
 * OP = div: a[i] / b[i] as cu_fdiv (cu_fmath.cuh:68-75): faithful, not
 * correctly rounded (Newton refinement at pb + 128 bits, bit-identical to the
 * reference); a zero divisor gives +0. */
int mpgfx_elementwise_real (int pb, int op, int64_t n,
                            const uint64_t *a_limbs, const int64_t *a_se,
                            const uint64_t *b_limbs, const int64_t *b_se,
                            uint64_t *o_limbs, int64_t *o_se, void *stream);

/* o[i] = sqrt(x[i]) for a real MP array, as cu_sqrt (cu_fmath.cuh:78-100):
 * faithful, bit-identical to the reference; +0 for zero and negative x. */
int mpgfx_sqrt_real (int pb, int64_t n, const uint64_t *x_limbs, const int64_t *x_se,
                     uint64_t *o_limbs, int64_t *o_se, void *stream);

/* Complex counterpart, MPC_RNDNN per component: cu_cmul / cu_cadd / cu_csub
 * applied elementwise (OP = div is not available); replaces
 * tools/test_fcomplex.cu:147-153 (kern<PB>), cu_fcomplex.cuh:115-124. */
int mpgfx_elementwise_complex (int pb, int op, int64_t n,
                               const uint64_t *a_limbs, const int64_t *a_se,
                               const uint64_t *b_limbs, const int64_t *b_se,
                               uint64_t *o_limbs, int64_t *o_se, void *stream);

/* o[i] = RNDN(RNDN(a*x[i]) + y[i]); a is a real MP array of ONE value.
 * Replaces demos/bench_3way.cu:44-50 (freal_axpy<P>) and
 * demos/sample_fixed.cu:20-30 (axpy<PB>) on full-limb values. */
int mpgfx_axpy_real (int pb, int64_t n,
                     const uint64_t *a_limbs, const int64_t *a_se,
                     const uint64_t *x_limbs, const int64_t *x_se,
                     const uint64_t *y_limbs, const int64_t *y_se,
                     uint64_t *o_limbs, int64_t *o_se, void *stream);

/* z[i] = a*x[i] + y[i] for complex MP arrays (a: ONE complex value), the
 * product and the sum each MPC_RNDNN.  Replaces demos/bench_3way.cu:51-59
 * (fcplx_axpy<P>) and tools/axpy_fast_complex.cu:32-40 (axpyc). */
int mpgfx_axpy_complex (int pb, int64_t n,
                        const uint64_t *a_limbs, const int64_t *a_se,
                        const uint64_t *x_limbs, const int64_t *x_se,
                        const uint64_t *y_limbs, const int64_t *y_se,
                        uint64_t *z_limbs, int64_t *z_se, void *stream);

/* y[i] = fold over k = 0..k-1 of acc = RNDN(acc + RNDN(A[i,k]*x[k])), acc
 * starting at +0 (per-operation rounding in sequential k order).  A is an
 * m x k real MP matrix, x a real MP array of k, y a real MP array of m.
 * Replaces tools/linalg_fast.cu:37-51 (matvec_kernel). */
int mpgfx_gemv_real (int pb, int64_t m, int64_t k,
                     const uint64_t *A_limbs, const int64_t *A_se,
                     const uint64_t *x_limbs, const int64_t *x_se,
                     uint64_t *y_limbs, int64_t *y_se, void *stream);

/* Complex per-operation GEMV: y[i] = fold over k = 0..k-1 of acc = acc +
 * A[i,k]*x[k], acc starting at +0, every product and sum MPC_RNDNN (each
 * component rounded once).  A is an m x k complex MP matrix, x a complex MP
 * array of k, y a complex MP array of m.  Replaces demos/matvec_mpc.cu:42-67
 * (cdot_row, matvec_kernel) on the fixed type. */
int mpgfx_gemv_complex (int pb, int64_t m, int64_t k,
                        const uint64_t *A_limbs, const int64_t *A_se,
                        const uint64_t *x_limbs, const int64_t *x_se,
                        uint64_t *y_limbs, int64_t *y_se, void *stream);

/* Bytes of device workspace with which mpgfx_gemm_real runs this shape
 * asynchronously (the worst case over exponent spreads on the MFMA path; 0
 * for the per-operation mode).  No counterpart in the reference (its kernels
 * need no workspace). */
int mpgfx_gemm_workspace_size (int pb, int mode, int64_t m, int64_t n, int64_t k,
                               size_t *bytes);

/* C = A*B, A m x k, B k x n, C m x n (real MP matrices).
 *  mode MPGFX_GEMM_PEROP: C[i,j] = fold over k of RNDN(acc + RNDN(A[i,k]*B[k,j])),
 *    acc from +0, k ascending -- the reference's matmul, tools/linalg_fast.cu:53-68
 *    (matmul_kernel).
 *  mode MPGFX_GEMM_EXACT: C[i,j] = RNDN(sum_k A[i,k]*B[k,j]) with the sum exact
 *    and ONE rounding (the semantics of mpfr_dot: exact products then
 *    mpfr_sum, src/mpfr_cuda/dot.c:47-54); slice products run on INT8 MFMA.
 * `ws`/`ws_bytes`: caller workspace.  With at least mpgfx_gemm_workspace_size()
 * bytes the exact mode is asynchronous: the digit count and the path are
 * decided on the device from the exponents, the call neither synchronises
 * `stream` nor allocates, and it can be captured in a HIP graph; its
 * data-dependent outcome is then the int64 at the start of `ws` once the
 * stream has run the call: MPGFX_OK, or MPGFX_EUNSUPPORTED when the exponent
 * spread needs more than 2^22 bits of accumulator (C is then +0).  (This mode
 * leaves out digit pairs far below the rounding position and recomputes
 * exactly the elements whose rounding they could change: C is the same bit
 * for bit.)  With a smaller workspace, or ws == NULL, it synchronises `stream` once (after the
 * exponent scan), allocates what the workspace lacks stream-ordered, and
 * returns that outcome directly. */
int mpgfx_gemm_real (int pb, int mode, int64_t m, int64_t n, int64_t k,
                     const uint64_t *A_limbs, const int64_t *A_se,
                     const uint64_t *B_limbs, const int64_t *B_se,
                     uint64_t *C_limbs, int64_t *C_se,
                     void *ws, size_t ws_bytes, void *stream);

/* The exact mode of mpgfx_gemm_real (C = RNDN(sum_k A[i,k]*B[k,j]), one
 * rounding per output) without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc).  Its workspace is a static device array
 * of 256 MiB per GPU, held from the library's first use on that GPU until the
 * process ends; the results are those of mpgfx_gemm_real(MPGFX_GEMM_EXACT).
 * A call whose workspace would exceed the array returns MPGFX_ENOMEM with the
 * sizes in mpgfx_last_error() (the size depends on the exponent spread, so
 * the check follows the exponent scan, which writes only the array); pb 256,
 * m = n = k = 1024 with exponent spreads below 8 needs 234 MB.  Like the exact
 * mode without a full-size workspace, it synchronises `stream` once and
 * returns the data-dependent outcome directly.  Every call on one GPU uses the same
 * array: calls that overlap on one GPU (other host threads, other streams)
 * race and must be serialised by the caller. */
int mpgfx_gemm_exact_noalloc (int pb, int64_t m, int64_t n, int64_t k,
                              const uint64_t *A_limbs, const int64_t *A_se,
                              const uint64_t *B_limbs, const int64_t *B_se,
                              uint64_t *C_limbs, int64_t *C_se, void *stream);

/* Complex per-operation GEMM (C = A*B, acc = +0, acc = acc + A[i,k]*B[k,j] with
 * MPC_RNDNN products and sums, k ascending).  Counterpart of the reference's
 * complex matmul demo order (demos/matmul_mpc.cu:33-52) on the fixed type. */
int mpgfx_gemm_complex (int pb, int64_t m, int64_t n, int64_t k,
                        const uint64_t *A_limbs, const int64_t *A_se,
                        const uint64_t *B_limbs, const int64_t *B_se,
                        uint64_t *C_limbs, int64_t *C_se, void *stream);

/* Complex exact GEMM: each component of C = A*B is the exact sum of the exact
 * products, rounded once (Re = RNDN(sum ar*br - ai*bi), Im = RNDN(sum ar*bi +
 * ai*br)).  No counterpart in the reference: the exact mode of
 * mpgfx_gemm_real (mpfr_dot semantics, src/mpfr_cuda/dot.c:47-54) carried to
 * the complex matmul of demos/matmul_mpc.cu:33-52.  Runs as two real exact
 * GEMMs over 2k terms on INT8 MFMA in their asynchronous mode, on one
 * stream-ordered allocation (the assembled operands and a full-size
 * workspace), and synchronises the stream once at its end to return the
 * data-dependent outcome (MPGFX_EUNSUPPORTED for a spread beyond 2^22
 * accumulator bits). */
int mpgfx_gemm_complex_exact (int pb, int64_t m, int64_t n, int64_t k,
                              const uint64_t *A_limbs, const int64_t *A_se,
                              const uint64_t *B_limbs, const int64_t *B_se,
                              uint64_t *C_limbs, int64_t *C_se, void *stream);

/* Convert between arrays of the reference's AoS structs (cu_freal<PB>:
 * {int sign; long exp; cu_limb m[N];}, sizeof 16+8N, cu_freal.cuh:75-77;
 * cu_fcomplex<PB>: {re, im}, cu_fcomplex.cuh:102) and MP arrays. `aos` is a
 * device pointer to n structs; `complex` selects cu_fcomplex<PB>. */
int mpgfx_pack (int pb, int complex, int64_t n, const void *aos,
                uint64_t *limbs, int64_t *se, void *stream);
int mpgfx_unpack (int pb, int complex, int64_t n, const uint64_t *limbs,
                  const int64_t *se, void *aos, void *stream);

/* Interop with host arrays of MPFR numbers (the on-wire format either side of
 * the path).  `x` points to n MPFR value structs (`__mpfr_struct`: prec,
 * sign, exp, limb pointer -- the layout MPFR's custom interface macros read,
 * mpfr.h) placed `stride` bytes apart: sizeof(mpfr_t) for an mpfr_t array;
 * for an mpc_t array use stride sizeof(mpc_t) with x = &z[0].re (real parts)
 * and x = &z[0].im (imaginary parts).  MPFR's significand limbs are the
 * storage of cu_freal<PB> (left-justified, little-endian limbs, cu_freal.cuh:12-15)
 * and its exponent is cu_freal's `exp`, so each value copies as-is; nothing
 * is rounded.
 * pb: a compiled precision or any precision of the run-time path (1 .. 2048).
 * import: every value must have precision pb and be zero or regular (NaN and
 * infinities give MPGFX_EINVAL); the limbs land in the device MP array.
 * export: every struct must have precision pb and a limb buffer of
 * ceil(pb/64) limbs (as mpfr_init2 leaves it); limbs, sign and exponent are
 * written (a zero keeps the sign of its sign field).
 * Both stage through host memory and synchronise `stream`.  No counterpart in
 * the reference (its host code converts through doubles or mpz,
 * tools/test_freal.cu:29-40). */
int mpgfx_import_mpfr (int pb, int64_t n, const void *x, size_t stride,
                       uint64_t *limbs, int64_t *se, void *stream);
int mpgfx_export_mpfr (int pb, int64_t n, const uint64_t *limbs, const int64_t *se,
                       void *x, size_t stride, void *stream);

/* Run-time precision (any prec in 1 .. 2048 bits, not only the compiled
 * ones): the role of the reference's runtime-precision MPFR / MPC device path
 * (cu_mpfr_* / cu_mpc_*, demos/axpy_mpfr.cu:36-50, demos/axpy_mpc.cu:40-53),
 * with MPFR's RNDN (MPC_RNDNN) results.  MP arrays of precision prec use
 * ceil(prec/64) limbs, the significand left-justified (MPFR's layout, low
 * bits zero).  Real zero signs follow MPFR (a product's zero has the sign of
 * the product of the signs; -0 + -0 = -0, any other exact zero sum is +0); an
 * exactly cancelling complex component is +0.
 * mpgfx_rt_elementwise_real: o[i] = a[i] OP b[i], OP in {mul, add, sub}.
 * mpgfx_rt_axpy_real: o[i] = RNDN(RNDN(a x[i]) + y[i]) (a: ONE value).
 * mpgfx_rt_axpy_complex: z[i] = a x[i] + y[i], MPC_RNDNN per operation. */
int mpgfx_rt_elementwise_real (int prec, int op, int64_t n,
                               const uint64_t *a_limbs, const int64_t *a_se,
                               const uint64_t *b_limbs, const int64_t *b_se,
                               uint64_t *o_limbs, int64_t *o_se, void *stream);
int mpgfx_rt_axpy_real (int prec, int64_t n,
                        const uint64_t *a_limbs, const int64_t *a_se,
                        const uint64_t *x_limbs, const int64_t *x_se,
                        const uint64_t *y_limbs, const int64_t *y_se,
                        uint64_t *o_limbs, int64_t *o_se, void *stream);
int mpgfx_rt_axpy_complex (int prec, int64_t n,
                           const uint64_t *a_limbs, const int64_t *a_se,
                           const uint64_t *x_limbs, const int64_t *x_se,
                           const uint64_t *y_limbs, const int64_t *y_se,
                           uint64_t *z_limbs, int64_t *z_se, void *stream);

/* Run-time precision per-operation folds (any prec in 1 .. 2048): the
 * reference's runtime-path matvec / matmul (demos/matvec_mpfr.cu,
 * demos/matmul_mpfr.cu: mpfr_mul then mpfr_add per k; demos/matvec_mpc.cu,
 * demos/matmul_mpc.cu: mpc_mul then mpc_add), acc from +0, k ascending, RNDN
 * (MPC_RNDNN) at prec with MPFR's zero signs.  cplx selects complex MP
 * arrays.  mpgfx_rt_gemv: y = A x (A m x k); a dot product is m = 1.
 * mpgfx_rt_gemm: C = A B (A m x k, B k x n).  k = 0 gives +0. */
int mpgfx_rt_gemv (int prec, int cplx, int64_t m, int64_t k,
                   const uint64_t *A_limbs, const int64_t *A_se,
                   const uint64_t *x_limbs, const int64_t *x_se,
                   uint64_t *y_limbs, int64_t *y_se, void *stream);
int mpgfx_rt_gemm (int prec, int cplx, int64_t m, int64_t n, int64_t k,
                   const uint64_t *A_limbs, const int64_t *A_se,
                   const uint64_t *B_limbs, const int64_t *B_se,
                   uint64_t *C_limbs, int64_t *C_se, void *stream);

/* n seeded synthetic nonzero values (value i = start + i of the stream keyed by
 * seed: significand limbs random with the top bit set and the low 64N-pb bits
 * clear, random sign, exponent uniform in [emin, emax]) -- the generator of the
 * tests and benchmarks, identical to mpc_cuda_amd/gen.py.  No counterpart in
 * the reference (its tests draw operands with xorshift64 on the host,
 * tools/test_freal.cu:15-27). */
int mpgfx_fill_random (int pb, int64_t n, int64_t start, uint64_t seed, int64_t emin, int64_t emax,
                       uint64_t *limbs, int64_t *se, void *stream);

/* ------------------------------------------------------------------------
 * Batched execution (no counterpart in the reference).  Many AXPY, dot and
 * GEMV instances of uneven lengths, mixed precision and real/complex run as one
 * hipGraph: per (precision, real/complex) class a persistent product kernel
 * (all AXPY outputs and every product of every dot and GEMV row, in parallel)
 * followed by a persistent fold kernel (the per-operation k-order additions,
 * one lane per dot / GEMV row, longest chains first).  Dot and GEMV semantics
 * are the reference's per-op fold (tools/linalg_fast.cu:37-51): acc = +0,
 * acc = RNDN(acc + RNDN(x_k*y_k)) for k ascending (complex: MPC_RNDNN).
 * ------------------------------------------------------------------------ */
enum mpgfx_task_kind { MPGFX_TASK_AXPY = 0, MPGFX_TASK_DOT = 1, MPGFX_TASK_GEMV = 2 };

/* One instance; all pointers are device MP arrays of precision pb (complex
 * layout when cplx != 0).
 *   AXPY: out[i] = in0 * in1[i] + in2[i], i < n; in0 is ONE value.
 *   DOT : out[0] = fold_k in0[k] * in1[k], k < n.
 *   GEMV: out[i] = fold_k in0[i*n + k] * in1[k], i < m, k < n (in0: m x n). */
typedef struct mpgfx_task
{
  int32_t kind, pb, cplx, reserved;
  int64_t m, n;
  const uint64_t *in0_limbs; const int64_t *in0_se;
  const uint64_t *in1_limbs; const int64_t *in1_se;
  const uint64_t *in2_limbs; const int64_t *in2_se;
  uint64_t *out_limbs; int64_t *out_se;
} mpgfx_task;

/* Mapping of one (phase, precision, real/complex) class: values per work item
 * of the product phase, lanes per workgroup, persistent workgroups per CU. */
typedef struct mpgfx_mapping
{
  int32_t chunk, block, wg_per_cu, reserved;
} mpgfx_mapping;

typedef struct mpgfx_batch mpgfx_batch;

/* Plan a batch: validate the `ntasks` host-side descriptors, allocate the
 * device work lists and product scratch (sized for the batch), and capture
 * the launches into a hipGraph on `stream`.  tuning: 0 = fixed default
 * mapping, 1 = the tuner's mapping (read from the JSON cache file `cache`,
 * or measured on this batch and written there when absent; cache may be NULL
 * to tune without a file).  The batch keeps its own copy of the descriptors.
 * Aliasing: an AXPY may run in place (out == in2 or out == in1); no other
 * output may overlap an input of the batch.  Tuning runs the batch's products
 * on its inputs with the AXPY outputs sent to scratch, so create leaves every
 * caller array as it found it apart from dot / GEMV outputs. */
int mpgfx_batch_create (const mpgfx_task *tasks, int64_t ntasks, int tuning, const char *cache,
                        void *stream, mpgfx_batch **out);

/* Launch the whole batch (one graph launch) on `stream`. */
int mpgfx_batch_run (mpgfx_batch *b, void *stream);

/* Bytes of device scratch the batch holds, and the number of classes. */
int mpgfx_batch_info (const mpgfx_batch *b, size_t *scratch_bytes, int *classes);

int mpgfx_batch_destroy (mpgfx_batch *b);

#ifdef __cplusplus
}
#endif
#endif
