/* oracle/mpfixed.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the fixed-precision arithmetic of tkouya/mpc_cuda
 * (include/mpc_cuda/cu_freal.cuh, include/mpc_cuda/cu_fcomplex.cuh) and of the
 * kernels built on it (demos/bench_3way.cu, tools/linalg_fast.cu), with the
 * precision PB as a RUNTIME parameter.  It is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
 * it, and never as the thing measured or shipped.
 *
 * Number model (cu_freal.cuh:12-15, 67-81):
 *   value = sign * M * 2^(exp - 64*N),  N = ceil(PB/64) u64 limbs, little endian,
 *   M's bit 64N-1 set, its low SB = 64N-PB bits zero; exp is the MPFR exponent;
 *   zero is exp == OR_ZERO_EXP with sign +1 (the reference's zero is always +0).
 *
 * Array layout shared with the HIP library (structure of arrays):
 *   limbs: N planes of n u64, limb l of value i at limbs[l*n + i]
 *   se   : n int64 sign/exponent words, se = exp*2 + (sign<0 ? 1 : 0)
 *   a complex array of n values is two real arrays of n back to back
 *   (re: limbs[0 .. N*n), se[0..n);  im: limbs[N*n .. 2*N*n), se[n..2n)).
 *
 * Parity pins (exercised by tests/): system MPFR/MPC through oracle/pin_mpfr.c,
 * the reference header compiled into oracle/_ref, committed golden fixtures
 * generated by MPFR, and an independent Python big-integer rounding oracle.
 */
#ifndef MPGFX_ORACLE_MPFIXED_H
#define MPGFX_ORACLE_MPFIXED_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXN 64                                  /* PB <= 4096            */
#define OR_ZERO_EXP (-0x7fffffffffffffLL)           /* cu_freal.cuh:37       */

typedef struct { int sign; int64_t exp; uint64_t m[OR_MAXN]; } or_real;

/* ---- scalar operations (pb = precision in bits, multiple of 32) ---- */
int  or_nlimbs (int pb);
void or_set_zero (int pb, or_real *r);
void or_finalize (int pb, const uint64_t *D, int len, long msb, int64_t ehi,
                  uint64_t brnd, uint64_t bsti, int sgn, or_real *r);
void or_mul (int pb, const or_real *a, const or_real *b, or_real *r);
int  or_cmpmag (int pb, const or_real *a, const or_real *b);
void or_add (int pb, const or_real *a, const or_real *b, or_real *r);
void or_sub (int pb, const or_real *a, const or_real *b, or_real *r);
void or_fmma (int pb, const or_real *a, const or_real *b, const or_real *c,
              const or_real *d, or_real *r);
void or_fmms (int pb, const or_real *a, const or_real *b, const or_real *c,
              const or_real *d, or_real *r);
void or_cmul (int pb, const or_real *ar, const or_real *ai, const or_real *br,
              const or_real *bi, or_real *rr, or_real *ri);
void or_from_double (int pb, double d, or_real *r);
double or_to_double (int pb, const or_real *x);

/* ---- the reference test generator (tools/test_freal.cu:15-27) ---- */
uint64_t or_xs (uint64_t *s);
void or_randf (int pb, uint64_t *s, int spread, or_real *r);

/* ---- SoA array front-ends (OpenMP over elements) ---- */
void or_load (int pb, const uint64_t *limbs, const int64_t *se, int64_t n,
              int64_t i, or_real *r);
void or_store (int pb, const or_real *r, uint64_t *limbs, int64_t *se,
               int64_t n, int64_t i);
/* division and square root at working precision pb + 128 (cu_fmath.cuh:53-100) */
void or_fdiv (int pb, const or_real *a, const or_real *b, or_real *r);
void or_fsqrt (int pb, const or_real *x, or_real *r);
/* op: 0 = mul, 1 = add, 2 = sub (tools/test_freal.cu:41-48), 3 = div (cu_fdiv) */
void or_elementwise_real (int pb, int op, int64_t n,
    const uint64_t *al, const int64_t *ase, const uint64_t *bl, const int64_t *bse,
    uint64_t *ol, int64_t *ose);
/* out = sqrt(x) per element (cu_sqrt) */
void or_sqrt_real (int pb, int64_t n, const uint64_t *xl, const int64_t *xse,
                   uint64_t *ol, int64_t *ose);
/* op on complex arrays (tools/test_fcomplex.cu:147-153) */
void or_elementwise_complex (int pb, int op, int64_t n,
    const uint64_t *al, const int64_t *ase, const uint64_t *bl, const int64_t *bse,
    uint64_t *ol, int64_t *ose);
/* out = RNDN(RNDN(a*x)+y); a is ONE value (demos/bench_3way.cu:44-50) */
void or_axpy_real (int pb, int64_t n, const uint64_t *al, const int64_t *ase,
    const uint64_t *xl, const int64_t *xse, const uint64_t *yl, const int64_t *yse,
    uint64_t *ol, int64_t *ose);
/* z = a*x + y, complex, MPC_RNDNN per op (demos/bench_3way.cu:51-59) */
void or_axpy_complex (int pb, int64_t n, const uint64_t *al, const int64_t *ase,
    const uint64_t *xl, const int64_t *xse, const uint64_t *yl, const int64_t *yse,
    uint64_t *ol, int64_t *ose);
/* C[i,j] = fold_k RNDN(acc + RNDN(A[i,k]*B[k,j])), acc = +0 (tools/linalg_fast.cu:53-68).
 * A is m x k row-major (element (i,kk) at i*k+kk), B is k x n row-major,
 * C is m x n row-major.  rows: only rows [r0, r1) of C are computed. */
void or_gemm_perop (int pb, int64_t m, int64_t n, int64_t k,
    const uint64_t *Al, const int64_t *Ase, const uint64_t *Bl, const int64_t *Bse,
    uint64_t *Cl, int64_t *Cse, int64_t r0, int64_t r1);
/* C[i,j] = RNDN(sum_k A[i,k]*B[k,j]) with every product and the sum exact and
 * ONE rounding (semantics of src/mpfr_cuda/dot.c:47-54: exact products then
 * mpfr_sum). */
void or_gemm_exact (int pb, int64_t m, int64_t n, int64_t k,
    const uint64_t *Al, const int64_t *Ase, const uint64_t *Bl, const int64_t *Bse,
    uint64_t *Cl, int64_t *Cse, int64_t r0, int64_t r1);
/* complex per-op GEMM: acc = +0; acc = cadd(acc, cmul(A,B)) */
void or_cgemm_perop (int pb, int64_t m, int64_t n, int64_t k,
    const uint64_t *Al, const int64_t *Ase, const uint64_t *Bl, const int64_t *Bse,
    uint64_t *Cl, int64_t *Cse, int64_t r0, int64_t r1);

/* fill n values with the reference generator, spread pattern as test_freal.cu:55
 * when spread < 0 ((i&3) ? 40 : pb+200), else uniform exponent in [-spread, spread] */
void or_fill_ref (int pb, int64_t n, uint64_t seed, int spread,
                  uint64_t *limbs, int64_t *se);

int or_num_threads (void);

#ifdef __cplusplus
}
#endif
#endif
