/* oracle/cpu_mpfr.c -- TEST INFRASTRUCTURE ONLY: the CPU baseline that
 * bench.py times beside the GPU (never the measured GPU path).
 *
 * The reference's CPU contender, restated on full-limb MP arrays: system
 * GNU MPFR / MPC in the kernels' operation order, OpenMP over the host's
 * cores with per-thread temporaries and a static schedule
 * (demos/bench3_cpu.cpp:24-31, 35-85):
 *   real AXPY        mpfr_mul + mpfr_add, RNDN          (bench3_cpu.cpp:49-53)
 *   complex AXPY     mpc_mul + mpc_add, MPC_RNDNN       (bench3_cpu.cpp:74-77)
 *   GEMV / GEMM      acc = +0; acc = acc + a*b per k    (tools/linalg_fast.cu:71-83)
 *   complex per-op   the same with mpc_mul / mpc_add    (demos/matmul_mpc.cu:33-52 order)
 *   exact GEMM       exact products at 2 pb bits, one mpfr_sum RNDN
 * Inputs and outputs are SoA MP arrays (oracle/mpfixed.h).  Conversion into
 * and out of mpfr_t happens outside the timed region; every function returns
 * the milliseconds of its timed loop.  GEMV / GEMM take a row range (a
 * documented sample of the output rows). */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <gmp.h>
#include <mpfr.h>
#include <mpc.h>
#include <omp.h>

#define ZERO_EXP (-0x7fffffffffffffLL)

static double now_ms (void)
{
  struct timespec t;
  clock_gettime (CLOCK_MONOTONIC, &t);
  return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

int cpu_num_threads (void) { return omp_get_max_threads (); }

/* the timed loops run this many times and report the average (the
 * conversions outside them run once) */
static int g_reps = 1;
void cpu_set_reps (int r) { g_reps = r > 0 ? r : 1; }

const char *cpu_versions (void)
{
  static char b[160];
  snprintf (b, sizeof b, "MPFR %s, MPC %s, GMP %s", mpfr_get_version (), mpc_get_version (), gmp_version);
  return b;
}

/* value i of a SoA array (N limbs per value, plane stride n) -> x, exactly */
static void set_val (mpfr_t x, int N, const uint64_t *l, const int64_t *se, int64_t n, int64_t i)
{
  const int64_t w = se[i], e = w >> 1;
  if (e == ZERO_EXP) { mpfr_set_zero (x, (w & 1) ? -1 : 1); return; }
  uint64_t m[64];
  for (int k = 0; k < N; k++) m[k] = l[(int64_t) k * n + i];
  mpz_t z;
  mpz_init (z);
  mpz_import (z, (size_t) N, -1, sizeof (uint64_t), 0, 0, m);
  mpfr_set_z_2exp (x, z, (mpfr_exp_t) (e - 64L * N), MPFR_RNDN);
  if (w & 1) mpfr_neg (x, x, MPFR_RNDN);
  mpz_clear (z);
}

/* x (precision pb = the array's) -> value i of a SoA array; zeros keep MPFR's sign */
static void get_val (const mpfr_t x, int N, uint64_t *l, int64_t *se, int64_t n, int64_t i)
{
  if (mpfr_zero_p (x))
    {
      for (int k = 0; k < N; k++) l[(int64_t) k * n + i] = 0;
      se[i] = (int64_t) ((uint64_t) ZERO_EXP << 1) | (mpfr_signbit (x) ? 1 : 0);
      return;
    }
  const mp_limb_t *d = (const mp_limb_t *) mpfr_custom_get_significand (x);
  const int xl = (int) ((mpfr_get_prec (x) + 63) / 64);
  for (int k = 0; k < N; k++) l[(int64_t) k * n + i] = (k + xl - N >= 0) ? d[k + xl - N] : 0;
  se[i] = (int64_t) ((uint64_t) (int64_t) mpfr_get_exp (x) << 1) | (mpfr_signbit (x) ? 1 : 0);
}

static mpfr_t *load (int pb, int N, const uint64_t *l, const int64_t *se, int64_t n, int64_t i0, int64_t cnt)
{
  mpfr_t *v = (mpfr_t *) malloc ((size_t) (cnt > 0 ? cnt : 1) * sizeof (mpfr_t));
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < cnt; j++)
    {
      mpfr_init2 (v[j], pb);
      set_val (v[j], N, l, se, n, i0 + j);
    }
  return v;
}

static void release (mpfr_t *v, int64_t cnt)
{
  for (int64_t j = 0; j < cnt; j++) mpfr_clear (v[j]);
  free (v);
}

/* o[i] = RNDN(RNDN(a*x[i]) + y[i]) */
double cpu_axpy_real (int pb, int64_t n, const uint64_t *al, const int64_t *ase, const uint64_t *xl,
                      const int64_t *xse, const uint64_t *yl, const int64_t *yse, uint64_t *ol, int64_t *ose)
{
  const int N = (pb + 63) / 64;
  mpfr_t *a = load (pb, N, al, ase, 1, 0, 1), *x = load (pb, N, xl, xse, n, 0, n), *y = load (pb, N, yl, yse, n, 0, n);
  mpfr_t *o = load (pb, N, yl, yse, n, 0, n);
  const double t0 = now_ms ();
  for (int rep = 0; rep < g_reps; rep++)
#pragma omp parallel
  {
    mpfr_t t;
    mpfr_init2 (t, pb);
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; i++)
      {
        mpfr_mul (t, a[0], x[i], MPFR_RNDN);
        mpfr_add (o[i], t, y[i], MPFR_RNDN);
      }
    mpfr_clear (t);
  }
  const double ms = (now_ms () - t0) / g_reps;
  for (int64_t i = 0; i < n; i++) get_val (o[i], N, ol, ose, n, i);
  release (a, 1); release (x, n); release (y, n); release (o, n);
  return ms;
}

/* z[i] = a*x[i] + y[i], mpc_mul then mpc_add (complex SoA: real parts, then
 * imaginary parts) */
double cpu_axpy_complex (int pb, int64_t n, const uint64_t *al, const int64_t *ase, const uint64_t *xl,
                         const int64_t *xse, const uint64_t *yl, const int64_t *yse, uint64_t *zl, int64_t *zse)
{
  const int N = (pb + 63) / 64;
  const int64_t o = (int64_t) N * n;
  mpc_t *a = (mpc_t *) malloc (sizeof (mpc_t)), *x = (mpc_t *) malloc ((size_t) n * sizeof (mpc_t));
  mpc_t *y = (mpc_t *) malloc ((size_t) n * sizeof (mpc_t)), *z = (mpc_t *) malloc ((size_t) n * sizeof (mpc_t));
  mpc_init2 (a[0], pb);
  set_val (mpc_realref (a[0]), N, al, ase, 1, 0);
  set_val (mpc_imagref (a[0]), N, al + N, ase + 1, 1, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    {
      mpc_init2 (x[i], pb); mpc_init2 (y[i], pb); mpc_init2 (z[i], pb);
      set_val (mpc_realref (x[i]), N, xl, xse, n, i);
      set_val (mpc_imagref (x[i]), N, xl + o, xse + n, n, i);
      set_val (mpc_realref (y[i]), N, yl, yse, n, i);
      set_val (mpc_imagref (y[i]), N, yl + o, yse + n, n, i);
    }
  const double t0 = now_ms ();
  for (int rep = 0; rep < g_reps; rep++)
#pragma omp parallel
  {
    mpc_t t;
    mpc_init2 (t, pb);
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; i++)
      {
        mpc_mul (t, a[0], x[i], MPC_RNDNN);
        mpc_add (z[i], t, y[i], MPC_RNDNN);
      }
    mpc_clear (t);
  }
  const double ms = (now_ms () - t0) / g_reps;
  for (int64_t i = 0; i < n; i++)
    {
      get_val (mpc_realref (z[i]), N, zl, zse, n, i);
      get_val (mpc_imagref (z[i]), N, zl + o, zse + n, n, i);
      mpc_clear (x[i]); mpc_clear (y[i]); mpc_clear (z[i]);
    }
  mpc_clear (a[0]);
  free (a); free (x); free (y); free (z);
  return ms;
}

/* rows r0 .. r1-1 of C = A B (A m x k, B k x n), per-op k order:
 * acc = +0; acc = RNDN(acc + RNDN(A[i,kk] B[kk,j])).  C holds m x n values;
 * only the sampled rows are written. */
double cpu_gemm_perop (int pb, int64_t m, int64_t n, int64_t k, const uint64_t *Al, const int64_t *Ase,
                       const uint64_t *Bl, const int64_t *Bse, uint64_t *Cl, int64_t *Cse, int64_t r0, int64_t r1)
{
  const int N = (pb + 63) / 64;
  const int64_t rows = r1 - r0;
  mpfr_t *A = load (pb, N, Al, Ase, m * k, r0 * k, rows * k), *B = load (pb, N, Bl, Bse, k * n, 0, k * n);
  mpfr_t *C = (mpfr_t *) malloc ((size_t) (rows * n > 0 ? rows * n : 1) * sizeof (mpfr_t));
  for (int64_t e = 0; e < rows * n; e++) mpfr_init2 (C[e], pb);
  const double t0 = now_ms ();
  for (int rep = 0; rep < g_reps; rep++)
#pragma omp parallel
  {
    mpfr_t t;
    mpfr_init2 (t, pb);
#pragma omp for schedule(static)
    for (int64_t e = 0; e < rows * n; e++)
      {
        const int64_t i = e / n, j = e % n;
        mpfr_set_zero (C[e], 1);
        for (int64_t kk = 0; kk < k; kk++)
          {
            mpfr_mul (t, A[i * k + kk], B[kk * n + j], MPFR_RNDN);
            mpfr_add (C[e], C[e], t, MPFR_RNDN);
          }
      }
    mpfr_clear (t);
  }
  const double ms = (now_ms () - t0) / g_reps;
  for (int64_t e = 0; e < rows * n; e++) get_val (C[e], N, Cl, Cse, m * n, r0 * n + e);
  release (A, rows * k); release (B, k * n); release (C, rows * n);
  return ms;
}

/* the same with MPC_RNDNN products and sums on complex SoA matrices */
double cpu_cgemm_perop (int pb, int64_t m, int64_t n, int64_t k, const uint64_t *Al, const int64_t *Ase,
                        const uint64_t *Bl, const int64_t *Bse, uint64_t *Cl, int64_t *Cse, int64_t r0, int64_t r1)
{
  const int N = (pb + 63) / 64;
  const int64_t rows = r1 - r0, nA = m * k, nB = k * n, nC = m * n;
  mpc_t *A = (mpc_t *) malloc ((size_t) (rows * k > 0 ? rows * k : 1) * sizeof (mpc_t));
  mpc_t *B = (mpc_t *) malloc ((size_t) (nB > 0 ? nB : 1) * sizeof (mpc_t));
  mpc_t *C = (mpc_t *) malloc ((size_t) (rows * n > 0 ? rows * n : 1) * sizeof (mpc_t));
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < rows * k; q++)
    {
      mpc_init2 (A[q], pb);
      set_val (mpc_realref (A[q]), N, Al, Ase, nA, r0 * k + q);
      set_val (mpc_imagref (A[q]), N, Al + (int64_t) N * nA, Ase + nA, nA, r0 * k + q);
    }
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < nB; q++)
    {
      mpc_init2 (B[q], pb);
      set_val (mpc_realref (B[q]), N, Bl, Bse, nB, q);
      set_val (mpc_imagref (B[q]), N, Bl + (int64_t) N * nB, Bse + nB, nB, q);
    }
  for (int64_t e = 0; e < rows * n; e++) mpc_init2 (C[e], pb);
  const double t0 = now_ms ();
  for (int rep = 0; rep < g_reps; rep++)
#pragma omp parallel
  {
    mpc_t t;
    mpc_init2 (t, pb);
#pragma omp for schedule(static)
    for (int64_t e = 0; e < rows * n; e++)
      {
        const int64_t i = e / n, j = e % n;
        mpc_set_ui (C[e], 0, MPC_RNDNN);
        for (int64_t kk = 0; kk < k; kk++)
          {
            mpc_mul (t, A[i * k + kk], B[kk * n + j], MPC_RNDNN);
            mpc_add (C[e], C[e], t, MPC_RNDNN);
          }
      }
    mpc_clear (t);
  }
  const double ms = (now_ms () - t0) / g_reps;
  for (int64_t e = 0; e < rows * n; e++)
    {
      get_val (mpc_realref (C[e]), N, Cl, Cse, nC, r0 * n + e);
      get_val (mpc_imagref (C[e]), N, Cl + (int64_t) N * nC, Cse + nC, nC, r0 * n + e);
      mpc_clear (C[e]);
    }
  for (int64_t q = 0; q < rows * k; q++) mpc_clear (A[q]);
  for (int64_t q = 0; q < nB; q++) mpc_clear (B[q]);
  free (A); free (B); free (C);
  return ms;
}

/* rows r0 .. r1-1 of C = RNDN(sum_k A B): exact products at 2 pb bits, one
 * mpfr_sum per output */
double cpu_gemm_exact (int pb, int64_t m, int64_t n, int64_t k, const uint64_t *Al, const int64_t *Ase,
                       const uint64_t *Bl, const int64_t *Bse, uint64_t *Cl, int64_t *Cse, int64_t r0, int64_t r1)
{
  const int N = (pb + 63) / 64;
  const int64_t rows = r1 - r0;
  mpfr_t *A = load (pb, N, Al, Ase, m * k, r0 * k, rows * k), *B = load (pb, N, Bl, Bse, k * n, 0, k * n);
  mpfr_t *C = (mpfr_t *) malloc ((size_t) (rows * n > 0 ? rows * n : 1) * sizeof (mpfr_t));
  for (int64_t e = 0; e < rows * n; e++) mpfr_init2 (C[e], pb);
  const double t0 = now_ms ();
  for (int rep = 0; rep < g_reps; rep++)
#pragma omp parallel
  {
    mpfr_t *P = (mpfr_t *) malloc ((size_t) (k > 0 ? k : 1) * sizeof (mpfr_t));
    mpfr_ptr *pp = (mpfr_ptr *) malloc ((size_t) (k > 0 ? k : 1) * sizeof (mpfr_ptr));
    for (int64_t kk = 0; kk < k; kk++) { mpfr_init2 (P[kk], 2 * pb); pp[kk] = P[kk]; }
#pragma omp for schedule(static)
    for (int64_t e = 0; e < rows * n; e++)
      {
        const int64_t i = e / n, j = e % n;
        for (int64_t kk = 0; kk < k; kk++) mpfr_mul (P[kk], A[i * k + kk], B[kk * n + j], MPFR_RNDN);
        mpfr_sum (C[e], pp, (unsigned long) k, MPFR_RNDN);
      }
    for (int64_t kk = 0; kk < k; kk++) mpfr_clear (P[kk]);
    free (P);
    free (pp);
  }
  const double ms = (now_ms () - t0) / g_reps;
  for (int64_t e = 0; e < rows * n; e++) get_val (C[e], N, Cl, Cse, m * n, r0 * n + e);
  release (A, rows * k); release (B, k * n); release (C, rows * n);
  return ms;
}
