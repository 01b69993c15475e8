/* oracle/pin_mpfr.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Computes the reference semantics of the fixed-precision path with GNU MPFR /
 * MPC (the libraries the reference's own bit-exactness tests compare against:
 * tools/test_freal.cu:69-76, tools/test_fcomplex.cu:176-183,
 * tools/linalg_fast.cu:71-83, demos/bench3_cpu.cpp:35-85).  Used to pin the C
 * oracle and to write the golden fixtures under tests/golden/.
 *
 * Usage:  pin_mpfr OP PB FILE_IN FILE_OUT [M N K]
 *   FILE_IN holds the operands as consecutive SoA arrays (see mpfixed.h);
 *   FILE_OUT receives the result array.  Zero results are written as
 *   se = ZERO*2 + (MPFR sign bit), limbs 0.
 *   OP: mul add sub div        (n = M)        a[n], b[n]        -> o[n]
 *       sqrt                   (n = M)        x[n]              -> o[n]  (x > 0)
 *       cmul cadd csub         (n = M)        complex a, b      -> o
 *       axpy                   (n = M)        a[1], x[n], y[n]  -> o[n]
 *       caxpy                  (n = M)        complex a[1], x, y -> o
 *       gemm_perop gemm_exact  (M N K)        A[M*K], B[K*N]    -> C[M*N]
 *       cgemm_perop            (M N K)        complex A, B      -> C
 *       cgemm_exact            (M N K)        complex A, B      -> C  (each component:
 *                                               exact products, one mpfr_sum) 
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <gmp.h>
#include <mpfr.h>
#include <mpc.h>

#define ZERO_EXP (-0x7fffffffffffffLL)

static int pb, N;

static void die (const char *m) { fprintf (stderr, "pin_mpfr: %s\n", m); exit (2); }

/* one value of a SoA array -> mpfr (tools/test_freal.cu:28-39) */
static void to_mpfr (mpfr_t out, const uint64_t *l, const int64_t *se, int64_t n, int64_t i)
{
  int64_t w = se[i], e = w >> 1;
  if (e == ZERO_EXP) { mpfr_set_zero (out, 1); return; }
  uint64_t m[128];
  for (int k = 0; k < N; k++) m[k] = l[(int64_t) k * n + i];
  mpz_t z; mpz_init (z);
  mpz_import (z, (size_t) N, -1, sizeof (uint64_t), 0, 0, m);
  mpfr_set_z (out, z, MPFR_RNDN);
  mpfr_mul_2si (out, out, (long) (e - 64L * N), MPFR_RNDN);
  if (w & 1) mpfr_neg (out, out, MPFR_RNDN);
  mpz_clear (z);
}

static void from_mpfr (const mpfr_t x, uint64_t *l, int64_t *se, int64_t n, int64_t i)
{
  if (mpfr_zero_p (x))
    {
      for (int k = 0; k < N; k++) l[(int64_t) k * n + i] = 0;
      se[i] = (int64_t) ((uint64_t) ZERO_EXP << 1) | (mpfr_signbit (x) ? 1 : 0);
      return;
    }
  if (!mpfr_number_p (x)) die ("non-finite result");
  long e = mpfr_get_exp (x);
  mpfr_t t; mpfr_init2 (t, mpfr_get_prec (x));
  mpfr_abs (t, x, MPFR_RNDN);
  mpfr_mul_2si (t, t, 64L * N - e, MPFR_RNDN);
  mpz_t z; mpz_init (z);
  mpfr_get_z (z, t, MPFR_RNDN);
  uint64_t m[128]; memset (m, 0, sizeof m);
  size_t cnt = 0;
  mpz_export (m, &cnt, -1, sizeof (uint64_t), 0, 0, z);
  if ((int) cnt > N) die ("export overflow");
  for (int k = 0; k < N; k++) l[(int64_t) k * n + i] = m[k];
  se[i] = (int64_t) ((uint64_t) (int64_t) e << 1) | (mpfr_signbit (x) ? 1 : 0);
  mpz_clear (z); mpfr_clear (t);
}

typedef struct { uint64_t *l; int64_t *se; int64_t n; } arr;

static arr read_arr (FILE *f, int64_t n)
{
  arr a; a.n = n;
  a.l = malloc (sizeof (uint64_t) * (size_t) (N * n));
  a.se = malloc (sizeof (int64_t) * (size_t) n);
  if (fread (a.l, sizeof (uint64_t), (size_t) (N * n), f) != (size_t) (N * n)) die ("short read");
  if (fread (a.se, sizeof (int64_t), (size_t) n, f) != (size_t) n) die ("short read");
  return a;
}
static arr new_arr (int64_t n)
{
  arr a; a.n = n;
  a.l = calloc ((size_t) (N * n), sizeof (uint64_t));
  a.se = calloc ((size_t) n, sizeof (int64_t));
  return a;
}
static void write_arr (FILE *f, arr a)
{
  fwrite (a.l, sizeof (uint64_t), (size_t) (N * a.n), f);
  fwrite (a.se, sizeof (int64_t), (size_t) a.n, f);
}
/* complex array of n = real array of 2n (re then im) */
static void to_mpc (mpc_t z, arr a, int64_t n, int64_t i)
{
  to_mpfr (mpc_realref (z), a.l, a.se, 2 * n, i);
  to_mpfr (mpc_imagref (z), a.l, a.se, 2 * n, n + i);
}
/* the complex layout is [re limbs plane-major over n][im ...]: re is the real
 * array (l, se) of n, im is (l + N*n, se + n).  Accessors: */
static void cget (mpc_t z, arr a, int64_t n, int64_t i)
{
  to_mpfr (mpc_realref (z), a.l, a.se, n, i);
  to_mpfr (mpc_imagref (z), a.l + (int64_t) N * n, a.se + n, n, i);
}
static void cput (const mpc_t z, arr a, int64_t n, int64_t i)
{
  from_mpfr (mpc_realref (z), a.l, a.se, n, i);
  from_mpfr (mpc_imagref (z), a.l + (int64_t) N * n, a.se + n, n, i);
}

int main (int argc, char **argv)
{
  if (argc < 6) die ("usage: pin_mpfr OP PB IN OUT M [N K]");
  const char *op = argv[1];
  pb = atoi (argv[2]); N = (pb + 63) / 64;
  if (pb <= 0 || N > 128) die ("bad PB");
  FILE *fi = fopen (argv[3], "rb"); if (!fi) die ("open in");
  FILE *fo = fopen (argv[4], "wb"); if (!fo) die ("open out");
  int64_t M = atoll (argv[5]);
  int64_t Nn = argc > 6 ? atoll (argv[6]) : 0, K = argc > 7 ? atoll (argv[7]) : 0;
  (void) to_mpc;

  if (!strcmp (op, "mul") || !strcmp (op, "add") || !strcmp (op, "sub") || !strcmp (op, "div"))
    {
      arr a = read_arr (fi, M), b = read_arr (fi, M), o = new_arr (M);
#pragma omp parallel
      {
        mpfr_t x, y, r; mpfr_inits2 (pb, x, y, r, (mpfr_ptr) 0);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < M; i++)
          {
            to_mpfr (x, a.l, a.se, M, i); to_mpfr (y, b.l, b.se, M, i);
            if (op[0] == 'm') mpfr_mul (r, x, y, MPFR_RNDN);
            else if (op[0] == 'a') mpfr_add (r, x, y, MPFR_RNDN);
            else if (op[0] == 'd') mpfr_div (r, x, y, MPFR_RNDN);
            else mpfr_sub (r, x, y, MPFR_RNDN);
            from_mpfr (r, o.l, o.se, M, i);
          }
        mpfr_clears (x, y, r, (mpfr_ptr) 0);
      }
      write_arr (fo, o);
    }
  else if (!strcmp (op, "sqrt"))
    {
      arr a = read_arr (fi, M), o = new_arr (M);
#pragma omp parallel
      {
        mpfr_t x, r; mpfr_inits2 (pb, x, r, (mpfr_ptr) 0);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < M; i++)
          {
            to_mpfr (x, a.l, a.se, M, i);
            mpfr_sqrt (r, x, MPFR_RNDN);
            from_mpfr (r, o.l, o.se, M, i);
          }
        mpfr_clears (x, r, (mpfr_ptr) 0);
      }
      write_arr (fo, o);
    }
  else if (!strcmp (op, "cmul") || !strcmp (op, "cadd") || !strcmp (op, "csub"))
    {
      arr a = read_arr (fi, 2 * M), b = read_arr (fi, 2 * M), o = new_arr (2 * M);
#pragma omp parallel
      {
        mpc_t x, y, r; mpc_init2 (x, pb); mpc_init2 (y, pb); mpc_init2 (r, pb);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < M; i++)
          {
            cget (x, a, M, i); cget (y, b, M, i);
            if (op[1] == 'm') mpc_mul (r, x, y, MPC_RNDNN);
            else if (op[1] == 'a') mpc_add (r, x, y, MPC_RNDNN);
            else mpc_sub (r, x, y, MPC_RNDNN);
            cput (r, o, M, i);
          }
        mpc_clear (x); mpc_clear (y); mpc_clear (r);
      }
      write_arr (fo, o);
    }
  else if (!strcmp (op, "axpy"))
    {
      arr a = read_arr (fi, 1), x = read_arr (fi, M), y = read_arr (fi, M), o = new_arr (M);
#pragma omp parallel
      {
        mpfr_t ma, mx, my, t, r; mpfr_inits2 (pb, ma, mx, my, t, r, (mpfr_ptr) 0);
        to_mpfr (ma, a.l, a.se, 1, 0);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < M; i++)
          {
            to_mpfr (mx, x.l, x.se, M, i); to_mpfr (my, y.l, y.se, M, i);
            mpfr_mul (t, ma, mx, MPFR_RNDN); mpfr_add (r, t, my, MPFR_RNDN);
            from_mpfr (r, o.l, o.se, M, i);
          }
        mpfr_clears (ma, mx, my, t, r, (mpfr_ptr) 0);
      }
      write_arr (fo, o);
    }
  else if (!strcmp (op, "caxpy"))
    {
      arr a = read_arr (fi, 2), x = read_arr (fi, 2 * M), y = read_arr (fi, 2 * M), o = new_arr (2 * M);
#pragma omp parallel
      {
        mpc_t ma, mx, my, t, r;
        mpc_init2 (ma, pb); mpc_init2 (mx, pb); mpc_init2 (my, pb); mpc_init2 (t, pb); mpc_init2 (r, pb);
        cget (ma, a, 1, 0);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < M; i++)
          {
            cget (mx, x, M, i); cget (my, y, M, i);
            mpc_mul (t, ma, mx, MPC_RNDNN); mpc_add (r, t, my, MPC_RNDNN);
            cput (r, o, M, i);
          }
        mpc_clear (ma); mpc_clear (mx); mpc_clear (my); mpc_clear (t); mpc_clear (r);
      }
      write_arr (fo, o);
    }
  else if (!strcmp (op, "gemm_perop") || !strcmp (op, "gemm_exact"))
    {
      int exact = !strcmp (op, "gemm_exact");
      arr A = read_arr (fi, M * K), B = read_arr (fi, K * Nn), C = new_arr (M * Nn);
#pragma omp parallel
      {
        mpfr_t acc, x, y, t; mpfr_inits2 (pb, acc, x, y, t, (mpfr_ptr) 0);
        mpfr_t *prod = malloc (sizeof (mpfr_t) * (size_t) (K ? K : 1));
        mpfr_ptr *tab = malloc (sizeof (mpfr_ptr) * (size_t) (K ? K : 1));
        for (int64_t k = 0; k < K; k++) { mpfr_init2 (prod[k], 2 * pb); tab[k] = prod[k]; }
#pragma omp for schedule(static) collapse(2)
        for (int64_t i = 0; i < M; i++)
          for (int64_t j = 0; j < Nn; j++)
            {
              if (!exact)
                {
                  /* tools/linalg_fast.cu:71-83: set_zero, then mul/add RNDN in k order */
                  mpfr_set_zero (acc, 1);
                  for (int64_t k = 0; k < K; k++)
                    {
                      to_mpfr (x, A.l, A.se, M * K, i * K + k);
                      to_mpfr (y, B.l, B.se, K * Nn, k * Nn + j);
                      mpfr_mul (t, x, y, MPFR_RNDN); mpfr_add (acc, acc, t, MPFR_RNDN);
                    }
                }
              else
                {
                  /* exact products at 2*PB bits, then one correctly rounded sum */
                  for (int64_t k = 0; k < K; k++)
                    {
                      to_mpfr (x, A.l, A.se, M * K, i * K + k);
                      to_mpfr (y, B.l, B.se, K * Nn, k * Nn + j);
                      if (mpfr_mul (prod[k], x, y, MPFR_RNDN) != 0) die ("inexact product");
                    }
                  if (K == 0) mpfr_set_zero (acc, 1);
                  else mpfr_sum (acc, tab, (unsigned long) K, MPFR_RNDN);
                }
              from_mpfr (acc, C.l, C.se, M * Nn, i * Nn + j);
            }
        for (int64_t k = 0; k < K; k++) mpfr_clear (prod[k]);
        free (prod); free (tab);
        mpfr_clears (acc, x, y, t, (mpfr_ptr) 0);
      }
      write_arr (fo, C);
    }
  else if (!strcmp (op, "cgemm_perop"))
    {
      arr A = read_arr (fi, 2 * M * K), B = read_arr (fi, 2 * K * Nn), C = new_arr (2 * M * Nn);
#pragma omp parallel
      {
        mpc_t acc, x, y, t; mpc_init2 (acc, pb); mpc_init2 (x, pb); mpc_init2 (y, pb); mpc_init2 (t, pb);
#pragma omp for schedule(static) collapse(2)
        for (int64_t i = 0; i < M; i++)
          for (int64_t j = 0; j < Nn; j++)
            {
              mpc_set_ui (acc, 0, MPC_RNDNN);
              for (int64_t k = 0; k < K; k++)
                {
                  cget (x, A, M * K, i * K + k); cget (y, B, K * Nn, k * Nn + j);
                  mpc_mul (t, x, y, MPC_RNDNN); mpc_add (acc, acc, t, MPC_RNDNN);
                }
              cput (acc, C, M * Nn, i * Nn + j);
            }
        mpc_clear (acc); mpc_clear (x); mpc_clear (y); mpc_clear (t);
      }
      write_arr (fo, C);
    }
  else if (!strcmp (op, "cgemm_exact"))
    {
      /* Re C = RNDN(sum_k ar*br - ai*bi), Im C = RNDN(sum_k ar*bi + ai*br): every
       * product exact at 2*PB bits, one correctly rounded sum per component */
      arr A = read_arr (fi, 2 * M * K), B = read_arr (fi, 2 * K * Nn), C = new_arr (2 * M * Nn);
#pragma omp parallel
      {
        mpfr_t acc, ar, ai, br, bi;
        mpfr_inits2 (pb, acc, ar, ai, br, bi, (mpfr_ptr) 0);
        const int64_t P = 2 * (K ? K : 1);
        mpfr_t *prod = malloc (sizeof (mpfr_t) * (size_t) P);
        mpfr_ptr *tab = malloc (sizeof (mpfr_ptr) * (size_t) P);
        for (int64_t q = 0; q < P; q++) { mpfr_init2 (prod[q], 2 * pb); tab[q] = prod[q]; }
#pragma omp for schedule(static) collapse(2)
        for (int64_t i = 0; i < M; i++)
          for (int64_t j = 0; j < Nn; j++)
            for (int part = 0; part < 2; part++)
              {
                for (int64_t k = 0; k < K; k++)
                  {
                    const int64_t ia = i * K + k, ib = k * Nn + j;
                    to_mpfr (ar, A.l, A.se, M * K, ia);
                    to_mpfr (ai, A.l + (int64_t) N * M * K, A.se + M * K, M * K, ia);
                    to_mpfr (br, B.l, B.se, K * Nn, ib);
                    to_mpfr (bi, B.l + (int64_t) N * K * Nn, B.se + K * Nn, K * Nn, ib);
                    if (part == 0)
                      {
                        if (mpfr_mul (prod[2 * k], ar, br, MPFR_RNDN) != 0) die ("inexact product");
                        if (mpfr_mul (prod[2 * k + 1], ai, bi, MPFR_RNDN) != 0) die ("inexact product");
                        mpfr_neg (prod[2 * k + 1], prod[2 * k + 1], MPFR_RNDN);
                      }
                    else
                      {
                        if (mpfr_mul (prod[2 * k], ar, bi, MPFR_RNDN) != 0) die ("inexact product");
                        if (mpfr_mul (prod[2 * k + 1], ai, br, MPFR_RNDN) != 0) die ("inexact product");
                      }
                  }
                if (K == 0) mpfr_set_zero (acc, 1);
                else mpfr_sum (acc, tab, (unsigned long) (2 * K), MPFR_RNDN);
                from_mpfr (acc, C.l + (int64_t) part * N * M * Nn, C.se + (int64_t) part * M * Nn, M * Nn, i * Nn + j);
              }
        for (int64_t q = 0; q < P; q++) mpfr_clear (prod[q]);
        free (prod); free (tab);
        mpfr_clears (acc, ar, ai, br, bi, (mpfr_ptr) 0);
      }
      write_arr (fo, C);
    }
  else die ("unknown op");
  fclose (fi); fclose (fo);
  return 0;
}
