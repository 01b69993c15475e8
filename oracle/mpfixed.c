/* oracle/mpfixed.c -- TEST INFRASTRUCTURE ONLY (the oracle; never shipped or
 * measured as the product).  A plain-C restatement, at runtime precision, of the
 * fixed-precision RNDN arithmetic of tkouya/mpc_cuda.  Every function cites the
 * reference lines whose algorithm it follows; the code is written independently.
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -shared).
 */
#include "mpfixed.h"
#include <string.h>
#include <stdlib.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

int or_nlimbs (int pb) { return (pb + 63) / 64; }

int or_num_threads (void)
{
#ifdef _OPENMP
  return omp_get_max_threads ();
#else
  return 1;
#endif
}

/* cu_freal.cuh:79-81 (set_zero: sign +1, exp = ZERO flag, limbs 0) */
void or_set_zero (int pb, or_real *r)
{
  int N = or_nlimbs (pb);
  r->sign = 1; r->exp = OR_ZERO_EXP;
  for (int i = 0; i < N; i++) r->m[i] = 0;
}

static int is_zero (const or_real *x) { return x->exp == OR_ZERO_EXP; }

/* bit i of a little-endian limb buffer, 0 outside [0, 64*len) (cu_freal.cuh:52-54) */
static uint64_t bit_at (const uint64_t *D, int len, long i)
{
  if (i < 0 || i >= 64L * len) return 0;
  return (D[i >> 6] >> (i & 63)) & 1u;
}

/* 1 iff any of bits [0, i) of D is set (cu_freal.cuh:56-64) */
static uint64_t any_below (const uint64_t *D, int len, long i)
{
  if (i <= 0) return 0;
  long full = i >> 6;
  uint64_t acc = 0;
  for (long j = 0; j < full && j < len; j++) acc |= D[j];
  int rem = (int) (i & 63);
  if (rem && full < len) acc |= D[full] & ((1ULL << rem) - 1);
  return acc != 0;
}

/* the 64 bits of D starting at bit position p (p may be negative: zero-fill) */
static uint64_t window64 (const uint64_t *D, int len, long p)
{
  if (p < 0)
    {
      long q = -p;
      if (q >= 64 || len <= 0) return 0;
      return D[0] << q;
    }
  long w = p >> 6; int b = (int) (p & 63);
  uint64_t lo = (w < len) ? D[w] : 0;
  if (!b) return lo;
  uint64_t hi = (w + 1 < len) ? D[w + 1] : 0;
  return (lo >> b) | (hi << (64 - b));
}

/* The single rounding point, cu_freal.cuh:89-138 and doc/algorithm.tex:332-351:
 * keep PB bits ending at `msb`, round bit just below, sticky = OR of the rest
 * (plus brnd/bsti, material below D bit 0), round half to even, renormalise on
 * carry-out, left-justify by SB, exp = Ehi + 1. */
void or_finalize (int pb, const uint64_t *D, int len, long msb, int64_t ehi,
                  uint64_t brnd, uint64_t bsti, int sgn, or_real *r)
{
  int N = or_nlimbs (pb), SB = 64 * N - pb;
  if (msb < 0) { or_set_zero (pb, r); return; }
  long lo = msb - pb + 1;
  uint64_t V[OR_MAXN + 1];
  uint64_t rbit, sticky;
  if (lo > 0)
    {
      rbit = bit_at (D, len, lo - 1);
      sticky = brnd | bsti | any_below (D, len, lo - 1);
    }
  else if (lo == 0)
    {
      rbit = brnd; sticky = bsti;
    }
  else
    {
      rbit = 0; sticky = 0;            /* fewer than PB bits: exact */
    }
  for (int i = 0; i < N; i++) V[i] = window64 (D, len, lo + 64L * i);
  if (rbit && (sticky || (V[0] & 1)))
    {
      uint64_t c = 1;
      for (int i = 0; i < N && c; i++) { V[i] += c; c = (V[i] == 0); }
      int ov = SB ? (int) ((V[pb >> 6] >> (pb & 63)) & 1) : (int) c;
      if (ov)
        {
          for (int i = 0; i < N; i++) V[i] = 0;
          V[(pb - 1) >> 6] = 1ULL << ((pb - 1) & 63);
          ehi++;
        }
    }
  if (SB)
    for (int i = N - 1; i >= 0; i--)
      r->m[i] = (V[i] << SB) | (i ? V[i - 1] >> (64 - SB) : 0);
  else
    for (int i = 0; i < N; i++) r->m[i] = V[i];
  r->sign = sgn; r->exp = ehi + 1;
}

/* exact N x N -> 2N limb product, cu_freal.cuh:174-200 */
static void mul_limbs (int N, const uint64_t *a, const uint64_t *b, uint64_t *p)
{
  for (int i = 0; i < 2 * N; i++) p[i] = 0;
  for (int j = 0; j < N; j++)
    {
      uint64_t carry = 0;
      for (int i = 0; i < N; i++)
        {
          u128 t = (u128) a[i] * b[j] + p[i + j] + carry;
          p[i + j] = (uint64_t) t;
          carry = (uint64_t) (t >> 64);
        }
      p[N + j] = carry;
    }
}

/* cu_fmul, cu_freal.cuh:203-213 */
void or_mul (int pb, const or_real *a, const or_real *b, or_real *r)
{
  int N = or_nlimbs (pb);
  if (is_zero (a) || is_zero (b)) { or_set_zero (pb, r); return; }
  uint64_t P[2 * OR_MAXN];
  mul_limbs (N, a->m, b->m, P);
  long msb = (P[2 * N - 1] >> 63) ? 128L * N - 1 : 128L * N - 2;
  int64_t ehi = a->exp + b->exp - 128L * N + msb;
  or_finalize (pb, P, 2 * N, msb, ehi, 0, 0, a->sign * b->sign, r);
}

/* cu_cmpmag, cu_freal.cuh:216-222 */
int or_cmpmag (int pb, const or_real *a, const or_real *b)
{
  int N = or_nlimbs (pb);
  if (a->exp != b->exp) return a->exp > b->exp ? 1 : -1;
  for (int i = N - 1; i >= 0; i--)
    if (a->m[i] != b->m[i]) return a->m[i] > b->m[i] ? 1 : -1;
  return 0;
}

/* shift a len-limb buffer right by d bits into out[len]; returns 1 iff any set
 * bit was shifted out (d may exceed the buffer) */
static uint64_t shr_sticky (const uint64_t *src, int len, long d, uint64_t *out)
{
  if (d >= 64L * len)
    {
      uint64_t any = 0;
      for (int i = 0; i < len; i++) { any |= src[i]; out[i] = 0; }
      return any != 0;
    }
  long ws = d >> 6; int bs = (int) (d & 63);
  for (int i = 0; i < len; i++)
    {
      long s = ws + i;
      uint64_t x = (s < len) ? src[s] : 0, y = (s + 1 < len) ? src[s + 1] : 0;
      out[i] = bs ? ((x >> bs) | (y << (64 - bs))) : x;
    }
  return any_below (src, len, d);
}

/* |a|+|b|, cu_freal.cuh:225-245: align the smaller-exponent operand by d into
 * N limbs, its shifted-out bits give round (bit d-1) and sticky (below), sum in
 * N+1 limbs, round once. */
static void addmag (int pb, const or_real *a0, const or_real *b0, int sgn, or_real *r)
{
  int N = or_nlimbs (pb);
  const or_real *a = a0, *b = b0;
  if (b->exp > a->exp) { a = b0; b = a0; }
  long d = (long) (a->exp - b->exp);
  uint64_t Bs[OR_MAXN], D[OR_MAXN + 1];
  if (d >= 64L * N) for (int i = 0; i < N; i++) Bs[i] = 0;
  else (void) shr_sticky (b->m, N, d, Bs);
  uint64_t brnd = bit_at (b->m, N, d - 1), bsti = any_below (b->m, N, d - 1);
  uint64_t c = 0;
  for (int i = 0; i < N; i++)
    {
      u128 s = (u128) a->m[i] + Bs[i] + c;
      D[i] = (uint64_t) s; c = (uint64_t) (s >> 64);
    }
  D[N] = c;
  long msb = c ? 64L * N : 64L * N - 1;
  int64_t ehi = a->exp - 64L * N + msb;
  or_finalize (pb, D, N + 1, msb, ehi, brnd, bsti, sgn, r);
}

/* |a|-|b| with |a|>|b|, cu_freal.cuh:248-277: N+2 limbs (two guard limbs), the
 * bits of b shifted below the window are replaced by subtracting one unit at the
 * window LSB and setting sticky (exact floor of the true difference). */
static void submag (int pb, const or_real *a, const or_real *b, int sgn, or_real *r)
{
  int N = or_nlimbs (pb), WL = N + 2;
  long d = (long) (a->exp - b->exp);
  uint64_t A[OR_MAXN + 2], T[OR_MAXN + 2], B[OR_MAXN + 2], D[OR_MAXN + 2];
  for (int i = 0; i < WL; i++) { A[i] = 0; T[i] = 0; }
  for (int i = 0; i < N; i++) { A[i + 2] = a->m[i]; T[i + 2] = b->m[i]; }
  uint64_t ext = shr_sticky (T, WL, d, B);
  uint64_t br = 0;
  for (int i = 0; i < WL; i++)
    {
      u128 s = (u128) A[i] - B[i] - br;
      D[i] = (uint64_t) s; br = (uint64_t) (s >> 64) & 1;
    }
  if (ext)
    for (int i = 0, bb = 1; i < WL && bb; i++) { bb = (D[i] == 0); D[i] -= 1; }
  int hl = -1;
  for (int i = WL - 1; i >= 0; i--) if (D[i]) { hl = i; break; }
  if (hl < 0) { or_set_zero (pb, r); return; }
  int top = 63 - __builtin_clzll (D[hl]);
  long msb = 64L * hl + top;
  int64_t ehi = a->exp - 64L * WL + msb;
  or_finalize (pb, D, WL, msb, ehi, 0, ext, sgn, r);
}

/* cu_fadd, cu_freal.cuh:280-289 (zero operand returns the other one as is;
 * exact cancellation gives +0) */
void or_add (int pb, const or_real *a, const or_real *b, or_real *r)
{
  if (is_zero (a)) { *r = *b; return; }
  if (is_zero (b)) { *r = *a; return; }
  if (a->sign == b->sign) { addmag (pb, a, b, a->sign, r); return; }
  int c = or_cmpmag (pb, a, b);
  if (c == 0) { or_set_zero (pb, r); return; }
  if (c > 0) submag (pb, a, b, a->sign, r);
  else submag (pb, b, a, b->sign, r);
}

/* cu_fsub, cu_freal.cuh:290-292 (negate the sign field of b, even for zero) */
void or_sub (int pb, const or_real *a, const or_real *b, or_real *r)
{
  or_real nb = *b;
  nb.sign = -nb.sign;
  or_add (pb, a, &nb, r);
}

/* ---------------- exact wide terms (cu_fcomplex.cuh:24-96) ---------------- */
typedef struct { int sign; int64_t exp; int zero; uint64_t M[2 * OR_MAXN]; } or_term;

/* cu_mk_term, cu_fcomplex.cuh:30-42: exact normalised 2N-limb product,
 * value = sign * M * 2^(exp - 128N), M's top bit set */
static void mk_term (int pb, const or_real *a, const or_real *b, or_term *t)
{
  int N = or_nlimbs (pb), WN = 2 * N;
  if (is_zero (a) || is_zero (b))
    {
      t->zero = 1; t->sign = 1; t->exp = 0;
      for (int i = 0; i < WN; i++) t->M[i] = 0;
      return;
    }
  t->zero = 0; t->sign = a->sign * b->sign;
  mul_limbs (N, a->m, b->m, t->M);
  if (t->M[WN - 1] >> 63) t->exp = a->exp + b->exp;
  else
    {
      for (int i = WN - 1; i > 0; i--) t->M[i] = (t->M[i] << 1) | (t->M[i - 1] >> 63);
      t->M[0] <<= 1;
      t->exp = a->exp + b->exp - 1;
    }
}

/* cu_wcmp, cu_fcomplex.cuh:44-50 */
static int wcmp (int WN, const or_term *a, const or_term *b)
{
  if (a->exp != b->exp) return a->exp > b->exp ? 1 : -1;
  for (int i = WN - 1; i >= 0; i--)
    if (a->M[i] != b->M[i]) return a->M[i] > b->M[i] ? 1 : -1;
  return 0;
}

/* cu_round_sum2, cu_fcomplex.cuh:53-88: RNDN(A+B) of two exact signed wide
 * terms; WN+3 limbs (two guard limbs below, one carry limb above), sticky
 * borrow as in submag, one finalize. */
static void round_sum2 (int pb, const or_term *A0, const or_term *B0, or_real *r)
{
  int N = or_nlimbs (pb), WN = 2 * N, WB = WN + 3;
  if (A0->zero && B0->zero) { or_set_zero (pb, r); return; }
  if (A0->zero) { or_finalize (pb, B0->M, WN, 64L * WN - 1, B0->exp - 1, 0, 0, B0->sign, r); return; }
  if (B0->zero) { or_finalize (pb, A0->M, WN, 64L * WN - 1, A0->exp - 1, 0, 0, A0->sign, r); return; }
  const or_term *A = A0, *B = B0;
  if (wcmp (WN, A, B) < 0) { A = B0; B = A0; }
  long d = (long) (A->exp - B->exp);
  uint64_t AB[2 * OR_MAXN + 3], T[2 * OR_MAXN + 3], BB[2 * OR_MAXN + 3], D[2 * OR_MAXN + 3];
  for (int i = 0; i < WB; i++) { AB[i] = 0; T[i] = 0; }
  for (int i = 0; i < WN; i++) { AB[i + 2] = A->M[i]; T[i + 2] = B->M[i]; }
  uint64_t sticky = shr_sticky (T, WB, d, BB);
  int sgn = A->sign;
  if (A->sign == B->sign)
    {
      uint64_t c = 0;
      for (int i = 0; i < WB; i++)
        {
          u128 s = (u128) AB[i] + BB[i] + c;
          D[i] = (uint64_t) s; c = (uint64_t) (s >> 64);
        }
    }
  else
    {
      uint64_t br = 0;
      for (int i = 0; i < WB; i++)
        {
          u128 s = (u128) AB[i] - BB[i] - br;
          D[i] = (uint64_t) s; br = (uint64_t) (s >> 64) & 1;
        }
      if (sticky)
        for (int i = 0, bb = 1; i < WB && bb; i++) { bb = (D[i] == 0); D[i] -= 1; }
    }
  int hl = -1;
  for (int i = WB - 1; i >= 0; i--) if (D[i]) { hl = i; break; }
  if (hl < 0) { or_set_zero (pb, r); return; }
  int top = 63 - __builtin_clzll (D[hl]);
  long msb = 64L * hl + top;
  int64_t ehi = A->exp - (int64_t) (WN + 2) * 64 + msb;
  or_finalize (pb, D, WB, msb, ehi, 0, sticky, sgn, r);
}

/* cu_fmma / cu_fmms, cu_fcomplex.cuh:91-96 */
void or_fmma (int pb, const or_real *a, const or_real *b, const or_real *c,
              const or_real *d, or_real *r)
{
  or_term t1, t2;
  mk_term (pb, a, b, &t1); mk_term (pb, c, d, &t2);
  round_sum2 (pb, &t1, &t2, r);
}
void or_fmms (int pb, const or_real *a, const or_real *b, const or_real *c,
              const or_real *d, or_real *r)
{
  or_term t1, t2;
  mk_term (pb, a, b, &t1); mk_term (pb, c, d, &t2);
  t2.sign = -t2.sign;
  round_sum2 (pb, &t1, &t2, r);
}

/* cu_cmul, cu_fcomplex.cuh:121-124 */
void or_cmul (int pb, const or_real *ar, const or_real *ai, const or_real *br,
              const or_real *bi, or_real *rr, or_real *ri)
{
  or_real tr, ti;
  or_fmms (pb, ar, br, ai, bi, &tr);
  or_fmma (pb, ar, bi, ai, br, &ti);
  *rr = tr; *ri = ti;
}

/* from_double, cu_freal.cuh:141-149 */
void or_from_double (int pb, double d, or_real *r)
{
  if (d == 0.0) { or_set_zero (pb, r); return; }
  int sgn = d < 0 ? -1 : 1;
  if (d < 0) d = -d;
  int k;
  double f = frexp (d, &k);
  uint64_t F = (uint64_t) ldexp (f, 53);
  or_finalize (pb, &F, 1, 52, (int64_t) k - 1, 0, 0, sgn, r);
}

/* to_double, cu_freal.cuh:150-165 (RNDN to 53 bits) */
double or_to_double (int pb, const or_real *x)
{
  int N = or_nlimbs (pb);
  if (is_zero (x)) return 0.0;
  long hb = 64L * N - 53;
  uint64_t F = window64 (x->m, N, hb) & ((1ULL << 53) - 1);
  uint64_t rbit = bit_at (x->m, N, hb - 1), st = any_below (x->m, N, hb - 1);
  int64_t e = x->exp;
  if (rbit && (st || (F & 1))) { F++; if (F == (1ULL << 53)) { F >>= 1; e++; } }
  return (double) x->sign * ldexp ((double) F, (int) (e - 53));
}

/* ---------------- division and square root (cu_fmath.cuh:25-100) ----------------
 * Faithful, not correctly rounded: Newton refinement at W = pb + 128 bits from
 * a double seed, every operation RNDN at W, the result rounded to pb. */
#define OR_FGUARD 128                   /* cu_fmath.cuh:21 */

/* exact small integer (cu_freal.cuh:326-334) */
static void or_from_si (int pb, long v, or_real *r)
{
  if (v == 0) { or_set_zero (pb, r); return; }
  int sgn = v < 0 ? -1 : 1;
  uint64_t a = v < 0 ? (uint64_t) (-(v + 1)) + 1u : (uint64_t) v;
  int nb = 64 - __builtin_clzll (a);
  or_finalize (pb, &a, 1, nb - 1, nb - 1, 0, 0, sgn, r);
}

/* re-round from precision `from` to `to` (cu_freal.cuh:317-323) */
static void or_convert (int to, int from, const or_real *x, or_real *r)
{
  if (is_zero (x)) { or_set_zero (to, r); return; }
  int N = or_nlimbs (from);
  or_finalize (to, x->m, N, 64L * N - 1, x->exp - 1, 0, 0, x->sign, r);
}

/* steps from a 52-bit seed to W bits, the seed counted (cu_fmath.cuh:25-26) */
static int or_newton_iters (int W)
{
  int k = 1;
  for (long p = 52; p < W; p += p) k++;
  return k;
}

/* 1/b at precision W (cu_fmath.cuh:53-66) */
static void or_recip_w (int W, const or_real *b, or_real *z)
{
  or_real m = *b, two, t, u;
  int sgn = b->sign;
  int64_t e = b->exp;
  m.sign = 1; m.exp = 0;
  or_from_double (W, 1.0 / or_to_double (W, &m), z);
  or_from_si (W, 2, &two);
  int it = or_newton_iters (W) + 1;
  for (int i = 0; i < it; i++)
    {
      or_mul (W, &m, z, &t);            /* m z       */
      or_sub (W, &two, &t, &u);         /* 2 - m z   */
      or_mul (W, z, &u, &t);            /* z (2-m z) */
      *z = t;
    }
  z->sign = sgn;
  z->exp -= e;
}

/* a / b (cu_fmath.cuh:68-75); a zero divisor gives +0 (the reference leaves it undefined) */
void or_fdiv (int pb, const or_real *a, const or_real *b, or_real *r)
{
  if (is_zero (a) || is_zero (b)) { or_set_zero (pb, r); return; }
  int W = pb + OR_FGUARD;
  or_real aw, bw, z, q;
  or_convert (W, pb, a, &aw);
  or_convert (W, pb, b, &bw);
  or_recip_w (W, &bw, &z);
  or_mul (W, &aw, &z, &q);
  or_convert (pb, W, &q, r);
}

/* sqrt(x), x > 0, at precision W (cu_fmath.cuh:78-92) */
static void or_sqrt_w (int W, const or_real *x, or_real *r)
{
  int64_t e = x->exp, t;
  or_real m = *x, y, three, y2, p, d, q;
  m.sign = 1;
  if (e & 1) { m.exp = 1; t = (e - 1) / 2; } else { m.exp = 0; t = e / 2; }
  or_from_double (W, 1.0 / sqrt (or_to_double (W, &m)), &y);
  or_from_si (W, 3, &three);
  int it = or_newton_iters (W) + 1;
  for (int i = 0; i < it; i++)
    {
      or_mul (W, &y, &y, &y2);          /* y^2               */
      or_mul (W, &m, &y2, &p);          /* m y^2             */
      or_sub (W, &three, &p, &d);       /* 3 - m y^2         */
      or_mul (W, &y, &d, &q);           /* y (3 - m y^2)     */
      if (!is_zero (&q)) q.exp -= 1;    /* / 2 (cu_scale2)   */
      y = q;
    }
  or_mul (W, &m, &y, r);
  r->exp += t;
}

/* sqrt(x); +0 for zero and negative x (cu_fmath.cuh:93-100) */
void or_fsqrt (int pb, const or_real *x, or_real *r)
{
  if (is_zero (x) || x->sign < 0) { or_set_zero (pb, r); return; }
  int W = pb + OR_FGUARD;
  or_real xw, s;
  or_convert (W, pb, x, &xw);
  or_sqrt_w (W, &xw, &s);
  or_convert (pb, W, &s, r);
}

/* ---------------- reference generator, tools/test_freal.cu:15-27 ---------------- */
uint64_t or_xs (uint64_t *s)
{
  uint64_t x = *s;
  x ^= x << 13; x ^= x >> 7; x ^= x << 17;
  return *s = x;
}

void or_randf (int pb, uint64_t *s, int spread, or_real *r)
{
  int N = or_nlimbs (pb), SB = 64 * N - pb;
  r->sign = (or_xs (s) & 1) ? 1 : -1;
  for (int i = 0; i < N; i++) r->m[i] = or_xs (s);
  r->m[N - 1] |= 1ULL << 63;
  if (SB) r->m[0] &= ~((1ULL << SB) - 1);
  r->exp = (int64_t) (or_xs (s) % (uint64_t) (2 * spread + 1)) - spread;
}

/* ---------------- SoA front-ends ---------------- */
void or_load (int pb, const uint64_t *limbs, const int64_t *se, int64_t n,
              int64_t i, or_real *r)
{
  int N = or_nlimbs (pb);
  for (int l = 0; l < N; l++) r->m[l] = limbs[(int64_t) l * n + i];
  int64_t w = se[i];
  r->sign = (w & 1) ? -1 : 1;
  r->exp = w >> 1;                      /* arithmetic shift */
}

void or_store (int pb, const or_real *r, uint64_t *limbs, int64_t *se,
               int64_t n, int64_t i)
{
  int N = or_nlimbs (pb);
  for (int l = 0; l < N; l++) limbs[(int64_t) l * n + i] = r->m[l];
  se[i] = (int64_t) ((uint64_t) r->exp << 1) | (r->sign < 0 ? 1 : 0);
}

void or_elementwise_real (int pb, int op, int64_t n,
    const uint64_t *al, const int64_t *ase, const uint64_t *bl, const int64_t *bse,
    uint64_t *ol, int64_t *ose)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    {
      or_real a, b, r;
      or_load (pb, al, ase, n, i, &a); or_load (pb, bl, bse, n, i, &b);
      if (op == 0) or_mul (pb, &a, &b, &r);
      else if (op == 1) or_add (pb, &a, &b, &r);
      else if (op == 2) or_sub (pb, &a, &b, &r);
      else or_fdiv (pb, &a, &b, &r);
      or_store (pb, &r, ol, ose, n, i);
    }
}

void or_sqrt_real (int pb, int64_t n, const uint64_t *xl, const int64_t *xse,
                   uint64_t *ol, int64_t *ose)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    {
      or_real x, r;
      or_load (pb, xl, xse, n, i, &x);
      or_fsqrt (pb, &x, &r);
      or_store (pb, &r, ol, ose, n, i);
    }
}

/* complex element i: re at index i of the first real array, im at index i of
 * the second (limbs + N*n, se + n) */
static void cload (int pb, const uint64_t *l, const int64_t *se, int64_t n, int64_t i,
                   or_real *re, or_real *im)
{
  int N = or_nlimbs (pb);
  or_load (pb, l, se, n, i, re);
  or_load (pb, l + (int64_t) N * n, se + n, n, i, im);
}
static void cstore (int pb, const or_real *re, const or_real *im, uint64_t *l,
                    int64_t *se, int64_t n, int64_t i)
{
  int N = or_nlimbs (pb);
  or_store (pb, re, l, se, n, i);
  or_store (pb, im, l + (int64_t) N * n, se + n, n, i);
}

void or_elementwise_complex (int pb, int op, int64_t n,
    const uint64_t *al, const int64_t *ase, const uint64_t *bl, const int64_t *bse,
    uint64_t *ol, int64_t *ose)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    {
      or_real ar, ai, br, bi, rr, ri;
      cload (pb, al, ase, n, i, &ar, &ai); cload (pb, bl, bse, n, i, &br, &bi);
      if (op == 0) or_cmul (pb, &ar, &ai, &br, &bi, &rr, &ri);
      else if (op == 1) { or_add (pb, &ar, &br, &rr); or_add (pb, &ai, &bi, &ri); }
      else { or_sub (pb, &ar, &br, &rr); or_sub (pb, &ai, &bi, &ri); }
      cstore (pb, &rr, &ri, ol, ose, n, i);
    }
}

void or_axpy_real (int pb, int64_t n, const uint64_t *al, const int64_t *ase,
    const uint64_t *xl, const int64_t *xse, const uint64_t *yl, const int64_t *yse,
    uint64_t *ol, int64_t *ose)
{
  or_real a;
  or_load (pb, al, ase, 1, 0, &a);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    {
      or_real x, y, t, r;
      or_load (pb, xl, xse, n, i, &x); or_load (pb, yl, yse, n, i, &y);
      or_mul (pb, &a, &x, &t);
      or_add (pb, &t, &y, &r);
      or_store (pb, &r, ol, ose, n, i);
    }
}

void or_axpy_complex (int pb, int64_t n, const uint64_t *al, const int64_t *ase,
    const uint64_t *xl, const int64_t *xse, const uint64_t *yl, const int64_t *yse,
    uint64_t *ol, int64_t *ose)
{
  or_real ar, ai;
  cload (pb, al, ase, 1, 0, &ar, &ai);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    {
      or_real xr, xi, yr, yi, tr, ti, rr, ri;
      cload (pb, xl, xse, n, i, &xr, &xi); cload (pb, yl, yse, n, i, &yr, &yi);
      or_cmul (pb, &ar, &ai, &xr, &xi, &tr, &ti);
      or_add (pb, &tr, &yr, &rr); or_add (pb, &ti, &yi, &ri);
      cstore (pb, &rr, &ri, ol, ose, n, i);
    }
}

void or_gemm_perop (int pb, int64_t m, int64_t n, int64_t k,
    const uint64_t *Al, const int64_t *Ase, const uint64_t *Bl, const int64_t *Bse,
    uint64_t *Cl, int64_t *Cse, int64_t r0, int64_t r1)
{
  (void) m;
  int64_t nA = m * k, nB = k * n, nC = m * n;
#pragma omp parallel for schedule(static) collapse(2)
  for (int64_t i = r0; i < r1; i++)
    for (int64_t j = 0; j < n; j++)
      {
        or_real acc, a, b, t, s;
        or_set_zero (pb, &acc);
        for (int64_t kk = 0; kk < k; kk++)
          {
            or_load (pb, Al, Ase, nA, i * k + kk, &a);
            or_load (pb, Bl, Bse, nB, kk * n + j, &b);
            or_mul (pb, &a, &b, &t);
            or_add (pb, &acc, &t, &s);
            acc = s;
          }
        or_store (pb, &acc, Cl, Cse, nC, i * n + j);
      }
}

void or_cgemm_perop (int pb, int64_t m, int64_t n, int64_t k,
    const uint64_t *Al, const int64_t *Ase, const uint64_t *Bl, const int64_t *Bse,
    uint64_t *Cl, int64_t *Cse, int64_t r0, int64_t r1)
{
  int64_t nA = m * k, nB = k * n, nC = m * n;
#pragma omp parallel for schedule(static) collapse(2)
  for (int64_t i = r0; i < r1; i++)
    for (int64_t j = 0; j < n; j++)
      {
        or_real accr, acci, ar, ai, br, bi, tr, ti, sr, si;
        or_set_zero (pb, &accr); or_set_zero (pb, &acci);
        for (int64_t kk = 0; kk < k; kk++)
          {
            cload (pb, Al, Ase, nA, i * k + kk, &ar, &ai);
            cload (pb, Bl, Bse, nB, kk * n + j, &br, &bi);
            or_cmul (pb, &ar, &ai, &br, &bi, &tr, &ti);
            or_add (pb, &accr, &tr, &sr); or_add (pb, &acci, &ti, &si);
            accr = sr; acci = si;
          }
        cstore (pb, &accr, &acci, Cl, Cse, nC, i * n + j);
      }
}

/* Exact dot: every product exact (2N limbs), all of them summed exactly in a
 * two's-complement fixed-point accumulator whose LSB is the lowest product bit,
 * then ONE RNDN rounding to PB bits (the mpfr_sum semantics that
 * src/mpfr_cuda/dot.c:47-54 relies on).  An exactly-zero sum gives +0. */
static void exact_dot (int pb, int64_t k, const or_real *const *a, const or_real *const *b,
                       or_real *r, uint64_t *acc, int acc_cap)
{
  int N = or_nlimbs (pb);
  int64_t emin = INT64_MAX, emax = INT64_MIN;
  for (int64_t kk = 0; kk < k; kk++)
    {
      if (is_zero (a[kk]) || is_zero (b[kk])) continue;
      int64_t e = a[kk]->exp + b[kk]->exp;         /* product < 2^e */
      if (e < emin) emin = e;
      if (e > emax) emax = e;
    }
  if (emin == INT64_MAX) { or_set_zero (pb, r); return; }
  /* product bit j has weight 2^(e - 128N + j); accumulator bit 0 = 2^(emin-128N) */
  long span = (long) (emax - emin) + 128L * N + 64 + 64;  /* headroom for k terms and sign */
  int L = (int) ((span + 63) / 64);
  if (L > acc_cap) L = acc_cap;                   /* caller sizes acc_cap >= L */
  for (int i = 0; i < L; i++) acc[i] = 0;
  uint64_t P[2 * OR_MAXN];
  for (int64_t kk = 0; kk < k; kk++)
    {
      if (is_zero (a[kk]) || is_zero (b[kk])) continue;
      mul_limbs (N, a[kk]->m, b[kk]->m, P);
      long off = (long) (a[kk]->exp + b[kk]->exp - emin);
      int ws = (int) (off >> 6), bs = (int) (off & 63);
      int neg = a[kk]->sign * b[kk]->sign < 0;
      /* add or subtract P << off */
      uint64_t c = 0, br = 0;
      for (int i = ws; i < L; i++)
        {
          int s = i - ws;
          uint64_t lo = (s < 2 * N) ? P[s] : 0;
          uint64_t hi = (s >= 1 && s - 1 < 2 * N) ? P[s - 1] : 0;
          uint64_t w = bs ? ((lo << bs) | (hi >> (64 - bs))) : lo;
          if (s > 2 * N && w == 0 && c == 0 && br == 0) break;
          if (!neg)
            {
              u128 t = (u128) acc[i] + w + c;
              acc[i] = (uint64_t) t; c = (uint64_t) (t >> 64);
            }
          else
            {
              u128 t = (u128) acc[i] - w - br;
              acc[i] = (uint64_t) t; br = (uint64_t) (t >> 64) & 1;
            }
        }
    }
  int sgn = 1;
  if (acc[L - 1] >> 63)
    {                                              /* negate */
      sgn = -1;
      uint64_t c = 1;
      for (int i = 0; i < L; i++) { u128 t = (u128) (~acc[i]) + c; acc[i] = (uint64_t) t; c = (uint64_t) (t >> 64); }
    }
  int hl = -1;
  for (int i = L - 1; i >= 0; i--) if (acc[i]) { hl = i; break; }
  if (hl < 0) { or_set_zero (pb, r); return; }
  long msb = 64L * hl + (63 - __builtin_clzll (acc[hl]));
  int64_t ehi = emin - 128L * N + msb;
  or_finalize (pb, acc, L, msb, ehi, 0, 0, sgn, r);
}

void or_gemm_exact (int pb, int64_t m, int64_t n, int64_t k,
    const uint64_t *Al, const int64_t *Ase, const uint64_t *Bl, const int64_t *Bse,
    uint64_t *Cl, int64_t *Cse, int64_t r0, int64_t r1)
{
  int64_t nA = m * k, nB = k * n, nC = m * n;
  int N = or_nlimbs (pb);
#pragma omp parallel
  {
    or_real *as = (or_real *) malloc (sizeof (or_real) * (size_t) (k ? k : 1));
    or_real *bs = (or_real *) malloc (sizeof (or_real) * (size_t) (k ? k : 1));
    const or_real **pa = (const or_real **) malloc (sizeof (void *) * (size_t) (k ? k : 1));
    const or_real **pbb = (const or_real **) malloc (sizeof (void *) * (size_t) (k ? k : 1));
    int cap = 0;
    uint64_t *acc = NULL;
#pragma omp for schedule(static) collapse(2)
    for (int64_t i = r0; i < r1; i++)
      for (int64_t j = 0; j < n; j++)
        {
          int64_t emin = INT64_MAX, emax = INT64_MIN;
          for (int64_t kk = 0; kk < k; kk++)
            {
              or_load (pb, Al, Ase, nA, i * k + kk, &as[kk]);
              or_load (pb, Bl, Bse, nB, kk * n + j, &bs[kk]);
              pa[kk] = &as[kk]; pbb[kk] = &bs[kk];
              if (!is_zero (&as[kk]) && !is_zero (&bs[kk]))
                {
                  int64_t e = as[kk].exp + bs[kk].exp;
                  if (e < emin) emin = e;
                  if (e > emax) emax = e;
                }
            }
          int need = 4;
          if (emin != INT64_MAX) need = (int) (((emax - emin) + 128L * N + 128 + 63) / 64);
          if (need > cap) { free (acc); cap = need; acc = (uint64_t *) malloc (sizeof (uint64_t) * (size_t) cap); }
          or_real res;
          exact_dot (pb, k, pa, pbb, &res, acc, cap);
          or_store (pb, &res, Cl, Cse, nC, i * n + j);
        }
    free (acc); free (as); free (bs); free (pa); free (pbb);
  }
}

void or_fill_ref (int pb, int64_t n, uint64_t seed, int spread,
                  uint64_t *limbs, int64_t *se)
{
  uint64_t s = seed;
  for (int64_t i = 0; i < n; i++)
    {
      or_real r;
      int sp = spread >= 0 ? spread : ((i & 3) ? 40 : pb + 200);
      or_randf (pb, &s, sp, &r);
      or_store (pb, &r, limbs, se, n, i);
    }
}
