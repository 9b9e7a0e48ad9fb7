/*
 * gpugo_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the GPUGO reference algorithm, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker
 * and as the host-core baseline.  Nothing in the product (gpugo_amd/,
 * libgpugo.so) links, imports or calls this file.
 *
 * What it restates (reference paths relative to GPUGO_1.1/):
 *   - interval ops of "Updated Interval Arithmetic Operations/IA.py":57-570.
 *     + - * / and sqrt use TRUE directed rounding: the FPU is held in
 *     FE_UPWARD and round-down results are formed as -RU(-x) (this replaces
 *     libdevice dadd_rd/ru, dmul_rd/ru, ddiv_rd/ru, IA.py:83-84,143-150,
 *     185-206).  exp/sin/cos call glibc in round-to-nearest and widen the
 *     result outward by ORACLE_W_TRANS ulps (glibc documents <= 1 ulp for
 *     these double functions on x86_64; the extra ulp is margin).  This
 *     follows the shape of IA.py:363-570 (RN value, widened on the bit
 *     pattern) without its CUDA-specific ulp constants.
 *   - deliberate departures from IA.py, following the mathematics:
 *       times() is sign-aware (IA.py:306-333 assumes a positive scalar);
 *       divide() by [l,0] uses the textbook endpoints (IA.py:193-202 divides
 *       by upper2 == 0); power(.,2) of [0,h] has lower bound exactly 0
 *       (IA.py:267-281 returns -2.5e-323); sin/cos extremum detection uses
 *       rigorous bounds of x/pi instead of IA.py:429-444's nudged pi/2 floor
 *       division.
 *   - subregion decode of selectVarInSubRegion / completeRegion
 *     (1-Ackley/ackley50.py:482-594, :820-948): base-`wid` digits of `pos`,
 *     least significant digit for the first coordinate of the window,
 *     child bounds L + (U-L)/wid * digit (computed here with one fma so the
 *     GPU can reproduce it bit for bit; the top child ends exactly at U).
 *   - the code lists <fn>Value_IA / <fn>Sample_IA / <fn>Derivative_IA
 *     (1-Ackley/ackley50.py:116-479, 7-Rastrigin/rastrigin50.py:110-372,
 *     11-Rosenbrock/rosenbrock50.py:115-451) for three families that contain
 *     those objectives as special cases, and Levy's (6-Levy/levy50.py:
 *     166-224) as a fourth (see family_* below).
 *   - the driver loop (1-Ackley/ackley50.py:1033-1380): sample children for
 *     the incumbent upper bound (sampling, :602-684), rule children out by
 *     value and by the monotonicity (KKT) test (ruleOutGPU, :686-766), prune
 *     the list when the upper bound drops (:1154-1220), pop the box with the
 *     smallest lower bound (:1270-1296), cycle the split window (:1362-1370).
 *     Generalised to K boxes per step (K = 1 is the reference), with the
 *     certified f-width stop (ub - lb <= tol_f) beside the reference's x-width
 *     output rule.
 */
#define _GNU_SOURCE
#include <fenv.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/gpugo.h"
#include "gpugo_oracle.h"

#pragma STDC FENV_ACCESS ON

#define ORACLE_W_TRANS 2

typedef struct { double lo, hi; } iv;

/* Bracketing doubles of irrational constants: X_LO < X < X_HI, adjacent. */
static const double INV_PI_LO = 0x1.45f306dc9c882p-2;
static const double INV_PI_HI = 0x1.45f306dc9c883p-2;
static const double TWO_PI_LO = 0x1.921fb54442d18p+2;
static const double E_LO      = 0x1.5bf0a8b145769p+1;
static const double E_HI      = 0x1.5bf0a8b14576ap+1;
/* above this magnitude the extremum search is skipped and [-1,1] returned */
static const double TRIG_BIG  = 0x1p40;
/* at or below this argument width one evaluation plus the Lipschitz bound */
static const double TRIG_NARROW = 0x1p-40;

/* ---------------------------------------------------------------------------
 * Directed rounding.  Every function in this section assumes FE_UPWARD.
 * ------------------------------------------------------------------------- */
static inline double ru_add(double a, double b) { return a + b; }
static inline double rd_add(double a, double b) { return -((-a) - b); }
static inline double ru_sub(double a, double b) { return a - b; }
static inline double rd_sub(double a, double b) { return -(b - a); }
static inline double ru_mul(double a, double b) { return (a == 0.0 || b == 0.0) ? 0.0 : a * b; }
static inline double rd_mul(double a, double b) { return (a == 0.0 || b == 0.0) ? 0.0 : -((-a) * b); }
static inline double ru_div(double a, double b) { return a / b; }
static inline double rd_div(double a, double b) { return -((-a) / b); }

static inline double ru_sqrt(double a) { return sqrt(a); }
static inline double rd_sqrt(double a)
{
  fesetround(FE_DOWNWARD);
  volatile double r = sqrt(a);
  fesetround(FE_UPWARD);
  return r;
}

static double down_ulps(double x, int k) { for (int i = 0; i < k; ++i) x = nextafter(x, -INFINITY); return x; }
static double up_ulps(double x, int k)   { for (int i = 0; i < k; ++i) x = nextafter(x, INFINITY); return x; }

static double exp_rn(double x) { fesetround(FE_TONEAREST); volatile double r = exp(x); fesetround(FE_UPWARD); return r; }
static double cos_rn(double x) { fesetround(FE_TONEAREST); volatile double r = cos(x); fesetround(FE_UPWARD); return r; }
static double sin_rn(double x) { fesetround(FE_TONEAREST); volatile double r = sin(x); fesetround(FE_UPWARD); return r; }

static inline iv mk(double lo, double hi) { iv r; r.lo = lo; r.hi = hi; return r; }
static inline iv entire(void) { return mk(-INFINITY, INFINITY); }

/* ---------------------------------------------------------------------------
 * Interval operations (FE_UPWARD held by the caller).
 * ------------------------------------------------------------------------- */
static iv iv_add(iv a, iv b)   { return mk(rd_add(a.lo, b.lo), ru_add(a.hi, b.hi)); }
static iv iv_sub(iv a, iv b)   { return mk(rd_sub(a.lo, b.hi), ru_sub(a.hi, b.lo)); }
static iv iv_neg(iv a)         { return mk(-a.hi, -a.lo); }
static iv iv_addc(iv a, double s)  { return mk(rd_add(a.lo, s), ru_add(a.hi, s)); }
static iv iv_rsubc(double s, iv a) { return mk(rd_sub(s, a.hi), ru_sub(s, a.lo)); }

static iv iv_scale(iv a, double s)
{
  if (s > 0) return mk(rd_mul(a.lo, s), ru_mul(a.hi, s));
  if (s < 0) return mk(rd_mul(a.hi, s), ru_mul(a.lo, s));
  return mk(0.0, 0.0);
}

static double min4(double a, double b, double c, double d)
{ double m = a; if (b < m) m = b; if (c < m) m = c; if (d < m) m = d; return m; }
static double max4(double a, double b, double c, double d)
{ double m = a; if (b > m) m = b; if (c > m) m = c; if (d > m) m = d; return m; }

static iv iv_mul(iv a, iv b)
{
  return mk(min4(rd_mul(a.lo, b.lo), rd_mul(a.lo, b.hi), rd_mul(a.hi, b.lo), rd_mul(a.hi, b.hi)),
            max4(ru_mul(a.lo, b.lo), ru_mul(a.lo, b.hi), ru_mul(a.hi, b.lo), ru_mul(a.hi, b.hi)));
}

static iv iv_div(iv a, iv b)
{
  if (b.lo > 0 || b.hi < 0)
    return mk(min4(rd_div(a.lo, b.lo), rd_div(a.lo, b.hi), rd_div(a.hi, b.lo), rd_div(a.hi, b.hi)),
              max4(ru_div(a.lo, b.lo), ru_div(a.lo, b.hi), ru_div(a.hi, b.lo), ru_div(a.hi, b.hi)));
  if (b.lo == 0 && b.hi > 0) {            /* b = [0, bh] */
    if (a.lo >= 0) return mk(rd_div(a.lo, b.hi), INFINITY);
    if (a.hi <= 0) return mk(-INFINITY, ru_div(a.hi, b.hi));
    return entire();
  }
  if (b.hi == 0 && b.lo < 0) {            /* b = [bl, 0] */
    if (a.lo >= 0) return mk(-INFINITY, ru_div(a.lo, b.lo));
    if (a.hi <= 0) return mk(rd_div(a.hi, b.lo), INFINITY);
    return entire();
  }
  return entire();                        /* 0 inside b, or b = [0,0] */
}

static iv iv_divc(iv a, double s)
{
  if (s > 0) return mk(rd_div(a.lo, s), ru_div(a.hi, s));
  return mk(rd_div(a.hi, s), ru_div(a.lo, s));
}

static iv iv_sqr(iv a)
{
  if (a.lo >= 0) return mk(rd_mul(a.lo, a.lo), ru_mul(a.hi, a.hi));
  if (a.hi <= 0) return mk(rd_mul(a.hi, a.hi), ru_mul(a.lo, a.lo));
  double u1 = ru_mul(a.lo, a.lo), u2 = ru_mul(a.hi, a.hi);
  return mk(0.0, u1 > u2 ? u1 : u2);
}

/* x >= 0: binary powering, most significant bit first */
static double pow_ru(double x, int n)
{
  int top = 31 - __builtin_clz((unsigned)n);
  double r = x;
  for (int b = top - 1; b >= 0; --b) { r = ru_mul(r, r); if ((n >> b) & 1) r = ru_mul(r, x); }
  return r;
}
static double pow_rd(double x, int n)
{
  int top = 31 - __builtin_clz((unsigned)n);
  double r = x;
  for (int b = top - 1; b >= 0; --b) { r = rd_mul(r, r); if ((n >> b) & 1) r = rd_mul(r, x); }
  return r;
}

static iv iv_pown(iv a, int n)
{
  if (n <= 0) return mk(1.0, 1.0);
  if (n == 1) return a;
  if ((n & 1) == 0) {
    if (a.lo >= 0) return mk(pow_rd(a.lo, n), pow_ru(a.hi, n));
    if (a.hi <= 0) return mk(pow_rd(-a.hi, n), pow_ru(-a.lo, n));
    double m = -a.lo > a.hi ? -a.lo : a.hi;
    return mk(0.0, pow_ru(m, n));
  }
  double lo = a.lo >= 0 ? pow_rd(a.lo, n) : -pow_ru(-a.lo, n);
  double hi = a.hi >= 0 ? pow_ru(a.hi, n) : -pow_rd(-a.hi, n);
  return mk(lo, hi);
}

static iv iv_sqrt(iv a)
{
  if (a.hi < 0) return mk(NAN, NAN);
  return mk(a.lo <= 0 ? 0.0 : rd_sqrt(a.lo), ru_sqrt(a.hi));
}

static iv iv_exp(iv a)
{
  double lo = down_ulps(exp_rn(a.lo), ORACLE_W_TRANS);
  if (lo < 0) lo = 0.0;
  return mk(lo, up_ulps(exp_rn(a.hi), ORACLE_W_TRANS));
}

/* Lower bound of a/pi and upper bound of b/pi (FE_UPWARD held). */
static double divpi_lo(double a) { return a >= 0 ? rd_mul(a, INV_PI_LO) : rd_mul(a, INV_PI_HI); }
static double divpi_hi(double b) { return b >= 0 ? ru_mul(b, INV_PI_HI) : ru_mul(b, INV_PI_LO); }

/* shift = 0: cos, extrema at x = m*pi (m even: max, odd: min).
 * shift = 1: sin, extrema at x = pi/2 + m*pi (m even: max, odd: min). */
static iv iv_trig(iv a, int shift)
{
  if (!(a.lo <= a.hi)) return mk(NAN, NAN);
  if (fabs(a.lo) > TRIG_BIG || fabs(a.hi) > TRIG_BIG) return mk(-1.0, 1.0);
  double w = ru_sub(a.hi, a.lo);
  if (w <= TRIG_NARROW) {
    /* narrow argument: sin and cos are 1-Lipschitz, so one evaluation
     * widened by the argument width encloses the range */
    double v = shift ? sin_rn(a.lo) : cos_rn(a.lo);
    double lo = rd_sub(down_ulps(v, ORACLE_W_TRANS), w);
    double hi = ru_add(up_ulps(v, ORACLE_W_TRANS), w);
    return mk(lo < -1.0 ? -1.0 : lo, hi > 1.0 ? 1.0 : hi);
  }
  if (w >= TWO_PI_LO) return mk(-1.0, 1.0);
  double va = shift ? sin_rn(a.lo) : cos_rn(a.lo);
  double vb = shift ? sin_rn(a.hi) : cos_rn(a.hi);
  double lo = down_ulps(va < vb ? va : vb, ORACLE_W_TRANS);
  double hi = up_ulps(va < vb ? vb : va, ORACLE_W_TRANS);
  if (lo < -1.0) lo = -1.0;
  if (hi > 1.0) hi = 1.0;
  double ta = divpi_lo(a.lo), tb = divpi_hi(a.hi);
  if (shift) { ta = rd_sub(ta, 0.5); tb = ru_sub(tb, 0.5); }
  double m0 = ceil(ta), m1 = floor(tb);
  for (double m = m0; m <= m1; m += 1.0) {
    if ((((int64_t)m) & 1) == 0) hi = 1.0; else lo = -1.0;
  }
  return mk(lo, hi);
}
static iv iv_cos(iv a) { return iv_trig(a, 0); }
static iv iv_sin(iv a) { return iv_trig(a, 1); }

/* ---------------------------------------------------------------------------
 * Batched op entry (tests).  Same op codes as gpugo_ia_eval.
 * ------------------------------------------------------------------------- */
int oracle_ia_eval(int op, int64_t n, const double *alo, const double *ahi,
                   const double *blo, const double *bhi, double s,
                   double *olo, double *ohi)
{
  int old = fegetround();
  fesetround(FE_UPWARD);
  for (int64_t i = 0; i < n; ++i) {
    iv a = mk(alo[i], ahi[i]);
    iv b = (blo && bhi) ? mk(blo[i], bhi[i]) : mk(0.0, 0.0);
    iv r;
    switch (op) {
      case GPUGO_OP_ADD:   r = iv_add(a, b); break;
      case GPUGO_OP_SUB:   r = iv_sub(a, b); break;
      case GPUGO_OP_MUL:   r = iv_mul(a, b); break;
      case GPUGO_OP_DIV:   r = iv_div(a, b); break;
      case GPUGO_OP_NEG:   r = iv_neg(a); break;
      case GPUGO_OP_SCALE: r = iv_scale(a, s); break;
      case GPUGO_OP_ADDC:  r = iv_addc(a, s); break;
      case GPUGO_OP_RSUBC: r = iv_rsubc(s, a); break;
      case GPUGO_OP_DIVC:  r = iv_divc(a, s); break;
      case GPUGO_OP_SQR:   r = iv_sqr(a); break;
      case GPUGO_OP_POWN:  r = iv_pown(a, (int)s); break;
      case GPUGO_OP_SQRT:  r = iv_sqrt(a); break;
      case GPUGO_OP_EXP:   r = iv_exp(a); break;
      case GPUGO_OP_SIN:   r = iv_sin(a); break;
      case GPUGO_OP_COS:   r = iv_cos(a); break;
      default: fesetround(old); return -1;
    }
    olo[i] = r.lo; ohi[i] = r.hi;
  }
  fesetround(old);
  return 0;
}

/* ---------------------------------------------------------------------------
 * Objective families (code lists).  FE_UPWARD held by the caller.
 *
 * family 0, separable:       f = sum_i a_i x_i^2 + b_i (1 - cos(c_i x_i))
 *     params [a(d), b(d), c(d)].  Rastrigin (7-Rastrigin/rastrigin50.py:
 *     167-188) is a=1, b=10, c=2pi: x^2 - 10cos(2pi x) + 10 per coordinate.
 * family 1, separable+chain: family 0 + sum_{i<d-1} w_i (x_{i+1} - x_i^2)^2
 *     + (1 - x_i)^2, params [a(d), b(d), c(d), w(d-1)].  Rosenbrock
 *     (11-Rosenbrock/rosenbrock50.py:170-198) is a=b=0, w=100.
 * family 2, Ackley-type:     f = -A exp(-B sqrt(mean x^2)) - exp(mean cos(C x))
 *     + A + e + sum_i q_i x_i^4 / 50, params [A, B, C, q(d)].  Ackley
 *     (1-Ackley/ackley50.py:171-212) is A=20, B=0.02, C=2pi, q=0.
 * family 3, Levy (6-Levy/levy50.py:166-224), no params:
 *     f = (pi/d) [10 sin^2(pi/4 x_0 + 3pi/4)
 *                 + sum_{i<d-1} ((x_i - 1)/4)^2 (1 + 10 sin^2(pi/4 x_{i+1} + 3pi/4))
 *                 + ((x_{d-1} - 1)/4)^2],  pi held as the interval [PI_LO, PI_HI].
 * Sums are left folds in increasing i starting from [0,0]; the GPU code
 * lists use the same order.
 * ------------------------------------------------------------------------- */
static iv fam0_term(iv x, double a, double b, double c)
{
  iv t1 = iv_scale(iv_sqr(x), a);
  iv t3 = iv_scale(iv_rsubc(1.0, iv_cos(iv_scale(x, c))), b);
  return iv_add(t1, t3);
}

static const double PI_LO = 0x1.921fb54442d18p+1, PI_HI = 0x1.921fb54442d19p+1;

/* k pi x + c pi */
static iv levy_arg(iv x, double k, double c)
{
  const iv pi = mk(PI_LO, PI_HI);
  return iv_add(iv_mul(x, iv_scale(pi, k)), iv_scale(pi, c));
}
/* ((x - 1)/4)^2 */
static iv levy_sq4(iv x) { return iv_sqr(iv_scale(iv_addc(x, -1.0), 0.25)); }
/* 1 + 10 sin^2(pi/4 x + 3pi/4) */
static iv levy_w(iv x) { return iv_addc(iv_scale(iv_sqr(iv_sin(levy_arg(x, 0.25, 0.75))), 10.0), 1.0); }
/* d/dx 10 sin^2(pi/4 x + 3pi/4) = (5pi/2) sin(pi/2 x + 3pi/2) */
static iv levy_dw(iv x) { return iv_mul(iv_scale(mk(PI_LO, PI_HI), 2.5), iv_sin(levy_arg(x, 0.5, 1.5))); }

static iv levy_value(int d, const iv *x)
{
  iv s = iv_scale(iv_sqr(iv_sin(levy_arg(x[0], 0.25, 0.75))), 10.0);
  for (int i = 0; i + 1 < d; ++i) s = iv_add(s, iv_mul(levy_sq4(x[i]), levy_w(x[i + 1])));
  s = iv_add(s, levy_sq4(x[d - 1]));
  return iv_mul(iv_divc(mk(PI_LO, PI_HI), (double)d), s);
}

static void levy_grad(int d, const iv *x, iv *g)
{
  const iv k = iv_divc(mk(PI_LO, PI_HI), (double)d);
  for (int j = 0; j < d; ++j) {
    iv gj = mk(0.0, 0.0);
    if (j == 0) gj = levy_dw(x[0]);
    if (j + 1 < d) gj = iv_add(gj, iv_mul(iv_scale(iv_addc(x[j], -1.0), 0.125), levy_w(x[j + 1])));
    if (j > 0) gj = iv_add(gj, iv_mul(levy_sq4(x[j - 1]), levy_dw(x[j])));
    if (j + 1 == d) gj = iv_add(gj, iv_scale(iv_addc(x[j], -1.0), 0.125));
    g[j] = iv_mul(k, gj);
  }
}

static iv family_value(int fam, int d, const double *P, const iv *x)
{
  if (fam == 3) return levy_value(d, x);
  if (fam == 0 || fam == 1) {
    const double *a = P, *b = P + d, *c = P + 2 * d;
    iv s = mk(0.0, 0.0);
    for (int i = 0; i < d; ++i) s = iv_add(s, fam0_term(x[i], a[i], b[i], c[i]));
    if (fam == 1) {
      const double *w = P + 3 * d;
      for (int i = 0; i + 1 < d; ++i) {
        iv u = iv_sub(x[i + 1], iv_sqr(x[i]));
        iv t = iv_scale(iv_sqr(u), w[i]);
        iv v = iv_sqr(iv_rsubc(1.0, x[i]));
        s = iv_add(s, iv_add(t, v));
      }
    }
    return s;
  }
  /* fam == 2 */
  double A = P[0], B = P[1], C = P[2];
  const double *q = P + 3;
  iv s1 = mk(0.0, 0.0), s2 = mk(0.0, 0.0), s3 = mk(0.0, 0.0);
  for (int i = 0; i < d; ++i) {
    s1 = iv_add(s1, iv_sqr(x[i]));
    s2 = iv_add(s2, iv_cos(iv_scale(x[i], C)));
    s3 = iv_add(s3, iv_scale(iv_pown(x[i], 4), q[i]));
  }
  iv r  = iv_sqrt(iv_divc(s1, (double)d));
  iv v1 = iv_scale(iv_exp(iv_scale(r, -B)), -A);
  iv t2 = iv_exp(iv_divc(s2, (double)d));
  iv ae = mk(rd_add(A, E_LO), ru_add(A, E_HI));
  return iv_add(iv_add(iv_sub(v1, t2), ae), iv_divc(s3, 50.0));
}

/* Interval gradient; g has d entries. */
static void family_grad(int fam, int d, const double *P, const iv *x, iv *g)
{
  if (fam == 3) {
    levy_grad(d, x, g);
    return;
  }
  if (fam == 0 || fam == 1) {
    const double *a = P, *b = P + d, *c = P + 2 * d;
    for (int j = 0; j < d; ++j) {
      iv t1 = iv_scale(x[j], 2.0 * a[j]);
      iv t2 = iv_scale(iv_scale(iv_sin(iv_scale(x[j], c[j])), c[j]), b[j]);
      g[j] = iv_add(t1, t2);
    }
    if (fam == 1) {
      const double *w = P + 3 * d;
      for (int j = 0; j < d; ++j) {
        iv gj = g[j];
        if (j + 1 < d) {
          iv u = iv_sub(x[j + 1], iv_sqr(x[j]));
          iv t = iv_sub(iv_scale(iv_mul(x[j], u), -4.0 * w[j]),
                        iv_scale(iv_rsubc(1.0, x[j]), 2.0));
          gj = iv_add(gj, t);
        }
        if (j > 0) {
          iv u = iv_sub(x[j], iv_sqr(x[j - 1]));
          gj = iv_add(gj, iv_scale(u, 2.0 * w[j - 1]));
        }
        g[j] = gj;
      }
    }
    return;
  }
  double A = P[0], B = P[1], C = P[2];
  const double *q = P + 3;
  iv s1 = mk(0.0, 0.0), s2 = mk(0.0, 0.0);
  for (int i = 0; i < d; ++i) {
    s1 = iv_add(s1, iv_sqr(x[i]));
    s2 = iv_add(s2, iv_cos(iv_scale(x[i], C)));
  }
  iv r  = iv_sqrt(iv_divc(s1, (double)d));
  iv t1 = iv_exp(iv_scale(r, -B));
  iv k1 = iv_divc(iv_scale(iv_scale(iv_div(t1, r), B), A), (double)d);
  iv k2 = iv_divc(iv_scale(iv_exp(iv_divc(s2, (double)d)), C), (double)d);
  for (int j = 0; j < d; ++j) {
    iv g1 = iv_mul(k1, x[j]);
    iv g2 = iv_mul(k2, iv_sin(iv_scale(x[j], C)));
    iv g3 = iv_divc(iv_scale(iv_pown(x[j], 3), 4.0 * q[j]), 50.0);
    g[j] = iv_add(iv_add(g1, g2), g3);
  }
}

/* Monotonicity test (1-Ackley/ackley50.py:458-477): prune when a gradient
 * component has constant sign and the box does not touch the initial box on
 * the side where the minimum over the box would lie. */
static int kkt_test(int d, const iv *x, const double *dlo, const double *dhi, const iv *g)
{
  for (int j = 0; j < d; ++j) {
    if (g[j].lo > 0 && x[j].lo != dlo[j]) return 1;
    if (g[j].hi < 0 && x[j].hi != dhi[j]) return 1;
  }
  return 0;
}

static int family_kkt(int fam, int d, const double *P, const iv *x,
                      const double *dlo, const double *dhi, iv *gbuf)
{
  family_grad(fam, d, P, x, gbuf);
  return kkt_test(d, x, dlo, dhi, gbuf);
}

/* Mean-value enclosure f(X) in f(c) + sum_j G_j (X_j - c_j), c the box
 * midpoint (computed in round-to-nearest), G an enclosure of the gradient
 * over X.  Returns its lower bound (FE_UPWARD held by the caller). */
static double family_mv_lower(int fam, int d, const double *P, const iv *x, const iv *g, iv *xs)
{
  fesetround(FE_TONEAREST);
  for (int j = 0; j < d; ++j) {
    double c = fma((x[j].hi - x[j].lo) / 2.0, 1.0, x[j].lo);
    xs[j] = mk(c, c);
  }
  fesetround(FE_UPWARD);
  iv s = family_value(fam, d, P, xs);
  for (int j = 0; j < d; ++j) {
    iv dx = mk(rd_sub(x[j].lo, xs[j].lo), ru_sub(x[j].hi, xs[j].lo));
    s = iv_add(s, iv_mul(g[j], dx));
  }
  return s.lo;
}

/* Round-to-nearest section: sample points along the diagonal
 * (1-Ackley/ackley50.py:288-290) -- x_j = lo_j + (hi_j-lo_j)/(S+1) * (t+1). */
static double family_sample_ub(int fam, int d, const double *P, const iv *x,
                               int S, iv *xs)
{
  double best = INFINITY;
  for (int t = 0; t < S; ++t) {
    fesetround(FE_TONEAREST);
    for (int j = 0; j < d; ++j) {
      double h = (x[j].hi - x[j].lo) / (double)(S + 1);
      double p = fma(h, (double)(t + 1), x[j].lo);
      xs[j] = mk(p, p);
    }
    fesetround(FE_UPWARD);
    iv f = family_value(fam, d, P, xs);
    if (f.hi < best) best = f.hi;
  }
  return best;
}

/* ---------------------------------------------------------------------------
 * Decode (round-to-nearest).
 * ------------------------------------------------------------------------- */
static void decode_child(int d, const double *plo, const double *phi, int win,
                         int wid, int c, uint32_t pos, double *clo, double *chi)
{
  memcpy(clo, plo, sizeof(double) * d);
  memcpy(chi, phi, sizeof(double) * d);
  for (int k = 0; k < c; ++k) {
    int j = (win + k) % d;
    uint32_t q = pos % (uint32_t)wid;
    pos /= (uint32_t)wid;
    double w = (phi[j] - plo[j]) / (double)wid;
    clo[j] = fma(w, (double)q, plo[j]);
    chi[j] = (q + 1 == (uint32_t)wid) ? phi[j] : fma(w, (double)(q + 1), plo[j]);
  }
}

void oracle_decode(int d, const double *plo, const double *phi, int win, int wid,
                   int c, int64_t n, const uint32_t *pos, double *clo, double *chi)
{
  int old = fegetround();
  fesetround(FE_TONEAREST);
  for (int64_t i = 0; i < n; ++i)
    decode_child(d, plo, phi, win, wid, c, pos[i], clo + i * d, chi + i * d);
  fesetround(old);
}

/* ---------------------------------------------------------------------------
 * Box evaluation entry (tests): value, gradient, KKT flag, sample ub.
 * ------------------------------------------------------------------------- */
int oracle_eval_boxes(int fam, int d, const double *P, int64_t n,
                      const double *lo, const double *hi,
                      const double *dlo, const double *dhi, int S,
                      double *flo, double *fhi, double *glo, double *ghi,
                      int32_t *kkt, double *ub)
{
  if (fam < 0 || fam > 3 || d < 1) return -1;
  int old = fegetround();
  iv *x = malloc(sizeof(iv) * d), *g = malloc(sizeof(iv) * d), *xs = malloc(sizeof(iv) * d);
  for (int64_t i = 0; i < n; ++i) {
    for (int j = 0; j < d; ++j) x[j] = mk(lo[i * d + j], hi[i * d + j]);
    fesetround(FE_UPWARD);
    if (flo || fhi) {
      iv f = family_value(fam, d, P, x);
      if (flo) flo[i] = f.lo;
      if (fhi) fhi[i] = f.hi;
    }
    if (glo || ghi) {
      family_grad(fam, d, P, x, g);
      for (int j = 0; j < d; ++j) {
        if (glo) glo[i * d + j] = g[j].lo;
        if (ghi) ghi[i * d + j] = g[j].hi;
      }
    }
    if (kkt) kkt[i] = family_kkt(fam, d, P, x, dlo, dhi, g);
    if (ub) ub[i] = family_sample_ub(fam, d, P, x, S, xs);
  }
  fesetround(old);
  free(x); free(g); free(xs);
  return 0;
}

/* Mean-value lower bound of n explicit boxes (the bound that
 * ORACLE_MEAN_VALUE solves use beside the natural extension): the exported
 * form of family_mv_lower, so tests can pin it on its own. */
int oracle_mv_lower(int fam, int d, const double *P, int64_t n,
                    const double *lo, const double *hi, double *out)
{
  if (fam < 0 || fam > 3 || d < 1) return -1;
  int old = fegetround();
  iv *x = malloc(sizeof(iv) * d), *g = malloc(sizeof(iv) * d), *xs = malloc(sizeof(iv) * d);
  for (int64_t i = 0; i < n; ++i) {
    for (int j = 0; j < d; ++j) x[j] = mk(lo[i * d + j], hi[i * d + j]);
    fesetround(FE_UPWARD);
    family_grad(fam, d, P, x, g);
    out[i] = family_mv_lower(fam, d, P, x, g, xs);
  }
  fesetround(old);
  free(x); free(g); free(xs);
  return 0;
}

/* ---------------------------------------------------------------------------
 * Branch and bound.
 * ------------------------------------------------------------------------- */
typedef struct { double lb; uint32_t parent, pos; } qent;

static int qless(const qent *a, const qent *b)
{
  if (a->lb != b->lb) return a->lb < b->lb;
  if (a->parent != b->parent) return a->parent < b->parent;
  return a->pos < b->pos;
}

typedef struct { qent *v; int64_t n, cap; } heap_t;

static int heap_push(heap_t *h, qent e)
{
  if (h->n == h->cap) {
    int64_t nc = h->cap ? h->cap * 2 : 1024;
    qent *nv = realloc(h->v, sizeof(qent) * nc);
    if (!nv) return -1;
    h->v = nv; h->cap = nc;
  }
  int64_t i = h->n++;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (!qless(&e, &h->v[p])) break;
    h->v[i] = h->v[p]; i = p;
  }
  h->v[i] = e;
  return 0;
}

static qent heap_pop(heap_t *h)
{
  qent top = h->v[0];
  qent last = h->v[--h->n];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1, r = l + 1, m = i;
    const qent *best = &last;
    if (l < h->n && qless(&h->v[l], best)) { m = l; best = &h->v[l]; }
    if (r < h->n && qless(&h->v[r], best)) { m = r; }
    if (m == i) break;
    h->v[i] = h->v[m]; i = m;
  }
  if (h->n > 0) h->v[i] = last;
  return top;
}

static double now_ms(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static double ipow_count(int wid, int c)
{
  double w = 1; for (int k = 0; k < c; ++k) w *= wid; return w;
}

int oracle_solve(const oracle_problem *pr, gpugo_result *res,
                 double *out_lo, double *out_hi, double *out_lb, int64_t out_max)
{
  const int d = pr->d, fam = pr->family;
  const int wid = pr->wid, c = pr->cycle < d ? pr->cycle : d;
  const int K = pr->k_select > 0 ? pr->k_select : 1;
  const int S = pr->n_samples > 0 ? pr->n_samples : 1;
  const double Wd = ipow_count(wid, c);
  if (fam < 0 || fam > 3 || d < 1 || wid < 2 || c < 1 || Wd > 4294967295.0) return -1;
  const uint32_t W = (uint32_t)Wd;
  int nthreads = pr->threads;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  double t0 = now_ms();
  int old = fegetround();
  fesetround(FE_TONEAREST);

  int64_t pcap = 1024, np = 0;
  double *plo = malloc(sizeof(double) * d * pcap), *phi = malloc(sizeof(double) * d * pcap);
  int *pwin = malloc(sizeof(int) * pcap);
  heap_t heap = {0};
  int64_t nout = 0;
  double out_min_lb = INFINITY;
  double ub = pr->init_ub;
  memset(res, 0, sizeof(*res));
  res->status = GPUGO_ST_MAX_ITERS;

  qent *sel = malloc(sizeof(qent) * K);
  double *clo = malloc(sizeof(double) * d), *chi = malloc(sizeof(double) * d);
  int64_t iter = 0;

  /* per-thread survivor buffers */
  qent **tbuf = calloc(nthreads, sizeof(qent *));
  int64_t *tcnt = calloc(nthreads, sizeof(int64_t)), *tcap = calloc(nthreads, sizeof(int64_t));
  int64_t pruned_v = 0, pruned_k = 0;
  int rc = 0;

  while (iter < pr->max_iters) {
    int nsel = 0;
    int64_t first_new = np;
    if (iter == 0) {
      memcpy(plo, pr->lo, sizeof(double) * d);
      memcpy(phi, pr->hi, sizeof(double) * d);
      pwin[0] = 0;
      np = 1;
      nsel = 1;
    } else {
      /* pop the K best boxes that survive the current upper bound */
      while (nsel < K && heap.n > 0) {
        qent e = heap_pop(&heap);
        if (e.lb > ub) continue;                       /* list prune */
        if (nsel == 0 && pr->tol_f > 0) {
          double lbg = e.lb < out_min_lb ? e.lb : out_min_lb;
          fesetround(FE_UPWARD);
          double wdt = ub - lbg;
          fesetround(FE_TONEAREST);
          if (wdt <= pr->tol_f) {
            res->status = GPUGO_ST_CERTIFIED;
            heap_push(&heap, e);
            goto done;
          }
        }
        /* materialise; reference x-width output rule */
        decode_child(d, plo + (int64_t)e.parent * d, phi + (int64_t)e.parent * d,
                     pwin[e.parent], wid, c, e.pos, clo, chi);
        if (pr->tol_x > 0) {
          double mw = 0;
          for (int j = 0; j < d; ++j) { double w = chi[j] - clo[j]; if (w > mw) mw = w; }
          if (mw < pr->tol_x) {
            if (nout < out_max && out_lo) {
              memcpy(out_lo + nout * d, clo, sizeof(double) * d);
              memcpy(out_hi + nout * d, chi, sizeof(double) * d);
              out_lb[nout] = e.lb;
            }
            ++nout;
            if (e.lb < out_min_lb) out_min_lb = e.lb;
            continue;
          }
        }
        if (np == pcap) {
          pcap *= 2;
          plo = realloc(plo, sizeof(double) * d * pcap);
          phi = realloc(phi, sizeof(double) * d * pcap);
          pwin = realloc(pwin, sizeof(int) * pcap);
        }
        memcpy(plo + np * d, clo, sizeof(double) * d);
        memcpy(phi + np * d, chi, sizeof(double) * d);
        pwin[np] = (pwin[e.parent] + c) % d;
        sel[nsel++] = e;
        ++np;
      }
      if (nsel == 0) {
        res->status = (pr->tol_x > 0 && nout > 0) ? GPUGO_ST_X_CONVERGED : GPUGO_ST_EXHAUSTED;
        goto done;
      }
    }
    const int64_t nchild = (int64_t)nsel * W;

    /* sampling pass: new incumbent upper bound */
    double ubs = ub;
#pragma omp parallel num_threads(nthreads) reduction(min:ubs)
    {
      iv *x = malloc(sizeof(iv) * d), *xs = malloc(sizeof(iv) * d);
      double *bl = malloc(sizeof(double) * d), *bh = malloc(sizeof(double) * d);
      int oldt = fegetround();
#pragma omp for schedule(static)
      for (int64_t t = 0; t < nchild; ++t) {
        int64_t pid = first_new + t / W;
        uint32_t pos = (uint32_t)(t % W);
        fesetround(FE_TONEAREST);
        decode_child(d, plo + pid * d, phi + pid * d, pwin[pid], wid, c, pos, bl, bh);
        for (int j = 0; j < d; ++j) x[j] = mk(bl[j], bh[j]);
        fesetround(FE_UPWARD);
        double u = family_sample_ub(fam, d, pr->params, x, S, xs);
        if (u < ubs) ubs = u;
      }
      fesetround(oldt);
      free(x); free(xs); free(bl); free(bh);
    }
    if (ubs < ub) ub = ubs;

    /* rule-out pass */
    for (int th = 0; th < nthreads; ++th) tcnt[th] = 0;
    int64_t pv = 0, pk = 0;
    int alloc_fail = 0;
#pragma omp parallel num_threads(nthreads) reduction(+:pv, pk)
    {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      iv *x = malloc(sizeof(iv) * d), *g = malloc(sizeof(iv) * d), *xs = malloc(sizeof(iv) * d);
      double *bl = malloc(sizeof(double) * d), *bh = malloc(sizeof(double) * d);
      int oldt = fegetround();
#pragma omp for schedule(static)
      for (int64_t t = 0; t < nchild; ++t) {
        int64_t pid = first_new + t / W;
        uint32_t pos = (uint32_t)(t % W);
        fesetround(FE_TONEAREST);
        decode_child(d, plo + pid * d, phi + pid * d, pwin[pid], wid, c, pos, bl, bh);
        for (int j = 0; j < d; ++j) x[j] = mk(bl[j], bh[j]);
        fesetround(FE_UPWARD);
        iv f = family_value(fam, d, pr->params, x);
        double lb = f.lo;
        if (lb != lb) lb = -INFINITY;
        if (lb > ub) { ++pv; continue; }
        if (pr->flags & ORACLE_MEAN_VALUE) {
          /* second-order lower bound from the gradient already needed for
           * the monotonicity test */
          family_grad(fam, d, pr->params, x, g);
          double lmv = family_mv_lower(fam, d, pr->params, x, g, xs);
          if (lmv > lb) lb = lmv;
          if (lb > ub) { ++pv; continue; }
          if (kkt_test(d, x, pr->lo, pr->hi, g)) { ++pk; continue; }
        } else if (family_kkt(fam, d, pr->params, x, pr->lo, pr->hi, g)) { ++pk; continue; }
        if (tcnt[tid] == tcap[tid]) {
          int64_t ncap = tcap[tid] ? tcap[tid] * 2 : 4096;
          qent *nb = realloc(tbuf[tid], sizeof(qent) * ncap);
          if (!nb) { alloc_fail = 1; continue; }
          tbuf[tid] = nb; tcap[tid] = ncap;
        }
        qent e; e.lb = lb; e.parent = (uint32_t)pid; e.pos = pos;
        tbuf[tid][tcnt[tid]++] = e;
      }
      fesetround(oldt);
      free(x); free(g); free(xs); free(bl); free(bh);
    }
    if (alloc_fail) { rc = -2; goto done; }
    pruned_v += pv; pruned_k += pk;
    for (int th = 0; th < nthreads; ++th)
      for (int64_t i = 0; i < tcnt[th]; ++i)
        if (heap_push(&heap, tbuf[th][i])) { rc = -2; goto done; }
    res->boxes_bounded += nchild;
    if (heap.n > res->queue_peak) res->queue_peak = heap.n;
    ++iter;
  }

done:
  /* final enclosure: smallest lower bound still alive (queue or outputs) */
  {
    double lbg = out_min_lb;
    /* drop entries above ub, find the minimum */
    for (int64_t i = 0; i < heap.n; ++i)
      if (heap.v[i].lb <= ub && heap.v[i].lb < lbg) lbg = heap.v[i].lb;
    int64_t live = 0;
    for (int64_t i = 0; i < heap.n; ++i) if (heap.v[i].lb <= ub) ++live;
    res->lb = lbg;
    res->ub = ub;
    fesetround(FE_UPWARD);
    res->width = ub - lbg;
    fesetround(FE_TONEAREST);
    res->iterations = iter;
    res->pruned_value = pruned_v;
    res->pruned_kkt = pruned_k;
    res->queue_size = live;
    res->n_out = nout;
    res->ms = now_ms() - t0;
  }
  for (int th = 0; th < nthreads; ++th) free(tbuf[th]);
  free(tbuf); free(tcnt); free(tcap);
  free(plo); free(phi); free(pwin); free(heap.v); free(sel); free(clo); free(chi);
  fesetround(old);
  return rc;
}

/* One split-and-bound step from explicit parent boxes (tests): sample pass,
 * then rule-out pass with the updated upper bound.  Survivors are returned
 * ordered by (parent index, pos).  Returns the number of survivors, or -1 if
 * out_max is too small. */
int64_t oracle_split_bound(int fam, int d, const double *P, int64_t nparents,
                           const double *plo, const double *phi, const int32_t *pwin,
                           int wid, int cycle, int S, const double *dlo, const double *dhi,
                           double *ub_inout, int64_t out_max, uint32_t *out_parent,
                           uint32_t *out_pos, double *out_lb)
{
  const int c = cycle < d ? cycle : d;
  const uint32_t W = (uint32_t)ipow_count(wid, c);
  int old = fegetround();
  iv *x = malloc(sizeof(iv) * d), *g = malloc(sizeof(iv) * d), *xs = malloc(sizeof(iv) * d);
  double *bl = malloc(sizeof(double) * d), *bh = malloc(sizeof(double) * d);
  double ub = *ub_inout;
  for (int64_t p = 0; p < nparents; ++p)
    for (uint32_t pos = 0; pos < W; ++pos) {
      fesetround(FE_TONEAREST);
      decode_child(d, plo + p * d, phi + p * d, pwin[p], wid, c, pos, bl, bh);
      for (int j = 0; j < d; ++j) x[j] = mk(bl[j], bh[j]);
      fesetround(FE_UPWARD);
      double u = family_sample_ub(fam, d, P, x, S, xs);
      if (u < ub) ub = u;
    }
  int64_t ns = 0;
  for (int64_t p = 0; p < nparents; ++p)
    for (uint32_t pos = 0; pos < W; ++pos) {
      fesetround(FE_TONEAREST);
      decode_child(d, plo + p * d, phi + p * d, pwin[p], wid, c, pos, bl, bh);
      for (int j = 0; j < d; ++j) x[j] = mk(bl[j], bh[j]);
      fesetround(FE_UPWARD);
      iv f = family_value(fam, d, P, x);
      double lb = f.lo;
      if (lb != lb) lb = -INFINITY;
      if (lb > ub) continue;
      if (family_kkt(fam, d, P, x, dlo, dhi, g)) continue;
      if (ns < out_max) { out_parent[ns] = (uint32_t)p; out_pos[ns] = pos; out_lb[ns] = lb; }
      ++ns;
    }
  fesetround(old);
  *ub_inout = ub;
  free(x); free(g); free(xs); free(bl); free(bh);
  return ns <= out_max ? ns : -1;
}

int oracle_rounding_selftest(void)
{
  /* 1 + 2^-60 rounds up to 1 + 2^-52 and down to 1 if directed rounding and
   * the -RU(-x) construction both work. */
  int old = fegetround();
  fesetround(FE_UPWARD);
  volatile double one = 1.0, tiny = 0x1p-60;
  double up = ru_add(one, tiny), dn = rd_add(one, tiny);
  double mu = ru_mul(1.0 + 0x1p-52, 1.0 + 0x1p-52), md = rd_mul(1.0 + 0x1p-52, 1.0 + 0x1p-52);
  fesetround(old);
  return (up == 1.0 + 0x1p-52 && dn == 1.0 && mu == 1.0 + 0x1p-51 + 0x1p-52 && md == 1.0 + 0x1p-51) ? 0 : 1;
}
