/* crlog.h -- natural logarithm rounded to nearest, for C and HIP.
 *
 * ln x for finite x > 0: x = 2^k m with m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(f) = 2 (f + f^3/3 + f^5/5 + ...), f = (m - 1)/(m + 1),
 * |f| <= 0.1716, evaluated in double-double arithmetic (Dekker's exact
 * products via fma, Knuth's two-sum) with 24 series terms (the omitted tail
 * is below 2^-110 of the sum), plus k ln 2 in double-double; the sum is
 * rounded once to double.  The result is the correctly rounded logarithm
 * unless ln x lies within ~2^-100 of a rounding midpoint (checked against
 * 60-digit decimal logarithms in tests/test_crlog.py).
 *
 * Used where a logarithm must be reproducible across CPU and GPU builds
 * (the reference-order FP64 mode's neo-Hookean ln J): libm logarithms are
 * accurate to about half an ulp but round differently from one library to
 * another.
 */
#ifndef MPM_CRLOG_H
#define MPM_CRLOG_H

#include <math.h>

#if defined(__HIPCC__)
#define MPM_HD __host__ __device__
#else
#define MPM_HD
#endif

typedef struct {
  double hi, lo;
} mpm_dd;

static MPM_HD inline mpm_dd mpm_dd_make(double hi, double lo) {
  mpm_dd r;
  r.hi = hi;
  r.lo = lo;
  return r;
}
static MPM_HD inline mpm_dd mpm_two_sum(double a, double b) {
  const double s = a + b, bb = s - a;
  return mpm_dd_make(s, (a - (s - bb)) + (b - bb));
}
static MPM_HD inline mpm_dd mpm_fast_two_sum(double a, double b) {
  const double s = a + b;
  return mpm_dd_make(s, b - (s - a));
}
static MPM_HD inline mpm_dd mpm_dd_add(mpm_dd a, mpm_dd b) {
  mpm_dd s = mpm_two_sum(a.hi, b.hi);
  const mpm_dd t = mpm_two_sum(a.lo, b.lo);
  s.lo = s.lo + t.hi;
  s = mpm_fast_two_sum(s.hi, s.lo);
  s.lo = s.lo + t.lo;
  return mpm_fast_two_sum(s.hi, s.lo);
}
static MPM_HD inline mpm_dd mpm_dd_mul(mpm_dd a, mpm_dd b) {
  const double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  e = e + (a.hi * b.lo + a.lo * b.hi);
  return mpm_fast_two_sum(p, e);
}
static MPM_HD inline mpm_dd mpm_dd_mul_d(mpm_dd a, double b) {
  const double p = a.hi * b;
  double e = fma(a.hi, b, -p);
  e = e + a.lo * b;
  return mpm_fast_two_sum(p, e);
}
/* a / b by long division: three double quotients, each remainder exact to
 * double-double precision */
static MPM_HD inline mpm_dd mpm_dd_div(mpm_dd a, mpm_dd b) {
  const double q1 = a.hi / b.hi;
  mpm_dd r = mpm_dd_add(a, mpm_dd_mul_d(b, -q1));
  const double q2 = r.hi / b.hi;
  r = mpm_dd_add(r, mpm_dd_mul_d(b, -q2));
  const double q3 = r.hi / b.hi;
  return mpm_dd_add(mpm_fast_two_sum(q1, q2), mpm_dd_make(q3, 0.0));
}

static MPM_HD inline double mpm_log_cr(double x) {
  if (!(x > 0.0) || x == HUGE_VAL) return log(x);   /* 0, negative, NaN, inf: libm's special values */
  int k;
  double m = frexp(x, &k);   /* m in [1/2, 1) */
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    k = k - 1;
  }
  const mpm_dd f = mpm_dd_div(mpm_two_sum(m, -1.0), mpm_two_sum(m, 1.0));
  const mpm_dd f2 = mpm_dd_mul(f, f);
  /* Horner over the coefficients 1/(2n+1), n = 23..0, as double-double */
  mpm_dd s = mpm_dd_make(0x1.5c9882b931057p-6, 0x1.310572620ae4cp-61);   /* 1/47 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.6c16c16c16c17p-6, -0x1.f49f49f49f49fp-61));   /* 1/45 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.7d05f417d05f4p-6, 0x1.7d05f417d05f4p-62));   /* 1/43 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.8f9c18f9c18fap-6, -0x1.f3831f3831f38p-61));   /* 1/41 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.a41a41a41a41ap-6, 0x1.0690690690690p-60));   /* 1/39 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.bacf914c1bad0p-6, -0x1.bacf914c1bad0p-60));   /* 1/37 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.d41d41d41d41dp-6, 0x1.0750750750750p-60));   /* 1/35 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.f07c1f07c1f08p-6, -0x1.f07c1f07c1f08p-61));   /* 1/33 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.0842108421084p-5, 0x1.0842108421084p-60));   /* 1/31 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.1a7b9611a7b96p-5, 0x1.1a7b9611a7b96p-61));   /* 1/29 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.2f684bda12f68p-5, 0x1.2f684bda12f68p-59));   /* 1/27 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.47ae147ae147bp-5, -0x1.eb851eb851eb8p-61));   /* 1/25 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.642c8590b2164p-5, 0x1.642c8590b2164p-60));   /* 1/23 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.8618618618618p-5, 0x1.8618618618618p-59));   /* 1/21 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.af286bca1af28p-5, 0x1.af286bca1af28p-59));   /* 1/19 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.e1e1e1e1e1e1ep-5, 0x1.e1e1e1e1e1e1ep-61));   /* 1/17 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.1111111111111p-4, 0x1.1111111111111p-60));   /* 1/15 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.3b13b13b13b14p-4, -0x1.3b13b13b13b14p-58));   /* 1/13 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.745d1745d1746p-4, -0x1.745d1745d1746p-59));   /* 1/11 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.c71c71c71c71cp-4, 0x1.c71c71c71c71cp-58));   /* 1/9 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.2492492492492p-3, 0x1.2492492492492p-57));   /* 1/7 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.999999999999ap-3, -0x1.999999999999ap-57));   /* 1/5 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.5555555555555p-2, 0x1.5555555555555p-56));   /* 1/3 */
  s = mpm_dd_add(mpm_dd_mul(s, f2), mpm_dd_make(0x1.0000000000000p+0, 0x0.0p+0));   /* 1/1 */
  const mpm_dd l = mpm_dd_mul_d(mpm_dd_mul(s, f), 2.0);
  const mpm_dd ln2 = mpm_dd_make(0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56);
  const mpm_dd r = mpm_dd_add(mpm_dd_mul_d(ln2, (double)k), l);
  return r.hi + r.lo;
}

#endif
