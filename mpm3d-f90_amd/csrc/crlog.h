/* crlog.h -- logarithm, exponential and power rounded to nearest, for C and HIP.
 *
 * ln x for finite x > 0: x = 2^k m with m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(f) = 2 (f + f^3/3 + f^5/5 + ...), f = (m - 1)/(m + 1),
 * |f| <= 0.1716, evaluated in double-double arithmetic (Dekker's exact
 * products via fma, Knuth's two-sum) with 24 series terms (the omitted tail
 * is below 2^-110 of the sum), plus k ln 2 in double-double.
 * e^t for a double-double t: t = k ln 2 + r with k = round(t / ln 2),
 * |r| <= ln(2)/2, r formed with a three-part ln 2; e^r by its Taylor series
 * to r^24/24! (tail below 2^-114) in double-double; times 2^k.
 * x^y = e^(y ln x) for x > 0 from the double-double logarithm, with C99's
 * special values elsewhere.
 * Each result is the double-double value rounded once to double: the
 * correctly rounded value unless the exact one lies within ~2^-95 (relative)
 * of a rounding midpoint (checked against 60-digit decimal values in
 * tests/test_crlog.py).  Results in the subnormal range are rounded twice.
 *
 * Used where a transcendental function must be reproducible across CPU and
 * GPU builds (the reference-order FP64 mode's neo-Hookean ln J, the
 * Johnson-Cook hardening powers and rate logarithm, the JWL exponentials):
 * libm versions are accurate to about half an ulp but round differently
 * from one library to another.
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

/* ln x as a double-double, finite x > 0 */
static MPM_HD inline mpm_dd mpm_log_dd(double x) {
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
  return mpm_dd_add(mpm_dd_mul_d(ln2, (double)k), l);
}

static MPM_HD inline double mpm_log_cr(double x) {
  if (!(x > 0.0) || x == HUGE_VAL) return log(x);   /* 0, negative, NaN, inf: libm's special values */
  const mpm_dd r = mpm_log_dd(x);
  return r.hi + r.lo;
}


/* e^t rounded to double, t a double-double */
static MPM_HD inline double mpm_exp_dd(mpm_dd t) {
  if (t.hi != t.hi) return t.hi;
  if (t.hi > 709.8) return HUGE_VAL;
  if (t.hi < -745.2) return 0.0;
  const double k = rint(t.hi * 0x1.71547652b82fep+0);   /* t / ln 2 */
  /* r = t - k ln 2, ln 2 = L1 + L2 + L3 (k L1 exact by fma) */
  const double L1 = 0x1.62e42fefa39efp-1, L2 = 0x1.abc9e3b39803fp-56, L3 = 0x1.7b57a079a1934p-111;
  const double p1 = k * L1;
  mpm_dd r = mpm_dd_add(t, mpm_dd_make(-p1, -fma(k, L1, -p1)));
  r = mpm_dd_add(r, mpm_dd_mul_d(mpm_dd_make(L2, L3), -k));
  mpm_dd s = mpm_dd_make(0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135);   /* 1/24! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.761b41316381ap-75, -0x1.3423c7d91404fp-130));   /* 1/23! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.0ce396db7f853p-70, -0x1.aebcdbd20331cp-124));   /* 1/22! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120));   /* 1/21! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120));   /* 1/20! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.2f49b46814157p-57, 0x1.2650f61dbdcb4p-112));   /* 1/19! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.6827863b97d97p-53, 0x1.eec01221a8b0bp-107));   /* 1/18! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103));   /* 1/17! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101));   /* 1/16! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.ae7f3e733b81fp-41, 0x1.1d8656b0ee8cbp-97));   /* 1/15! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.93974a8c07c9dp-37, 0x1.05d6f8a2efd1fp-92));   /* 1/14! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87));   /* 1/13! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83));   /* 1/12! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80));   /* 1/11! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76));   /* 1/10! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73));   /* 1/9! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76));   /* 1/8! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73));   /* 1/7! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65));   /* 1/6! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.1111111111111p-7, 0x1.1111111111111p-63));   /* 1/5! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.5555555555555p-5, 0x1.5555555555555p-59));   /* 1/4! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.5555555555555p-3, 0x1.5555555555555p-57));   /* 1/3! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(0x1.0000000000000p-1, 0x0.0p+0));   /* 1/2! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(1.0, 0.0));   /* 1/1! */
  s = mpm_dd_add(mpm_dd_mul(s, r), mpm_dd_make(1.0, 0.0));   /* 1/0! */
  return ldexp(s.hi + s.lo, (int)k);
}

static MPM_HD inline double mpm_exp_cr(double x) { return mpm_exp_dd(mpm_dd_make(x, 0.0)); }

/* x^y (C99 pow's special values; x^y = e^(y ln x) for finite x > 0) */
static MPM_HD inline double mpm_pow_cr(double x, double y) {
  if (y == 0.0 || x == 1.0) return 1.0;
  if (x != x || y != y) return x + y;
  if (!(x > 0.0) || x == HUGE_VAL || y == HUGE_VAL || y == -HUGE_VAL) return pow(x, y);
  const mpm_dd l = mpm_log_dd(x);
  return mpm_exp_dd(mpm_dd_mul_d(l, y));
}

#endif
