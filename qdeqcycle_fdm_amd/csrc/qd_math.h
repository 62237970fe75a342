// qd_math.h -- FP64 elementary functions of the fault solver (exp, sinh,
// cosh, log, log1p, asinh), written with +,-,*,/, sqrt, floor, frexp and
// ldexp only, so that they give the same bits on the GPU and on a host CPU
// (with -ffp-contract=off on both).  Methods:
//   exp   Cody-Waite reduction x = k ln2 + r (ln2 split in a 32-bit head and a
//         tail), e^r - 1 by its Taylor series to r^13, scale by 2^k;
//   sinh  Taylor series to x^17 for |x| < 1/2, (e^x - e^-x)/2 above;
//   cosh  (e^x + e^-x)/2;
//   log   x = 2^e m, m in [1/sqrt2, sqrt2), log m = 2 atanh(s), s = (m-1)/(m+1),
//         atanh by its odd series to s^25;
//   log1p Goldberg's form x log(1+x)/((1+x)-1);
//   asinh log(x + sqrt(x^2+1)) in the cancellation-free forms for small and
//         large |x|.
// Accuracy against glibc is a few ulp (checked by the tests).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace qdeq {

#define QD_HD __host__ __device__ inline

constexpr double QD_LN2_HI = 6.93147180369123816490e-01;
constexpr double QD_LN2_LO = 1.90821492927058770002e-10;
constexpr double QD_INV_LN2 = 1.44269504088896338700e+00;
constexpr double QD_LN2 = 6.93147180559945309417e-01;

QD_HD double qd_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893383973096) return HUGE_VAL;
  if (x < -745.133219101941108420) return 0.0;
  const double kd = floor(x * QD_INV_LN2 + 0.5);
  const int k = (int)kd;
  const double r = (x - kd * QD_LN2_HI) - kd * QD_LN2_LO;
  double p = 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  const double y = 1.0 + p * r;
  if (k > 1020) return ldexp(ldexp(y, k - 1020), 1020);
  if (k < -1020) return ldexp(ldexp(y, k + 1020), -1020);
  return ldexp(y, k);
}

QD_HD double qd_sinh(double x) {
  if (x != x) return x;
  const double ax = fabs(x);
  double r;
  if (ax < 0.5) {
    const double x2 = ax * ax;
    double p = 1.0 / 355687428096000.0;
    p = p * x2 + 1.0 / 1307674368000.0;
    p = p * x2 + 1.0 / 6227020800.0;
    p = p * x2 + 1.0 / 39916800.0;
    p = p * x2 + 1.0 / 362880.0;
    p = p * x2 + 1.0 / 5040.0;
    p = p * x2 + 1.0 / 120.0;
    p = p * x2 + 1.0 / 6.0;
    r = ax + ax * (x2 * p);
  } else if (ax < 22.0) {
    const double e = qd_exp(ax);
    r = 0.5 * (e - 1.0 / e);
  } else if (ax < 709.0) {
    r = 0.5 * qd_exp(ax);
  } else {
    const double e = qd_exp(0.5 * ax);
    r = (0.5 * e) * e;
  }
  return x < 0.0 ? -r : r;
}

QD_HD double qd_cosh(double x) {
  if (x != x) return x;
  const double ax = fabs(x);
  if (ax < 22.0) {
    const double e = qd_exp(ax);
    return 0.5 * (e + 1.0 / e);
  }
  if (ax < 709.0) return 0.5 * qd_exp(ax);
  const double e = qd_exp(0.5 * ax);
  return (0.5 * e) * e;
}

// qd_sinh(x) and qd_cosh(x) together: each branch does the operations of the
// two functions above for its range, sharing the e^|x| (and 1/e^|x|) they
// both form, so both results have the same bits as the separate calls.
QD_HD void qd_sinh_cosh(double x, double *s, double *c) {
  if (x != x) {
    *s = x;
    *c = x;
    return;
  }
  const double ax = fabs(x);
  double r, ch;
  if (ax < 22.0) {
    const double e = qd_exp(ax);
    const double ie = 1.0 / e;
    ch = 0.5 * (e + ie);
    if (ax < 0.5) {
      const double x2 = ax * ax;
      double p = 1.0 / 355687428096000.0;
      p = p * x2 + 1.0 / 1307674368000.0;
      p = p * x2 + 1.0 / 6227020800.0;
      p = p * x2 + 1.0 / 39916800.0;
      p = p * x2 + 1.0 / 362880.0;
      p = p * x2 + 1.0 / 5040.0;
      p = p * x2 + 1.0 / 120.0;
      p = p * x2 + 1.0 / 6.0;
      r = ax + ax * (x2 * p);
    } else {
      r = 0.5 * (e - ie);
    }
  } else if (ax < 709.0) {
    r = 0.5 * qd_exp(ax);
    ch = r;
  } else {
    const double e = qd_exp(0.5 * ax);
    r = (0.5 * e) * e;
    ch = r;
  }
  *s = x < 0.0 ? -r : r;
  *c = ch;
}

QD_HD double qd_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return (x - x) / (x - x);
  if (x == 0.0) return -HUGE_VAL;
  if (x > 1.79769313486231570815e+308) return x;
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double s2 = s * s;
  double p = 1.0 / 25.0;
  p = p * s2 + 1.0 / 23.0;
  p = p * s2 + 1.0 / 21.0;
  p = p * s2 + 1.0 / 19.0;
  p = p * s2 + 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  const double lm = 2.0 * s + 2.0 * s * (s2 * p);
  const double ed = (double)e;
  return ed * QD_LN2_HI + (ed * QD_LN2_LO + lm);
}

QD_HD double qd_log1p(double x) {
  const double u = 1.0 + x;
  if (u == 1.0) return x;
  return qd_log(u) * (x / (u - 1.0));
}

QD_HD double qd_asinh(double x) {
  if (x != x) return x;
  const double ax = fabs(x);
  double r;
  if (ax < 3.7252902984619140625e-09) return x;  // 2^-28
  if (ax > 268435456.0) {                        // 2^28
    r = qd_log(ax) + QD_LN2;
  } else if (ax > 2.0) {
    r = qd_log(2.0 * ax + 1.0 / (sqrt(ax * ax + 1.0) + ax));
  } else {
    const double t = ax * ax;
    r = qd_log1p(ax + t / (1.0 + sqrt(1.0 + t)));
  }
  return x < 0.0 ? -r : r;
}

#undef QD_HD

}  // namespace qdeq
