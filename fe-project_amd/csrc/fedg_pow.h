/*
 * Portable double-double pow(x, y) for x > 0, y finite.
 *
 * The same source is compiled for the GPU (HIP) and for the host (C99), and it
 * uses only IEEE-754 operations that are correctly rounded on both (+, -, *, /,
 * fma, rint) plus exact bit manipulation, so both produce the same bits for the
 * same inputs.  The result is the correctly rounded x^y except when x^y lies
 * within about 2^-70 relative of a rounding midpoint.
 *
 * Method (textbook table-driven log and exp, e.g. P.T.P. Tang, ACM TOMS 15
 * (1989) and 16 (1990); J.-M. Muller, Elementary Functions, ch. 6), carried
 * in double-double:
 *   x = 2^e m, m in [1,2); i = top 7 bits of m; r = m*invc_i - 1 exactly in
 *   double-double (|r| <= 2^-8);
 *   ln x = e ln2 - ln(invc_i) + ln(1 + r), ln(1+r) by its Taylor series;
 *   z = y ln x (double-double); z = (128 q + j) ln2/128 + t, |t| <= ln2/256;
 *   x^y = 2^q 2^(j/128) exp(t), exp(t) by its Taylor series.
 * Tables: fedg_pow_tables.h (gen_pow_tables.py).
 *
 * Inputs outside x > 0 (finite, normal or subnormal) with |y ln x| < 708 go
 * to the platform pow.
 */
#ifndef FEDG_POW_H
#define FEDG_POW_H

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define FEDG_POW_FN __device__ static inline
#define FEDG_POW_TABLE __device__ __constant__ static const
#else
#include <math.h>
#include <stdint.h>
#include <string.h>
#define FEDG_POW_FN static inline
#define FEDG_POW_TABLE static const
#endif

#include "fedg_pow_tables.h"

FEDG_POW_FN uint64_t fedg_pow_bits(double x) {
  uint64_t u;
  memcpy(&u, &x, sizeof(u));
  return u;
}
FEDG_POW_FN double fedg_pow_from_bits(uint64_t u) {
  double x;
  memcpy(&x, &u, sizeof(x));
  return x;
}

/* s + err = a + b exactly */
FEDG_POW_FN double fedg_two_sum(double a, double b, double *err) {
  double s = a + b;
  double bb = s - a;
  *err = (a - (s - bb)) + (b - bb);
  return s;
}
/* s + err = a + b exactly, |a| >= |b| */
FEDG_POW_FN double fedg_fast_two_sum(double a, double b, double *err) {
  double s = a + b;
  *err = b - (s - a);
  return s;
}

/* The core with the five 128-entry tables passed in, so that a caller can
 * read them from a faster memory (the GPU kernels copy them to LDS); the
 * arithmetic does not depend on where the tables live. */
FEDG_POW_FN double fedg_pow_core_t(double x, double y, const double *tinvc, const double *tlogc_hi,
                                   const double *tlogc_lo, const double *texp_hi, const double *texp_lo) {
  /* ---- ln x in double-double ---- */
  uint64_t ux = fedg_pow_bits(x);
  int e = (int)((ux >> 52) & 0x7ff);
  if (e == 0) { /* subnormal: scale by 2^52 */
    ux = fedg_pow_bits(x * 4503599627370496.0);
    e = (int)((ux >> 52) & 0x7ff) - 52;
  }
  e -= 1023;
  const int i = (int)((ux >> 45) & 127);
  const double m = fedg_pow_from_bits((ux & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  const double invc = tinvc[i];
  const double p_hi = m * invc;
  const double p_lo = fma(m, invc, -p_hi);
  const double r_hi = p_hi - 1.0; /* exact: p_hi in [1-2^-8, 1+2^-8] */
  double r_lo;
  const double r = fedg_two_sum(r_hi, p_lo, &r_lo);
  /* ln(1+r) = r - r^2/2 + r^3 (1/3 - r/4 + ...) with r = (r, r_lo) */
  const double h_hi = 0.5 * (r * r);
  const double h_lo = 0.5 * fma(r, r, -(r * r));
  const double poly = 0.33333333333333333 +
                      r * (-0.25 + r * (0.2 + r * (-0.16666666666666666 + r * (0.14285714285714285 +
                      r * (-0.125 + r * (0.1111111111111111 + r * (-0.1)))))));
  const double tail = (r * r) * r * poly;
  const double corr = r_lo - r_lo * r + r_lo * (r * r);
  double a_lo;
  const double a_hi = fedg_fast_two_sum(r, -h_hi, &a_lo);
  const double l1_lo = a_lo - h_lo + tail + corr;
  /* e ln2 + (-ln invc) + ln(1+r) */
  const double ed = (double)e;
  double s1_lo, s2_lo;
  const double s1 = fedg_two_sum(ed * FEDG_POW_LN2_HI, tlogc_hi[i], &s1_lo);
  const double s2 = fedg_two_sum(s1, a_hi, &s2_lo);
  const double lo = ed * FEDG_POW_LN2_LO + tlogc_lo[i] + s1_lo + s2_lo + l1_lo;
  double L_lo;
  const double L_hi = fedg_two_sum(s2, lo, &L_lo);

  /* ---- z = y ln x ---- */
  const double z_hi = y * L_hi;
  const double z_lo = fma(y, L_hi, -z_hi) + y * L_lo;

  /* ---- exp(z) ---- */
  const double kd = rint(z_hi * FEDG_POW_INVLN2N);
  const int k = (int)kd;
  const double t_hi = z_hi - kd * FEDG_POW_LN2N_HI; /* exact */
  const double t_lo = z_lo - kd * FEDG_POW_LN2N_LO;
  double t_lo2;
  const double t = fedg_two_sum(t_hi, t_lo, &t_lo2);
  /* expm1(t) = t + t^2/2 + t^3 Q(t) plus the first-order effect of t_lo2 */
  const double q = 0.16666666666666666 +
                   t * (0.041666666666666664 + t * (0.008333333333333333 + t * (0.001388888888888889 +
                   t * (0.0001984126984126984 + t * 2.48015873015873e-05))));
  const double s_hi = 0.5 * (t * t);
  const double s_lo = 0.5 * fma(t, t, -(t * t));
  double em_lo0;
  const double em_hi = fedg_fast_two_sum(t, s_hi, &em_lo0);
  const double em_lo = em_lo0 + s_lo + ((t * t) * t) * q + (t_lo2 + t_lo2 * t);
  const int j = k & 127;
  const int qexp = (k - j) / 128;
  const double Eh = texp_hi[j], El = texp_lo[j];
  const double pr_hi = Eh * em_hi;
  const double pr_lo = fma(Eh, em_hi, -pr_hi);
  double u_lo;
  const double u_hi = fedg_fast_two_sum(Eh, pr_hi, &u_lo);
  const double res = u_hi + (u_lo + pr_lo + El + El * em_hi + Eh * em_lo);
  /* 2^qexp exactly (qexp within the normal range here) */
  return res * fedg_pow_from_bits((uint64_t)(qexp + 1023) << 52);
}

FEDG_POW_FN double fedg_pow_core(double x, double y) {
  return fedg_pow_core_t(x, y, fedg_pow_invc, fedg_pow_logc_hi, fedg_pow_logc_lo, fedg_pow_exp_hi, fedg_pow_exp_lo);
}

/* Whether (x, y) takes the portable path: x > 0 finite and |y log2 x| well
 * inside the double range. */
FEDG_POW_FN int fedg_pow_in_range(double x, double y) {
  if (x > 0.0 && x < 1.0e308 && y == y && y > -1.0e300 && y < 1.0e300) {
    const double lx = 710.0 / 0.6931471805599453; /* bound on |y * log2 x| */
    uint64_t ux = fedg_pow_bits(x);
    int e = (int)((ux >> 52) & 0x7ff) - 1023;
    const double ae = (double)(e < 0 ? -e : e) + 1.0;
    const double ay = y < 0.0 ? -y : y;
    return ay * ae < lx - 2.0;
  }
  return 0;
}

/* pow(x, y): the portable path for x > 0 finite and moderate results, the
 * platform pow otherwise (never reached by the dynamical core's pressure). */
FEDG_POW_FN double fedg_pow(double x, double y) {
  if (fedg_pow_in_range(x, y)) return fedg_pow_core(x, y);
  return pow(x, y);
}

/* fedg_pow with the tables at t = {invc, logc_hi, logc_lo, exp_hi, exp_lo},
 * 128 entries each, contiguous. */
FEDG_POW_FN double fedg_pow_tab(double x, double y, const double *t) {
  if (fedg_pow_in_range(x, y)) return fedg_pow_core_t(x, y, t, t + 128, t + 256, t + 384, t + 512);
  return pow(x, y);
}

#endif /* FEDG_POW_H */
