// Natural logarithm and power in FP64, for results that are rounded once to
// FP32 (calc_p_rho_phi's log(pfd / pfu) and temp^(cp/cv)).
//
// Written from the series, not a vendor library: the device library's FP64
// log and pow are full FP64-accuracy routines (extra-precision logarithm
// inside pow) and made the kernel that calls them bound by vector ALU issue.
// An FP64 result within a few ulp of the true value already rounds to the
// same FP32 value as a correctly rounded one except when the true value
// lies within ~1e-15 (relative) of an FP32 rounding boundary, about 1 value
// in 10^7; such a value then differs by one FP32 ulp.
//
//  f64_log(x), x > 0 finite and normal:
//    x = 2^e * m with m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact), and
//    log(m) = 2 atanh(s), s = f / (2 + f), |s| <= 0.1716.  With hfsq = f^2 / 2
//    one has 2s = f - hfsq + s * hfsq, so
//      log(m) = f - hfsq + s * (hfsq + R),   R = sum_{n>=1} 2 s^(2n) / (2n+1)
//    (the arrangement of Sun's fdlibm e_log.c; the coefficients here are the
//    series' own 2 / (2n+1), n = 1..11: the first term left out is below
//    2^-60 of the result).  e * ln2 is added as e * ln2_hi (exact product:
//    ln2_hi has 11 trailing zero bits) plus e * ln2_lo.
//  f64_exp(y), |y| < 700:
//    y = n * ln2 + r, n = rint(y / ln2), r reduced with the same split of
//    ln2 (|r| <= 0.347), exp(r) = sum_{k=0}^{13} r^k / k! by Horner
//    (remainder below 4e-18), scaled by 2^n.
//  f64_pow(t, c) = f64_exp(c * f64_log(t)), t > 0: the logarithm's absolute
//    error (a few 1e-17 for the |log t| < 3 of the callers) times c stays
//    far below an FP64 ulp of the product.
//
// Only FMA-explicit arithmetic (__builtin_fma) and IEEE operations, so the
// host and the device evaluate it to the same bits; tests/test_f64math.py
// compiles this header for the host and compares the FP32-rounded results
// with the C library's log and pow over every FP32 input of the callers'
// ranges.
#pragma once

#ifdef __HIPCC__
#define F64M_FN __host__ __device__ __forceinline__
#else
#define F64M_FN static inline
#endif

// ln 2 = kLn2Hi + kLn2Lo; kLn2Hi has its low 11 bits zero
#define F64M_LN2_HI 6.93147180369123816490e-01
#define F64M_LN2_LO 1.90821492927058770002e-10

F64M_FN double f64_log(double x) {
  int e;
  double m = __builtin_frexp(x, &e);  // [0.5, 1)
  if (m < 0.70710678118654752440) {
    m = m + m;
    e = e - 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  // R = z * (2/3 + z * (2/5 + ... + z * 2/23))
  double p = 2.0 / 23.0;
  p = __builtin_fma(p, z, 2.0 / 21.0);
  p = __builtin_fma(p, z, 2.0 / 19.0);
  p = __builtin_fma(p, z, 2.0 / 17.0);
  p = __builtin_fma(p, z, 2.0 / 15.0);
  p = __builtin_fma(p, z, 2.0 / 13.0);
  p = __builtin_fma(p, z, 2.0 / 11.0);
  p = __builtin_fma(p, z, 2.0 / 9.0);
  p = __builtin_fma(p, z, 2.0 / 7.0);
  p = __builtin_fma(p, z, 2.0 / 5.0);
  p = __builtin_fma(p, z, 2.0 / 3.0);
  const double R = z * p;
  const double hfsq = 0.5 * f * f;
  const double de = (double)e;
  // e*ln2_hi + (f - (hfsq - (s*(hfsq + R) + e*ln2_lo)))
  const double corr = __builtin_fma(s, hfsq + R, de * F64M_LN2_LO);
  return de * F64M_LN2_HI + (f - (hfsq - corr));
}

F64M_FN double f64_exp(double y) {
  const double n = __builtin_rint(y * 1.44269504088896338700);  // y / ln2
  const double r = __builtin_fma(-n, F64M_LN2_LO, __builtin_fma(-n, F64M_LN2_HI, y));
  // sum_{k=0}^{13} r^k / k!
  double p = 1.0 / 6227020800.0;  // 1/13!
  p = __builtin_fma(p, r, 1.0 / 479001600.0);
  p = __builtin_fma(p, r, 1.0 / 39916800.0);
  p = __builtin_fma(p, r, 1.0 / 3628800.0);
  p = __builtin_fma(p, r, 1.0 / 362880.0);
  p = __builtin_fma(p, r, 1.0 / 40320.0);
  p = __builtin_fma(p, r, 1.0 / 5040.0);
  p = __builtin_fma(p, r, 1.0 / 720.0);
  p = __builtin_fma(p, r, 1.0 / 120.0);
  p = __builtin_fma(p, r, 1.0 / 24.0);
  p = __builtin_fma(p, r, 1.0 / 6.0);
  p = __builtin_fma(p, r, 0.5);
  p = __builtin_fma(p, r, 1.0);
  p = __builtin_fma(p, r, 1.0);
  return __builtin_ldexp(p, (int)n);
}

F64M_FN double f64_pow(double t, double c) { return f64_exp(c * f64_log(t)); }
