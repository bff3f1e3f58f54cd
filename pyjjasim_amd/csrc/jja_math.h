// jja_math.h -- double-precision sine for the step kernel, usable on host and
// device (host builds test it against the C library).
//
// Method (published, not derived from any GPU vendor library):
//   argument reduction  x = n*(pi/2) + r, n = rint(x*2/pi), with pi/2 split
//     into P1 + P2 (P1 = pi/2 rounded to double): n*P1 is formed exactly as a
//     product pair (hi + lo by FMA), x - hi is exact (Sterbenz: x and hi are
//     within a factor two whenever n != 0), the small terms -lo - n*P2 are
//     added with one rounding, and the rounding error of that addition is kept
//     as a tail rl. Accurate for |x| up to about 1e15; beyond, the neglected
//     n*P3 term grows like 1e-33 |x|.
//   kernels on |r| <= pi/4: the sine and cosine polynomials of Sun's fdlibm in
//     the branch-free form of FreeBSD msun (k_sin.c, k_cos.c), which take the
//     tail rl as a first-order correction; the polynomials are evaluated with
//     fused multiply-adds (one rounding per step instead of two).
// Measured against glibc's sin (tests/test_jja_math.py): at most 1 ulp for
// |x| <= 1e9; near the zeros k*pi/2, where the result is tiny, the absolute
// error stays below 1e-24.
//
// The kernel polynomials jja_ksin and jja_kcos follow fdlibm's k_sin.c and
// k_cos.c, which carry this notice:
//
//   Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//
//   Developed at SunPro, a Sun Microsystems, Inc. business.
//   Permission to use, copy, modify, and distribute this
//   software is freely granted, provided that this notice
//   is preserved.
#ifndef JJA_MATH_H
#define JJA_MATH_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define JJA_HD __host__ __device__ __forceinline__
#else
#define JJA_HD static inline
#endif

// sin(x + y) for |x| <= pi/4, |y| << ulp(x) (fdlibm __kernel_sin, iy = 1)
JJA_HD double jja_ksin(double x, double y) {
  const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
               S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
               S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
  const double z = x * x;
  const double w = z * z;
  const double r = fma(z * w, fma(z, S6, S5), fma(z, fma(z, S4, S3), S2));
  const double v = z * x;
  return x - fma(-v, S1, fma(z, fma(-v, r, 0.5 * y), -y));
}

// cos(x + y) for |x| <= pi/4 (fdlibm __kernel_cos, msun form)
JJA_HD double jja_kcos(double x, double y) {
  const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
               C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
               C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
  const double z = x * x;
  const double w = z * z;
  const double r = fma(w * w, fma(z, fma(z, C6, C5), C4), z * fma(z, fma(z, C3, C2), C1));
  const double hz = 0.5 * z;
  const double ww = 1.0 - hz;
  return ww + (((1.0 - ww) - hz) + fma(z, r, -(x * y)));
}

JJA_HD double jja_sin(double x) {
  const double kTwoOverPi = 6.36619772367581382433e-01;
  const double P1 = 1.57079632679489655800e+00;  // pi/2 rounded to double
  const double P2 = 6.12323399573676603587e-17;  // pi/2 - P1 rounded to double
  const double n = rint(x * kTwoOverPi);
  const double hi = n * P1;
  const double lo = fma(n, P1, -hi);  // n*P1 = hi + lo exactly
  const double rh = x - hi;           // exact
  const double s = fma(-n, P2, -lo);
  const double r = rh + s;
  const double rl = (rh - r) + s;
  const double q = n - 4.0 * floor(n * 0.25);  // n mod 4, exact
  const double v = (q == 0.0 || q == 2.0) ? jja_ksin(r, rl) : jja_kcos(r, rl);
  return q >= 2.0 ? -v : v;
}

// sin(x) and cos(x) with one shared argument reduction (same method and
// accuracy as jja_sin; cos(x) = sin(x + pi/2) is quadrant n + 1).
JJA_HD void jja_sincos(double x, double* sn, double* cs) {
  const double kTwoOverPi = 6.36619772367581382433e-01;
  const double P1 = 1.57079632679489655800e+00;
  const double P2 = 6.12323399573676603587e-17;
  const double n = rint(x * kTwoOverPi);
  const double hi = n * P1;
  const double lo = fma(n, P1, -hi);
  const double rh = x - hi;
  const double s = fma(-n, P2, -lo);
  const double r = rh + s;
  const double rl = (rh - r) + s;
  const double q = n - 4.0 * floor(n * 0.25);
  const double ks = jja_ksin(r, rl), kc = jja_kcos(r, rl);
  const double vs = (q == 0.0 || q == 2.0) ? ks : kc;
  const double vc = (q == 0.0 || q == 2.0) ? kc : ks;
  *sn = q >= 2.0 ? -vs : vs;
  *cs = (q == 1.0 || q == 2.0) ? -vc : vc;
}

#endif  // JJA_MATH_H
