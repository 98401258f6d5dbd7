/*
 * pmx_trig.h -- sine, cosine and arc cosine from +, -, *, / and sqrt only.
 *
 * libm and the device math library may differ in the last place, and on the
 * step path such a difference reaches every later state: the exponential
 * map beyond 10 degrees (macros.h:1488-1493) and the extremum eigenvalue of
 * a symmetric 3x3 matrix (macros.h:1538-1593), which decides the critical
 * and adaptive time steps and the equivalent mass of a critically damped
 * spring.  These functions are compiled from this one header on the device
 * and, with -ffp-contract=off, on the host, so both give the same bits.
 *
 * sin, cos: x = k pi/2 + r with k = rint(2x/pi); pi/2 is split into parts of
 * 33, 33 and 53 bits (Cody and Waite, "Software Manual for the Elementary
 * Functions", 1980), so k times the first two parts is exact for |k| < 2^20;
 * then the Taylor polynomials of sin r (degree 21) and cos r (degree 20) on
 * |r| <= pi/4, whose first omitted terms are below 2^-60 of the result.
 * acos: pi/2 - asin x for |x| <= 1/2, 2 asin(sqrt((1-x)/2)) above and
 * pi - 2 asin(sqrt((1+x)/2)) below, the sqrt argument then <= 1/4 and exact
 * (Sterbenz); asin by its Taylor series sum_n (2n)!/(4^n n!^2 (2n+1)) x^(2n+1)
 * to degree 49 (first omitted term below 2^-56 of the sum at |x| = 1/2); pi
 * and pi/2 carry a low part added last.
 */
#ifndef PMX_TRIG_H
#define PMX_TRIG_H

#include <cmath>

#if defined(__HIPCC__)
#define PMX_HD __host__ __device__ __forceinline__
#else
#define PMX_HD inline
#endif

namespace pmxt {

/* 1/(2n+1)! with alternating signs, n = 1..10 */
PMX_HD double sin_taylor (double r)
{
  double z = r * r;
  double p = 1.9572941063391263e-20;
  p = -8.22063524662433e-18 + p * z;
  p = 2.8114572543455206e-15 + p * z;
  p = -7.647163731819816e-13 + p * z;
  p = 1.6059043836821613e-10 + p * z;
  p = -2.505210838544172e-08 + p * z;
  p = 2.7557319223985893e-06 + p * z;
  p = -0.0001984126984126984 + p * z;
  p = 0.008333333333333333 + p * z;
  p = -0.16666666666666666 + p * z;
  return r + r * (z * p);
}

/* 1/(2n)! with alternating signs, n = 1..10 */
PMX_HD double cos_taylor (double r)
{
  double z = r * r;
  double p = 4.110317623312165e-19;
  p = -1.5619206968586225e-16 + p * z;
  p = 4.779477332387385e-14 + p * z;
  p = -1.1470745597729725e-11 + p * z;
  p = 2.08767569878681e-09 + p * z;
  p = -2.755731922398589e-07 + p * z;
  p = 2.48015873015873e-05 + p * z;
  p = -0.001388888888888889 + p * z;
  p = 0.041666666666666664 + p * z;
  p = -0.5 + p * z;
  return 1.0 + z * p;
}

/* x = k pi/2 + r, |r| <= pi/4 (up to rounding); returns k mod 4 */
PMX_HD int reduce_pio2 (double x, double *r)
{
  const double twoopi = 0.6366197723675814;
  const double p1 = 1.5707963267341256, p2 = 6.077100506303966e-11, p3 = 2.0222662487959506e-21;
  double k = rint (x * twoopi);
  *r = ((x - k * p1) - k * p2) - k * p3;
  double q = k - 4.0 * floor (k * 0.25);
  return (int) q;
}

PMX_HD double sin (double x)
{
  if (!(x - x == 0.0)) return x - x;   /* inf, nan */
  double r;
  switch (reduce_pio2 (x, &r))
  {
  case 0: return sin_taylor (r);
  case 1: return cos_taylor (r);
  case 2: return -sin_taylor (r);
  default: return -cos_taylor (r);
  }
}

PMX_HD double cos (double x)
{
  if (!(x - x == 0.0)) return x - x;
  double r;
  switch (reduce_pio2 (x, &r))
  {
  case 0: return cos_taylor (r);
  case 1: return -sin_taylor (r);
  case 2: return -cos_taylor (r);
  default: return sin_taylor (r);
  }
}

/* asin on |x| <= 1/2: x + x z P(z), z = x^2, coefficients of x^3 .. x^49 */
PMX_HD double asin_taylor (double x)
{
  double z = x * x;
  double p = 0.0024894486782468836;
  p = 0.00265787063820729 + p * z;
  p = 0.002846178401108942 + p * z;
  p = 0.0030578216492580306 + p * z;
  p = 0.003297059503473485 + p * z;
  p = 0.0035692053938259347 + p * z;
  p = 0.003880964558837669 + p * z;
  p = 0.004240907093679363 + p * z;
  p = 0.004660143486915096 + p * z;
  p = 0.005153309682319905 + p * z;
  p = 0.005740037670841924 + p * z;
  p = 0.006447210311889649 + p * z;
  p = 0.0073125258735988454 + p * z;
  p = 0.008390335809616815 + p * z;
  p = 0.009761609529194078 + p * z;
  p = 0.011551800896139705 + p * z;
  p = 0.01396484375 + p * z;
  p = 0.017352764423076924 + p * z;
  p = 0.022372159090909092 + p * z;
  p = 0.030381944444444444 + p * z;
  p = 0.044642857142857144 + p * z;
  p = 0.075 + p * z;
  p = 0.16666666666666666 + p * z;
  return x + x * (z * p);
}

PMX_HD double acos (double x)
{
  const double pio2_hi = 1.5707963267948966, pio2_lo = 6.123233995736766e-17;
  const double pi_hi = 3.141592653589793, pi_lo = 1.2246467991473532e-16;
  if (x != x) return x;
  if (x > 1.0 || x < -1.0) return (x - x) / (x - x);
  if (x <= 0.5 && x >= -0.5) return pio2_hi - (asin_taylor (x) - pio2_lo);
  if (x > 0.5) return 2.0 * asin_taylor (sqrt ((1.0 - x) * 0.5));
  return pi_hi - (2.0 * asin_taylor (sqrt ((1.0 + x) * 0.5)) - pi_lo);
}

} /* namespace pmxt */

#endif
