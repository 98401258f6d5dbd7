/*
 * pmx_device.h -- FP64 device math of the step path.  Operation order follows
 * the reference macros (macros.h) so that, compiled with -ffp-contract=off,
 * every value is the IEEE result of the same sequence of +, -, *, / and sqrt.
 */
#ifndef PMX_DEVICE_H
#define PMX_DEVICE_H

#include <hip/hip_runtime.h>
#include "pmx_trig.h"

namespace pmxd {

/* exponential map, macros.h:1453-1534, including the reference's
 * v02 = Omega2*Omega3 term (macros.h:1505) */
__device__ __forceinline__ void expmap (double O1, double O2, double O3, double L[9])
{
  double angsq, sx, cx, v0, v1, v2, v01, v02, v12, s0, s1, s2;
  v0 = O1 * O1;
  v1 = O2 * O2;
  v2 = O3 * O3;
  angsq = v0 + v1 + v2;
  if (angsq < 3.0461741978671E-02)
  {
    sx = 1.0 + (-1.666666666666667E-1 + (8.333333333333333E-3 + (-1.984126984126984E-4 +
         (2.755731922398589E-6 + (-2.505210838544172E-8 + 1.605904383682161E-10 * angsq
         ) * angsq) * angsq) * angsq) * angsq) * angsq;
    cx = 0.5 + (-4.166666666666667E-2 + (1.388888888888889E-3 + (-2.480158730158730E-5 +
         (2.755731922398589E-7 + (-2.087675698786810E-9 + 1.147074559772972E-11 * angsq
         ) * angsq) * angsq) * angsq) * angsq) * angsq;
  }
  else
  {
    double t = angsq;
    angsq = sqrt (angsq);
    double s = pmxt::sin (angsq), c = pmxt::cos (angsq);
    sx = s / angsq;
    cx = (1.0 - c) / t;
  }
  v01 = O1 * O2;
  v02 = O2 * O3;
  v12 = O2 * O3;
  s0 = sx * O1;
  s1 = sx * O2;
  s2 = sx * O3;
  L[0] = -cx * (v2 + v1);
  L[1] = cx * v01;
  L[2] = cx * v02;
  L[3] = L[1];
  L[4] = -cx * (v2 + v0);
  L[5] = cx * v12;
  L[6] = L[2];
  L[7] = L[5];
  L[8] = -cx * (v1 + v0);
  L[0] += 1.0;
  L[1] += s2;
  L[2] -= s1;
  L[3] -= s2;
  L[4] += 1.0;
  L[5] += s0;
  L[6] += s1;
  L[7] -= s0;
  L[8] += 1.0;
}

__device__ __forceinline__ void nvmul (const double A[9], const double B[3], double C[3])
{
  C[0] = A[0]*B[0] + A[3]*B[1] + A[6]*B[2];
  C[1] = A[1]*B[0] + A[4]*B[1] + A[7]*B[2];
  C[2] = A[2]*B[0] + A[5]*B[1] + A[8]*B[2];
}

__device__ __forceinline__ void tvmul (const double A[9], const double B[3], double C[3])
{
  C[0] = A[0]*B[0] + A[1]*B[1] + A[2]*B[2];
  C[1] = A[3]*B[0] + A[4]*B[1] + A[5]*B[2];
  C[2] = A[6]*B[0] + A[7]*B[1] + A[8]*B[2];
}

__device__ __forceinline__ void nnmul (const double A[9], const double B[9], double C[9])
{
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int r = 0; r < 3; r++)
      C[3*c+r] = A[r]*B[3*c] + A[3+r]*B[3*c+1] + A[6+r]*B[3*c+2];
}

__device__ __forceinline__ double clamp01 (double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

/* closest point of a triangle to a sphere centre (Eberly regions),
 * condet.ispc:294-400 and 519-624; returns r - distance */
__device__ __forceinline__ double triangle_sphere (double ax, double ay, double az, double bx, double by, double bz,
                                                    double cx, double cy, double cz, double px, double py, double pz,
                                                    double r, double point[3], double normal[3])
{
  double e0x = bx-ax, e0y = by-ay, e0z = bz-az;
  double e1x = cx-ax, e1y = cy-ay, e1z = cz-az;
  double v0x = ax-px, v0y = ay-py, v0z = az-pz;
  double a = e0x*e0x + e0y*e0y + e0z*e0z;
  double b = e0x*e1x + e0y*e1y + e0z*e1z;
  double c = e1x*e1x + e1y*e1y + e1z*e1z;
  double d = e0x*v0x + e0y*v0y + e0z*v0z;
  double e = e1x*v0x + e1y*v0y + e1z*v0z;
  double det = a*c - b*b;
  double s = b*e - c*d;
  double t = b*d - a*e;
  if (s + t < det)
  {
    if (s < 0.0)
    {
      if (t < 0.0)
      {
        if (d < 0.0) { s = clamp01 (-d/a); t = 0.0; }
        else { s = 0.0; t = clamp01 (-e/c); }
      }
      else { s = 0.0; t = clamp01 (-e/c); }
    }
    else if (t < 0.0) { s = clamp01 (-d/a); t = 0.0; }
    else
    {
      double invDet = 1.0 / det;
      s *= invDet;
      t *= invDet;
    }
  }
  else
  {
    if (s < 0.0)
    {
      double tmp0 = b+d;
      double tmp1 = c+e;
      if (tmp1 > tmp0)
      {
        double numer = tmp1 - tmp0;
        double denom = a-2*b+c;
        s = clamp01 (numer/denom);
        t = 1-s;
      }
      else { t = clamp01 (-e/c); s = 0.0; }
    }
    else if (t < 0.0)
    {
      if (a+d > b+e)
      {
        double numer = c+e-b-d;
        double denom = a-2*b+c;
        s = clamp01 (numer/denom);
        t = 1-s;
      }
      else { s = clamp01 (-e/c); t = 0.0; }
    }
    else
    {
      double numer = c+e-b-d;
      double denom = a-2*b+c;
      s = clamp01 (numer/denom);
      t = 1.0 - s;
    }
  }
  point[0] = ax + s*e0x + t*e1x;
  point[1] = ay + s*e0y + t*e1y;
  point[2] = az + s*e0z + t*e1z;
  double qx = px-point[0], qy = py-point[1], qz = pz-point[2];
  double len = sqrt (qx*qx + qy*qy + qz*qz);
  double ilen = 1.0/len;
  normal[0] = ilen*qx;
  normal[1] = ilen*qy;
  normal[2] = ilen*qz;
  return r - len;
}

/* pairing of colours, forces.ispc:98-122 (loop exits where the reference
 * would spin forever) */
__device__ __forceinline__ int pairing (int nummat, const int *pairs, int i, int j)
{
  int p0 = i, p1 = j, start = 1, end = nummat;
  if (i > j) { p0 = j; p1 = i; }
  while (start < end)
  {
    int mid = ((end-start)>>1) + start;
    int q0 = pairs[2*mid], q1 = pairs[2*mid+1];
    if (p0 > q0) { if (start == mid) break; start = mid; }
    else if (p0 == q0)
    {
      if (p1 > q1) { if (start == mid) break; start = mid; }
      else if (p1 == q1) return mid;
      else end = mid;
    }
    else end = mid;
  }
  return 0;
}

/* granular spring-dashpot with Coulomb friction, forces.ispc:125-192; R is
 * the arithmetic of the law (double: the FP64 path; float: mixed precision) */
template <class R>
__device__ __forceinline__ void granular (const R n[3], const R vij[3], R depth, R ma, R kn, R damper, R kskn, R mus,
                                          R mud, R state[3], R step, R f[3])
{
  R en = damper * R (2.0) * sqrt (kn*ma);
  R vn = vij[0]*n[0] + vij[1]*n[1] + vij[2]*n[2];
  R fn = kn*depth + en*vn;
  f[0] = fn*n[0];
  f[1] = fn*n[1];
  f[2] = fn*n[2];
  R kt = kn * kskn;
  R et = en * kskn;
  R vt0 = vij[0] - vn*n[0];
  R vt1 = vij[1] - vn*n[1];
  R vt2 = vij[2] - vn*n[2];
  state[0] += step*vt0;
  state[1] += step*vt1;
  state[2] += step*vt2;
  R ft0 = kt*state[0] + et*vt0;
  R ft1 = kt*state[1] + et*vt1;
  R ft2 = kt*state[2] + et*vt2;
  R len = sqrt (ft0*ft0 + ft1*ft1 + ft2*ft2);
  R lim = mus*fn;
  /* len == 0: the tangential force is the zero vector and stays so (the
   * reference would divide by zero here when f_n < 0) */
  if (len > lim && len > R (0.0))
  {
    lim = mud*fn/len;
    ft0 *= lim;
    ft1 *= lim;
    ft2 *= lim;
  }
  f[0] += ft0;
  f[1] += ft1;
  f[2] += ft2;
}

/* NTMUL, macros.h:932-943: C = A B^T */
__device__ __forceinline__ void ntmul (const double A[9], const double B[9], double C[9])
{
  C[0] = A[0]*B[0]+A[3]*B[3]+A[6]*B[6];
  C[1] = A[1]*B[0]+A[4]*B[3]+A[7]*B[6];
  C[2] = A[2]*B[0]+A[5]*B[3]+A[8]*B[6];
  C[3] = A[0]*B[1]+A[3]*B[4]+A[6]*B[7];
  C[4] = A[1]*B[1]+A[4]*B[4]+A[7]*B[7];
  C[5] = A[2]*B[1]+A[5]*B[4]+A[8]*B[7];
  C[6] = A[0]*B[2]+A[3]*B[5]+A[6]*B[8];
  C[7] = A[1]*B[2]+A[4]*B[5]+A[7]*B[8];
  C[8] = A[2]*B[2]+A[5]*B[5]+A[8]*B[8];
}

/* INVERT, macros.h:1033-1051: the inverse where the determinant is not
 * zero; otherwise INV keeps what it held */
__device__ __forceinline__ void invert3 (const double F[9], double INV[9])
{
  double det = F[0]*(F[4]*F[8]-F[5]*F[7]) + F[3]*(F[2]*F[7]-F[1]*F[8]) + (F[1]*F[5]-F[2]*F[4])*F[6];
  if (det != 0.0)
  {
    det = 1.0 / det;
    INV[0] = (F[4]*F[8]-F[5]*F[7])*det;
    INV[1] = (F[2]*F[7]-F[1]*F[8])*det;
    INV[2] = (F[1]*F[5]-F[2]*F[4])*det;
    INV[3] = (F[5]*F[6]-F[3]*F[8])*det;
    INV[4] = (F[0]*F[8]-F[2]*F[6])*det;
    INV[5] = (F[2]*F[3]-F[0]*F[5])*det;
    INV[6] = (F[3]*F[7]-F[4]*F[6])*det;
    INV[7] = (F[1]*F[6]-F[0]*F[7])*det;
    INV[8] = (F[0]*F[4]-F[1]*F[3])*det;
  }
}

/* extremum_eigenvalue, macros.h:1538-1593: the largest eigenvalue of a
 * symmetric 3x3 matrix, or the smallest with `minimum`, by the
 * trigonometric solution of the characteristic cubic.  A diagonal matrix
 * returns its largest diagonal entry whether or not `minimum` is set, as
 * the reference does. */
__device__ __forceinline__ double extremum_eigenvalue (const double A[9], bool minimum)
{
  const double pi = 3.14159265358979323846;
  double p1 = A[3]*A[3] + A[6]*A[6] + A[7]*A[7];
  if (p1 == 0.0)
  {
    double r = A[0] > A[4] ? A[0] : A[4];
    return r > A[8] ? r : A[8];
  }
  double q = (A[0]+A[4]+A[8])/3.0;
  double p2 = (A[0]-q)*(A[0]-q) + (A[4]-q)*(A[4]-q) + (A[8]-q)*(A[8]-q) + 2*p1;
  double p = sqrt (p2 / 6.0);
  double invp = 1.0 / p;
  double B[9] = {invp*(A[0]-q), invp*A[1], invp*A[2], invp*A[3], invp*(A[4]-q), invp*A[5],
                 invp*A[6], invp*A[7], invp*(A[8]-q)};
  double r = 0.5 * (B[0]*(B[4]*B[8]-B[5]*B[7]) + B[3]*(B[2]*B[7]-B[1]*B[8]) + B[6]*(B[1]*B[5]-B[2]*B[4]));
  double phi;
  if (r <= -1.0) phi = pi / 3.0;
  else if (r >= 1.0) phi = 0.0;
  else phi = pmxt::acos (r) / 3.0;
  if (minimum) return q + 2.0 * p * pmxt::cos (phi + (2*pi/3));
  return q + 2.0 * p * pmxt::cos (phi);
}

/* one particle's critical step, time_step_task (parmec.ispc:275-420): the
 * translational estimate from the largest summed |direction| component,
 * the rotational one from the largest eigenvalue of L^-1 K L^-T (J = L L^T
 * by Cholesky), each shortened for damping (h = 2/w (sqrt(1+x^2) - x));
 * returns the smaller, with the larger frequency in *o and damping ratio
 * in *r.  A particle whose summed directions are all zero returns DBL_MAX
 * and leaves *o and *r alone.  min and max are a < b ? a : b and
 * a > b ? a : b throughout (so a NaN second operand is taken). */
__device__ __forceinline__ double critical_step (const double kact[3], double kmax, double emax, const double krot[6],
                                                 double m, const double J[9], double *o, double *r)
{
  double nxy = kact[0] > kact[1] ? kact[0] : kact[1];
  double nxyz = nxy > kact[2] ? nxy : kact[2];
  if (!(nxyz > 0.0)) return 1.7976931348623157e308;
  double k = kmax;
  double omegaxyz = sqrt (nxyz*k/m);
  double K[9];
  K[0] = k*krot[0];
  K[1] = k*krot[3];
  K[2] = k*krot[4];
  K[3] = K[1];
  K[4] = k*krot[1];
  K[5] = k*krot[5];
  K[6] = K[2];
  K[7] = K[5];
  K[8] = k*krot[2];
  double mrot = extremum_eigenvalue (J, true);
  double L[9], inv, invL[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, temp[9], A[9];
  L[0] = sqrt (J[0]);
  inv = 1.0/L[0];
  L[1] = inv * J[1];
  L[2] = inv * J[2];
  L[3] = 0.0;
  L[4] = sqrt (J[4] - L[1]*L[1]);
  inv = 1.0/L[4];
  L[5] = inv * (J[5] - L[2]*L[1]);
  L[6] = 0.0;
  L[7] = 0.0;
  L[8] = sqrt (J[8] - (L[2]*L[2] + L[5]*L[5]));
  invert3 (L, invL);
  ntmul (K, invL, temp);
  nnmul (invL, temp, A);
  double lambda = extremum_eigenvalue (A, false);
  double omegarot = sqrt (lambda);
  double damping = emax;
  double ratioxyz = damping / (2.0*m*omegaxyz);
  double ratiorot = damping / (2.0*mrot*omegarot);
  double hxyz = (2.0/omegaxyz)*(sqrt (1.0+ratioxyz*ratioxyz) - ratioxyz);
  double hrot = (2.0/omegarot)*(sqrt (1.0+ratiorot*ratiorot) - ratiorot);
  *o = omegaxyz > omegarot ? omegaxyz : omegarot;
  *r = ratioxyz > ratiorot ? ratioxyz : ratiorot;
  return hxyz < hrot ? hxyz : hrot;
}

/* compensated (Kahan) float accumulation */
__device__ __forceinline__ void kahan_add (float &sum, float &comp, float x)
{
  float y = x - comp;
  float t = sum + y;
  comp = (t - sum) - y;
  sum = t;
}

} /* namespace pmxd */

#endif
