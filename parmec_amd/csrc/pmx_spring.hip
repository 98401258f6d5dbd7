/*
 * pmx_spring.hip -- SPRING elements (springs_task, forces.ispc:641-1412) and
 * the geometry update of meshed particles (nodes_task/faces_task,
 * shapes.ispc:101-211).
 *
 * k_springs: one thread per spring computes the current points, direction,
 *   stroke, curve force (elastic or yield/unload), dashpot, friction, and the
 *   spring's force and the torques it applies to both particles; spring state
 *   (points, direction, stroke, yield limits, flags, friction history) is
 *   updated in place.
 * k_spring_acc: one thread per particle adds the spring contributions in the
 *   order the reference's single-task loop flushes them: a local sum per run
 *   of consecutive springs sharing part1 (added) or part2 (subtracted),
 *   flushed where the run ends (forces.ispc:1203-1411).  The run table is
 *   built on the host at upload; springs are sorted by (part1, part2) as
 *   sort_springs leaves them (parmec.cpp:1683-1966).
 */
#include "pmx_internal.h"
#include "pmx_device.h"
#include <algorithm>

using namespace pmxd;

namespace {

inline dim3 grid (int n) { return dim3 ((n + PMX_TPB - 1) / PMX_TPB); }

/* forces.ispc:410-426 */
__device__ __forceinline__ double look_up (const double *x, const double *y, int lo, int hi, double arg, double *slope)
{
  int h0 = hi;
  while (hi > lo+1)
  {
    int mid = lo + (hi-lo)/2;
    if (arg > x[mid]) lo = mid;
    else hi = mid;
  }
  if (lo+1 == h0) lo--;
  *slope = (y[lo+1]-y[lo])/(x[lo+1]-x[lo]);
  return y[lo] + (*slope)*(arg-x[lo]);
}

/* forces.ispc:429-450, searching along y (the unload curves) */
__device__ __forceinline__ double look_up_slope_y (const double *x, const double *y, int lo, int hi, double arg, double *at)
{
  int h0 = hi;
  while (hi > lo+1)
  {
    int mid = lo + (hi-lo)/2;
    if (arg > y[mid]) lo = mid;
    else hi = mid;
  }
  if (lo+1 == h0) lo--;
  if (at) at[0] = x[lo] + (x[lo+1]-x[lo])*(arg-y[lo])/(y[lo+1]-y[lo]);
  return (y[lo+1]-y[lo])/(x[lo+1]-x[lo]);
}

__device__ __forceinline__ void skew (const double v[3], double A[9])
{
  A[0] = 0; A[1] = v[2]; A[2] = -v[1];
  A[3] = -v[2]; A[4] = 0; A[5] = v[0];
  A[6] = v[1]; A[7] = -v[0]; A[8] = 0;
}

__device__ __forceinline__ void tnmul (const double A[9], const double B[9], double C[9])
{
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int r = 0; r < 3; r++)
      C[3*c+r] = A[3*r]*B[3*c] + A[3*r+1]*B[3*c+1] + A[3*r+2]*B[3*c+2];
}

/* eqm, parmec.ispc:423-482 */
__device__ double eqm (int N, const double *inverse, const double *invm, const double *position, int part1,
                       const double point1[3], int part2, const double point2[3], const double dir[3])
{
  double Jiv[9], im, A[3], Ask[9], tmp[9], Miv[9];
#pragma unroll
  for (int k = 0; k < 9; k++) Jiv[k] = inverse[(size_t) k*N + part1];
  im = invm[part1];
#pragma unroll
  for (int k = 0; k < 3; k++) A[k] = position[(size_t) (3+k)*N + part1] - point1[k];
  skew (A, Ask);
  tnmul (Ask, Jiv, tmp);
  nnmul (tmp, Ask, Miv);
  Miv[0] += im; Miv[4] += im; Miv[8] += im;
  if (part2 >= 0)
  {
#pragma unroll
    for (int k = 0; k < 9; k++) Jiv[k] = inverse[(size_t) k*N + part2];
    im = invm[part2];
#pragma unroll
    for (int k = 0; k < 3; k++) A[k] = position[(size_t) (3+k)*N + part2] - point2[k];
    skew (A, Ask);
    tnmul (Ask, Jiv, tmp);
    nnmul (tmp, Ask, Jiv);
#pragma unroll
    for (int k = 0; k < 9; k++) Miv[k] = Miv[k] + Jiv[k];
    Miv[0] += im; Miv[4] += im; Miv[8] += im;
  }
  if (dir[0]*dir[0] + dir[1]*dir[1] + dir[2]*dir[2] > 0.0)
  {
    nvmul (Miv, dir, tmp);
    return 1.0/(tmp[0]*dir[0] + tmp[1]*dir[1] + tmp[2]*dir[2]);
  }
  return 1.0/extremum_eigenvalue (Miv, true);
}

struct SprArgs
{
  int S, N;
  const int *type, *part, *spridx, *dashidx, *unidx;
  int *unspring, *flg;
  const int *offset, *lcidx;          /* stroke offset curves (-1: none), load curve ranges */
  const double *lcx, *lcy, *ttab;     /* load curves; curtime of each step of the call */
  double *pnt, *dir, *yield, *sdsp, *stroke, *frc;
  const double *stroke0, *fric, *kskn;
  const double *sx, *sy, *dx, *dy, *ux, *uy;
  const double *angular, *linear, *rotation, *position, *inverse, *invm;
  double *out;   /* [9][S]: force, torque on part1, torque on part2 */
  double *adp;   /* [17][S] terms of the adaptive step, or null (see k_springs) */
  int *flags;
  const int *ctl;
  const int *stop, *par_inst;   /* frozen instances (staggered dem calls) */
};

__global__ void k_springs (SprArgs A, double step)
{
  if (A.ctl[CTL_ABORT]) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.S || frozen (A.stop, A.par_inst[A.part[i]], A.ctl)) return;
  size_t S = A.S, N = A.N;
  auto P = [&] (int p, int c) -> double & { return A.pnt[(size_t) (6*p+c) * S + i]; };
  int j = A.part[i], k = A.part[S + i];
  double oj[3], vj[3], xj[3], Xj[3], Lj[9];
#pragma unroll
  for (int c = 0; c < 3; c++)
  {
    oj[c] = A.angular[(3+c)*N + j]; vj[c] = A.linear[c*N + j];
    xj[c] = A.position[c*N + j]; Xj[c] = A.position[(3+c)*N + j];
  }
#pragma unroll
  for (int c = 0; c < 9; c++) Lj[c] = A.rotation[c*N + j];
  double len = 0.0, inv, velocity, spring_force = 0.0, dashpot_force = 0.0, total_force;
  double Yj[3], Yk[3] = {0.0, 0.0, 0.0}, p[3], q[3], z[3] = {0.0, 0.0, 0.0}, w[3], frc[3], u[3], a[3], b[3] = {0.0, 0.0, 0.0};
  double stiffness = 0.0, damping = 0.0;
#pragma unroll
  for (int c = 0; c < 3; c++) Yj[c] = P (0, 3+c) - Xj[c];
  p[0] = xj[0] + (Lj[0]*Yj[0] + Lj[3]*Yj[1] + Lj[6]*Yj[2]);
  p[1] = xj[1] + (Lj[1]*Yj[0] + Lj[4]*Yj[1] + Lj[7]*Yj[2]);
  p[2] = xj[2] + (Lj[2]*Yj[0] + Lj[5]*Yj[1] + Lj[8]*Yj[2]);
#pragma unroll
  for (int c = 0; c < 3; c++) { P (0, c) = p[c]; u[c] = vj[c] * -1.0; a[c] = p[c] - xj[c]; }
  u[0] -= oj[1]*a[2] - oj[2]*a[1];
  u[1] -= oj[2]*a[0] - oj[0]*a[2];
  u[2] -= oj[0]*a[1] - oj[1]*a[0];
  int flg = A.flg[i];
  if (k >= 0)
  {
    double ok[3], vk[3], xk[3], Xk[3], Lk[9];
#pragma unroll
    for (int c = 0; c < 3; c++)
    {
      ok[c] = A.angular[(3+c)*N + k]; vk[c] = A.linear[c*N + k];
      xk[c] = A.position[c*N + k]; Xk[c] = A.position[(3+c)*N + k];
    }
#pragma unroll
    for (int c = 0; c < 9; c++) Lk[c] = A.rotation[c*N + k];
    if (flg & PMX_SPRDIR_PROJECT)
    {
#pragma unroll
      for (int c = 0; c < 3; c++) w[c] = A.dir[(3+c)*S + i];
      nvmul (Lk, w, z);
#pragma unroll
      for (int c = 0; c < 3; c++) A.dir[c*S + i] = z[c];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) Yk[c] = P (1, 3+c) - Xk[c];
    q[0] = xk[0] + (Lk[0]*Yk[0] + Lk[3]*Yk[1] + Lk[6]*Yk[2]);
    q[1] = xk[1] + (Lk[1]*Yk[0] + Lk[4]*Yk[1] + Lk[7]*Yk[2]);
    q[2] = xk[2] + (Lk[2]*Yk[0] + Lk[5]*Yk[1] + Lk[8]*Yk[2]);
#pragma unroll
    for (int c = 0; c < 3; c++) { P (1, c) = q[c]; u[c] += vk[c]; b[c] = q[c] - xk[c]; }
    u[0] += ok[1]*b[2] - ok[2]*b[1];
    u[1] += ok[2]*b[0] - ok[0]*b[2];
    u[2] += ok[0]*b[1] - ok[1]*b[0];
  }
  else
  {
#pragma unroll
    for (int c = 0; c < 3; c++) q[c] = P (1, c);
    if (flg & PMX_SPRDIR_PROJECT)
    {
#pragma unroll
      for (int c = 0; c < 3; c++) z[c] = A.dir[c*S + i];
    }
  }
  switch (flg & (PMX_SPRDIR_FOLLOWER|PMX_SPRDIR_CONSTANT|PMX_SPRDIR_PLANAR|PMX_SPRDIR_PROJECT))
  {
  case PMX_SPRDIR_FOLLOWER:
    for (int c = 0; c < 3; c++) w[c] = q[c] - p[c];
    len = sqrt (w[0]*w[0] + w[1]*w[1] + w[2]*w[2]);
    inv = len > 0.0 ? 1.0/len : 0.0;
    for (int c = 0; c < 3; c++) { z[c] = inv*w[c]; A.dir[c*S + i] = z[c]; }
    break;
  case PMX_SPRDIR_CONSTANT:
    for (int c = 0; c < 3; c++) { w[c] = q[c] - p[c]; z[c] = A.dir[c*S + i]; }
    len = w[0]*z[0] + w[1]*z[1] + w[2]*z[2];
    break;
  case PMX_SPRDIR_PLANAR:
    for (int c = 0; c < 3; c++) w[c] = q[c] - p[c];
    inv = w[0]*A.dir[3*S + i] + w[1]*A.dir[4*S + i] + w[2]*A.dir[5*S + i];
    for (int c = 0; c < 3; c++) w[c] -= inv*A.dir[(3+c)*S + i];
    len = sqrt (w[0]*w[0] + w[1]*w[1] + w[2]*w[2]);
    inv = len > 0.0 ? 1.0/len : 0.0;
    for (int c = 0; c < 3; c++) { z[c] = inv*w[c]; A.dir[c*S + i] = z[c]; }
    break;
  case PMX_SPRDIR_PROJECT:
    for (int c = 0; c < 3; c++) w[c] = p[c] - q[c];
    len = w[0]*z[0] + w[1]*z[1] + w[2]*z[2];
    break;
  }
  velocity = z[0]*u[0] + z[1]*u[1] + z[2]*u[2];
  /* the stroke offset: a load curve of time at the step's curtime
   * (forces.ispc:915-919) */
  double offset = 0.0;
  int ilc = A.offset ? A.offset[i] : -1;
  if (ilc >= 0)
  {
    double slope;
    offset = look_up (A.lcx, A.lcy, A.lcidx[ilc], A.lcidx[ilc+1], A.ttab[A.ctl[CTL_KSTEP]], &slope);
  }
  int unspr = A.unspring[i];
  double *stroke = A.stroke, *yld = A.yield;
  if (unspr <= -2)
  {
    if (A.type[i] == PMX_SPRING_NONLINEAR_ELASTIC)
    {
      stroke[i] = len - A.stroke0[i];
      spring_force = look_up (A.sx, A.sy, A.spridx[i], A.spridx[i+1], stroke[i] - offset, &stiffness);
    }
    else if ((flg & PMX_SPRING_YIELDED) == 0)
    {
      stroke[i] = len - A.stroke0[i];
      spring_force = look_up (A.sx, A.sy, A.spridx[i], A.spridx[i+1], stroke[i] - offset, &stiffness);
      if (spring_force < yld[i] || spring_force > yld[S + i])
      {
        flg |= PMX_SPRING_YIELDED;
        if (spring_force < 0) { stroke[S + i] = stroke[i]; stroke[2*S + i] = 0.0; yld[S + i] = 0.0; }
        else { stroke[S + i] = 0.0; stroke[2*S + i] = stroke[i]; yld[i] = 0.0; }
      }
    }
    else
    {
      double s0 = stroke[i];
      double s1 = len - A.stroke0[i];
      double ds = s1 - s0;
      double f0 = A.frc[S + i];
      bool loading = false;
      stroke[i] = s1;
      if (f0 * ds > 0)
      {
        if (flg & PMX_SPRING_UNLOADING)
        {
          double dfds = look_up_slope_y (A.ux, A.uy, A.unidx[i], A.unidx[i+1], f0, nullptr);
          spring_force = f0 + dfds*ds;
          stiffness = dfds;
          if (spring_force < yld[i] || spring_force > yld[S + i]) { flg &= ~PMX_SPRING_UNLOADING; loading = true; }
        }
        else loading = true;
      }
      else
      {
        double at;
        double dfds = look_up_slope_y (A.ux, A.uy, A.unidx[i], A.unidx[i+1], f0, &at);
        stiffness = dfds;
        spring_force = f0 + dfds*ds;
        if (spring_force * f0 < 0.0)
          spring_force = look_up (A.ux, A.uy, A.unidx[i], A.unidx[i+1], ds+at, &stiffness);
        if (spring_force < yld[i] || spring_force > yld[S + i]) { flg &= ~PMX_SPRING_UNLOADING; loading = true; }
        else flg |= PMX_SPRING_UNLOADING;
      }
      if (loading)
      {
        if (ds < 0) stroke[S + i] += ds;
        else stroke[2*S + i] += ds;
        spring_force = look_up (A.sx, A.sy, A.spridx[i], A.spridx[i+1], ds < 0 ? stroke[S + i] : stroke[2*S + i], &stiffness);
        if (ds < 0) yld[i] = spring_force;
        else yld[S + i] = spring_force;
      }
    }
    if (A.dashidx[i]+1 == A.dashidx[i+1])
    {
      double mass = eqm (A.N, A.inverse, A.invm, A.position, j, Yj, k, Yk, z);
      double ratio = A.dx[A.dashidx[i]];
      damping = ratio * 2.0 * sqrt (fabs (stiffness) * mass);
      dashpot_force = velocity * damping;
    }
    else dashpot_force = look_up (A.dx, A.dy, A.dashidx[i], A.dashidx[i+1], velocity, &damping);
    total_force = spring_force + (spring_force == 0.0 ? 0.0 : dashpot_force);
  }
  else if (unspr == -1)
  {
    /* an inactive spring (zero force) */
    stroke[i] = len - A.stroke0[i];
    stiffness = 0.;
    damping = 0.;
    spring_force = dashpot_force = total_force = 0.;
  }
  else
  {
    /* switched off by UNSPRING onto load curve unspr: incremental unloading
     * to zero, then zero for ever (forces.ispc:1080-1115) */
    double s0 = stroke[i];
    double s1 = len - A.stroke0[i];
    double ds = s1 - s0;
    double f0 = A.frc[S + i];
    double sgn = f0 > 0. ? 1. : -1.;
    stroke[i] = s1;
    double at;
    double dfds = look_up_slope_y (A.lcx, A.lcy, A.lcidx[unspr], A.lcidx[unspr+1], f0, &at);
    stiffness = dfds;
    spring_force = f0 - sgn*fabs (dfds*ds);
    if (spring_force * f0 < 0.0)
    {
      A.unspring[i] = -1;
      spring_force = 0.;
    }
    damping = 0.;
    dashpot_force = 0.;
    total_force = spring_force;
  }
  A.flg[i] = flg;
#pragma unroll
  for (int c = 0; c < 3; c++) frc[c] = z[c] * total_force;
  A.frc[i] = total_force;
  A.frc[S + i] = spring_force;
  double fric = A.fric[i];
  if (fric > 0.0)
  {
    double us[3] = {u[0] - velocity*z[0], u[1] - velocity*z[1], u[2] - velocity*z[2]};
    double flim = fric*fabs (total_force);
    double fs[3] = {0.0, 0.0, 0.0};
    double kskn = A.kskn[i];
    double *sd = A.sdsp;
    if (kskn == 0.0)
    {
      double ulen = sqrt (us[0]*us[0] + us[1]*us[1] + us[2]*us[2]);
      if (ulen > 0.0)
      {
        double coef = flim/ulen;
        for (int c = 0; c < 3; c++) fs[c] = coef*us[c];
      }
    }
    else
    {
      if (spring_force == 0.0) { sd[i] = 0.0; sd[S + i] = 0.0; sd[2*S + i] = 0.0; }
      double ds[3] = {sd[i] + us[0]*step, sd[S + i] + us[1]*step, sd[2*S + i] + us[2]*step};
      double c0 = kskn * fabs (stiffness);
      double c1 = spring_force == 0.0 ? 0.0 : kskn * fabs (damping);
      for (int c = 0; c < 3; c++) fs[c] = c0 * ds[c] + c1 * us[c];
      double flen = sqrt (fs[0]*fs[0] + fs[1]*fs[1] + fs[2]*fs[2]);
      if (flen > flim)
      {
        double coef = flim/flen;
        for (int c = 0; c < 3; c++) fs[c] *= coef;
        double dl0 = sqrt (ds[0]*ds[0] + ds[1]*ds[1] + ds[2]*ds[2]);
        coef = flim/(c0*dl0);
        for (int c = 0; c < 3; c++) sd[c*S + i] = coef*ds[c];
      }
      else for (int c = 0; c < 3; c++) sd[c*S + i] = ds[c];
    }
    A.frc[2*S + i] = sqrt (fs[0]*fs[0] + fs[1]*fs[1] + fs[2]*fs[2]);
    for (int c = 0; c < 3; c++) frc[c] += fs[c];
  }
  if (A.adp)
  {
    /* the spring's part of the adaptive step's per-particle sums
     * (forces.ispc:1249-1270, 1329-1350), zero where the total force is:
     * |z|, |stiffness|, |damping|, then the rotational stiffness terms of
     * part1's arm a and part2's arm b, both rotated by part1's L^T as the
     * reference does */
    double *d = A.adp;
    double t[17] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (total_force != 0.0)
    {
      t[0] = fabs (z[0]); t[1] = fabs (z[1]); t[2] = fabs (z[2]);
      t[3] = fabs (stiffness); t[4] = fabs (damping);
      double R[3], dot;
      tvmul (Lj, a, R);
      dot = R[0]*R[0] + R[1]*R[1] + R[2]*R[2];
      t[5] = dot-R[0]*R[0]; t[6] = dot-R[1]*R[1]; t[7] = dot-R[2]*R[2];
      t[8] = -R[0]*R[1]; t[9] = -R[0]*R[2]; t[10] = -R[1]*R[2];
      if (k >= 0)
      {
        tvmul (Lj, b, R);
        dot = R[0]*R[0] + R[1]*R[1] + R[2]*R[2];
        t[11] = dot-R[0]*R[0]; t[12] = dot-R[1]*R[1]; t[13] = dot-R[2]*R[2];
        t[14] = -R[0]*R[1]; t[15] = -R[0]*R[2]; t[16] = -R[1]*R[2];
      }
    }
#pragma unroll
    for (int c = 0; c < 17; c++) d[c*S + i] = t[c];
  }
  double *o = A.out;
  o[i] = frc[0]; o[S + i] = frc[1]; o[2*S + i] = frc[2];
  o[3*S + i] = a[1]*frc[2] - a[2]*frc[1];
  o[4*S + i] = a[2]*frc[0] - a[0]*frc[2];
  o[5*S + i] = a[0]*frc[1] - a[1]*frc[0];
  o[6*S + i] = b[1]*frc[2] - b[2]*frc[1];
  o[7*S + i] = b[2]*frc[0] - b[0]*frc[2];
  o[8*S + i] = b[0]*frc[1] - b[1]*frc[0];
}

/* per particle: its run flushes in the reference's order */
__global__ void k_spring_acc (int N, int S, const int *evoff, const int *evrun, const int *runoff, const int *runspr,
                              const double *out, double *force, double *torque, const int *ctl, const int *stop,
                              const int *par_inst)
{
  if (ctl[CTL_ABORT]) return;
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= N || frozen (stop, par_inst[b], ctl)) return;
  int e0 = evoff[b], e1 = evoff[b+1];
  if (e0 == e1) return;
  double f[3] = {force[b], force[(size_t) N + b], force[2*(size_t) N + b]};
  double t[3] = {torque[b], torque[(size_t) N + b], torque[2*(size_t) N + b]};
  for (int ev = e0; ev < e1; ev++)
  {
    int code = evrun[ev];
    int run = code >> 2;          /* run number */
    bool kside = code & 1;        /* part2 run: subtracted, torque from column 6-8 */
    bool zero = code & 2;         /* first run of its side: starts from 0.0 */
    int r0 = runoff[run], r1 = runoff[run+1];
    int tc = kside ? 6 : 3;
    double af[3], at[3];
    int s = runspr[r0];
    if (zero)
    {
      for (int c = 0; c < 3; c++) { af[c] = 0.0 + out[(size_t) c*S + s]; at[c] = 0.0 + out[(size_t) (tc+c)*S + s]; }
    }
    else
    {
      for (int c = 0; c < 3; c++) { af[c] = out[(size_t) c*S + s]; at[c] = out[(size_t) (tc+c)*S + s]; }
    }
    for (int r = r0 + 1; r < r1; r++)
    {
      s = runspr[r];
      for (int c = 0; c < 3; c++) { af[c] += out[(size_t) c*S + s]; at[c] += out[(size_t) (tc+c)*S + s]; }
    }
    if (kside) for (int c = 0; c < 3; c++) { f[c] -= af[c]; t[c] -= at[c]; }
    else for (int c = 0; c < 3; c++) { f[c] += af[c]; t[c] += at[c]; }
  }
  force[b] = f[0]; force[(size_t) N + b] = f[1]; force[2*(size_t) N + b] = f[2];
  torque[b] = t[0]; torque[(size_t) N + b] = t[1]; torque[2*(size_t) N + b] = t[2];
}

/* per particle: the adaptive step's spring terms added run by run in the
 * flush order of k_spring_acc (forces.ispc:1203-1411): each run sums its
 * springs' terms from 0.0 (inactive springs hold +0.0, which changes no
 * such sum), then adds them to the particle's contact sums; kmax and emax
 * take the larger of the run's maximum and the particle's */
__global__ void k_spring_adp (int N, int S, const int *evoff, const int *evrun, const int *runoff, const int *runspr,
                              const double *t, double *adp, const int *ctl)
{
  if (ctl[CTL_ABORT]) return;
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= N) return;
  int e0 = evoff[b], e1 = evoff[b+1];
  if (e0 == e1) return;
  double v[11];
#pragma unroll
  for (int c = 0; c < 11; c++) v[c] = adp[(size_t) c*N + b];
  for (int ev = e0; ev < e1; ev++)
  {
    int code = evrun[ev];
    int run = code >> 2;
    int rc = (code & 1) ? 11 : 5;   /* part2 runs take the arm-b terms */
    double ka[3] = {0.0, 0.0, 0.0}, km = 0.0, em = 0.0, kr[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int r = runoff[run]; r < runoff[run+1]; r++)
    {
      size_t s = runspr[r];
      for (int c = 0; c < 3; c++) ka[c] += t[c*(size_t) S + s];
      double st = t[3*(size_t) S + s], dm = t[4*(size_t) S + s];
      km = st > km ? st : km;
      em = dm > em ? dm : em;
      for (int c = 0; c < 6; c++) kr[c] += t[(rc+c)*(size_t) S + s];
    }
    for (int c = 0; c < 3; c++) v[c] += ka[c];
    v[3] = km > v[3] ? km : v[3];
    v[4] = em > v[4] ? em : v[4];
    for (int c = 0; c < 6; c++) v[5+c] += kr[c];
  }
#pragma unroll
  for (int c = 0; c < 11; c++) adp[(size_t) c*N + b] = v[c];
}

/* ---------------------------------------------- critical step estimates */

/* the largest |slope| of curve points [lo, hi) (parmec.ispc:510-515) */
__device__ __forceinline__ double max_slope (const double *x, const double *y, int lo, int hi)
{
  double m = 0.0;
  for (int q = lo; q < hi-1; q++)
  {
    double s = fabs ((y[q+1]-y[q]) / (x[q+1]-x[q]));
    m = s > m ? s : m;
  }
  return m;
}

/* per spring, the bounds CRITICAL's per-particle estimate adds to both of
 * its particles (critical_perparticle_springs, parmec.ispc:717-750): kspr,
 * the largest |slope| of the spring curve, and espr, 2 ratio sqrt(kspr m)
 * when the dashpot entry is a damping ratio, else the largest |slope| of
 * the spring's dashpot curve.  The reference tests for the ratio form, and
 * takes the ratio, at dashidx[i] with i the spring's first PARTICLE index
 * (parmec.ispc:728, 742); that is reproduced where i < S, and the spring's
 * own entry is used where dashidx[i+1] would lie beyond the array */
__global__ void k_crit_spring_terms (SprArgs A, double *out)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= A.S) return;
  size_t S = A.S;
  int i = A.part[j], k = A.part[S + j];
  double kspr = max_slope (A.sx, A.sy, A.spridx[j], A.spridx[j+1]);
  double espr = 0.0;
  int di = i < A.S ? i : j;
  if (A.dashidx[di]+1 == A.dashidx[di+1])
  {
    double Pi[3], Pk[3], d[3];
    for (int c = 0; c < 3; c++)
    {
      Pi[c] = A.pnt[(size_t) (3+c) * S + j];
      Pk[c] = A.pnt[(size_t) (9+c) * S + j];
      d[c] = A.dir[(size_t) c * S + j];
    }
    double mass = eqm (A.N, A.inverse, A.invm, A.position, i, Pi, k, Pk, d);
    double ratio = A.dx[A.dashidx[di]];
    espr = ratio * 2.0 * sqrt (kspr * mass);
  }
  else espr = max_slope (A.dx, A.dy, A.dashidx[j], A.dashidx[j+1]);
  out[j] = kspr;
  out[S + j] = espr;
}

/* per particle, its springs in spring order (the reference updates each
 * particle under a lock from parallel tasks; here in ascending spring
 * number, side 1 before side 2 of one spring): the summed |direction| (a
 * zero direction adds 1 to each component; a planar one adds |t1| + |t2|
 * of two unit vectors normal to it), the rotational stiffness terms of the
 * arm L^T (point - x), and the larger of the current and the spring's
 * kspr, espr (parmec.ispc:753-857).  The planar test reads sprflg at the
 * PARTICLE index, as the reference does; a particle index >= S reads no
 * flag */
__global__ void k_crit_springs (SprArgs A, const int *pofs, const int *plist, const double *terms, double *adp)
{
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= A.N) return;
  int q0 = pofs[b], q1 = pofs[b+1];
  if (q0 == q1) return;
  size_t S = A.S, N = A.N;
  double v[11], L[9], x[3];
#pragma unroll
  for (int c = 0; c < 11; c++) v[c] = adp[c*N + b];
#pragma unroll
  for (int c = 0; c < 9; c++) L[c] = A.rotation[c*N + b];
#pragma unroll
  for (int c = 0; c < 3; c++) x[c] = A.position[c*N + b];
  int flg = b < A.S ? A.flg[b] : 0;
  for (int q = q0; q < q1; q++)
  {
    int s = plist[q] >> 1, side = plist[q] & 1;
    double d[3] = {A.dir[s], A.dir[S + s], A.dir[2*S + s]};
    if (d[0]*d[0] + d[1]*d[1] + d[2]*d[2] > 0.0)
    {
      if (flg & PMX_SPRDIR_PLANAR)
      {
        double t1[3] = {d[0] + 0.123, d[1] + 0.346, d[2] + 0.789}, t2[3];
        t2[0] = d[1]*t1[2] - d[2]*t1[1];
        t2[1] = d[2]*t1[0] - d[0]*t1[2];
        t2[2] = d[0]*t1[1] - d[1]*t1[0];
        double l2 = sqrt (t2[0]*t2[0] + t2[1]*t2[1] + t2[2]*t2[2]);
        double i2 = 1.0/l2;
        t2[0] *= i2; t2[1] *= i2; t2[2] *= i2;
        t1[0] = t2[1]*d[2] - t2[2]*d[1];
        t1[1] = t2[2]*d[0] - t2[0]*d[2];
        t1[2] = t2[0]*d[1] - t2[1]*d[0];
        for (int c = 0; c < 3; c++) v[c] += fabs (t1[c]) + fabs (t2[c]);
      }
      else for (int c = 0; c < 3; c++) v[c] += fabs (d[c]);
    }
    else for (int c = 0; c < 3; c++) v[c] += 1.0;
    double a[3], R[3];
    for (int c = 0; c < 3; c++) a[c] = A.pnt[(size_t) (6*side+c) * S + s] - x[c];
    tvmul (L, a, R);
    double dot = R[0]*R[0] + R[1]*R[1] + R[2]*R[2];
    v[5] += dot-R[0]*R[0];
    v[6] += dot-R[1]*R[1];
    v[7] += dot-R[2]*R[2];
    v[8] += -R[0]*R[1];
    v[9] += -R[0]*R[2];
    v[10] += -R[1]*R[2];
    double ks = terms[s], es = terms[S + s];
    v[3] = v[3] > ks ? v[3] : ks;
    v[4] = v[4] > es ? v[4] : es;
  }
#pragma unroll
  for (int c = 0; c < 11; c++) adp[c*N + b] = v[c];
}

/* critical_perspring (parmec.ispc:485-535): each spring as a single degree
 * of freedom of its equivalent mass along its direction, stiffness the
 * largest |slope| of its curve; h = 2/w (sqrt(1+x^2) - x), or 1 where w is
 * not positive */
__global__ void k_crit_perspring (SprArgs A, double *h, double *o, double *r)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= A.S) return;
  size_t S = A.S;
  int i = A.part[j], k = A.part[S + j];
  double Pi[3], Pk[3], d[3];
  for (int c = 0; c < 3; c++)
  {
    Pi[c] = A.pnt[(size_t) (3+c) * S + j];
    Pk[c] = A.pnt[(size_t) (9+c) * S + j];
    d[c] = A.dir[(size_t) c * S + j];
  }
  double kspr = max_slope (A.sx, A.sy, A.spridx[j], A.spridx[j+1]);
  double mspr = eqm (A.N, A.inverse, A.invm, A.position, i, Pi, k, Pk, d);
  double omega = sqrt (kspr/mspr), ksi = 0.0;
  if (A.dashidx[j]+1 == A.dashidx[j+1]) ksi = A.dx[A.dashidx[j]];
  else
  {
    double espr = max_slope (A.dx, A.dy, A.dashidx[j], A.dashidx[j+1]);
    if (omega > 0.0) ksi = espr / (2*mspr*omega);
  }
  h[j] = omega > 0.0 ? (2.0/omega)*(sqrt (1.0+ksi*ksi) - ksi) : 1.0;
  o[j] = omega;
  r[j] = ksi;
}

/* nodes_task, shapes.ispc:101-186 */
__global__ void k_nodes (int NOD, int N, const int *nodpart, double *nodes, const double *rotation, const double *position,
                         const int *ctl, const int *stop, const int *par_inst, const int *list, int nlist)
{
  if (ctl && ctl[CTL_ABORT]) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlist) return;
  if (list) i = list[i];
  int j = nodpart[i];
  if (ctl && frozen (stop, par_inst[j], ctl)) return;
  double L[9], C[3];
#pragma unroll
  for (int k = 0; k < 9; k++) L[k] = rotation[(size_t) k*N + j];
#pragma unroll
  for (int k = 0; k < 3; k++) C[k] = nodes[(size_t) (3+k)*NOD + i] - position[(size_t) (3+k)*N + j];
  nodes[i] = position[j] + (L[0]*C[0] + L[3]*C[1] + L[6]*C[2]);
  nodes[(size_t) NOD + i] = position[(size_t) N + j] + (L[1]*C[0] + L[4]*C[1] + L[7]*C[2]);
  nodes[2*(size_t) NOD + i] = position[2*(size_t) N + j] + (L[2]*C[0] + L[5]*C[1] + L[8]*C[2]);
}

/* faces_task, shapes.ispc:189-211 */
__global__ void k_faces (int FAC, int NOD, int T, const int *facnod, const int *factri, const double *nodes, double *tri,
                         const int *ctl, const int *stop, const int *tri_inst, const int *list, int nlist)
{
  if (ctl && ctl[CTL_ABORT]) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlist) return;
  if (list) i = list[i];
  int t = factri[i];
  if (ctl && frozen (stop, tri_inst[t], ctl)) return;
#pragma unroll
  for (int v = 0; v < 3; v++)
  {
    int k = facnod[(size_t) v*FAC + i];
#pragma unroll
    for (int c = 0; c < 3; c++) tri[(size_t) (3*v+c)*T + t] = nodes[(size_t) c*NOD + k];
  }
}

/* ------------------------------------------------------- torsion springs */

struct TsArgs
{
  int T, N;
  const int *part, *kidx, *didx, *cone;   /* kidx, didx: [3][T+1] into krpy, drpy */
  const double *z0, *x0, *kx, *ky, *dx, *dy;
  double *z1, *x1, *rpy, *tot, *spr;
  double *out;                            /* [6][T]: torque, adaptive terms */
  const double *angular, *rotation, *inverse;
  const int *ctl;
};

/* the torque of a spring curve of a torsion spring, or zero torque and
 * stiffness when the curve is empty (forces.ispc:1625-1637) */
__device__ __forceinline__ double trq_curve (const TsArgs &A, int o, int i, double arg, double *k)
{
  int lo = A.kidx[o*(A.T+1) + i], hi = A.kidx[o*(A.T+1) + i + 1];
  if (hi > lo) return look_up (A.kx, A.ky, lo, hi, arg, k);
  *k = 0.0;
  return 0.0;
}

/* the damper of axis o: a critical damping ratio (one table point) against
 * the local inertia `I`, or a curve of the angular velocity
 * (forces.ispc:1641-1675); `ratio_axis` is the table whose first point is
 * read as the ratio (the pitch of the pitch-yaw cone reads roll's, as the
 * reference does) */
__device__ __forceinline__ double trq_damper (const TsArgs &A, int o, int ratio_axis, int i, double k, double I,
                                              double vel, double *d)
{
  int lo = A.didx[o*(A.T+1) + i], hi = A.didx[o*(A.T+1) + i + 1];
  if (lo + 1 == hi)
  {
    double ratio = A.dx[A.didx[ratio_axis*(A.T+1) + i]];
    *d = - ratio * 2.0 * sqrt (fabs (k) * I);
    return vel * *d;
  }
  return look_up (A.dx, A.dy, lo, hi, vel, d);
}

/* trqspr_task, forces.ispc:1415-2055, one thread per torsion spring: the
 * current axes (part1's rotation of the reference axes), the roll, pitch
 * and yaw rates of the relative angular velocity and their integrals, the
 * curve torques with dampers (per axis, or on the cone of two or three axes
 * when `cone` joins them), and the global torque; the flush into the
 * particles' torques follows in k_torsion_acc */
__global__ void k_torsion (TsArgs A, double step)
{
  if (A.ctl[CTL_ABORT]) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.T) return;
  size_t T = A.T, N = A.N;
  int j = A.part[i], k = A.part[T + i];
  double oj[3], Lj[9], ok[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < 3; c++) oj[c] = A.angular[(3+c)*N + j];
#pragma unroll
  for (int c = 0; c < 9; c++) Lj[c] = A.rotation[c*N + j];
  double zdir0[3], xdir0[3], ydir0[3], rot1j[9];
#pragma unroll
  for (int c = 0; c < 3; c++) { zdir0[c] = A.z0[c*T + i]; xdir0[c] = A.x0[c*T + i]; }
  ydir0[0] = zdir0[1]*xdir0[2] - zdir0[2]*xdir0[1];
  ydir0[1] = zdir0[2]*xdir0[0] - zdir0[0]*xdir0[2];
  ydir0[2] = zdir0[0]*xdir0[1] - zdir0[1]*xdir0[0];
  nvmul (Lj, xdir0, rot1j);
  nvmul (Lj, ydir0, rot1j+3);
  nvmul (Lj, zdir0, rot1j+6);
#pragma unroll
  for (int c = 0; c < 3; c++) { A.z1[c*T + i] = rot1j[6+c]; A.x1[c*T + i] = rot1j[c]; }
  /* part2's angular velocity; zero for a spring to the ground, and also
   * where the reference's loop would still hold the zeros of such a spring
   * (it reloads part2's velocity only when part2 changes; bit 8 of cone) */
  if (k >= 0 && !(A.cone[i] & 256))
  {
#pragma unroll
    for (int c = 0; c < 3; c++) ok[c] = A.angular[(3+c)*N + k];
  }
  double ojk[3] = {oj[0]-ok[0], oj[1]-ok[1], oj[2]-ok[2]};
  double rolldot = rot1j[0]*ojk[0] + rot1j[1]*ojk[1] + rot1j[2]*ojk[2];
  double pitchdot = rot1j[3]*ojk[0] + rot1j[4]*ojk[1] + rot1j[5]*ojk[2];
  double yawdot = rot1j[6]*ojk[0] + rot1j[7]*ojk[1] + rot1j[8]*ojk[2];
  double roll = A.rpy[i], pitch = A.rpy[T + i], yaw = A.rpy[2*T + i];
  roll += rolldot*step;
  pitch += pitchdot*step;
  yaw += yawdot*step;
  A.rpy[i] = roll; A.rpy[T + i] = pitch; A.rpy[2*T + i] = yaw;
  /* local equivalent inertia inv (R (inverse_j + inverse_k) R') where a
   * damper is a critical damping ratio */
  double Ixx = 0.0, Iyy = 0.0, Izz = 0.0;
  const int *D = A.didx;
  if (D[i]+1 == D[i+1] || D[(T+1) + i]+1 == D[(T+1) + i+1] || D[2*(T+1) + i]+1 == D[2*(T+1) + i+1])
  {
    double I[9], tmp[9];
#pragma unroll
    for (int c = 0; c < 9; c++) I[c] = A.inverse[c*N + j];
    if (k >= 0)
#pragma unroll
      for (int c = 0; c < 9; c++) I[c] += A.inverse[c*N + k];
    ntmul (I, rot1j, tmp);
    nnmul (rot1j, tmp, I);
    invert3 (I, tmp);
    Ixx = tmp[0]; Iyy = tmp[4]; Izz = tmp[8];
  }
  double kroll = 0.0, kpitch = 0.0, kyaw = 0.0, droll, dpitch, dyaw;
  double trqroll, trqpitch, trqyaw, trqrolldot, trqpitchdot, trqyawdot;
  double trqrolltot = 0.0, trqpitchtot = 0.0, trqyawtot = 0.0;
  switch (A.cone[i] & 255)
  {
  case 0:
    trqroll = trq_curve (A, 0, i, roll, &kroll);
    trqpitch = trq_curve (A, 1, i, pitch, &kpitch);
    trqyaw = trq_curve (A, 2, i, yaw, &kyaw);
    trqrolldot = trq_damper (A, 0, 0, i, kroll, Ixx, rolldot, &droll);
    trqpitchdot = trq_damper (A, 1, 1, i, kpitch, Iyy, pitchdot, &dpitch);
    trqyawdot = trq_damper (A, 2, 2, i, kyaw, Izz, yawdot, &dyaw);
    trqrolltot = trqroll + (trqroll == 0.0 ? 0.0 : trqrolldot);
    trqpitchtot = trqpitch + (trqpitch == 0.0 ? 0.0 : trqpitchdot);
    trqyawtot = trqyaw + (trqyaw == 0.0 ? 0.0 : trqyawdot);
    break;
  case 3:   /* roll and pitch on one cone */
  {
    double angle = sqrt (roll*roll + pitch*pitch);
    double angledot = angle > 0.0 ? (rolldot*roll+pitchdot*pitch)/angle : 0.0;
    trqroll = trq_curve (A, 0, i, angle, &kroll);
    kpitch = kroll;
    trqrolldot = trq_damper (A, 0, 0, i, kroll, sqrt (Ixx*Ixx + Iyy*Iyy), angledot, &droll);
    trqyaw = trq_curve (A, 2, i, yaw, &kyaw);
    trqyawdot = trq_damper (A, 2, 2, i, kyaw, Izz, yawdot, &dyaw);
    trqyawtot = trqyaw + (trqyaw == 0.0 ? 0.0 : trqyawdot);
    if (angle > 0.)
    {
      double xcoef = roll/angle, ycoef = pitch/angle;
      trqrolltot = xcoef*(trqroll + (trqroll == 0.0 ? 0.0 : trqrolldot));
      trqpitchtot = ycoef*(trqroll + (trqroll == 0.0 ? 0.0 : trqrolldot));
      trqpitch = trqroll*ycoef;
      trqroll *= xcoef;
    }
    else { trqrolltot = trqpitchtot = 0.0; trqroll = trqpitch = 0.0; }
    break;
  }
  case 5:   /* roll and yaw */
  {
    double angle = sqrt (roll*roll + yaw*yaw);
    double angledot = angle > 0.0 ? (rolldot*roll+yawdot*yaw)/angle : 0.0;
    trqroll = trq_curve (A, 0, i, angle, &kroll);
    kyaw = kroll;
    trqrolldot = trq_damper (A, 0, 0, i, kroll, sqrt (Ixx*Ixx + Izz*Izz), angledot, &droll);
    trqpitch = trq_curve (A, 1, i, pitch, &kpitch);
    trqpitchdot = trq_damper (A, 1, 1, i, kpitch, Iyy, pitchdot, &dpitch);
    trqpitchtot = trqpitch + (trqpitch == 0.0 ? 0.0 : trqpitchdot);
    if (angle > 0.)
    {
      double xcoef = roll/angle, zcoef = yaw/angle;
      trqrolltot = xcoef*(trqroll + (trqroll == 0.0 ? 0.0 : trqrolldot));
      trqyawtot = zcoef*(trqroll + (trqroll == 0.0 ? 0.0 : trqrolldot));
      trqyaw = trqroll*zcoef;
      trqroll *= xcoef;
    }
    else { trqrolltot = trqyawtot = 0.0; trqroll = trqyaw = 0.0; }
    break;
  }
  case 6:   /* pitch and yaw */
  {
    trqroll = trq_curve (A, 0, i, roll, &kroll);
    trqrolldot = trq_damper (A, 0, 0, i, kroll, Ixx, rolldot, &droll);
    double angle = sqrt (pitch*pitch + yaw*yaw);
    double angledot = angle > 0.0 ? (pitchdot*pitch+yawdot*yaw)/angle : 0.0;
    trqpitch = trq_curve (A, 1, i, angle, &kpitch);
    kyaw = kpitch;
    trqpitchdot = trq_damper (A, 1, 0, i, kpitch, sqrt (Iyy*Iyy + Izz*Izz), angledot, &dpitch);
    trqrolltot = trqroll + (trqroll == 0.0 ? 0.0 : trqrolldot);
    if (angle > 0.)
    {
      double ycoef = pitch/angle, zcoef = yaw/angle;
      trqpitchtot = ycoef*(trqpitch + (trqpitch == 0.0 ? 0.0 : trqpitchdot));
      trqyawtot = zcoef*(trqpitch + (trqpitch == 0.0 ? 0.0 : trqpitchdot));
      trqyaw = zcoef*trqpitch;
      trqpitch *= ycoef;
    }
    else { trqpitchtot = trqyawtot = 0.0; trqpitch = trqyaw = 0.0; }
    break;
  }
  default:  /* 7: roll, pitch and yaw */
  {
    double angle = sqrt (roll*roll + pitch*pitch + yaw*yaw);
    double angledot = angle > 0.0 ? (rolldot*roll+pitchdot*pitch+yawdot*yaw)/angle : 0.0;
    trqroll = trq_curve (A, 0, i, angle, &kroll);
    kpitch = kyaw = kroll;
    trqrolldot = trq_damper (A, 0, 0, i, kroll, sqrt (Ixx*Ixx + Iyy*Iyy + Izz*Izz), angledot, &droll);
    if (angle > 0.)
    {
      double xcoef = roll/angle, ycoef = pitch/angle, zcoef = yaw/angle;
      trqrolltot = xcoef*(trqroll + (trqroll == 0.0 ? 0.0 : trqrolldot));
      trqpitchtot = ycoef*(trqroll + (trqroll == 0.0 ? 0.0 : trqrolldot));
      trqyawtot = zcoef*(trqroll + (trqroll == 0.0 ? 0.0 : trqrolldot));
      trqpitch = ycoef*trqroll;
      trqyaw = zcoef*trqroll;
      trqroll *= xcoef;
    }
    else { trqrolltot = trqpitchtot = trqyawtot = 0.0; trqroll = trqpitch = trqyaw = 0.0; }
    break;
  }
  }
  A.tot[i] = trqrolltot; A.tot[T + i] = trqpitchtot; A.tot[2*T + i] = trqyawtot;
  A.spr[i] = trqroll; A.spr[T + i] = trqpitch; A.spr[2*T + i] = trqyaw;
  double *o = A.out;
  o[i] = trqrolltot*rot1j[0] + trqpitchtot*rot1j[3] + trqyawtot*rot1j[6];
  o[T + i] = trqrolltot*rot1j[1] + trqpitchtot*rot1j[4] + trqyawtot*rot1j[7];
  o[2*T + i] = trqrolltot*rot1j[2] + trqpitchtot*rot1j[5] + trqyawtot*rot1j[8];
  /* the adaptive step's rotational stiffness terms K(1,1), K(2,2), K(3,3)
   * (forces.ispc:1968-1977) */
  o[3*T + i] = fabs (kroll)*rot1j[0];
  o[4*T + i] = fabs (kpitch)*rot1j[4];
  o[5*T + i] = fabs (kyaw)*rot1j[8];
}

/* per particle: the torsion springs' torques in the flush order of the
 * single-task loop (forces.ispc:1942-2054), as k_spring_acc for forces */
__global__ void k_torsion_acc (int N, int T, const int *evoff, const int *evrun, const int *runoff, const int *runspr,
                               const double *out, double *torque, const int *ctl)
{
  if (ctl[CTL_ABORT]) return;
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= N) return;
  int e0 = evoff[b], e1 = evoff[b+1];
  if (e0 == e1) return;
  double t[3] = {torque[b], torque[(size_t) N + b], torque[2*(size_t) N + b]};
  for (int ev = e0; ev < e1; ev++)
  {
    int code = evrun[ev];
    int run = code >> 2;
    int r0 = runoff[run], r1 = runoff[run+1];
    double at[3];
    int s = runspr[r0];
    if (code & 2) for (int c = 0; c < 3; c++) at[c] = 0.0 + out[(size_t) c*T + s];
    else for (int c = 0; c < 3; c++) at[c] = out[(size_t) c*T + s];
    for (int r = r0 + 1; r < r1; r++)
    {
      s = runspr[r];
      for (int c = 0; c < 3; c++) at[c] += out[(size_t) c*T + s];
    }
    if (code & 1) for (int c = 0; c < 3; c++) t[c] -= at[c];
    else for (int c = 0; c < 3; c++) t[c] += at[c];
  }
  torque[b] = t[0]; torque[(size_t) N + b] = t[1]; torque[2*(size_t) N + b] = t[2];
}

/* per particle: the torsion springs' rotational stiffness terms added run
 * by run (each run summed from 0.0; the off-diagonal terms add 0.0, as the
 * reference's zero buffers do) */
__global__ void k_torsion_adp (int N, int T, const int *evoff, const int *evrun, const int *runoff, const int *runspr,
                               const double *out, double *adp, const int *ctl)
{
  if (ctl[CTL_ABORT]) return;
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= N) return;
  int e0 = evoff[b], e1 = evoff[b+1];
  if (e0 == e1) return;
  double v[6];
#pragma unroll
  for (int c = 0; c < 6; c++) v[c] = adp[(size_t) (5+c)*N + b];
  for (int ev = e0; ev < e1; ev++)
  {
    int run = evrun[ev] >> 2;
    double kr[3] = {0.0, 0.0, 0.0};
    for (int r = runoff[run]; r < runoff[run+1]; r++)
    {
      size_t s = runspr[r];
      for (int c = 0; c < 3; c++) kr[c] += out[(3+c)*(size_t) T + s];
    }
    for (int c = 0; c < 3; c++) v[c] += kr[c];
    for (int c = 3; c < 6; c++) v[c] += 0.0;
  }
#pragma unroll
  for (int c = 0; c < 6; c++) adp[(size_t) (5+c)*N + b] = v[c];
}

/* unspring_task, forces.ispc:2058-2127, one thread per condition: on the
 * dem calls whose number is a multiple of nsteps, the sum, minimum or
 * maximum of an entity over the tested springs; outside the limits the
 * failure count falls, and at zero the condition switches its springs off
 * (to zero force or onto an unloading curve) and activates its inactive
 * ones.  min and max as ISPC's (a < b ? a : b, a > b ? a : b) */
__global__ void k_unspring (int U, int S, int stepnum, const int *tidx, const int *tspr, const int *midx, const int *mspr,
                            const int *aidx, const int *aspr, const int *ent, const int *op, const int *uabs,
                            const int *nsteps, int *nfreq, const int *action, const double *lim, const double *stroke,
                            const double *frc, int *unspring, const int *ctl)
{
  if (ctl[CTL_ABORT]) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= U) return;
  if (stepnum % nsteps[i]) return;
  if (nfreq[i] == 0) return;
  double result = op[i] == PMX_OP_MIN ? 1.7976931348623157e308 : op[i] == PMX_OP_MAX ? -1.7976931348623157e308 : 0.;
  for (int q = tidx[i]; q < tidx[i+1]; q++)
  {
    int s = tspr[q];
    double value;
    switch (ent[i])
    {
    case PMX_HIS_STROKE: value = stroke[s]; break;
    case PMX_HIS_F: value = frc[s]; break;
    case PMX_HIS_FF: value = frc[2*(size_t) S + s]; break;
    default: value = frc[(size_t) S + s]; break;
    }
    if (uabs[i]) value = fabs (value);
    if (op[i] == PMX_OP_MIN) result = value < result ? value : result;
    else if (op[i] == PMX_OP_MAX) result = value > result ? value : result;
    else result += value;
  }
  if (result < lim[i] || result > lim[U + i])
  {
    int f = --nfreq[i];
    if (f == 0)
    {
      for (int q = midx[i]; q < midx[i+1]; q++) unspring[mspr[q]] = action[i];
      for (int q = aidx[i]; q < aidx[i+1]; q++) unspring[aspr[q]] = -3;
    }
  }
}

} /* namespace */

int springs_setup (pmx_ctx &c, const pmx_model *m)
{
  int S = m->sprnum;
  c.S = S;
  if (S == 0) return PMX_OK;
  /* run table of the single-task flush order (forces.ispc:1203-1411); in a
   * batch every instance's springs form their own sequence */
  struct Ev { int pos, sub, body, code; };
  std::vector<Ev> ev;
  std::vector<int> runoff (1, 0), runspr;
  int nrun = 0;
  const int *p1 = m->sprpart, *p2 = m->sprpart + S;
  for (int inst = 0; inst < c.I; inst++)
  {
    int s0 = c.off[inst].spr, s1 = c.off[inst+1].spr;
    bool firstj = true, firstk = true;
    for (int i = s0; i < s1; )
    {
      int e = i;
      while (e < s1 && p1[e] == p1[i]) e++;
      for (int q = i; q < e; q++) runspr.push_back (q);
      runoff.push_back ((int) runspr.size ());
      ev.push_back ({e, 0, p1[i], (nrun << 2) | (firstj ? 2 : 0)});
      firstj = false;
      nrun++;
      i = e;
    }
    int cur = -1;
    for (int i = s0; i <= s1; i++)
    {
      if (i < s1 && p2[i] < 0) continue;
      int k = i < s1 ? p2[i] : -2;
      if (cur >= 0 && k != cur)
      {
        runoff.push_back ((int) runspr.size ());
        ev.push_back ({i, 1, cur, (nrun << 2) | 1 | (firstk ? 2 : 0)});
        firstk = false;
        nrun++;
      }
      if (i < s1) { runspr.push_back (i); cur = k; }
    }
  }
  std::stable_sort (ev.begin (), ev.end (), [] (const Ev &a, const Ev &b) {
    return a.body < b.body || (a.body == b.body && (a.pos < b.pos || (a.pos == b.pos && a.sub < b.sub))); });
  std::vector<int> evoff (c.N + 1, 0), evrun (ev.size ());
  for (auto &x : ev) evoff[x.body + 1]++;
  for (int b = 0; b < c.N; b++) evoff[b+1] += evoff[b];
  for (size_t q = 0; q < ev.size (); q++) evrun[q] = ev[q].code;
  /* each particle's springs in ascending spring number, (spring << 1) | side
   * (CRITICAL's per-particle spring part) */
  std::vector<int> pofs (c.N + 1, 0), plist;
  for (int q = 0; q < S; q++) { pofs[p1[q] + 1]++; if (p2[q] >= 0) pofs[p2[q] + 1]++; }
  for (int b = 0; b < c.N; b++) pofs[b+1] += pofs[b];
  plist.resize (pofs[c.N]);
  {
    std::vector<int> at (pofs.begin (), pofs.end () - 1);
    for (int q = 0; q < S; q++)
    {
      plist[at[p1[q]]++] = q << 1;
      if (p2[q] >= 0) plist[at[p2[q]]++] = (q << 1) | 1;
    }
  }
  auto upi = [&] (dvec<int> &d, const int *src, size_t n) -> int {
    if (d.alloc (std::max (n, (size_t) 1)) != hipSuccess) { pmx_set_error ("out of device memory (springs)"); return PMX_ENOMEM; }
    if (n) PMX_HIP (hipMemcpyAsync (d.p, src, n * sizeof (int), hipMemcpyHostToDevice, c.stream));
    PMX_HIP (hipStreamSynchronize (c.stream));   /* the source may be freed once this returns */
    return PMX_OK;
  };
  auto upd = [&] (dvec<double> &d, const double *src, size_t n) -> int {
    if (d.alloc (std::max (n, (size_t) 1)) != hipSuccess) { pmx_set_error ("out of device memory (springs)"); return PMX_ENOMEM; }
    if (n) PMX_HIP (hipMemcpyAsync (d.p, src, n * sizeof (double), hipMemcpyHostToDevice, c.stream));
    PMX_HIP (hipStreamSynchronize (c.stream));
    return PMX_OK;
  };
  int rc;
  size_t SS = S;
  if ((rc = upi (c.sp_evoff, evoff.data (), evoff.size ())) || (rc = upi (c.sp_evrun, evrun.data (), evrun.size ())) ||
      (rc = upi (c.sp_runoff, runoff.data (), runoff.size ())) || (rc = upi (c.sp_runspr, runspr.data (), runspr.size ())) ||
      (rc = upi (c.sp_pofs, pofs.data (), pofs.size ())) || (rc = upi (c.sp_plist, plist.data (), plist.size ())) ||
      (rc = upi (c.sp_type, m->sprtype, SS)) || (rc = upi (c.sp_part, m->sprpart, 2*SS)) ||
      (rc = upi (c.sp_spridx, m->spridx, SS+1)) || (rc = upi (c.sp_dashidx, m->dashidx, SS+1)) ||
      (rc = upi (c.sp_unidx, m->unidx, SS+1)) || (rc = upi (c.sp_unspring, m->unspring, SS)) ||
      (rc = upi (c.sp_flg, m->sprflg, SS)) || (rc = upd (c.sp_pnt, m->sprpnt, 12*SS)) ||
      (rc = upd (c.sp_dir, m->sprdir, 6*SS)) || (rc = upd (c.sp_yield, m->yield, 2*SS)) ||
      (rc = upd (c.sp_sdsp, m->sprsdsp, 3*SS)) || (rc = upd (c.sp_stroke, m->stroke, 3*SS)) ||
      (rc = upd (c.sp_frc, m->sprfrc, 3*SS)) || (rc = upd (c.sp_stroke0, m->stroke0, SS)) ||
      (rc = upd (c.sp_fric, m->sprfric, SS)) || (rc = upd (c.sp_kskn, m->sprkskn, SS)) ||
      (rc = upd (c.sp_curve, m->spring, 2*(size_t) m->sprcrv)) || (rc = upd (c.sp_dash, m->dashpot, 2*(size_t) m->dashcrv)) ||
      (rc = upd (c.sp_unload, m->unload, 2*(size_t) m->uncrv)))
    return rc;
  c.sp_offsets = false;
  if (m->sproffset)
  {
    for (int q = 0; q < S; q++) c.sp_offsets = c.sp_offsets || m->sproffset[q] >= 0;
    if (c.sp_offsets && (rc = upi (c.sp_offset, m->sproffset, SS))) return rc;
  }
  c.sprcrv = m->sprcrv; c.dashcrv = m->dashcrv; c.uncrv = m->uncrv;
  if (c.sp_out.alloc (9*SS) != hipSuccess) { pmx_set_error ("out of device memory (springs)"); return PMX_ENOMEM; }
  return PMX_OK;
}

int springs_download (pmx_ctx &c, pmx_model *m)
{
  size_t S = c.S;
  if (S == 0) return PMX_OK;
  auto dn = [&] (double *dst, const dvec<double> &d, size_t n) -> int {
    if (dst && n) PMX_HIP (hipMemcpyAsync (dst, d.p, n * sizeof (double), hipMemcpyDeviceToHost, c.stream));
    return PMX_OK;
  };
  int rc;
  if ((rc = dn (m->sprpnt, c.sp_pnt, 12*S)) || (rc = dn (m->sprdir, c.sp_dir, 6*S)) || (rc = dn (m->yield, c.sp_yield, 2*S)) ||
      (rc = dn (m->sprsdsp, c.sp_sdsp, 3*S)) || (rc = dn (m->stroke, c.sp_stroke, 3*S)) || (rc = dn (m->sprfrc, c.sp_frc, 3*S)))
    return rc;
  if (m->sprflg) PMX_HIP (hipMemcpyAsync (m->sprflg, c.sp_flg.p, S * sizeof (int), hipMemcpyDeviceToHost, c.stream));
  /* UNSPRING and load-curve unloading change a spring's state */
  if (m->unspring) PMX_HIP (hipMemcpyAsync (m->unspring, c.sp_unspring.p, S * sizeof (int), hipMemcpyDeviceToHost, c.stream));
  return PMX_OK;
}

/* the flush order of the single-task loop over springs sorted by
 * (part1, part2) (forces.ispc:1203-1411, 1942-2054): runs of consecutive
 * springs sharing part1 (added) and, among those with a part2, sharing
 * part2 (subtracted), each flushed where it ends; per particle its flushes
 * in loop order.  evrun: (run << 2) | (first run of its side ? 2 : 0) |
 * (part2 run ? 1 : 0) */
static void flush_runs (int N, int n, const int *p1, const int *p2, std::vector<int> &evoff, std::vector<int> &evrun,
                        std::vector<int> &runoff, std::vector<int> &runspr)
{
  struct Ev { int pos, sub, body, code; };
  std::vector<Ev> ev;
  runoff.assign (1, 0);
  runspr.clear ();
  int nrun = 0;
  bool firstj = true, firstk = true;
  for (int i = 0; i < n; )
  {
    int e = i;
    while (e < n && p1[e] == p1[i]) e++;
    for (int q = i; q < e; q++) runspr.push_back (q);
    runoff.push_back ((int) runspr.size ());
    ev.push_back ({e, 0, p1[i], (nrun << 2) | (firstj ? 2 : 0)});
    firstj = false;
    nrun++;
    i = e;
  }
  int cur = -1;
  for (int i = 0; i <= n; i++)
  {
    if (i < n && p2[i] < 0) continue;
    int k = i < n ? p2[i] : -2;
    if (cur >= 0 && k != cur)
    {
      runoff.push_back ((int) runspr.size ());
      ev.push_back ({i, 1, cur, (nrun << 2) | 1 | (firstk ? 2 : 0)});
      firstk = false;
      nrun++;
    }
    if (i < n) { runspr.push_back (i); cur = k; }
  }
  std::stable_sort (ev.begin (), ev.end (), [] (const Ev &a, const Ev &b) {
    return a.body < b.body || (a.body == b.body && (a.pos < b.pos || (a.pos == b.pos && a.sub < b.sub))); });
  evoff.assign (N + 1, 0);
  evrun.resize (ev.size ());
  for (auto &x : ev) evoff[x.body + 1]++;
  for (int b = 0; b < N; b++) evoff[b+1] += evoff[b];
  for (size_t q = 0; q < ev.size (); q++) evrun[q] = ev[q].code;
}

int torsion_setup (pmx_ctx &c, const pmx_model *m)
{
  auto upi = [&] (dvec<int> &d, const int *src, size_t n) -> int {
    if (d.alloc (std::max (n, (size_t) 1)) != hipSuccess) { pmx_set_error ("out of device memory (torsion springs)"); return PMX_ENOMEM; }
    if (n) PMX_HIP (hipMemcpyAsync (d.p, src, n * sizeof (int), hipMemcpyHostToDevice, c.stream));
    PMX_HIP (hipStreamSynchronize (c.stream));   /* the source may be freed once this returns */
    return PMX_OK;
  };
  auto upd = [&] (dvec<double> &d, const double *src, size_t n) -> int {
    if (d.alloc (std::max (n, (size_t) 1)) != hipSuccess) { pmx_set_error ("out of device memory (torsion springs)"); return PMX_ENOMEM; }
    if (n) PMX_HIP (hipMemcpyAsync (d.p, src, n * sizeof (double), hipMemcpyHostToDevice, c.stream));
    PMX_HIP (hipStreamSynchronize (c.stream));
    return PMX_OK;
  };
  int rc;
  int T = c.TS = m->trqsprnum;
  if (T > 0)
  {
    size_t TT = T;
    std::vector<int> evoff, evrun, runoff, runspr;
    flush_runs (c.N, T, m->trqsprpart, m->trqsprpart + T, evoff, evrun, runoff, runspr);
    /* trqspr_task reloads part2's angular velocity only where part2 differs
     * from the last spring's with a part2, and zeroes it for a spring to
     * the ground (forces.ispc:1500-1526): a spring whose part2 repeats after
     * such a spring reads zeros; bit 8 of its cone word marks it */
    std::vector<int> cone (m->trqcone, m->trqcone + T);
    {
      int k0 = -1;
      bool zero = false;
      for (int i = 0; i < T; i++)
      {
        int k = m->trqsprpart[T + i];
        if (k >= 0)
        {
          if (k != k0) zero = false;
          if (zero) cone[i] |= 256;
          k0 = k;
        }
        else zero = true;
      }
    }
    c.krpycrv = m->krpycrv; c.drpycrv = m->drpycrv;
    if ((rc = upi (c.ts_part, m->trqsprpart, 2*TT)) || (rc = upi (c.ts_kidx, m->krpyidx, 3*(TT+1))) ||
        (rc = upi (c.ts_didx, m->drpyidx, 3*(TT+1))) || (rc = upi (c.ts_cone, cone.data (), TT)) ||
        (rc = upd (c.ts_z0, m->trqzdir0, 3*TT)) || (rc = upd (c.ts_x0, m->trqxdir0, 3*TT)) ||
        (rc = upd (c.ts_krpy, m->krpy, 2*(size_t) m->krpycrv)) || (rc = upd (c.ts_drpy, m->drpy, 2*(size_t) m->drpycrv)) ||
        (rc = upd (c.ts_z1, m->trqzdir1, 3*TT)) || (rc = upd (c.ts_x1, m->trqxdir1, 3*TT)) ||
        (rc = upd (c.ts_rpy, m->trqrpy, 3*TT)) || (rc = upd (c.ts_tot, m->trqrpytot, 3*TT)) ||
        (rc = upd (c.ts_spr, m->trqrpyspr, 3*TT)) ||
        (rc = upi (c.ts_evoff, evoff.data (), evoff.size ())) || (rc = upi (c.ts_evrun, evrun.data (), evrun.size ())) ||
        (rc = upi (c.ts_runoff, runoff.data (), runoff.size ())) || (rc = upi (c.ts_runspr, runspr.data (), runspr.size ())))
      return rc;
    if (c.ts_out.alloc (6*TT) != hipSuccess) { pmx_set_error ("out of device memory (torsion springs)"); return PMX_ENOMEM; }
  }
  int U = c.US = m->unsprnum;
  if (U > 0)
  {
    size_t UU = U;
    if ((rc = upi (c.un_tidx, m->tspridx, UU+1)) || (rc = upi (c.un_t, m->tsprings, m->tspridx[U])) ||
        (rc = upi (c.un_midx, m->mspridx, UU+1)) || (rc = upi (c.un_m, m->msprings, m->mspridx[U])) ||
        (rc = upi (c.un_aidx, m->actidx, UU+1)) || (rc = upi (c.un_a, m->activate, m->actidx[U])) ||
        (rc = upi (c.un_ent, m->unent, UU)) || (rc = upi (c.un_op, m->unop, UU)) || (rc = upi (c.un_abs, m->unabs, UU)) ||
        (rc = upi (c.un_nsteps, m->nsteps, UU)) || (rc = upi (c.un_nfreq, m->nfreq, UU)) ||
        (rc = upi (c.un_action, m->unaction, UU)) || (rc = upd (c.un_lim, m->unlim, 2*UU)))
      return rc;
  }
  return PMX_OK;
}

int torsion_download (pmx_ctx &c, pmx_model *m)
{
  size_t T = c.TS;
  auto dn = [&] (double *dst, const dvec<double> &d, size_t n) -> int {
    if (dst && n) PMX_HIP (hipMemcpyAsync (dst, d.p, n * sizeof (double), hipMemcpyDeviceToHost, c.stream));
    return PMX_OK;
  };
  int rc;
  if (T && ((rc = dn (m->trqzdir1, c.ts_z1, 3*T)) || (rc = dn (m->trqxdir1, c.ts_x1, 3*T)) ||
            (rc = dn (m->trqrpy, c.ts_rpy, 3*T)) || (rc = dn (m->trqrpytot, c.ts_tot, 3*T)) ||
            (rc = dn (m->trqrpyspr, c.ts_spr, 3*T))))
    return rc;
  if (c.US && m->nfreq)
    PMX_HIP (hipMemcpyAsync (m->nfreq, c.un_nfreq.p, c.US * sizeof (int), hipMemcpyDeviceToHost, c.stream));
  return PMX_OK;
}

static SprArgs spr_args (pmx_ctx &c)
{
  SprArgs A;
  A.S = c.S; A.N = c.N;
  A.type = c.sp_type.p; A.part = c.sp_part.p; A.spridx = c.sp_spridx.p; A.dashidx = c.sp_dashidx.p;
  A.unidx = c.sp_unidx.p; A.unspring = c.sp_unspring.p; A.flg = c.sp_flg.p;
  A.pnt = c.sp_pnt.p; A.dir = c.sp_dir.p; A.yield = c.sp_yield.p; A.sdsp = c.sp_sdsp.p; A.stroke = c.sp_stroke.p;
  A.frc = c.sp_frc.p; A.stroke0 = c.sp_stroke0.p; A.fric = c.sp_fric.p; A.kskn = c.sp_kskn.p;
  A.sx = c.sp_curve.p; A.sy = c.sp_curve.p + c.sprcrv;
  A.dx = c.sp_dash.p; A.dy = c.sp_dash.p + c.dashcrv;
  A.ux = c.sp_unload.p; A.uy = c.sp_unload.p + c.uncrv;
  A.angular = c.angular.p; A.linear = c.linear.p; A.rotation = c.rotation.p; A.position = c.position.p;
  A.inverse = c.inverse.p; A.invm = c.invm.p; A.out = c.sp_out.p; A.flags = c.dflags.p; A.ctl = c.ctl.p;
  A.stop = c.staggered ? c.istop.p : nullptr; A.par_inst = c.par_inst.p;
  A.adp = c.adp_on ? c.sp_adp.p : nullptr;
  A.offset = c.sp_offsets ? (const int*) c.sp_offset.p : nullptr;
  A.lcidx = c.lcidx.p; A.lcx = c.lcurve.p; A.lcy = c.lcurve.p + c.lccrv; A.ttab = c.ttab.p;
  return A;
}

int step_springs (pmx_ctx &c, double h)
{
  if (c.S == 0) return PMX_OK;
  SprArgs A = spr_args (c);
  {
    ProfScope ps (c, K_SPRINGS, 420.0 * c.S);
    hipLaunchKernelGGL (k_springs, grid (c.S), dim3 (PMX_TPB), 0, c.stream, A, h);
  }
  PMX_HIP (hipGetLastError ());
  {
    ProfScope ps (c, K_SPRING_ACC, 96.0 * c.N + 2 * 72.0 * c.S);
    hipLaunchKernelGGL (k_spring_acc, grid (c.N), dim3 (PMX_TPB), 0, c.stream, c.N, c.S, (const int*) c.sp_evoff.p,
                        (const int*) c.sp_evrun.p, (const int*) c.sp_runoff.p, (const int*) c.sp_runspr.p,
                        (const double*) c.sp_out.p, c.force.p, c.torque.p, (const int*) c.ctl.p,
                        c.staggered ? (const int*) c.istop.p : (const int*) nullptr, (const int*) c.par_inst.p);
  }
  if (c.adp_on)
    hipLaunchKernelGGL (k_spring_adp, grid (c.N), dim3 (PMX_TPB), 0, c.stream, c.N, c.S, (const int*) c.sp_evoff.p,
                        (const int*) c.sp_evrun.p, (const int*) c.sp_runoff.p, (const int*) c.sp_runspr.p,
                        (const double*) c.sp_adp.p, c.adp.p, (const int*) c.ctl.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

int step_torsion (pmx_ctx &c, double h)
{
  if (c.TS == 0) return PMX_OK;
  TsArgs A;
  A.T = c.TS; A.N = c.N;
  A.part = c.ts_part.p; A.kidx = c.ts_kidx.p; A.didx = c.ts_didx.p; A.cone = c.ts_cone.p;
  A.z0 = c.ts_z0.p; A.x0 = c.ts_x0.p;
  A.kx = c.ts_krpy.p; A.ky = c.ts_krpy.p + c.krpycrv; A.dx = c.ts_drpy.p; A.dy = c.ts_drpy.p + c.drpycrv;
  A.z1 = c.ts_z1.p; A.x1 = c.ts_x1.p; A.rpy = c.ts_rpy.p; A.tot = c.ts_tot.p; A.spr = c.ts_spr.p; A.out = c.ts_out.p;
  A.angular = c.angular.p; A.rotation = c.rotation.p; A.inverse = c.inverse.p; A.ctl = c.ctl.p;
  hipLaunchKernelGGL (k_torsion, grid (c.TS), dim3 (PMX_TPB), 0, c.stream, A, h);
  hipLaunchKernelGGL (k_torsion_acc, grid (c.N), dim3 (PMX_TPB), 0, c.stream, c.N, c.TS, (const int*) c.ts_evoff.p,
                      (const int*) c.ts_evrun.p, (const int*) c.ts_runoff.p, (const int*) c.ts_runspr.p,
                      (const double*) c.ts_out.p, c.torque.p, (const int*) c.ctl.p);
  if (c.adp_on)
    hipLaunchKernelGGL (k_torsion_adp, grid (c.N), dim3 (PMX_TPB), 0, c.stream, c.N, c.TS, (const int*) c.ts_evoff.p,
                        (const int*) c.ts_evrun.p, (const int*) c.ts_runoff.p, (const int*) c.ts_runspr.p,
                        (const double*) c.ts_out.p, c.adp.p, (const int*) c.ctl.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

int step_unspring (pmx_ctx &c)
{
  if (c.US == 0) return PMX_OK;
  hipLaunchKernelGGL (k_unspring, grid (c.US), dim3 (PMX_TPB), 0, c.stream, c.US, c.S, c.stepnum,
                      (const int*) c.un_tidx.p, (const int*) c.un_t.p, (const int*) c.un_midx.p, (const int*) c.un_m.p,
                      (const int*) c.un_aidx.p, (const int*) c.un_a.p, (const int*) c.un_ent.p, (const int*) c.un_op.p,
                      (const int*) c.un_abs.p, (const int*) c.un_nsteps.p, c.un_nfreq.p, (const int*) c.un_action.p,
                      (const double*) c.un_lim.p, (const double*) c.sp_stroke.p, (const double*) c.sp_frc.p,
                      c.sp_unspring.p, (const int*) c.ctl.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

/* CRITICAL's spring part of the per-particle sums in c.adp (the contact
 * part already there); c.sp_adp holds kspr, espr per spring meanwhile */
int crit_springs (pmx_ctx &c)
{
  if (c.S == 0) return PMX_OK;
  if (c.sp_adp.alloc (17 * (size_t) c.S)) { pmx_set_error ("out of device memory"); return PMX_ENOMEM; }
  SprArgs A = spr_args (c);
  hipLaunchKernelGGL (k_crit_spring_terms, grid (c.S), dim3 (PMX_TPB), 0, c.stream, A, c.sp_adp.p);
  hipLaunchKernelGGL (k_crit_springs, grid (c.N), dim3 (PMX_TPB), 0, c.stream, A, (const int*) c.sp_pofs.p,
                      (const int*) c.sp_plist.p, (const double*) c.sp_adp.p, c.adp.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

/* critical_perspring into device arrays h, o, r of S entries */
int crit_perspring (pmx_ctx &c, double *h, double *o, double *r)
{
  if (c.S == 0) return PMX_OK;
  SprArgs A = spr_args (c);
  hipLaunchKernelGGL (k_crit_perspring, grid (c.S), dim3 (PMX_TPB), 0, c.stream, A, h, o, r);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

/* the nodes and faces of meshed particles from their particles' current
 * pose; gated (inside the step): only those that spheres can touch
 * (mnod, mfac where not all), the others formed ungated where the state is
 * read */
int step_mesh_shapes (pmx_ctx &c, bool gated)
{
  if (c.NOD == 0) return PMX_OK;
  const int *ctl = gated ? (const int*) c.ctl.p : nullptr;
  const bool ln = gated && c.nmnod >= 0, lf = gated && c.nmfac >= 0;
  const int nn = ln ? c.nmnod : c.NOD, nf = lf ? c.nmfac : c.FAC;
  if (nn > 0)
    hipLaunchKernelGGL (k_nodes, grid (nn), dim3 (PMX_TPB), 0, c.stream, c.NOD, c.N, (const int*) c.nodpart.p,
                        c.nodes.p, (const double*) c.rotation.p, (const double*) c.position.p, ctl,
                        c.staggered ? (const int*) c.istop.p : (const int*) nullptr, (const int*) c.par_inst.p,
                        ln ? (const int*) c.mnod.p : (const int*) nullptr, nn);
  if (nf > 0)
    hipLaunchKernelGGL (k_faces, grid (nf), dim3 (PMX_TPB), 0, c.stream, c.FAC, c.NOD, c.T, (const int*) c.facnod.p,
                        (const int*) c.factri.p, (const double*) c.nodes.p, c.tri.p, ctl,
                        c.staggered ? (const int*) c.istop.p : (const int*) nullptr, (const int*) c.tri_inst.p,
                        lf ? (const int*) c.mfac.p : (const int*) nullptr, nf);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}
