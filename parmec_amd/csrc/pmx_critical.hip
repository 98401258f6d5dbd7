/*
 * pmx_critical.hip -- critical time step estimates (CRITICAL,
 * input.cpp:3034-3154; critical_perparticle and critical_perspring,
 * parmec.ispc:485-888) and the per-step minimum of the adaptive step
 * (adaptive_timestep, parmec.ispc:391-420).
 *
 * Per particle the estimate needs four sums over what loads it: the summed
 * |direction| components kact, the largest stiffness kmax and damping
 * coefficient emax, and the rotational stiffness terms krot, held as
 * adp [11][N] (kact 0-2, kmax 3, emax 4, krot 5-10).  The adaptive step
 * forms them inside the step (k_rows<R, true>, k_spring_adp); CRITICAL
 * forms them from the contacts the last step recorded (k_crit_contacts) and
 * from the springs' curve bounds (k_crit_springs).  k_time_step turns them
 * into the estimate (critical_step) and, for the adaptive step, reduces the
 * minimum over particles: a minimum is exact in any order, so one atomic
 * minimum per wave gives the reference's value.
 */
#include "pmx_internal.h"
#include "pmx_device.h"
#include <string>

using namespace pmxd;

namespace {

int fail (int code, const std::string &msg) { pmx_set_error (msg); return code; }

inline dim3 grid (int n) { return dim3 ((n + PMX_TPB - 1) / PMX_TPB); }

__global__ void k_adp_zero (size_t n, double *a, unsigned long long *hmin, const int *ctl)
{
  if (ctl && ctl[CTL_ABORT]) return;
  size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
  if (i == 0 && hmin) *hmin = (unsigned long long) __double_as_longlong (1.7976931348623157e308);
}

/* time_step_task over every particle (parmec.ispc:275-420); with h, the
 * estimate, frequency and damping ratio of each particle (frequency and
 * ratio 0 where the estimate is DBL_MAX: the reference leaves them unset);
 * with hmin, the minimum estimate below DBL_MAX, as the bits of a
 * non-negative double (whose unsigned order is the numeric order): NaN and
 * infinite estimates never win against the initial DBL_MAX, as in the
 * reference's min (hcri[i], hmin) */
__global__ void k_time_step (int N, const double *adp, const double *mass, const double *inertia, double *h,
                             double *o, double *r, unsigned long long *hmin, const int *ctl)
{
  if (ctl && ctl[CTL_ABORT]) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 1.7976931348623157e308;
  if (i < N)
  {
    double ka[3], kr[6], J[9];
    size_t n = N;
#pragma unroll
    for (int c = 0; c < 3; c++) ka[c] = adp[c*n + i];
#pragma unroll
    for (int c = 0; c < 6; c++) kr[c] = adp[(5+c)*n + i];
#pragma unroll
    for (int c = 0; c < 9; c++) J[c] = inertia[c*n + i];
    double oo = 0.0, rr = 0.0;
    double hh = critical_step (ka, adp[3*n + i], adp[4*n + i], kr, mass[i], J, &oo, &rr);
    if (h) { h[i] = hh; o[i] = oo; r[i] = rr; }
    if (hh < v) v = hh == 0.0 ? 0.0 : hh;
  }
  if (!hmin) return;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
  {
    double w = __shfl_xor (v, off, 64);
    v = w < v ? w : v;
  }
  if ((threadIdx.x & 63) == 0 && v < 1.7976931348623157e308)
    atomicMin (hmin, (unsigned long long) __double_as_longlong (v));
}

/* critical_perparticle_contacts (parmec.ispc:547-656) from the contacts the
 * last step recorded: per sphere its row entries in row order, arms taken
 * from the particle's current position and rotated by its current L */
__global__ void k_crit_contacts (int E, int N, const int *part, const int *cnt, int cap, const double *crec,
                                 const double *rotation, const double *position, double *adp)
{
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  int n = cnt[e];
  if (n == 0) return;
  size_t NN = N, EE = E;
  int pe = part[e];
  double L[9], x[3];
#pragma unroll
  for (int c = 0; c < 9; c++) L[c] = rotation[c*NN + pe];
#pragma unroll
  for (int c = 0; c < 3; c++) x[c] = position[c*NN + pe];
  double ka[3] = {0.0, 0.0, 0.0}, km = 0.0, em = 0.0, kr[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int q = 0; q < n && q < cap; q++)
  {
    double rec[8];
#pragma unroll
    for (int c = 0; c < 8; c++) rec[c] = crec[(size_t) (c*cap + q) * EE + e];
    for (int c = 0; c < 3; c++) ka[c] += rec[3+c];
    km = rec[6] > km ? rec[6] : km;
    em = rec[7] > em ? rec[7] : em;
    double a[3] = {rec[0]-x[0], rec[1]-x[1], rec[2]-x[2]}, A[3];
    nvmul (L, a, A);
    double dot = A[0]*A[0] + A[1]*A[1] + A[2]*A[2];
    kr[0] += dot - A[0]*A[0];
    kr[1] += dot - A[1]*A[1];
    kr[2] += dot - A[2]*A[2];
    kr[3] += -A[0]*A[1];
    kr[4] += -A[0]*A[2];
    kr[5] += -A[1]*A[2];
  }
  double v[11] = {ka[0], ka[1], ka[2], km, em, kr[0], kr[1], kr[2], kr[3], kr[4], kr[5]};
  for (int c = 0; c < 11; c++) adp[c*NN + pe] = v[c];
}

} /* namespace */

int adp_begin (pmx_ctx &c)
{
  size_t n = 11 * (size_t) c.N;
  hipLaunchKernelGGL (k_adp_zero, dim3 ((n + PMX_TPB - 1) / PMX_TPB), dim3 (PMX_TPB), 0, c.stream, n, c.adp.p,
                      c.hmin.p, (const int*) c.ctl.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

int adp_time_step (pmx_ctx &c, double *h, double *o, double *r)
{
  hipLaunchKernelGGL (k_time_step, grid (c.N), dim3 (PMX_TPB), 0, c.stream, c.N, (const double*) c.adp.p,
                      (const double*) c.mass.p, (const double*) c.inertia.p, h, o, r,
                      h ? (unsigned long long*) nullptr : c.hmin.p, (const int*) c.ctl.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

int crit_contacts (pmx_ctx &c)
{
  if (c.E == 0) return PMX_OK;
  hipLaunchKernelGGL (k_crit_contacts, grid (c.E), dim3 (PMX_TPB), 0, c.stream, c.E, c.N, (const int*) c.part.p,
                      (const int*) c.rowcnt[c.cur].p, c.cap, (const double*) c.crec.p, (const double*) c.rotation.p,
                      (const double*) c.position.p, c.adp.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

extern "C" {

int pmx_critical (pmx_ctx *c, int which, double *h, double *o, double *r)
{
  if (!c) return fail (PMX_EINVAL, "null ctx");
  if (!c->uploaded) return fail (PMX_EINVAL, "pmx_upload must precede pmx_critical");
  if (c->I != 1) return fail (PMX_EUNSUPPORTED, "pmx_critical takes single-model contexts");
  if (!h) return fail (PMX_EINVAL, "null output");
  if (which != PMX_CRIT_PARTICLE && which != PMX_CRIT_SPRING) return fail (PMX_EINVAL, "unknown estimate");
  (void) hipSetDevice (c->device);
  int rc, n = which == PMX_CRIT_SPRING ? c->S : c->N;
  if (n == 0) return PMX_OK;
  if (c->crit.alloc (3 * (size_t) n)) return fail (PMX_ENOMEM, "out of device memory (critical step)");
  double *dh = c->crit.p, *dov = dh + n, *dr = dh + 2 * (size_t) n;
  if (which == PMX_CRIT_SPRING)
  {
    if ((rc = crit_perspring (*c, dh, dov, dr))) return rc;
  }
  else
  {
    if (c->crec_stale)
      return fail (PMX_EUNSUPPORTED, "the contacts of the last step were not recorded (pmx_options.critical = 0)");
    if (c->adp.alloc (11 * (size_t) c->N)) return fail (PMX_ENOMEM, "out of device memory (critical step)");
    size_t m = 11 * (size_t) c->N;
    hipLaunchKernelGGL (k_adp_zero, dim3 ((m + PMX_TPB - 1) / PMX_TPB), dim3 (PMX_TPB), 0, c->stream, m, c->adp.p,
                        (unsigned long long*) nullptr, (const int*) nullptr);
    if (c->crec_valid && ((rc = crit_contacts (*c)) || (rc = step_react (*c, true)))) return rc;
    if ((rc = crit_springs (*c))) return rc;
    hipLaunchKernelGGL (k_time_step, grid (c->N), dim3 (PMX_TPB), 0, c->stream, c->N, (const double*) c->adp.p,
                        (const double*) c->mass.p, (const double*) c->inertia.p, dh, dov, dr,
                        (unsigned long long*) nullptr, (const int*) nullptr);
    PMX_HIP (hipGetLastError ());
  }
  PMX_HIP (hipMemcpyAsync (h, dh, sizeof (double) * n, hipMemcpyDeviceToHost, c->stream));
  if (o) PMX_HIP (hipMemcpyAsync (o, dov, sizeof (double) * n, hipMemcpyDeviceToHost, c->stream));
  if (r) PMX_HIP (hipMemcpyAsync (r, dr, sizeof (double) * n, hipMemcpyDeviceToHost, c->stream));
  PMX_HIP (hipStreamSynchronize (c->stream));
  return PMX_OK;
}

} /* extern "C" */
