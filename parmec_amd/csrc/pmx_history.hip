/*
 * pmx_history.hip -- time histories sampled from the resident state.
 *
 * HISTORY (input.cpp:3157-3555) registers an entity over a source: a list
 * of particles (or of springs), or one particle's body point.  dem()
 * appends the entity's value at the end of every step that crosses the
 * history interval (parmec.cpp:2509-2514) and once before the first step
 * of the first call (parmec.cpp:2386-2394); output_history
 * (output.cpp:5143-5520) forms the value.  Here the host marks which steps
 * are sampled (pmx_forcing.hisrow) and k_history, the last kernel of the
 * captured step, writes the sampled rows on the device; the rows go back
 * to the host once per dem call.  The arithmetic is output_history's:
 * a list value is the sum over the list in list order divided by its
 * length; a point value transforms the referential point with the
 * particle's rotation (NVADDMUL, macros.h:973-978).
 */
#include "pmx_internal.h"
#include <mutex>

/* Scratch of pmx_history_sample_noalloc: one sample of at most
 * PMX_HIS_NOALLOC_MAX histories (32 KiB), held by every GPU the library runs
 * on until the process ends.  Host code never names it: each device's
 * address of it is taken once, at context creation, and kept in
 * scratch_dev. */
static __device__ __align__(256) double scratch[PMX_HIS_NOALLOC_MAX];

namespace {

constexpr int MAX_DEVICES = 64;
double *scratch_dev[MAX_DEVICES] = {};
std::mutex scratch_mu;

inline dim3 grid (int n) { return dim3 ((n + PMX_TPB - 1) / PMX_TPB); }

__device__ __forceinline__ double len3 (double a, double b, double c) { return sqrt (a*a + b*b + c*c); }

/* the state of the torsion springs, UNSPRING and joints a history reads */
struct HisMore
{
  int T, J;
  const int *unspring;
  const double *tz1, *tx1, *rpy, *tot, *spr, *jreac;
};

/* one history's value (entity e over list[o0..o1)) from the state arrays */
__device__ double his_value (int e, const int *list, int o0, int o1, bool point, const double *pt, int N, int S,
                             const double *position, const double *linear, const double *angular,
                             const double *rotation, const double *force, const double *torque, const double *spnt,
                             const double *stroke, const double *sfrc, const HisMore &M)
{
  if (point)
  {
    int k = list[o0];
    double x[3] = {position[k], position[(size_t) N + k], position[2*(size_t) N + k]};
    double X[3] = {position[3*(size_t) N + k], position[4*(size_t) N + k], position[5*(size_t) N + k]};
    double v[3] = {linear[k], linear[(size_t) N + k], linear[2*(size_t) N + k]};
    double o[3] = {angular[3*(size_t) N + k], angular[4*(size_t) N + k], angular[5*(size_t) N + k]};
    double L[9];
    for (int r = 0; r < 9; r++) L[r] = rotation[(size_t) r*N + k];
    double P[3] = {pt[0], pt[1], pt[2]};
    double Q[3] = {P[0] - X[0], P[1] - X[1], P[2] - X[2]};
    double p[3], a[3];
    for (int r = 0; r < 3; r++) p[r] = x[r] + (L[r]*Q[0] + L[r+3]*Q[1] + L[r+6]*Q[2]);
    for (int r = 0; r < 3; r++) a[r] = p[r] - x[r];
    double w[3] = {v[0] + a[1]*o[2] - a[2]*o[1], v[1] + a[2]*o[0] - a[0]*o[2], v[2] + a[0]*o[1] - a[1]*o[0]};
    switch (e)
    {
      case PMX_HIS_PX: return p[0];
      case PMX_HIS_PY: return p[1];
      case PMX_HIS_PZ: return p[2];
      case PMX_HIS_PL: return len3 (p[0], p[1], p[2]);
      case PMX_HIS_DX: return p[0] - P[0];
      case PMX_HIS_DY: return p[1] - P[1];
      case PMX_HIS_DZ: return p[2] - P[2];
      case PMX_HIS_DL: return len3 (p[0] - P[0], p[1] - P[1], p[2] - P[2]);
      case PMX_HIS_VX: return w[0];
      case PMX_HIS_VY: return w[1];
      case PMX_HIS_VZ: return w[2];
      case PMX_HIS_VL: return len3 (w[0], w[1], w[2]);
      case PMX_HIS_OX: return o[0];
      case PMX_HIS_OY: return o[1];
      case PMX_HIS_OZ: return o[2];
      case PMX_HIS_OL: return len3 (o[0], o[1], o[2]);
      case PMX_HIS_FX: return force[k];
      case PMX_HIS_FY: return force[(size_t) N + k];
      case PMX_HIS_FZ: return force[2*(size_t) N + k];
      case PMX_HIS_FL: return len3 (force[k], force[(size_t) N + k], force[2*(size_t) N + k]);
      case PMX_HIS_TX: return torque[k];
      case PMX_HIS_TY: return torque[(size_t) N + k];
      case PMX_HIS_TZ: return torque[2*(size_t) N + k];
      case PMX_HIS_TL: return len3 (torque[k], torque[(size_t) N + k], torque[2*(size_t) N + k]);
      default: return 0.0;
    }
  }
  double value = 0.0;
  for (int j = o0; j < o1; j++)
  {
    int k = list[j];
    size_t n1 = (size_t) N + k, n2 = 2*(size_t) N + k;
    switch (e)
    {
      case PMX_HIS_PX: value += position[k]; break;
      case PMX_HIS_PY: value += position[n1]; break;
      case PMX_HIS_PZ: value += position[n2]; break;
      case PMX_HIS_PL: value += len3 (position[k], position[n1], position[n2]); break;
      case PMX_HIS_DX: value += position[k] - position[3*(size_t) N + k]; break;
      case PMX_HIS_DY: value += position[n1] - position[4*(size_t) N + k]; break;
      case PMX_HIS_DZ: value += position[n2] - position[5*(size_t) N + k]; break;
      case PMX_HIS_DL:
        value += len3 (position[k] - position[3*(size_t) N + k], position[n1] - position[4*(size_t) N + k],
                       position[n2] - position[5*(size_t) N + k]);
        break;
      case PMX_HIS_VX: value += linear[k]; break;
      case PMX_HIS_VY: value += linear[n1]; break;
      case PMX_HIS_VZ: value += linear[n2]; break;
      case PMX_HIS_VL: value += len3 (linear[k], linear[n1], linear[n2]); break;
      case PMX_HIS_OX: value += angular[3*(size_t) N + k]; break;
      case PMX_HIS_OY: value += angular[4*(size_t) N + k]; break;
      case PMX_HIS_OZ: value += angular[5*(size_t) N + k]; break;
      case PMX_HIS_OL:
        value += len3 (angular[3*(size_t) N + k], angular[4*(size_t) N + k], angular[5*(size_t) N + k]);
        break;
      case PMX_HIS_FX: value += force[k]; break;
      case PMX_HIS_FY: value += force[n1]; break;
      case PMX_HIS_FZ: value += force[n2]; break;
      case PMX_HIS_FL: value += len3 (force[k], force[n1], force[n2]); break;
      case PMX_HIS_TX: value += torque[k]; break;
      case PMX_HIS_TY: value += torque[n1]; break;
      case PMX_HIS_TZ: value += torque[n2]; break;
      case PMX_HIS_TL: value += len3 (torque[k], torque[n1], torque[n2]); break;
      /* springs: point p component c of spring k at (6p+c) S + k */
      case PMX_HIS_LENGTH:
        value += len3 (spnt[(size_t) 6*S + k] - spnt[k], spnt[(size_t) 7*S + k] - spnt[(size_t) S + k],
                       spnt[(size_t) 8*S + k] - spnt[(size_t) 2*S + k]);
        break;
      case PMX_HIS_STROKE: value += stroke[k]; break;
      case PMX_HIS_F: value += sfrc[k]; break;
      case PMX_HIS_SF: value += sfrc[(size_t) S + k]; break;
      case PMX_HIS_FF: value += sfrc[2*(size_t) S + k]; break;
      case PMX_HIS_SS: value += (double) M.unspring[k]; break;
      /* torsion springs, component c of spring k at c T + k */
      case PMX_HIS_ZDIR_X: value += M.tz1[k]; break;
      case PMX_HIS_ZDIR_Y: value += M.tz1[(size_t) M.T + k]; break;
      case PMX_HIS_ZDIR_Z: value += M.tz1[2*(size_t) M.T + k]; break;
      case PMX_HIS_XDIR_X: value += M.tx1[k]; break;
      case PMX_HIS_XDIR_Y: value += M.tx1[(size_t) M.T + k]; break;
      case PMX_HIS_XDIR_Z: value += M.tx1[2*(size_t) M.T + k]; break;
      case PMX_HIS_ROLL: value += M.rpy[k]; break;
      case PMX_HIS_PITCH: value += M.rpy[(size_t) M.T + k]; break;
      case PMX_HIS_YAW: value += M.rpy[2*(size_t) M.T + k]; break;
      case PMX_HIS_TRQTOT_R: value += M.tot[k]; break;
      case PMX_HIS_TRQTOT_P: value += M.tot[(size_t) M.T + k]; break;
      case PMX_HIS_TRQTOT_Y: value += M.tot[2*(size_t) M.T + k]; break;
      case PMX_HIS_TRQSPR_R: value += M.spr[k]; break;
      case PMX_HIS_TRQSPR_P: value += M.spr[(size_t) M.T + k]; break;
      case PMX_HIS_TRQSPR_Y: value += M.spr[2*(size_t) M.T + k]; break;
      /* joint reactions, joint k in creation order */
      case PMX_HIS_JREAC_X: value += M.jreac[k]; break;
      case PMX_HIS_JREAC_Y: value += M.jreac[(size_t) M.J + k]; break;
      case PMX_HIS_JREAC_Z: value += M.jreac[2*(size_t) M.J + k]; break;
      case PMX_HIS_JREAC_L: value += len3 (M.jreac[k], M.jreac[(size_t) M.J + k], M.jreac[2*(size_t) M.J + k]); break;
      default: break;
    }
  }
  return value / (double) (o1 - o0);
}

/* thread h writes history h's value for the step's row (row < 0 in the
 * step table: not sampled); row0 >= 0 (no table): that row, now */
__global__ void k_history (const int *ctl, const int *hisrow, int row0, int nhis, const int *ent, const int *off,
                           const int *list, const signed char *haspt, const double *pts, int N, int S,
                           const double *position, const double *linear, const double *angular,
                           const double *rotation, const double *force, const double *torque, const double *spnt,
                           const double *stroke, const double *sfrc, HisMore M, double *out)
{
  int row = row0;
  if (hisrow)
  {
    if (ctl[CTL_ABORT]) return;
    row = hisrow[ctl[CTL_KSTEP]];
  }
  int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (row < 0 || h >= nhis) return;
  int e = ent[h];
  double v = e == PMX_HIS_TIME ? 0.0
           : his_value (e, list, off[h], off[h+1], haspt[h] != 0, pts + 3*(size_t) h, N, S, position, linear,
                        angular, rotation, force, torque, spnt, stroke, sfrc, M);
  out[(size_t) row * nhis + h] = v;
}

/* particles whose force and torque a history reads keep them stored every
 * step (not formed lazily in k_dynamics) */
__global__ void k_iso_keep (int n, const int *list, signed char *iso)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) iso[list[j]] = (signed char) (iso[list[j]] & ~4);
}

int fail (int code, const std::string &msg) { pmx_set_error (msg); return code; }

void launch_history (pmx_ctx &c, const int *hisrow, int row0, double *out)
{
  HisMore M;
  M.T = c.TS; M.J = c.J;
  M.unspring = c.sp_unspring.p; M.tz1 = c.ts_z1.p; M.tx1 = c.ts_x1.p; M.rpy = c.ts_rpy.p; M.tot = c.ts_tot.p;
  M.spr = c.ts_spr.p; M.jreac = c.jt_reac.p;
  hipLaunchKernelGGL (k_history, grid (c.nhis), dim3 (PMX_TPB), 0, c.stream, (const int*) c.ctl.p, hisrow, row0,
                      c.nhis, (const int*) c.his_ent.p, (const int*) c.his_off.p, (const int*) c.his_list.p,
                      (const signed char*) c.his_pt.p, (const double*) c.his_pts.p, c.N, c.S,
                      (const double*) c.position.p, (const double*) c.linear.p, (const double*) c.angular.p,
                      (const double*) c.rotation.p, (const double*) c.force.p, (const double*) c.torque.p,
                      (const double*) c.sp_pnt.p, (const double*) c.sp_stroke.p, (const double*) c.sp_frc.p, M, out);
}

} /* namespace */

int history_set (pmx_ctx &c, const pmx_history *h)
{
  c.nhis = 0;
  c.his_active = false;
  c.graph_valid = false;
  if (!h || h->count <= 0) return PMX_OK;
  if (c.I != 1) return fail (PMX_EUNSUPPORTED, "time histories need a single-model context");
  int n = h->count;
  if (!h->entity || !h->off) return fail (PMX_EINVAL, "history entity and offsets required");
  int nl = h->off[n];
  std::vector<signed char> pt (n, 0);
  std::vector<double> pts (3*(size_t) n, 0.0);
  std::vector<int> keep;
  for (int i = 0; i < n; i++)
  {
    int e = h->entity[i], o0 = h->off[i], o1 = h->off[i+1];
    if (e < 0 || e >= PMX_HIS_COUNT) return fail (PMX_EINVAL, "invalid history entity");
    if (o0 < 0 || o1 < o0 || o1 > nl) return fail (PMX_EINVAL, "invalid history offsets");
    bool spring = e >= PMX_HIS_LENGTH;
    /* the count a source number must stay below: particles, springs,
     * torsion springs or joints */
    int lim = e >= PMX_HIS_JREAC_X ? c.J : e >= PMX_HIS_ZDIR_X ? c.TS : spring ? c.S : c.N;
    const char *what = e >= PMX_HIS_JREAC_X ? "Joint index out of range" : e >= PMX_HIS_ZDIR_X ?
                       "Torsion spring index out of range" : spring ? "Spring index out of range" : "Particle index out of range";
    if (e != PMX_HIS_TIME && o1 == o0) return fail (PMX_EINVAL, "empty history source");
    for (int j = o0; j < o1; j++)
    {
      int k = h->list[j];
      if (k < 0 || k >= lim) return fail (PMX_EINVAL, what);
      if (e >= PMX_HIS_FX && e <= PMX_HIS_TL) keep.push_back (k);
    }
    if (h->haspoint && h->haspoint[i])
    {
      if (spring || o1 - o0 != 1 || !h->point) return fail (PMX_EINVAL, "a point history takes one particle");
      pt[i] = 1;
      for (int d = 0; d < 3; d++) pts[3*(size_t) i + d] = h->point[3*(size_t) i + d];
    }
  }
  std::vector<int> list (h->list, h->list + std::max (nl, 0));
  if (list.empty ()) list.push_back (0);
  auto upi = [&] (dvec<int> &d, const int *src, size_t cnt) -> int {
    if (d.alloc (std::max (cnt, (size_t) 1)) != hipSuccess) return fail (PMX_ENOMEM, "out of device memory (histories)");
    PMX_HIP (hipMemcpyAsync (d.p, src, sizeof (int) * cnt, hipMemcpyHostToDevice, c.stream));
    return PMX_OK;
  };
  int rc;
  if ((rc = upi (c.his_ent, h->entity, n)) || (rc = upi (c.his_off, h->off, n + 1)) ||
      (rc = upi (c.his_list, list.data (), list.size ())))
    return rc;
  if (c.his_pt.alloc (n) != hipSuccess || c.his_pts.alloc (3*(size_t) n) != hipSuccess)
    return fail (PMX_ENOMEM, "out of device memory (histories)");
  PMX_HIP (hipMemcpyAsync (c.his_pt.p, pt.data (), n, hipMemcpyHostToDevice, c.stream));
  PMX_HIP (hipMemcpyAsync (c.his_pts.p, pts.data (), sizeof (double) * 3 * n, hipMemcpyHostToDevice, c.stream));
  if (!keep.empty () && c.iso.p)
  {
    dvec<int> kl;
    if ((rc = upi (kl, keep.data (), keep.size ()))) return rc;
    hipLaunchKernelGGL (k_iso_keep, grid ((int) keep.size ()), dim3 (PMX_TPB), 0, c.stream, (int) keep.size (),
                        (const int*) kl.p, c.iso.p);
    PMX_HIP (hipGetLastError ());
    PMX_HIP (hipStreamSynchronize (c.stream));
  }
  PMX_HIP (hipStreamSynchronize (c.stream));
  c.nhis = n;
  return PMX_OK;
}

int history_now (pmx_ctx &c, double *out)
{
  if (c.nhis == 0) return PMX_OK;
  if (c.hisbuf.alloc (std::max ((size_t) c.nhis, c.hisbuf.n)) != hipSuccess) return fail (PMX_ENOMEM, "out of device memory (histories)");
  c.graph_valid = false;   /* the buffer may have moved */
  launch_history (c, nullptr, 0, c.hisbuf.p);
  PMX_HIP (hipGetLastError ());
  PMX_HIP (hipMemcpyAsync (out, c.hisbuf.p, sizeof (double) * c.nhis, hipMemcpyDeviceToHost, c.stream));
  PMX_HIP (hipStreamSynchronize (c.stream));
  return PMX_OK;
}

/* Energy and momentum of particles [p0, p1): translational and rotational
 * kinetic energy, gravitational potential energy -m g.x, linear momentum m v
 * and angular momentum about the origin x cross m v + L J Omega.  Each block
 * sums a fixed stride of particles per thread and reduces its threads in a
 * fixed tree; k_energy_sum adds the blocks in order: the result is a
 * function of the state alone. */
constexpr int EN_TPB = 256, EN_BLOCKS = 1024, EN_Q = 9;

__global__ void k_energy_part (int N, int p0, int p1, const double *position, const double *linear,
                               const double *angular, const double *rotation, const double *inertia,
                               const double *mass, double gx, double gy, double gz, double *part)
{
  __shared__ double sh[EN_Q][EN_TPB];
  double q[EN_Q] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = p0 + blockIdx.x * EN_TPB + threadIdx.x; i < p1; i += EN_TPB * gridDim.x)
  {
    double m = mass[i];
    double v[3] = {linear[i], linear[(size_t) N + i], linear[2*(size_t) N + i]};
    double x[3] = {position[i], position[(size_t) N + i], position[2*(size_t) N + i]};
    double O[3] = {angular[i], angular[(size_t) N + i], angular[2*(size_t) N + i]};
    double J[9], L[9], JO[3], h[3];
    for (int k = 0; k < 9; k++) { J[k] = inertia[(size_t) k*N + i]; L[k] = rotation[(size_t) k*N + i]; }
    for (int r = 0; r < 3; r++) JO[r] = J[r]*O[0] + J[3+r]*O[1] + J[6+r]*O[2];
    for (int r = 0; r < 3; r++) h[r] = L[r]*JO[0] + L[3+r]*JO[1] + L[6+r]*JO[2];
    double p[3] = {m*v[0], m*v[1], m*v[2]};
    q[0] += 0.5*m*(v[0]*v[0] + v[1]*v[1] + v[2]*v[2]);
    q[1] += 0.5*(O[0]*JO[0] + O[1]*JO[1] + O[2]*JO[2]);
    q[2] += -m*(gx*x[0] + gy*x[1] + gz*x[2]);
    q[3] += p[0]; q[4] += p[1]; q[5] += p[2];
    q[6] += x[1]*p[2] - x[2]*p[1] + h[0];
    q[7] += x[2]*p[0] - x[0]*p[2] + h[1];
    q[8] += x[0]*p[1] - x[1]*p[0] + h[2];
  }
  for (int k = 0; k < EN_Q; k++) sh[k][threadIdx.x] = q[k];
  __syncthreads ();
  for (int w = EN_TPB / 2; w > 0; w >>= 1)
  {
    if ((int) threadIdx.x < w)
      for (int k = 0; k < EN_Q; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads ();
  }
  if (threadIdx.x < EN_Q) part[(size_t) blockIdx.x * EN_Q + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_energy_sum (int nb, const double *part, double *out)
{
  int k = threadIdx.x;
  if (k >= EN_Q) return;
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += part[(size_t) b * EN_Q + k];
  out[k] = s;
}

/* the device address of scratch on the current device (once per device);
 * called by pmx_create with the context's device current */
int history_scratch_init (int device)
{
  if (device < 0 || device >= MAX_DEVICES) return fail (PMX_EINVAL, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock (scratch_mu);
  if (scratch_dev[device]) return PMX_OK;
  void *p = nullptr;
  PMX_HIP (hipGetSymbolAddress (&p, HIP_SYMBOL (scratch)));
  if (!p) return fail (PMX_EDEVICE, "no device address for the history scratch");
  scratch_dev[device] = (double*) p;
  return PMX_OK;
}

/* the step's sample, the last kernel of the step before its counter moves */
int step_history (pmx_ctx &c)
{
  if (!c.his_active || c.nhis == 0) return PMX_OK;
  launch_history (c, (const int*) c.hisrow.p, -1, c.hisbuf.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

extern "C" int pmx_history_sample_noalloc (pmx_ctx *c, double *out)
{
  if (!c || !out) return fail (PMX_EINVAL, "null argument");
  if (c->nhis > PMX_HIS_NOALLOC_MAX)
    return fail (PMX_ECAPACITY, "pmx_history_sample_noalloc: " + std::to_string (c->nhis) + " histories, at most " +
                 std::to_string (PMX_HIS_NOALLOC_MAX));
  if (c->nhis == 0) return PMX_OK;
  if (c->device < 0 || c->device >= MAX_DEVICES || !scratch_dev[c->device])
    return fail (PMX_EDEVICE, "history scratch not set up for this device");
  (void) hipSetDevice (c->device);
  double *buf = scratch_dev[c->device];
  launch_history (*c, nullptr, 0, buf);
  PMX_HIP (hipGetLastError ());
  PMX_HIP (hipMemcpyAsync (out, buf, sizeof (double) * c->nhis, hipMemcpyDeviceToHost, c->stream));
  PMX_HIP (hipStreamSynchronize (c->stream));
  return PMX_OK;
}

extern "C" int pmx_energy (pmx_ctx *c, int instance, const double *g, double *out)
{
  if (!c || !out) return fail (PMX_EINVAL, "null argument");
  if (!c->uploaded) return fail (PMX_EINVAL, "no model uploaded");
  if (instance < -1 || instance >= c->I) return fail (PMX_EINVAL, "instance out of range");
  if (c->staggered && instance < 0)
    return fail (PMX_EINVAL, "instances are at different times: ask for one instance");
  (void) hipSetDevice (c->device);
  int p0 = instance < 0 ? 0 : c->off[instance].par, p1 = instance < 0 ? c->N : c->off[instance + 1].par;
  int nb = std::max (1, std::min (EN_BLOCKS, (p1 - p0 + EN_TPB - 1) / EN_TPB));
  if (c->energy.alloc ((size_t) EN_BLOCKS * EN_Q + EN_Q) != hipSuccess) return fail (PMX_ENOMEM, "out of device memory");
  double gg[3] = {g ? g[0] : 0.0, g ? g[1] : 0.0, g ? g[2] : 0.0};
  hipLaunchKernelGGL (k_energy_part, dim3 (nb), dim3 (EN_TPB), 0, c->stream, c->N, p0, p1,
                      (const double*) c->position.p, (const double*) c->linear.p, (const double*) c->angular.p,
                      (const double*) c->rotation.p, (const double*) c->inertia.p, (const double*) c->mass.p,
                      gg[0], gg[1], gg[2], c->energy.p);
  hipLaunchKernelGGL (k_energy_sum, dim3 (1), dim3 (64), 0, c->stream, nb, (const double*) c->energy.p,
                      c->energy.p + (size_t) EN_BLOCKS * EN_Q);
  PMX_HIP (hipGetLastError ());
  PMX_HIP (hipMemcpyAsync (out, c->energy.p + (size_t) EN_BLOCKS * EN_Q, sizeof (double) * EN_Q, hipMemcpyDeviceToHost,
                           c->stream));
  PMX_HIP (hipStreamSynchronize (c->stream));
  return PMX_OK;
}
