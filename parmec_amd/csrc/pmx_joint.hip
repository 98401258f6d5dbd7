/*
 * pmx_joint.hip -- restrained particle motion (RESTRAIN; restrain.ispc:30-154)
 * and algebraic ball joints (BALL_JOINT; joints.cpp:83-567), with the set-up
 * of the torsion springs, UNSPRING conditions and load curves.
 *
 * Restraints: k_restrain_vel (once per dem call, before the steps) removes
 * the velocity components along up to three linear and three angular
 * directions; k_restrain_frc (every step, after the joints) removes the
 * force and torque components.  One thread per particle applies its entries
 * in entry order.
 *
 * Ball joints: the joints that share particles form independent sets
 * (joints.cpp:83-177: a depth-first walk from the lowest unvisited joint,
 * neighbours in ascending order).  Each step, per set, the joint reactions R
 * solve W (h R) = B (joints.cpp:181-567), W the dense 3n x 3n matrix of the
 * set's joints (blocks H J^-1 H' + 1/m on the joints of a particle) and B
 * the joint point velocities at t and t+h that the free motion would give:
 *   k_joint_assemble  one thread per joint forms its block row of W and its
 *                     part of B, in the reference's order of additions;
 *   k_joint_solve     one thread per set factorises W as amgcl's skyline_lu
 *                     does (Cuthill-McKee order, which is the identity for a
 *                     dense pattern; Crout factorisation on the profile of
 *                     non-zero entries; the diagonal held as inverses) and
 *                     solves; R = x / h;
 *   k_joint_apply     one thread per jointed particle adds +-R and the
 *                     moments of R in the set's joint order.
 */
#include "pmx_internal.h"
#include "pmx_device.h"
#include <algorithm>
#include <map>
#include <set>
#include <string>
#include <vector>

using namespace pmxd;

namespace {

int fail (int code, const std::string &msg) { pmx_set_error (msg); return code; }

inline dim3 grid (int n) { return dim3 ((n + PMX_TPB - 1) / PMX_TPB); }

/* the most joints one independent set may hold: its dense system is
 * (3n)^2 doubles and is factorised by one thread, as the reference's is by
 * one core */
constexpr int PMX_JOINT_SET_MAX = 2048;

/* v - (v.d) d, restrain.ispc:55-60 */
__device__ __forceinline__ void unproject (double v[3], const double d[3])
{
  double dot = v[0]*d[0] + v[1]*d[1] + v[2]*d[2];
  v[0] = v[0] - dot*d[0];
  v[1] = v[1] - dot*d[1];
  v[2] = v[2] - dot*d[2];
}

/* restrain_velocities_task, restrain.ispc:30-89 */
__global__ void k_restrain_vel (int RR, int N, int R, const int *runoff, const int *part, const double *lin,
                                const double *ang, double *linear, double *angular, const double *rotation)
{
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= RR) return;
  size_t NN = N, RS = R;
  int j = part[runoff[r]];
  double v[3], O[3], o[3], L[9];
#pragma unroll
  for (int c = 0; c < 3; c++) { v[c] = linear[c*NN + j]; O[c] = angular[c*NN + j]; o[c] = angular[(3+c)*NN + j]; }
#pragma unroll
  for (int c = 0; c < 9; c++) L[c] = rotation[c*NN + j];
  for (int i = runoff[r]; i < runoff[r+1]; i++)
  {
    double l[9], a[9], A0[3], A1[3], A2[3];
#pragma unroll
    for (int c = 0; c < 9; c++) { l[c] = lin[c*RS + i]; a[c] = ang[c*RS + i]; }
    unproject (v, l); unproject (v, l+3); unproject (v, l+6);
    unproject (o, a); unproject (o, a+3); unproject (o, a+6);
    tvmul (L, a, A0); tvmul (L, a+3, A1); tvmul (L, a+6, A2);
    unproject (O, A0); unproject (O, A1); unproject (O, A2);
  }
#pragma unroll
  for (int c = 0; c < 3; c++) { linear[c*NN + j] = v[c]; angular[c*NN + j] = O[c]; angular[(3+c)*NN + j] = o[c]; }
}

/* restrain_forces_task, restrain.ispc:92-133 */
__global__ void k_restrain_frc (int RR, int N, int R, const int *runoff, const int *part, const double *lin,
                                const double *ang, double *force, double *torque, const int *ctl)
{
  if (ctl[CTL_ABORT]) return;
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= RR) return;
  size_t NN = N, RS = R;
  int j = part[runoff[r]];
  double f[3], t[3];
#pragma unroll
  for (int c = 0; c < 3; c++) { f[c] = force[c*NN + j]; t[c] = torque[c*NN + j]; }
  for (int i = runoff[r]; i < runoff[r+1]; i++)
  {
    double l[9], a[9];
#pragma unroll
    for (int c = 0; c < 9; c++) { l[c] = lin[c*RS + i]; a[c] = ang[c*RS + i]; }
    unproject (f, l); unproject (f, l+3); unproject (f, l+6);
    unproject (t, a); unproject (t, a+3); unproject (t, a+6);
  }
#pragma unroll
  for (int c = 0; c < 3; c++) { force[c*NN + j] = f[c]; torque[c*NN + j] = t[c]; }
}

/* Rot * skew (X - point), the map of a body's angular velocity to its
 * point velocity about the joint (joints.cpp:247-250) */
__device__ __forceinline__ void joint_h (const double Rot[9], const double *position, size_t N, int part,
                                         const double P[3], double H[9])
{
  double A[3], Ask[9];
#pragma unroll
  for (int c = 0; c < 3; c++) A[c] = position[(3+c)*N + part] - P[c];
  Ask[0] = 0; Ask[1] = A[2]; Ask[2] = -A[1];
  Ask[3] = -A[2]; Ask[4] = 0; Ask[5] = A[0];
  Ask[6] = A[1]; Ask[7] = -A[0]; Ask[8] = 0;
  nnmul (Rot, Ask, H);
}

struct JtArgs
{
  int J, N;
  const int *part, *set, *pos, *adjoff, *adj, *soff;
  const double *point;
  const long long *moff;
  double *mat, *rhs;
  const double *rotation, *inverse, *inertia, *invm, *mass, *position, *angular, *linear, *force, *torque;
  const double *dtab;
  const int *ctl;
};

/* joints.cpp:209-293 (block row of W) and 347-449 (part of B), one thread
 * per joint */
__global__ void k_joint_assemble (JtArgs A, double step0, double step1)
{
  if (A.ctl[CTL_ABORT]) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.J) return;
  size_t N = A.N, J = A.J;
  const double half = 0.5*step0, step = 0.5*(step0+step1);
  int s = A.set[i], i0 = A.pos[i];
  int n = 3*(A.soff[s+1] - A.soff[s]);
  double *W = A.mat + A.moff[s];
  double Pi[3] = {A.point[i], A.point[J + i], A.point[2*J + i]};
  /* this joint's block row starts from zero */
  for (int r = 0; r < 3; r++)
    for (int q = 0; q < n; q++) W[(size_t) (3*i0 + r) * n + q] = 0.0;
  double Wii[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double B[3] = {0.0, 0.0, 0.0};
  const double *damping = A.dtab + 6*(size_t) A.ctl[CTL_KSTEP];
  for (int k = 0; k < 2; k++)
  {
    int part = A.part[k*J + i];
    if (part < 0) continue;
    double Rot[9], Jiv[9], Hi[9], Hj[9], C[9], D[9], Wij[9];
#pragma unroll
    for (int c = 0; c < 9; c++) { Rot[c] = A.rotation[c*N + part]; Jiv[c] = A.inverse[c*N + part]; }
    double im = A.invm[part];
    joint_h (Rot, A.position, N, part, Pi, Hi);
    ntmul (Jiv, Hi, C);
    nnmul (Hi, C, D);
#pragma unroll
    for (int c = 0; c < 9; c++) Wii[c] = Wii[c] + D[c];
    Wii[0] += im; Wii[4] += im; Wii[8] += im;
    for (int q = A.adjoff[2*i + k]; q < A.adjoff[2*i + k + 1]; q++)
    {
      int code = A.adj[q], j = code >> 1;
      int j0 = A.pos[j];
      double Pj[3] = {A.point[j], A.point[J + j], A.point[2*J + j]};
      joint_h (Rot, A.position, N, part, Pj, Hj);
      ntmul (Jiv, Hj, C);
      nnmul (Hi, C, Wij);
      Wij[0] += im; Wij[4] += im; Wij[8] += im;
      if (code & 1)
      {
        for (int c = 0; c < 9; c++) Wij[c] *= -1.0;
      }
      for (int i1 = 0; i1 < 3; i1++)
        for (int j1 = 0; j1 < 3; j1++) W[(size_t) (3*i0 + i1) * n + 3*j0 + j1] += Wij[i1*3+j1];
    }
    /* the joint point's velocity now (V) and after a free step (C),
     * joints.cpp:383-447; the linear damping reads the mass of the particle
     * numbered like the joint, as the reference does */
    double J9[9], O[3], o[3], v[3], f[3], t[3], V[3], T[3], dRot[9], a3[3], c3[3];
#pragma unroll
    for (int c = 0; c < 9; c++) J9[c] = A.inertia[c*N + part];
#pragma unroll
    for (int c = 0; c < 3; c++)
    {
      O[c] = A.angular[c*N + part]; f[c] = A.force[c*N + part]; t[c] = A.torque[c*N + part];
      v[c] = A.linear[c*N + part];
    }
    nvmul (Hi, O, V);
    V[0] += v[0]; V[1] += v[1]; V[2] += v[2];
    if (damping[0] != 0.0 || damping[1] != 0.0 || damping[2] != 0.0)
    {
      double ma = A.mass[i];
      f[0] -= ma * damping[0] * v[0];
      f[1] -= ma * damping[1] * v[1];
      f[2] -= ma * damping[2] * v[2];
    }
    if (damping[3] != 0.0 || damping[4] != 0.0 || damping[5] != 0.0)
    {
#pragma unroll
      for (int c = 0; c < 3; c++) o[c] = damping[3+c]*A.angular[(3+c)*N + part];
      tvmul (Rot, o, c3);
      nvmul (J9, c3, T);
      nvmul (Rot, T, c3);
      t[0] -= c3[0]; t[1] -= c3[1]; t[2] -= c3[2];
    }
    expmap (-half*O[0], -half*O[1], -half*O[2], dRot);
    tvmul (Rot, t, T);
    nvmul (J9, O, a3);
    nvmul (dRot, a3, c3);
#pragma unroll
    for (int c = 0; c < 3; c++) c3[c] = c3[c] + half*T[c];
    nvmul (Jiv, c3, a3);                     /* O(t+h/2) */
    nvmul (J9, a3, c3);
    T[0] -= a3[1]*c3[2] - a3[2]*c3[1];
    T[1] -= a3[2]*c3[0] - a3[0]*c3[2];
    T[2] -= a3[0]*c3[1] - a3[1]*c3[0];
    T[0] *= step; T[1] *= step; T[2] *= step;
    {
      double Ot[3];
      Ot[0] = O[0] + (Jiv[0]*T[0]+Jiv[3]*T[1]+Jiv[6]*T[2]);
      Ot[1] = O[1] + (Jiv[1]*T[0]+Jiv[4]*T[1]+Jiv[7]*T[2]);
      Ot[2] = O[2] + (Jiv[2]*T[0]+Jiv[5]*T[1]+Jiv[8]*T[2]);   /* O(t+h) */
      nvmul (Hi, Ot, c3);
    }
    c3[0] += v[0]; c3[1] += v[1]; c3[2] += v[2];
    double sim = step*im;
#pragma unroll
    for (int c = 0; c < 3; c++) c3[c] = c3[c] + sim*f[c];
    if (part == A.part[i])
      for (int c = 0; c < 3; c++) B[c] -= V[c]+c3[c];
    else
      for (int c = 0; c < 3; c++) B[c] += V[c]+c3[c];
  }
  for (int i1 = 0; i1 < 3; i1++)
    for (int j1 = 0; j1 < 3; j1++) W[(size_t) (3*i0 + i1) * n + 3*i0 + j1] += Wii[i1*3+j1];
  double *rhs = A.rhs + 3*(size_t) A.soff[s];
  for (int c = 0; c < 3; c++) rhs[3*i0 + c] = B[c];
}

/* one independent set per thread: amgcl::solver::skyline_lu on the dense
 * W (skyline_lu.hpp:74-303) and R = x / h (joints.cpp:514-521) */
__global__ void k_joint_solve (int JS, const int *soff, const int *sk, const long long *moff, const long long *woff,
                               const double *mat, const double *rhs, double *work, int *iwork, double *jreac, int J,
                               double step0, double step1, int *flags, int *ctl)
{
  if (ctl[CTL_ABORT]) return;
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= JS) return;
  int n = 3*(soff[s+1] - soff[s]);
  const double *W = mat + moff[s];
  const double *b = rhs + 3*(size_t) soff[s];
  int *ptr = iwork + 3*(size_t) soff[s] + s;   /* [n+1] */
  /* profile: the longest row of L and column of U below and above the
   * diagonal over the non-zero entries */
  for (int i = 0; i <= n; i++) ptr[i] = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
    {
      double v = W[(size_t) i * n + j];
      if (v != 0.0)
      {
        if (i > j) { if (ptr[i] < i - j) ptr[i] = i - j; }
        else if (i < j) { if (ptr[j] < j - i) ptr[j] = j - i; }
      }
    }
  {
    int last = 0;
    for (int i = 1; i <= n; i++)
    {
      int tmp = ptr[i];
      ptr[i] = ptr[i-1] + last;
      last = tmp;
    }
  }
  size_t m = ptr[n];
  double *L = work + woff[s], *U = L + m, *D = U + m, *y = D + n;
  for (size_t q = 0; q < m; q++) { L[q] = 0.0; U[q] = 0.0; }
  for (int i = 0; i < n; i++) D[i] = 0.0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
    {
      double v = W[(size_t) i * n + j];
      if (v != 0.0)
      {
        if (i < j) U[ptr[j + 1] + i - j] = v;
        else if (i == j) D[i] = v;
        else L[ptr[i + 1] + j - i] = v;
      }
    }
  /* Crout factorisation, the diagonal of U is 1 */
  if (D[0] == 0.0) { atomicAdd (&flags[FLAG_JOINT_PIVOT], 1); abort_step (ctl); return; }
  D[0] = 1.0 / D[0];
  for (int k = 0; k < n - 1; ++k)
  {
    if (ptr[k + 1] + k + 1 == ptr[k + 2]) U[ptr[k+1]] = D[0] * U[ptr[k+1]];
    int indexEntry = ptr[k + 1];
    int iBeginCol = k + 1 - ptr[k + 2] + ptr[k + 1];
    for (int i = iBeginCol; i <= k; ++indexEntry, ++i)
    {
      if (i == 0) continue;
      double sum = U[indexEntry];
      int jBeginRow = i - ptr[i + 1] + ptr[i];
      int jBeginMult = iBeginCol > jBeginRow ? iBeginCol : jBeginRow;
      int indexL = ptr[i] + jBeginMult - jBeginRow;
      int indexU = ptr[k+1] + jBeginMult - iBeginCol;
      for (int j = jBeginMult; j < i; ++j, ++indexL, ++indexU) sum -= L[indexL] * U[indexU];
      U[indexEntry] = D[i] * sum;
    }
    indexEntry = ptr[k+1];
    int jBeginRow = k + 1 - ptr[k + 2] + ptr[k + 1];
    for (int i = iBeginCol; i <= k; ++indexEntry, ++i)
    {
      if (i == 0) continue;
      double sum = L[indexEntry];
      int jBeginCol = i - ptr[i+1] + ptr[i];
      int jBeginMult = jBeginCol > jBeginRow ? jBeginCol : jBeginRow;
      int indexL = ptr[k+1] + jBeginMult - jBeginRow;
      int indexU = ptr[i] + jBeginMult - jBeginCol;
      for (int j = jBeginMult; j < i; ++j, ++indexL, ++indexU) sum -= L[indexL] * U[indexU];
      L[indexEntry] = sum;
    }
    double sum = D[k+1];
    for (int j = ptr[k+1]; j < ptr[k+2]; ++j) sum -= L[j] * U[j];
    if (sum == 0.0) { atomicAdd (&flags[FLAG_JOINT_PIVOT], 1); abort_step (ctl); return; }
    D[k+1] = 1.0 / sum;
  }
  /* y = L^-1 b, then y = U^-1 y */
  for (int i = 0; i < n; ++i)
  {
    double sum = b[i];
    for (int k = ptr[i], j = i - ptr[i+1] + k; k < ptr[i+1]; ++k, ++j) sum -= L[k] * y[j];
    y[i] = D[i] * sum;
  }
  for (int j = n - 1; j >= 0; --j)
    for (int k = ptr[j], i = j - ptr[j+1] + k; k < ptr[j+1]; ++k, ++i) y[i] -= U[k] * y[j];
  double invstep = 1./(0.5*(step0+step1));
  for (int q = soff[s]; q < soff[s+1]; q++)
  {
    int i = sk[q], i0 = q - soff[s];
    for (int c = 0; c < 3; c++) jreac[(size_t) c*J + i] = y[3*i0 + c] * invstep;
  }
}

/* joints.cpp:517-563: per particle, its joints in their set's order */
__global__ void k_joint_apply (int JP, int N, int J, const int *ppart, const int *poff, const int *plist,
                               const int *jpart, const double *jpoint, const double *jreac, const double *rotation,
                               const double *position, double *force, double *torque, const int *ctl)
{
  if (ctl[CTL_ABORT]) return;
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= JP) return;
  size_t NN = N, JJ = J;
  int part = ppart[p];
  double Rot[9], f[3], t[3];
#pragma unroll
  for (int c = 0; c < 9; c++) Rot[c] = rotation[c*NN + part];
#pragma unroll
  for (int c = 0; c < 3; c++) { f[c] = force[c*NN + part]; t[c] = torque[c*NN + part]; }
  for (int q = poff[p]; q < poff[p+1]; q++)
  {
    int i = plist[q] >> 1;
    double R[3], A[3], a[3], m[3];
#pragma unroll
    for (int c = 0; c < 3; c++) { R[c] = jreac[c*JJ + i]; A[c] = position[(3+c)*NN + part] - jpoint[c*JJ + i]; }
    nvmul (Rot, A, a);
    m[0] = R[1]*a[2] - R[2]*a[1];
    m[1] = R[2]*a[0] - R[0]*a[2];
    m[2] = R[0]*a[1] - R[1]*a[0];
    if (part == jpart[i])
      for (int c = 0; c < 3; c++) { t[c] += m[c]; f[c] += R[c]; }
    else
      for (int c = 0; c < 3; c++) { t[c] -= m[c]; f[c] -= R[c]; }
  }
#pragma unroll
  for (int c = 0; c < 3; c++) { force[c*NN + part] = f[c]; torque[c*NN + part] = t[c]; }
}

} /* namespace */

int jt_validate (const pmx_model *m)
{
  const int N = m->parnum, S = m->sprnum;
  if (m->lcnum < 0 || (m->lcnum > 0 && (!m->lcidx || !m->lcurve))) return fail (PMX_EINVAL, "load curve arrays missing");
  for (int l = 0; l < m->lcnum; l++)
    if (m->lcidx[l] < 0 || m->lcidx[l+1] < m->lcidx[l] + 2 || m->lcidx[l+1] > m->lccrv)
      return fail (PMX_EINVAL, "load curve index range out of bounds");
  for (int i = 0; i < S; i++)
  {
    if (m->sproffset && m->sproffset[i] != -1 && (m->sproffset[i] < 0 || m->sproffset[i] >= m->lcnum))
      return fail (PMX_EINVAL, "spring stroke offset load curve out of range");
    if (m->unspring[i] >= m->lcnum) return fail (PMX_EINVAL, "spring unloading load curve out of range");
  }
  const int T = m->trqsprnum;
  if (T < 0) return fail (PMX_EINVAL, "negative torsion spring count");
  if (T > 0 && (!m->trqsprpart || !m->trqzdir0 || !m->trqxdir0 || !m->krpyidx || !m->krpy || !m->drpyidx ||
                !m->drpy || !m->trqcone || !m->trqzdir1 || !m->trqxdir1 || !m->trqrpy || !m->trqrpytot || !m->trqrpyspr))
    return fail (PMX_EINVAL, "torsion spring arrays missing");
  for (int i = 0; i < T; i++)
  {
    int j = m->trqsprpart[i], k = m->trqsprpart[T + i];
    if (j < 0 || j >= N || k < -1 || k >= N) return fail (PMX_EINVAL, "torsion spring particle index out of range");
    if (i && (m->trqsprpart[i-1] > j || (m->trqsprpart[i-1] == j && m->trqsprpart[T + i-1] > k)))
      return fail (PMX_EINVAL, "torsion springs must be sorted by (part1, part2)");
    int cone = m->trqcone[i];
    if (cone != 0 && cone != 3 && cone != 5 && cone != 6 && cone != 7) return fail (PMX_EINVAL, "invalid torsion spring cone");
    for (int o = 0; o < 3; o++)
    {
      const int *ki = m->krpyidx + o*(T+1), *di = m->drpyidx + o*(T+1);
      if (ki[i] < 0 || ki[i+1] > m->krpycrv || (ki[i+1] != ki[i] && ki[i+1] < ki[i] + 2))
        return fail (PMX_EINVAL, "torsion spring curve index range out of bounds");
      if (di[i] < 0 || di[i+1] > m->drpycrv || di[i+1] < di[i] + 1)
        return fail (PMX_EINVAL, "torsion damper curve index range out of bounds");
    }
  }
  const int U = m->unsprnum;
  if (U < 0) return fail (PMX_EINVAL, "negative UNSPRING count");
  if (U > 0 && (!m->tspridx || !m->mspridx || !m->actidx || !m->unlim || !m->unent || !m->unop || !m->unabs ||
                !m->nsteps || !m->nfreq || !m->unaction))
    return fail (PMX_EINVAL, "UNSPRING arrays missing");
  for (int i = 0; i < U; i++)
  {
    if (m->tspridx[i] < 0 || m->tspridx[i+1] < m->tspridx[i] || m->mspridx[i] < 0 || m->mspridx[i+1] < m->mspridx[i] ||
        m->actidx[i] < 0 || m->actidx[i+1] < m->actidx[i])
      return fail (PMX_EINVAL, "UNSPRING list ranges invalid");
    if (m->nsteps[i] <= 0) return fail (PMX_EINVAL, "UNSPRING nsteps must be positive");
    if (m->unaction[i] < -1 || m->unaction[i] >= m->lcnum) return fail (PMX_EINVAL, "UNSPRING unloading curve out of range");
    int e = m->unent[i];
    if (e != PMX_HIS_STROKE && e != PMX_HIS_F && e != PMX_HIS_SF && e != PMX_HIS_FF) return fail (PMX_EINVAL, "invalid UNSPRING entity");
    if (m->unop[i] != PMX_OP_SUM && m->unop[i] != PMX_OP_MIN && m->unop[i] != PMX_OP_MAX)
      return fail (PMX_EINVAL, "invalid UNSPRING operator");
  }
  if (U > 0)
  {
    auto in = [&] (const int *l, int n) { for (int q = 0; q < n; q++) if (l[q] < 0 || l[q] >= S) return false; return true; };
    if ((m->tspridx[U] && (!m->tsprings || !in (m->tsprings, m->tspridx[U]))) ||
        (m->mspridx[U] && (!m->msprings || !in (m->msprings, m->mspridx[U]))) ||
        (m->actidx[U] && (!m->activate || !in (m->activate, m->actidx[U]))))
      return fail (PMX_EINVAL, "UNSPRING spring index out of range");
  }
  const int R = m->rstnum;
  if (R < 0) return fail (PMX_EINVAL, "negative restraint count");
  if (R > 0 && (!m->rstpart || !m->rstlin || !m->rstang)) return fail (PMX_EINVAL, "restraint arrays missing");
  for (int i = 0; i < R; i++)
    if (m->rstpart[i] < 0 || m->rstpart[i] >= N) return fail (PMX_EINVAL, "restrained particle index out of range");
  const int J = m->jnum;
  if (J < 0) return fail (PMX_EINVAL, "negative joint count");
  if (J > 0 && (!m->jpart || !m->jpoint || !m->jreac)) return fail (PMX_EINVAL, "joint arrays missing");
  for (int i = 0; i < J; i++)
    if (m->jpart[i] < 0 || m->jpart[i] >= N || m->jpart[J + i] < -1 || m->jpart[J + i] >= N)
      return fail (PMX_EINVAL, "joint particle index out of range");
  if (J > N)
    return fail (PMX_EUNSUPPORTED, "more ball joints than particles (the joint solve reads the mass of the particle "
                                   "numbered like each joint, joints.cpp:399)");
  return PMX_OK;
}

int jt_setup (pmx_ctx &c, const pmx_model *m)
{
  int rc;
  /* load curves */
  c.LC = m->lcnum; c.lccrv = m->lccrv;
  if ((rc = [&] () -> int {
        if (c.lcidx.alloc (std::max (m->lcnum + 1, 1)) || c.lcurve.alloc (std::max (2*m->lccrv, 1)))
          return fail (PMX_ENOMEM, "out of device memory (load curves)");
        if (m->lcnum > 0)
        {
          PMX_HIP (hipMemcpyAsync (c.lcidx.p, m->lcidx, (m->lcnum + 1) * sizeof (int), hipMemcpyHostToDevice, c.stream));
          PMX_HIP (hipMemcpyAsync (c.lcurve.p, m->lcurve, 2 * m->lccrv * sizeof (double), hipMemcpyHostToDevice, c.stream));
        }
        return PMX_OK; } ())) return rc;
  if ((rc = torsion_setup (c, m))) return rc;
  auto upi = [&] (dvec<int> &d, const int *src, size_t n) -> int {
    if (d.alloc (std::max (n, (size_t) 1)) != hipSuccess) return fail (PMX_ENOMEM, "out of device memory (joints)");
    if (n) PMX_HIP (hipMemcpyAsync (d.p, src, n * sizeof (int), hipMemcpyHostToDevice, c.stream));
    PMX_HIP (hipStreamSynchronize (c.stream));   /* the source may be freed once this returns */
    return PMX_OK;
  };
  auto upd = [&] (dvec<double> &d, const double *src, size_t n) -> int {
    if (d.alloc (std::max (n, (size_t) 1)) != hipSuccess) return fail (PMX_ENOMEM, "out of device memory (joints)");
    if (n) PMX_HIP (hipMemcpyAsync (d.p, src, n * sizeof (double), hipMemcpyHostToDevice, c.stream));
    PMX_HIP (hipStreamSynchronize (c.stream));
    return PMX_OK;
  };
  /* restraints: entries grouped by particle, entry order kept */
  c.RS = m->rstnum; c.RR = 0;
  if (c.RS > 0)
  {
    int R = c.RS;
    std::vector<int> ord (R);
    for (int i = 0; i < R; i++) ord[i] = i;
    std::stable_sort (ord.begin (), ord.end (), [&] (int a, int b) { return m->rstpart[a] < m->rstpart[b]; });
    std::vector<int> part (R), runoff (1, 0);
    std::vector<double> lin (9 * (size_t) R), ang (9 * (size_t) R);
    for (int q = 0; q < R; q++)
    {
      int i = ord[q];
      part[q] = m->rstpart[i];
      for (int d = 0; d < 9; d++) { lin[d*(size_t) R + q] = m->rstlin[d*(size_t) R + i]; ang[d*(size_t) R + q] = m->rstang[d*(size_t) R + i]; }
      if (q + 1 == R || m->rstpart[ord[q+1]] != part[q]) runoff.push_back (q + 1);
    }
    c.RR = (int) runoff.size () - 1;
    if ((rc = upi (c.rs_part, part.data (), R)) || (rc = upi (c.rs_runoff, runoff.data (), runoff.size ())) ||
        (rc = upd (c.rs_lin, lin.data (), lin.size ())) || (rc = upd (c.rs_ang, ang.data (), ang.size ())))
      return rc;
  }
  /* ball joints: independent sets (joints.cpp:83-177) */
  const int J = c.J = m->jnum;
  c.JS = c.JP = 0;
  if (J > 0)
  {
    const int *p1 = m->jpart, *p2 = m->jpart + J;
    std::map<int, std::set<int>> partadj;
    for (int i = 0; i < J; i++)
    {
      partadj[p1[i]].insert (i);
      if (p2[i] >= 0) partadj[p2[i]].insert (i);
    }
    std::vector<std::set<int>> jadj (J);
    for (auto &pa : partadj)
      for (int a : pa.second)
        for (int b : pa.second)
          if (a != b) jadj[a].insert (b);
    std::vector<std::vector<int>> adjv (J);
    for (int i = 0; i < J; i++) adjv[i].assign (jadj[i].begin (), jadj[i].end ());
    std::vector<char> left (J, 1);
    std::vector<int> soff (1, 0), sk, set (J), pos (J);
    for (int start = 0; start < J; start++)
    {
      if (!left[start]) continue;
      /* jmark: depth-first, pre-order, neighbours ascending */
      int s = (int) soff.size () - 1, n0 = (int) sk.size ();
      std::vector<std::pair<int, size_t>> st;
      auto visit = [&] (int j) { left[j] = 0; set[j] = s; pos[j] = (int) sk.size () - n0; sk.push_back (j); st.push_back ({j, 0}); };
      visit (start);
      while (!st.empty ())
      {
        int top = st.back ().first;
        size_t at = st.back ().second;
        if (at < adjv[top].size ())
        {
          st.back ().second = at + 1;
          int nb = adjv[top][at];
          if (left[nb]) visit (nb);
        }
        else st.pop_back ();
      }
      if ((int) sk.size () - n0 > PMX_JOINT_SET_MAX)
        return fail (PMX_EUNSUPPORTED, "a set of connected ball joints larger than " + std::to_string (PMX_JOINT_SET_MAX));
      soff.push_back ((int) sk.size ());
    }
    int JS = c.JS = (int) soff.size () - 1;
    /* per (joint, side): the other joints on that side's particle,
     * ascending, as (j << 1) | 1 where their block enters with a minus sign
     * (the particle is part1 of one and part2 of the other, joints.cpp:281-285) */
    std::vector<int> adjoff (1, 0), adj;
    for (int i = 0; i < J; i++)
      for (int k = 0; k < 2; k++)
      {
        int part = k ? p2[i] : p1[i];
        if (part >= 0)
          for (int j : partadj[part])
            if (j != i)
            {
              bool flip = (part == p1[i] && part == p2[j]) || (part == p2[i] && part == p1[j]);
              adj.push_back ((j << 1) | (flip ? 1 : 0));
            }
        adjoff.push_back ((int) adj.size ());
      }
    /* dense matrices and skyline work spaces per set */
    std::vector<long long> moff (JS), woff (JS);
    long long mt = 0, wt = 0;
    for (int s = 0; s < JS; s++)
    {
      long long n = 3LL * (soff[s+1] - soff[s]);
      moff[s] = mt; mt += n * n;
      woff[s] = wt; wt += n * (n - 1) + 2 * n;   /* L, U (at most n(n-1)/2 each), D, y */
    }
    /* per jointed particle: (joint << 1) | side in set order */
    std::map<int, std::vector<int>> plist;
    for (int q = 0; q < J; q++)
    {
      int i = sk[q];
      if (p1[i] >= 0) plist[p1[i]].push_back (i << 1);
      if (p2[i] >= 0) plist[p2[i]].push_back ((i << 1) | 1);
    }
    std::vector<int> ppart, poff (1, 0), pl;
    for (auto &e : plist)
    {
      ppart.push_back (e.first);
      pl.insert (pl.end (), e.second.begin (), e.second.end ());
      poff.push_back ((int) pl.size ());
    }
    c.JP = (int) ppart.size ();
    if ((rc = upi (c.jt_part, m->jpart, 2*(size_t) J)) || (rc = upi (c.jt_set, set.data (), J)) ||
        (rc = upi (c.jt_pos, pos.data (), J)) || (rc = upi (c.jt_adjoff, adjoff.data (), adjoff.size ())) ||
        (rc = upi (c.jt_adj, adj.data (), adj.size ())) || (rc = upi (c.jt_soff, soff.data (), soff.size ())) ||
        (rc = upi (c.jt_sk, sk.data (), sk.size ())) || (rc = upi (c.jt_ppart, ppart.data (), ppart.size ())) ||
        (rc = upi (c.jt_poff, poff.data (), poff.size ())) || (rc = upi (c.jt_plist, pl.data (), pl.size ())) ||
        (rc = upd (c.jt_point, m->jpoint, 3*(size_t) J)) || (rc = upd (c.jt_reac, m->jreac, 3*(size_t) J)))
      return rc;
    if (c.jt_moff.alloc (JS) || c.jt_woff.alloc (JS) || c.jt_mat.alloc (std::max (mt, 1LL)) ||
        c.jt_work.alloc (std::max (wt, 1LL)) || c.jt_rhs.alloc (3*(size_t) J) || c.jt_iw.alloc (3*(size_t) J + JS))
      return fail (PMX_ENOMEM, "out of device memory (joints)");
    PMX_HIP (hipMemcpyAsync (c.jt_moff.p, moff.data (), JS * sizeof (long long), hipMemcpyHostToDevice, c.stream));
    PMX_HIP (hipMemcpyAsync (c.jt_woff.p, woff.data (), JS * sizeof (long long), hipMemcpyHostToDevice, c.stream));
    PMX_HIP (hipStreamSynchronize (c.stream));
  }
  PMX_HIP (hipStreamSynchronize (c.stream));
  return PMX_OK;
}

int jt_download (pmx_ctx &c, pmx_model *m)
{
  int rc;
  if ((rc = torsion_download (c, m))) return rc;
  if (c.J > 0 && m->jreac)
    PMX_HIP (hipMemcpyAsync (m->jreac, c.jt_reac.p, 3*(size_t) c.J * sizeof (double), hipMemcpyDeviceToHost, c.stream));
  return PMX_OK;
}

int step_restrain_vel (pmx_ctx &c)
{
  if (c.RR == 0) return PMX_OK;
  hipLaunchKernelGGL (k_restrain_vel, grid (c.RR), dim3 (PMX_TPB), 0, c.stream, c.RR, c.N, c.RS, (const int*) c.rs_runoff.p,
                      (const int*) c.rs_part.p, (const double*) c.rs_lin.p, (const double*) c.rs_ang.p, c.linear.p,
                      c.angular.p, (const double*) c.rotation.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

int step_joints (pmx_ctx &c, double step0, double step1)
{
  if (c.J > 0)
  {
    JtArgs A;
    A.J = c.J; A.N = c.N;
    A.part = c.jt_part.p; A.set = c.jt_set.p; A.pos = c.jt_pos.p; A.adjoff = c.jt_adjoff.p; A.adj = c.jt_adj.p;
    A.soff = c.jt_soff.p; A.point = c.jt_point.p; A.moff = c.jt_moff.p; A.mat = c.jt_mat.p; A.rhs = c.jt_rhs.p;
    A.rotation = c.rotation.p; A.inverse = c.inverse.p; A.inertia = c.inertia.p; A.invm = c.invm.p; A.mass = c.mass.p;
    A.position = c.position.p; A.angular = c.angular.p; A.linear = c.linear.p; A.force = c.force.p; A.torque = c.torque.p;
    A.dtab = c.dtab.p; A.ctl = c.ctl.p;
    hipLaunchKernelGGL (k_joint_assemble, grid (c.J), dim3 (PMX_TPB), 0, c.stream, A, step0, step1);
    hipLaunchKernelGGL (k_joint_solve, grid (c.JS), dim3 (PMX_TPB), 0, c.stream, c.JS, (const int*) c.jt_soff.p,
                        (const int*) c.jt_sk.p, (const long long*) c.jt_moff.p, (const long long*) c.jt_woff.p,
                        (const double*) c.jt_mat.p, (const double*) c.jt_rhs.p, c.jt_work.p, c.jt_iw.p, c.jt_reac.p,
                        c.J, step0, step1, c.dflags.p, c.ctl.p);
    hipLaunchKernelGGL (k_joint_apply, grid (c.JP), dim3 (PMX_TPB), 0, c.stream, c.JP, c.N, c.J, (const int*) c.jt_ppart.p,
                        (const int*) c.jt_poff.p, (const int*) c.jt_plist.p, (const int*) c.jt_part.p,
                        (const double*) c.jt_point.p, (const double*) c.jt_reac.p, (const double*) c.rotation.p,
                        (const double*) c.position.p, c.force.p, c.torque.p, (const int*) c.ctl.p);
    PMX_HIP (hipGetLastError ());
  }
  if (c.RR > 0)
  {
    hipLaunchKernelGGL (k_restrain_frc, grid (c.RR), dim3 (PMX_TPB), 0, c.stream, c.RR, c.N, c.RS, (const int*) c.rs_runoff.p,
                        (const int*) c.rs_part.p, (const double*) c.rs_lin.p, (const double*) c.rs_ang.p, c.force.p,
                        c.torque.p, (const int*) c.ctl.p);
    PMX_HIP (hipGetLastError ());
  }
  return PMX_OK;
}

int step_times_table (pmx_ctx &c, const double *times, int n)
{
  if (c.ttab.alloc (std::max (n, 1)) != hipSuccess) return fail (PMX_ENOMEM, "out of device memory (step times)");
  if (n > 0) PMX_HIP (hipMemcpyAsync (c.ttab.p, times, n * sizeof (double), hipMemcpyHostToDevice, c.stream));
  PMX_HIP (hipStreamSynchronize (c.stream));
  return PMX_OK;
}
