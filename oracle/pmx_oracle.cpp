/*
 * pmx_oracle.cpp -- CPU oracle of Parmec's DEM step path.
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library as the checker; the product library
 * (parmec_amd/csrc) never includes, links or calls it.
 *
 * Scalar restatement of the reference kernels, following their structure:
 * k-d partitioning tree with leaf buffers (partition.ispc), per-particle
 * master contact lists with find-or-insert and per-step slave copies
 * (condet.ispc, forces.ispc), explicit integrator (dynamics.ispc), geometry
 * update (shapes.ispc), first half step and inertia inversion (parmec.ispc)
 * and the dem() stage order (parmec.cpp:2315-2530).  Floating point follows
 * the reference macros' operation order (macros.h) and is compiled with
 * -ffp-contract=off.  Canonical choices are listed in pmx_oracle.h.
 *
 * Built with OpenMP: the loops that are independent per item run in parallel;
 * every result is independent of the thread count.
 */
#include "pmx_oracle.h"
/* sin, cos and acos from +, -, *, / and sqrt, shared with the device path so
 * that the two agree bit for bit where the reference calls them (pinned
 * against libm in tests/test_oracle_trig.py) */
#include "../parmec_amd/csrc/pmx_trig.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

int fail (int code, const char *msg)
{
  g_err = msg;
  return code;
}

const double REAL_MAX_ = 1.7976931348623157E+308;

/* ---------------------------------------------------------------- math */

/* vectorisable exponential map, macros.h:1453-1534, including the reference's
 * v02 = Omega2*Omega3 term (macros.h:1505) */
void expmap (double O1, double O2, double O3, double L[9])
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
    angsq = std::sqrt (angsq);
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

/* column-major products, macros.h:906-1000 */
inline void nvmul (const double A[9], const double B[3], double C[3])
{
  C[0] = A[0]*B[0] + A[3]*B[1] + A[6]*B[2];
  C[1] = A[1]*B[0] + A[4]*B[1] + A[7]*B[2];
  C[2] = A[2]*B[0] + A[5]*B[1] + A[8]*B[2];
}
inline void tvmul (const double A[9], const double B[3], double C[3])
{
  C[0] = A[0]*B[0] + A[1]*B[1] + A[2]*B[2];
  C[1] = A[3]*B[0] + A[4]*B[1] + A[5]*B[2];
  C[2] = A[6]*B[0] + A[7]*B[1] + A[8]*B[2];
}
inline void nnmul (const double A[9], const double B[9], double C[9])
{
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++)
      C[3*c+r] = A[r]*B[3*c] + A[3+r]*B[3*c+1] + A[6+r]*B[3*c+2];
}
/* NTMUL: C = A B^T */
inline void ntmul (const double A[9], const double B[9], double C[9])
{
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++)
      C[3*c+r] = A[r]*B[c] + A[3+r]*B[3+c] + A[6+r]*B[6+c];
}
/* INVERT, macros.h:1033-1051: INV is left as it was where det = 0 */
inline void invert (const double F[9], double INV[9])
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

/* Eberly closest point on a triangle, condet.ispc:294-400 (detection form) and
 * 519-624 (update form; the two have the same operation sequence) */
double triangle_sphere (double ax, double ay, double az, double bx, double by, double bz,
                        double cx, double cy, double cz, double px, double py, double pz,
                        double r, double point[3], double normal[3])
{
  double edge0[3] = {bx-ax, by-ay, bz-az};
  double edge1[3] = {cx-ax, cy-ay, cz-az};
  double v0[3] = {ax-px, ay-py, az-pz};
  double a = edge0[0]*edge0[0] + edge0[1]*edge0[1] + edge0[2]*edge0[2];
  double b = edge0[0]*edge1[0] + edge0[1]*edge1[1] + edge0[2]*edge1[2];
  double c = edge1[0]*edge1[0] + edge1[1]*edge1[1] + edge1[2]*edge1[2];
  double d = edge0[0]*v0[0] + edge0[1]*v0[1] + edge0[2]*v0[2];
  double e = edge1[0]*v0[0] + edge1[1]*v0[1] + edge1[2]*v0[2];
  double det = a*c - b*b;
  double s = b*e - c*d;
  double t = b*d - a*e;
  auto clamp01 = [] (double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); };

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

  point[0] = ax + s*edge0[0] + t*edge1[0];
  point[1] = ay + s*edge0[1] + t*edge1[1];
  point[2] = az + s*edge0[2] + t*edge1[2];
  double q[3] = {px-point[0], py-point[1], pz-point[2]};
  double len = std::sqrt (q[0]*q[0] + q[1]*q[1] + q[2]*q[2]);
  double ilen = 1.0/len;
  normal[0] = ilen*q[0];
  normal[1] = ilen*q[1];
  normal[2] = ilen*q[2];
  return r - len;
}

/* ------------------------------------------------------- partitioning */

/* 10-bit to 30-bit spread, partition.ispc:58-66 */
inline unsigned expandbits (unsigned v)
{
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

/* Morton code, partition.ispc:69-94.  A zero extent along an axis maps that
 * axis to cell 0 (the reference divides by zero there). */
unsigned morton (double x, double y, double z, const double e[6])
{
  double w[3] = {e[3]-e[0], e[4]-e[1], e[5]-e[2]};
  double p[3] = {x, y, z};
  unsigned q[3];
  for (int k = 0; k < 3; k++)
  {
    double pk = w[k] > 0.0 ? (p[k]-e[k])/w[k] : 0.0;
    double s = pk * 1024.0;
    s = s > 0.0 ? s : 0.0;
    s = s < 1023.0 ? s : 1023.0;
    q[k] = expandbits ((unsigned)(int)s);
  }
  return q[0]*4 + q[1]*2 + q[2];
}

inline int lzc (unsigned x) { return x ? __builtin_clz (x) : 32; }

/* generalised common prefix length, partition.ispc:272-281 */
inline int delta (int i, unsigned codei, int j, int n, const unsigned *code)
{
  if (j < 0 || j > n-1) return -1;
  unsigned codej = code[j];
  if (codei == codej) return 32 + lzc ((unsigned)(i ^ j));
  return lzc (codei ^ codej);
}

struct RNode { int first, size, split, dimension; double coord; };
struct PNode { double coord; int dimension, left, right, leaf; };

struct Tree
{
  std::vector<PNode> node;
  std::vector<std::vector<int>> leaf;   /* stored ellipsoid indices (<= LSIZE) */
  std::vector<int> leafsize;            /* drop counts (may exceed LSIZE) */
};

/* recursive conversion, partition.ispc:525-569 */
void convert (const std::vector<RNode> &rt, int rnode, Tree &T, int pnode)
{
  T.node[pnode].coord = rt[rnode].coord;
  T.node[pnode].dimension = rt[rnode].dimension;
  auto mkleaf = [&] (int p) {
    T.node[p].coord = 0.0; T.node[p].dimension = -1; T.node[p].left = T.node[p].right = -1;
    T.node[p].leaf = (int) T.leaf.size ();
    T.leaf.emplace_back ();
  };
  if (rt[rnode].dimension >= 0)
  {
    int l = (int) T.node.size (); T.node.push_back (PNode ());
    int r = (int) T.node.size (); T.node.push_back (PNode ());
    T.node[pnode].left = l; T.node[pnode].right = r; T.node[pnode].leaf = -1;
    int j = rt[rnode].split;
    if (rt[rnode].first != j) convert (rt, j, T, l); else mkleaf (l);
    if (rt[rnode].first + rt[rnode].size - 1 != j + 1) convert (rt, j+1, T, r); else mkleaf (r);
  }
  else
  {
    T.node[pnode].left = T.node[pnode].right = -1;
    T.node[pnode].leaf = (int) T.leaf.size ();
    T.leaf.emplace_back ();
  }
}

/* partitioning_create, partition.ispc:663-746, with a stable sort */
void tree_create (int n, const double *cx, const double *cy, const double *cz, Tree &T)
{
  T.node.clear (); T.leaf.clear (); T.leafsize.clear ();
  if (n == 0) return;
  double e[6] = {REAL_MAX_, REAL_MAX_, REAL_MAX_, -REAL_MAX_, -REAL_MAX_, -REAL_MAX_};
  for (int i = 0; i < n; i++)
  {
    if (cx[i] < e[0]) e[0] = cx[i];
    if (cy[i] < e[1]) e[1] = cy[i];
    if (cz[i] < e[2]) e[2] = cz[i];
    if (cx[i] > e[3]) e[3] = cx[i];
    if (cy[i] > e[4]) e[4] = cy[i];
    if (cz[i] > e[5]) e[5] = cz[i];
  }
  std::vector<unsigned> code0 (n), code (n);
  std::vector<int> order (n);
  for (int i = 0; i < n; i++) { code0[i] = morton (cx[i], cy[i], cz[i], e); order[i] = i; }
  std::stable_sort (order.begin (), order.end (), [&] (int a, int b) { return code0[a] < code0[b]; });
  for (int i = 0; i < n; i++) code[i] = code0[order[i]];

  const double *point[3] = {cx, cy, cz};
  std::vector<RNode> rt (std::max (n-1, 1));
  if (n == 1)
  {
    rt[0].first = 0; rt[0].size = 1; rt[0].dimension = -1; rt[0].split = 0; rt[0].coord = 0.0;
  }
  /* radix_tree_create, partition.ispc:311-394 */
  for (int i = 0; i < n-1; i++)
  {
    unsigned codei = code[i];
    int dd = delta (i, codei, i+1, n, code.data ()) - delta (i, codei, i-1, n, code.data ());
    int d = (dd > 0) - (dd < 0);
    int dmin = delta (i, codei, i-d, n, code.data ());
    int lmax = 2;
    while (delta (i, codei, i+lmax*d, n, code.data ()) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax/2; t >= 1; t /= 2)
      if (delta (i, codei, i+(l+t)*d, n, code.data ()) > dmin) l = l+t;
    int j = i+l*d;
    rt[i].first = d > 0 ? i : j;
    rt[i].size = l+1;
    if (rt[i].size <= PMX_CUTOFF)
    {
      rt[i].split = 0; rt[i].coord = 0.0; rt[i].dimension = -1;
    }
    else
    {
      int dnode = delta (i, codei, j, n, code.data ());
      int s = 0;
      for (int t = (l+1)/2;; t = (t+1)/2)
      {
        if (delta (i, codei, i+(s+t)*d, n, code.data ()) > dnode) s = s+t;
        if (t == 1) break;
      }
      rt[i].split = i+s*d+std::min (d, 0);
      int dimension = (dnode-2)%3;
      double ret = REAL_MAX_;   /* mincoord, partition.ispc:297-308 */
      for (int k = rt[i].split+1; k < rt[i].first+rt[i].size; k++)
      {
        double val = point[dimension][order[k]];
        if (val < ret) ret = val;
      }
      rt[i].coord = ret;
      rt[i].dimension = dimension;
    }
  }
  T.node.push_back (PNode ());
  convert (rt, 0, T, 0);
  T.leafsize.assign (T.leaf.size (), 0);
}

/* partitioning_store / drop_ellipsoid, partition.ispc:600-660, 749-764;
 * returns the overflow count */
int tree_store (Tree &T, int n, const double *cx, const double *cy, const double *cz)
{
  if (n == 0) return 0;
  for (auto &l : T.leaf) l.clear ();
  std::fill (T.leafsize.begin (), T.leafsize.end (), 0);
  const double *c[3] = {cx, cy, cz};
  int repart = 0;
  for (int i = 0; i < n; i++)
  {
    int node = 0;
    while (T.node[node].dimension >= 0)
    {
      int d = T.node[node].dimension;
      node = c[d][i] < T.node[node].coord ? T.node[node].left : T.node[node].right;
    }
    int L = T.node[node].leaf;
    int j = T.leafsize[L]++;
    if (j < PMX_LSIZE) T.leaf[L].push_back (i);
    else repart++;
  }
  return repart;
}

/* leaves reached by an AABB, condet.ispc:98-113 and 412-426 */
void tree_reach (const Tree &T, int node, const double lo[3], const double hi[3], std::vector<int> &out)
{
  int d = T.node[node].dimension;
  if (d >= 0)
  {
    if (lo[d] <= T.node[node].coord) tree_reach (T, T.node[node].left, lo, hi, out);
    if (hi[d] >= T.node[node].coord) tree_reach (T, T.node[node].right, lo, hi, out);
  }
  else out.push_back (T.node[node].leaf);
}

/* -------------------------------------------------------------- contacts */

struct Contact
{
  int master, slave0, slave1, color0, color1;
  double point[3], normal[3], depth, force[3], state[3];
  double kcur = 0.0, ecur = 0.0;   /* stiffness and damping coefficient of the last force (forces.ispc:318-319) */
};

struct Slave
{
  int master0, master1;   /* particle, ellipsoid */
  int slave1;             /* the contact's slave ellipsoid or -(triangle+1) */
  double point[3], force[3];
};

/* partner-ordered identity of a contact entry (see pmx_oracle.h) */
inline unsigned long long side_key (int ell_or_negtri)
{
  if (ell_or_negtri >= 0) return (unsigned long long) ell_or_negtri;
  return 0x80000000ull | (unsigned long long) (-ell_or_negtri - 1);
}

/* pairing, forces.ispc:98-122 */
int pairing (int nummat, const int *pairs, int i, int j)
{
  int p[2] = {i, j}, start = 1, end = nummat;
  if (i > j) { p[0] = j; p[1] = i; }
  while (start < end)
  {
    int mid = ((end-start)>>1) + start;
    if (p[0] > pairs[2*mid]) { if (start == mid) break; start = mid; }
    else if (p[0] == pairs[2*mid])
    {
      if (p[1] > pairs[2*mid+1]) { if (start == mid) break; start = mid; }
      else if (p[1] == pairs[2*mid+1]) return mid;
      else end = mid;
    }
    else end = mid;
  }
  return 0;
}

} /* namespace */

/* ------------------------------------------------------------- context */

struct orc_ctx
{
  int nthreads = 1;
  int parnum = 0, ellnum = 0, trinum = 0, nodnum = 0, facnum = 0, pairnum = 0;
  std::vector<double> angular, linear, rotation, position, inertia, inverse, mass, invm, force, torque;
  std::vector<int> flags;
  std::vector<double> center, radii;
  std::vector<int> part, ellcol;
  std::vector<double> tri;
  std::vector<int> tricol, triobs;
  int prsnum = 0;
  std::vector<int> prspart, prskind, prscol;  /* [prsnum], [2][prsnum], [prsnum] */
  int prscols = 0;                            /* history columns */
  int obsnum = 0;
  std::vector<int> trirng;                    /* [2*obsnum] */
  std::vector<double> obspnt, obsang, obslin;  /* [3*obsnum] */
  std::vector<double> nodes;
  std::vector<int> nodpart, facnod, factri;
  std::vector<int> pairs, ikind;
  std::vector<double> iparam;
  /* springs (sorted by part1, part2) */
  int sprnum = 0;
  std::vector<int> sprtype, sprpart, spridx, dashidx, unidx, sprflg, unspring, sproffset;
  std::vector<double> sprpnt, spring, dashpot, unload, yield, sprdir, sprfric, sprkskn, sprsdsp, stroke0, stroke, sprfrc;
  int sprcrv = 0, dashcrv = 0, uncrv = 0;
  /* load curves (lcurve_from_time_series, parmec.cpp:1239-1269) */
  int lcnum = 0, lccrv = 0;
  std::vector<int> lcidx;
  std::vector<double> lcurve;
  /* torsion springs, sorted by (part1, part2); curve tables shared by the
   * roll, pitch and yaw curves */
  int trqsprnum = 0, krpycrv = 0, drpycrv = 0;
  std::vector<int> trqsprpart, krpyidx, drpyidx, trqcone;
  std::vector<double> trqzdir0, trqxdir0, krpy, drpy, trqzdir1, trqxdir1, trqrpy, trqrpytot, trqrpyspr;
  /* UNSPRING conditions (spring numbers of the sorted springs) */
  int unsprnum = 0;
  std::vector<int> tspridx, tsprings, mspridx, msprings, actidx, activate, unent, unop, unabs, nsteps, nfreq, unaction;
  std::vector<double> unlim;
  /* RESTRAIN */
  int rstnum = 0;
  std::vector<int> rstpart;
  std::vector<double> rstlin, rstang;
  /* BALL_JOINT */
  int jnum = 0;
  std::vector<int> jpart;
  std::vector<double> jpoint, jreac;
  double curtime = 0.0, curstep = 0.0;
  int stepnum = 0;
  std::vector<std::vector<Contact>> master;
  std::vector<std::vector<Slave>> slave;
  /* adaptive step: forces and springs form per particle kact[3], kmax,
   * emax, krot[6] (forces.ispc:453-596, 1203-1411) as adp [11][parnum] */
  bool adaptive = false;
  std::vector<double> adp;
  pmx_stats stats {};

  double *ang (int k) { return angular.data () + (size_t) k * parnum; }
  double *lin (int k) { return linear.data () + (size_t) k * parnum; }
  double *rot (int k) { return rotation.data () + (size_t) k * parnum; }
  double *pos (int k) { return position.data () + (size_t) k * parnum; }
  double *cen (int k) { return center.data () + (size_t) k * ellnum; }
  double *rad (int k) { return radii.data () + (size_t) k * ellnum; }
  double *trv (int v, int c) { return tri.data () + (size_t) (3*v+c) * trinum; }
  double *nod (int k) { return nodes.data () + (size_t) k * nodnum; }
};

namespace {

template <class T> void take (std::vector<T> &dst, const T *src, size_t n)
{
  dst.assign (src ? src : (const T*) nullptr, src ? src + n : (const T*) nullptr);
  if (!src) dst.assign (n, T ());
}

/* euler_task, parmec.ispc:224-266 */
void euler (orc_ctx &s, double step)
{
  #pragma omp parallel for num_threads(s.nthreads)
  for (int i = 0; i < s.parnum; i++)
  {
    double O[3], L1[9], DL[9], L2[9], o[3];
    O[0] = step*s.ang(0)[i];
    O[1] = step*s.ang(1)[i];
    O[2] = step*s.ang(2)[i];
    expmap (O[0], O[1], O[2], DL);
    for (int k = 0; k < 9; k++) L1[k] = s.rot(k)[i];
    nnmul (L1, DL, L2);
    for (int k = 0; k < 9; k++) s.rot(k)[i] = L2[k];
    nvmul (L2, O, o);
    s.ang(3)[i] = o[0]; s.ang(4)[i] = o[1]; s.ang(5)[i] = o[2];
    s.pos(0)[i] += step * s.lin(0)[i];
    s.pos(1)[i] += step * s.lin(1)[i];
    s.pos(2)[i] += step * s.lin(2)[i];
  }
}

/* invert, parmec.ispc:62-87 with INVERT, macros.h:1033-1051 */
void invert_inertia (orc_ctx &s)
{
  for (int i = 0; i < s.parnum; i++)
  {
    double F[9], INV[9], det;
    for (int k = 0; k < 9; k++) F[k] = s.inertia[(size_t) k*s.parnum + i];
    det = F[0]*(F[4]*F[8]-F[5]*F[7]) + F[3]*(F[2]*F[7]-F[1]*F[8]) + (F[1]*F[5]-F[2]*F[4])*F[6];
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
      for (int k = 0; k < 9; k++) s.inverse[(size_t) k*s.parnum + i] = INV[k];
    }
    s.invm[i] = 1.0 / s.mass[i];
  }
}

/* obstacles, obstacles.ispc:28-78: every moving obstacle advances its
 * triangles by the rigid motion of the step, rotation expmap(h*omega) about
 * its point x and translation h*v, then x += h*v.  Canonical: angular and
 * linear velocities are read with stride 3, as they are stored (the
 * reference strides them by 6, which reads another obstacle's values from
 * index 1 on). */
void obstacles (orc_ctx &s, double step)
{
  for (int i = 0; i < s.obsnum; i++)
  {
    double *x = &s.obspnt[3*i], *o = &s.obsang[3*i], *v = &s.obslin[3*i];
    double y[3], DL[9];
    y[0] = x[0] + step*v[0];
    y[1] = x[1] + step*v[1];
    y[2] = x[2] + step*v[2];
    expmap (step*o[0], step*o[1], step*o[2], DL);
    for (int j = s.trirng[2*i]; j < s.trirng[2*i+1]; j++)
      for (int a = 0; a < 3; a++)
      {
        double z[3] = {s.trv (a, 0)[j] - x[0], s.trv (a, 1)[j] - x[1], s.trv (a, 2)[j] - x[2]};
        s.trv (a, 0)[j] = DL[0]*z[0]+DL[3]*z[1]+DL[6]*z[2] + y[0];
        s.trv (a, 1)[j] = DL[1]*z[0]+DL[4]*z[1]+DL[7]*z[2] + y[1];
        s.trv (a, 2)[j] = DL[2]*z[0]+DL[5]*z[1]+DL[8]*z[2] + y[2];
      }
    x[0] = y[0];
    x[1] = y[1];
    x[2] = y[2];
  }
}

/* prescribe_acceleration, input.cpp:4247-4341: before dynamics, a prescribed
 * acceleration sets f = m a (angular: t = J alpha with the referential
 * inertia), a prescribed velocity zeroes the force (torque).  row: the
 * step's [prsnum][6] values (linear 0-2, angular 3-5), zeros when null */
void prescribe_acceleration (orc_ctx &s, const double *row)
{
  size_t N = s.parnum;
  for (int i = 0; i < s.prsnum; i++)
  {
    int j = s.prspart[i], lk = s.prskind[i], ak = s.prskind[s.prsnum + i];
    double a[3], b[3];
    int c = s.prscol[i];
    for (int d = 0; d < 3; d++) { a[d] = row ? row[6*c+d] : 0.0; b[d] = row ? row[6*c+3+d] : 0.0; }
    if (lk == 1) for (int d = 0; d < 3; d++) s.force[d*N + j] = s.mass[j] * a[d];
    else if (lk == 0) for (int d = 0; d < 3; d++) s.force[d*N + j] = 0.0;
    if (ak == 1)
    {
      double J[9], t[3];
      for (int k = 0; k < 9; k++) J[k] = s.inertia[k*N + j];
      nvmul (J, b, t);
      for (int d = 0; d < 3; d++) s.torque[d*N + j] = t[d];
    }
    else if (ak == 0) for (int d = 0; d < 3; d++) s.torque[d*N + j] = 0.0;
  }
}

/* prescribe_velocity, input.cpp:4344-4421: after dynamics, a prescribed
 * velocity is set (angular: spatial omega, body Omega = L^T omega with the
 * updated rotation) */
void prescribe_velocity (orc_ctx &s, const double *row)
{
  size_t N = s.parnum;
  for (int i = 0; i < s.prsnum; i++)
  {
    int j = s.prspart[i], lk = s.prskind[i], ak = s.prskind[s.prsnum + i];
    int c = s.prscol[i];
    if (lk == 0) for (int d = 0; d < 3; d++) s.linear[d*N + j] = row ? row[6*c+d] : 0.0;
    if (ak == 0)
    {
      double o[3], L[9], O[3];
      for (int d = 0; d < 3; d++) o[d] = row ? row[6*c+3+d] : 0.0;
      for (int k = 0; k < 9; k++) L[k] = s.rotation[k*N + j];
      tvmul (L, o, O);
      for (int d = 0; d < 3; d++) { s.angular[d*N + j] = O[d]; s.angular[(3+d)*N + j] = o[d]; }
    }
  }
}

/* obstaclev, input.cpp:4199-4244: obstacle velocities at a time, here read
 * from a host-evaluated table row ([obsnum][6]: angular 0-2, linear 3-5;
 * zero without a table) */
void obstaclev (orc_ctx &s, const double *row)
{
  for (int i = 0; i < s.obsnum; i++)
    for (int d = 0; d < 3; d++)
    {
      s.obsang[3*i+d] = row ? row[6*i+d] : 0.0;
      s.obslin[3*i+d] = row ? row[6*i+3+d] : 0.0;
    }
}

/* ellipsoids_task / nodes_task / faces_task, shapes.ispc:29-211 (spheres:
 * centre only; the orientation update applies to ellipsoids, which this path
 * does not support) */
void shapes (orc_ctx &s)
{
  #pragma omp parallel for num_threads(s.nthreads)
  for (int i = 0; i < s.ellnum; i++)
  {
    int j = s.part[i];
    double L[9], X[3], x[3], C[3], c[3];
    for (int k = 0; k < 9; k++) L[k] = s.rot(k)[j];
    for (int k = 0; k < 3; k++) { X[k] = s.pos(3+k)[j]; x[k] = s.pos(k)[j]; C[k] = s.cen(3+k)[i]; }
    C[0] -= X[0]; C[1] -= X[1]; C[2] -= X[2];
    c[0] = x[0] + (L[0]*C[0] + L[3]*C[1] + L[6]*C[2]);
    c[1] = x[1] + (L[1]*C[0] + L[4]*C[1] + L[7]*C[2]);
    c[2] = x[2] + (L[2]*C[0] + L[5]*C[1] + L[8]*C[2]);
    s.cen(0)[i] = c[0]; s.cen(1)[i] = c[1]; s.cen(2)[i] = c[2];
  }
  #pragma omp parallel for num_threads(s.nthreads)
  for (int i = 0; i < s.nodnum; i++)
  {
    int j = s.nodpart[i];
    double L[9], X[3], x[3], C[3], c[3];
    for (int k = 0; k < 9; k++) L[k] = s.rot(k)[j];
    for (int k = 0; k < 3; k++) { X[k] = s.pos(3+k)[j]; x[k] = s.pos(k)[j]; C[k] = s.nod(3+k)[i]; }
    C[0] -= X[0]; C[1] -= X[1]; C[2] -= X[2];
    c[0] = x[0] + (L[0]*C[0] + L[3]*C[1] + L[6]*C[2]);
    c[1] = x[1] + (L[1]*C[0] + L[4]*C[1] + L[7]*C[2]);
    c[2] = x[2] + (L[2]*C[0] + L[5]*C[1] + L[8]*C[2]);
    s.nod(0)[i] = c[0]; s.nod(1)[i] = c[1]; s.nod(2)[i] = c[2];
  }
  for (int i = 0; i < s.facnum; i++)
  {
    int j = s.factri[i];
    for (int v = 0; v < 3; v++)
    {
      int k = s.facnod[(size_t) v*s.facnum + i];
      for (int c = 0; c < 3; c++) s.trv(v, c)[j] = s.nod(c)[k];
    }
  }
}

/* update_existing, condet.ispc:627-700 */
int update_existing (orc_ctx &s)
{
  int bad = 0;
  #pragma omp parallel for num_threads(s.nthreads) reduction(+:bad)
  for (int l = 0; l < s.parnum; l++)
  {
    for (Contact &con : s.master[l])
    {
      int i = con.master, j = con.slave1;
      if (!(s.rad(1)[i] < 0.0)) { bad++; continue; }
      if (j < 0)
      {
        int u = -j-1;
        con.depth = triangle_sphere (s.trv(0,0)[u], s.trv(0,1)[u], s.trv(0,2)[u],
                                     s.trv(1,0)[u], s.trv(1,1)[u], s.trv(1,2)[u],
                                     s.trv(2,0)[u], s.trv(2,1)[u], s.trv(2,2)[u],
                                     s.cen(0)[i], s.cen(1)[i], s.cen(2)[i], s.rad(0)[i],
                                     con.point, con.normal);
      }
      else if (s.rad(1)[j] < 0.0)
      {
        double p[3] = {s.cen(0)[i], s.cen(1)[i], s.cen(2)[i]};
        double c[3] = {s.cen(0)[j], s.cen(1)[j], s.cen(2)[j]};
        double q[3] = {p[0]-c[0], p[1]-c[1], p[2]-c[2]};
        double len = std::sqrt (q[0]*q[0] + q[1]*q[1] + q[2]*q[2]);
        double ilen = 1.0/len;
        con.point[0] = 0.5*(p[0]+c[0]);
        con.point[1] = 0.5*(p[1]+c[1]);
        con.point[2] = 0.5*(p[2]+c[2]);
        con.normal[0] = ilen*q[0];
        con.normal[1] = ilen*q[1];
        con.normal[2] = ilen*q[2];
        con.depth = s.rad(0)[i]+s.rad(0)[j] - len;
      }
      else bad++;
    }
  }
  return bad;
}

struct Cand { int j; double point[3], normal[3], depth; };

/* test_ellipsoids / drop_ellipsoid, condet.ispc:98-290; the traversal of each
 * sphere runs in parallel, insertions run sequentially afterwards */
int test_ellipsoids (orc_ctx &s, const Tree &T)
{
  int n = s.ellnum, bad = 0;
  std::vector<std::vector<Cand>> cand (n);
  #pragma omp parallel for num_threads(s.nthreads) schedule(dynamic, 256) reduction(+:bad)
  for (int i = 0; i < n; i++)
  {
    double p[3] = {s.cen(0)[i], s.cen(1)[i], s.cen(2)[i]};
    double r[3] = {s.rad(0)[i], s.rad(1)[i], s.rad(2)[i]};
    double rx = r[0] > r[1] ? r[0] : r[1];
    rx = rx > r[2] ? rx : r[2];
    double lo[3] = {p[0]-rx, p[1]-rx, p[2]-rx};
    double hi[3] = {p[0]+rx, p[1]+rx, p[2]+rx};
    if (!(r[1] < 0.0)) { bad++; continue; }
    std::vector<int> leaves;
    tree_reach (T, 0, lo, hi, leaves);
    for (int L : leaves)
      for (int j : T.leaf[L])
      {
        if (!(s.rad(1)[j] < 0.0)) { bad++; continue; }
        double c[3] = {s.cen(0)[j], s.cen(1)[j], s.cen(2)[j]};
        double q[3] = {p[0]-c[0], p[1]-c[1], p[2]-c[2]};
        double len = std::sqrt (q[0]*q[0] + q[1]*q[1] + q[2]*q[2]);
        double ilen = len > 0.0 ? 1.0/len : 1.0;
        Cand k;
        k.j = j;
        k.point[0] = 0.5*(p[0]+c[0]);
        k.point[1] = 0.5*(p[1]+c[1]);
        k.point[2] = 0.5*(p[2]+c[2]);
        k.normal[0] = ilen*q[0];
        k.normal[1] = ilen*q[1];
        k.normal[2] = ilen*q[2];
        k.depth = rx + s.rad(0)[j] - len;
        if (k.depth > 0.0 && s.part[i] != s.part[j]) cand[i].push_back (k);
      }
  }
  for (int i = 0; i < n; i++)
  {
    int part = s.part[i];
    for (const Cand &k : cand[i])
    {
      int lpart = s.part[k.j];
      int index = std::min (part, lpart);
      std::vector<Contact> &ml = s.master[index];
      bool found = false;
      for (const Contact &con : ml)
      {
        if (part == index) { if (con.master == i && con.slave0 == lpart && con.slave1 == k.j) { found = true; break; } }
        else { if (con.master == k.j && con.slave0 == part && con.slave1 == i) { found = true; break; } }
      }
      if (found) continue;
      Contact con;
      double sg = part == index ? 1.0 : -1.0;
      con.master = part == index ? i : k.j;
      con.slave0 = part == index ? lpart : part;
      con.slave1 = part == index ? k.j : i;
      con.color0 = part == index ? s.ellcol[i] : s.ellcol[k.j];
      con.color1 = part == index ? s.ellcol[k.j] : s.ellcol[i];
      for (int d = 0; d < 3; d++)
      {
        con.point[d] = k.point[d];
        con.normal[d] = sg > 0 ? k.normal[d] : -k.normal[d];
        con.state[d] = 0.0;  /* canonical: zero tangential history at creation */
        con.force[d] = 0.0;
      }
      con.depth = k.depth;
      ml.push_back (con);
    }
  }
  return bad;
}

/* test_triangles / drop_triangle, condet.ispc:412-516 */
int test_triangles (orc_ctx &s, const Tree &T)
{
  int bad = 0;
  struct TC { int t, j; double point[3], normal[3], depth; };
  std::vector<std::vector<TC>> cand (s.trinum);
  #pragma omp parallel for num_threads(s.nthreads) schedule(dynamic, 16) reduction(+:bad)
  for (int t = 0; t < s.trinum; t++)
  {
    double ax = s.trv(0,0)[t], ay = s.trv(0,1)[t], az = s.trv(0,2)[t];
    double bx = s.trv(1,0)[t], by = s.trv(1,1)[t], bz = s.trv(1,2)[t];
    double cx = s.trv(2,0)[t], cy = s.trv(2,1)[t], cz = s.trv(2,2)[t];
    auto mn = [] (double a, double b, double c) { double d = a < b ? a : b; return d < c ? d : c; };
    auto mx = [] (double a, double b, double c) { double d = a > b ? a : b; return d > c ? d : c; };
    double lo[3] = {mn (ax,bx,cx), mn (ay,by,cy), mn (az,bz,cz)};
    double hi[3] = {mx (ax,bx,cx), mx (ay,by,cy), mx (az,bz,cz)};
    std::vector<int> leaves;
    tree_reach (T, 0, lo, hi, leaves);
    for (int L : leaves)
      for (int j : T.leaf[L])
      {
        if (!(s.rad(1)[j] < 0.0)) { bad++; continue; }
        TC k; k.t = t; k.j = j;
        k.depth = triangle_sphere (ax, ay, az, bx, by, bz, cx, cy, cz,
                                   s.cen(0)[j], s.cen(1)[j], s.cen(2)[j], s.rad(0)[j], k.point, k.normal);
        if (k.depth > 0.0) cand[t].push_back (k);
      }
  }
  for (int t = 0; t < s.trinum; t++)
    for (const TC &k : cand[t])
    {
      std::vector<Contact> &ml = s.master[s.part[k.j]];
      bool found = false;
      for (const Contact &con : ml)
        if (con.master == k.j && con.slave1 == -(t+1)) { found = true; break; }
      if (found) continue;
      Contact con;
      con.master = k.j;
      con.slave0 = s.triobs[t];
      con.slave1 = -(t+1);
      con.color0 = s.ellcol[k.j];
      con.color1 = s.tricol[t];
      for (int d = 0; d < 3; d++) { con.point[d] = k.point[d]; con.normal[d] = k.normal[d]; con.state[d] = 0.0; con.force[d] = 0.0; }
      con.depth = k.depth;
      ml.push_back (con);
    }
  return bad;
}

/* granural_force, forces.ispc:125-192 */
int granular_force (const double n[3], const double vij[3], double depth, int i, int j,
                    const orc_ctx &s, int ij, double state[3], double step, double f[5])
{
  double ma;
  if (j >= 0) ma = 1.0 / (s.invm[i] + s.invm[j]);
  else ma = s.mass[i];
  const double *ip = s.iparam.data ();
  int np = s.pairnum;
  double kn = ip[PMX_SPRING*np + ij];
  double en = ip[PMX_DAMPER*np + ij] * 2.0 * std::sqrt (kn*ma);
  double vn = vij[0]*n[0] + vij[1]*n[1] + vij[2]*n[2];
  double fn = kn*depth + en*vn;
  f[0] = fn*n[0];
  f[1] = fn*n[1];
  f[2] = fn*n[2];
  f[3] = kn;
  f[4] = en;
  double kt = kn * ip[PMX_KSKN*np + ij];
  double et = en * ip[PMX_KSKN*np + ij];
  double mus = ip[PMX_FRISTAT*np + ij];
  double mud = ip[PMX_FRIDYN*np + ij];
  double vt[3], ft[3];
  vt[0] = vij[0] - vn*n[0];
  vt[1] = vij[1] - vn*n[1];
  vt[2] = vij[2] - vn*n[2];
  state[0] += step*vt[0];
  state[1] += step*vt[1];
  state[2] += step*vt[2];
  ft[0] = kt*state[0] + et*vt[0];
  ft[1] = kt*state[1] + et*vt[1];
  ft[2] = kt*state[2] + et*vt[2];
  double len = std::sqrt (ft[0]*ft[0] + ft[1]*ft[1] + ft[2]*ft[2]);
  double lim = mus*fn;
  /* the reference divides by len below even when the tangential force is
   * exactly zero (f_n < 0 on a damped rebound): 0*(-inf) = NaN there.  With
   * len == 0 the tangential force is the zero vector, which is kept. */
  if (len > lim && len > 0.0)
  {
    lim = mud*fn/len;
    ft[0] *= lim;
    ft[1] *= lim;
    ft[2] *= lim;
  }
  f[0] += ft[0];
  f[1] += ft[1];
  f[2] += ft[2];
  return depth < 0.0 ? 1 : 0;
}

/* forces: contacts_task (forces.ispc:195-407), then contacts_acc_task
 * (forces.ispc:453-596) with the canonical summation order */
int forces (orc_ctx &s, double step, const double g[3])
{
  int np = s.parnum, bad = 0;
  #pragma omp parallel for num_threads(s.nthreads) schedule(dynamic, 256) reduction(+:bad)
  for (int i = 0; i < np; i++)
  {
    double oi[3] = {s.ang(3)[i], s.ang(4)[i], s.ang(5)[i]};
    double v[3] = {s.lin(0)[i], s.lin(1)[i], s.lin(2)[i]};
    double x[3] = {s.pos(0)[i], s.pos(1)[i], s.pos(2)[i]};
    std::vector<Contact> &ml = s.master[i];
    std::vector<Contact> kept;
    kept.reserve (ml.size ());
    for (Contact &con : ml)
    {
      double z[3], vi[3], vj[3], oj[3], vij[3];
      z[0] = con.point[0]-x[0];
      z[1] = con.point[1]-x[1];
      z[2] = con.point[2]-x[2];
      vi[0] = oi[1]*z[2]-oi[2]*z[1] + v[0];
      vi[1] = oi[2]*z[0]-oi[0]*z[2] + v[1];
      vi[2] = oi[0]*z[1]-oi[1]*z[0] + v[2];
      int j = con.slave0;
      if (j >= 0)
      {
        z[0] = con.point[0]-s.pos(0)[j];
        z[1] = con.point[1]-s.pos(1)[j];
        z[2] = con.point[2]-s.pos(2)[j];
        oj[0] = s.ang(3)[j]; oj[1] = s.ang(4)[j]; oj[2] = s.ang(5)[j];
        vj[0] = oj[1]*z[2]-oj[2]*z[1] + s.lin(0)[j];
        vj[1] = oj[2]*z[0]-oj[0]*z[2] + s.lin(1)[j];
        vj[2] = oj[0]*z[1]-oj[1]*z[0] + s.lin(2)[j];
      }
      else if (-j-2 >= 0)   /* moving obstacle, forces.ispc:265-280 */
      {
        int l = -j-2;
        z[0] = con.point[0]-s.obspnt[3*l];
        z[1] = con.point[1]-s.obspnt[3*l+1];
        z[2] = con.point[2]-s.obspnt[3*l+2];
        oj[0] = s.obsang[3*l]; oj[1] = s.obsang[3*l+1]; oj[2] = s.obsang[3*l+2];
        vj[0] = oj[1]*z[2]-oj[2]*z[1] + s.obslin[3*l];
        vj[1] = oj[2]*z[0]-oj[0]*z[2] + s.obslin[3*l+1];
        vj[2] = oj[0]*z[1]-oj[1]*z[0] + s.obslin[3*l+2];
      }
      else { vj[0] = vj[1] = vj[2] = 0.0; }
      vij[0] = vj[0] - vi[0];
      vij[1] = vj[1] - vi[1];
      vij[2] = vj[2] - vi[2];
      int ij = pairing (s.pairnum, s.pairs.data (), con.color0, con.color1);
      if (s.ikind[ij] != PMX_GRANULAR_FORCE) { bad++; continue; }
      double f[5];
      int gone = granular_force (con.normal, vij, con.depth, i, j, s, ij, con.state, step, f);
      con.force[0] = f[0]; con.force[1] = f[1]; con.force[2] = f[2];
      con.kcur = f[3]; con.ecur = f[4];
      if (!gone) kept.push_back (con);   /* canonical: state moves with its pair */
    }
    ml.swap (kept);
  }

  /* symmetric copy into slave lists, forces.ispc:383-405 */
  for (auto &sl : s.slave) sl.clear ();
  for (int i = 0; i < np; i++)
    for (const Contact &con : s.master[i])
      if (con.slave0 >= 0)
      {
        Slave sv;
        sv.master0 = i; sv.master1 = con.master; sv.slave1 = con.slave1;
        for (int d = 0; d < 3; d++) { sv.point[d] = con.point[d]; sv.force[d] = -con.force[d]; }
        s.slave[con.slave0].push_back (sv);
      }

  /* accumulation; each particle sums its contacts ordered by (partner identity,
   * own identity), master- and slave-side entries interleaved */
  #pragma omp parallel for num_threads(s.nthreads) schedule(dynamic, 256)
  for (int i = 0; i < np; i++)
  {
    struct E { unsigned long long k0, k1; const double *p; double f[3]; const double *n; double kc, ec; };
    std::vector<E> es;
    for (const Contact &con : s.master[i])
      es.push_back ({side_key (con.slave1), side_key (con.master), con.point, {con.force[0], con.force[1], con.force[2]},
                     con.normal, con.kcur, con.ecur});
    /* slave kcur, ecur: never written by the reference, 0 here; the reactions
     * on a meshed particle's triangles go triangle by triangle, spheres in
     * index order within a triangle */
    for (const Slave &sv : s.slave[i])
      es.push_back ({side_key (sv.slave1 < 0 ? sv.slave1 : sv.master1), side_key (sv.slave1 < 0 ? sv.master1 : sv.slave1),
                     sv.point, {sv.force[0], sv.force[1], sv.force[2]}, nullptr, 0.0, 0.0});
    std::sort (es.begin (), es.end (), [] (const E &a, const E &b) { return a.k0 < b.k0 || (a.k0 == b.k0 && a.k1 < b.k1); });
    double po[3] = {s.pos(0)[i], s.pos(1)[i], s.pos(2)[i]};
    double fs[3] = {0.0, 0.0, 0.0}, ts[3] = {0.0, 0.0, 0.0};
    double kact0[3] = {0.0, 0.0, 0.0}, kmax0 = 0.0, emax0 = 0.0, krot0[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, L[9];
    if (s.adaptive) for (int c = 0; c < 9; c++) L[c] = s.rot(c)[i];
    for (const E &e : es)
    {
      double a[3] = {e.p[0]-po[0], e.p[1]-po[1], e.p[2]-po[2]};
      fs[0] += e.f[0]; fs[1] += e.f[1]; fs[2] += e.f[2];
      ts[0] += a[1]*e.f[2] - a[2]*e.f[1];
      ts[1] += a[2]*e.f[0] - a[0]*e.f[2];
      ts[2] += a[0]*e.f[1] - a[1]*e.f[0];
      if (s.adaptive)   /* forces.ispc:505-533 (master), 542-577 (slave) */
      {
        double d[3];
        if (e.n) for (int c = 0; c < 3; c++) d[c] = e.n[c];
        else
        {
          double f[3] = {e.f[0], e.f[1], e.f[2]};
          double len = std::sqrt (f[0]*f[0] + f[1]*f[1] + f[2]*f[2]);
          if (len != 0.0)
          {
            len = 1.0/len;
            f[0] *= len; f[1] *= len; f[2] *= len;
          }
          else f[0] = f[1] = f[2] = 0.0;
          for (int c = 0; c < 3; c++) d[c] = f[c];
        }
        kact0[0] += std::fabs (d[0]);
        kact0[1] += std::fabs (d[1]);
        kact0[2] += std::fabs (d[2]);
        kmax0 = e.kc > kmax0 ? e.kc : kmax0;
        emax0 = e.ec > emax0 ? e.ec : emax0;
        double A[3];
        nvmul (L, a, A);
        double dot = A[0]*A[0] + A[1]*A[1] + A[2]*A[2];
        krot0[0] += dot - A[0]*A[0];
        krot0[1] += dot - A[1]*A[1];
        krot0[2] += dot - A[2]*A[2];
        krot0[3] += -A[0]*A[1];
        krot0[4] += -A[0]*A[2];
        krot0[5] += -A[1]*A[2];
      }
    }
    if (s.adaptive)
    {
      double v[11] = {kact0[0], kact0[1], kact0[2], kmax0, emax0, krot0[0], krot0[1], krot0[2], krot0[3], krot0[4], krot0[5]};
      for (int c = 0; c < 11; c++) s.adp[(size_t) c*np + i] = v[c];
    }
    double ma = s.mass[i];
    s.force[(size_t) 0*np + i] = fs[0] + ma * g[0];
    s.force[(size_t) 1*np + i] = fs[1] + ma * g[1];
    s.force[(size_t) 2*np + i] = fs[2] + ma * g[2];
    s.torque[(size_t) 0*np + i] = ts[0];
    s.torque[(size_t) 1*np + i] = ts[1];
    s.torque[(size_t) 2*np + i] = ts[2];
  }
  return bad;
}

/* dynamics_task, dynamics.ispc:30-169 */
void dynamics (orc_ctx &s, const double damping[6], double step0, double step1)
{
  int np = s.parnum;
  double half = 0.5*step0;
  double step = 0.5*(step0+step1);
  #pragma omp parallel for num_threads(s.nthreads)
  for (int i = 0; i < np; i++)
  {
    double O[3], o[3], v[3], L1[9], J[9], I[9], ma, im, f[3], t[3], T[3], DL[9], L2[9], A[3], B[3];
    double *force[3] = {&s.force[i], &s.force[(size_t) np + i], &s.force[(size_t) 2*np + i]};
    double *torque[3] = {&s.torque[i], &s.torque[(size_t) np + i], &s.torque[(size_t) 2*np + i]};
    for (int k = 0; k < 3; k++) { O[k] = s.ang(k)[i]; v[k] = s.lin(k)[i]; }
    for (int k = 0; k < 9; k++)
    {
      L1[k] = s.rot(k)[i];
      J[k] = s.inertia[(size_t) k*np + i];
      I[k] = s.inverse[(size_t) k*np + i];
    }
    im = s.invm[i];
    if (damping[0] != 0.0 || damping[1] != 0.0 || damping[2] != 0.0)
    {
      ma = s.mass[i];
      *force[0] -= ma * damping[0] * v[0];
      *force[1] -= ma * damping[1] * v[1];
      *force[2] -= ma * damping[2] * v[2];
    }
    if (damping[3] != 0.0 || damping[4] != 0.0 || damping[5] != 0.0)
    {
      o[0] = damping[3]*s.ang(3)[i];
      o[1] = damping[4]*s.ang(4)[i];
      o[2] = damping[5]*s.ang(5)[i];
      tvmul (L1, o, t);
      nvmul (J, t, T);
      nvmul (L1, T, t);
      *torque[0] -= t[0];
      *torque[1] -= t[1];
      *torque[2] -= t[2];
    }
    f[0] = *force[0]; f[1] = *force[1]; f[2] = *force[2];
    t[0] = *torque[0]; t[1] = *torque[1]; t[2] = *torque[2];
    tvmul (L1, t, T);
    expmap (-half*O[0], -half*O[1], -half*O[2], DL);
    nvmul (J, O, A);
    nvmul (DL, A, B);
    B[0] = B[0] + half*T[0];
    B[1] = B[1] + half*T[1];
    B[2] = B[2] + half*T[2];
    nvmul (I, B, A);
    nvmul (J, A, B);
    T[0] -= A[1]*B[2] - A[2]*B[1];
    T[1] -= A[2]*B[0] - A[0]*B[2];
    T[2] -= A[0]*B[1] - A[1]*B[0];
    T[0] *= step0; T[1] *= step0; T[2] *= step0;
    O[0] = O[0] + (I[0]*T[0] + I[3]*T[1] + I[6]*T[2]);
    O[1] = O[1] + (I[1]*T[0] + I[4]*T[1] + I[7]*T[2]);
    O[2] = O[2] + (I[2]*T[0] + I[5]*T[1] + I[8]*T[2]);
    im *= step0;
    v[0] = v[0] + im*f[0];
    v[1] = v[1] + im*f[1];
    v[2] = v[2] + im*f[2];
    expmap (step*O[0], step*O[1], step*O[2], DL);
    nnmul (L1, DL, L2);
    for (int k = 0; k < 9; k++) s.rot(k)[i] = L2[k];
    nvmul (L2, O, o);
    s.pos(0)[i] += step*v[0];
    s.pos(1)[i] += step*v[1];
    s.pos(2)[i] += step*v[2];
    for (int k = 0; k < 3; k++) { s.ang(k)[i] = O[k]; s.ang(3+k)[i] = o[k]; s.lin(k)[i] = v[k]; }
  }
}


/* ------------------------------------------------------------- springs */

/* look-up table linear spline with last-interval extrapolation,
 * forces.ispc:410-426 */
double look_up (const double *x, const double *y, int lo, int hi, double arg, double *slope)
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

/* slope of the spline searched along x (i == 0) or y (i == 1),
 * forces.ispc:429-450 */
double look_up_slope (int i, const double *x, const double *y, int lo, int hi, double arg, double *at)
{
  const double *tab[2] = {x, y};
  int h0 = hi;
  while (hi > lo+1)
  {
    int mid = lo + (hi-lo)/2;
    if (arg > tab[i][mid]) lo = mid;
    else hi = mid;
  }
  if (lo+1 == h0) lo--;
  if (at)
  {
    int j = (i == 0 ? 1 : 0);
    at[0] = tab[j][lo] + (tab[j][lo+1]-tab[j][lo])*(arg-tab[i][lo])/(tab[i][lo+1]-tab[i][lo]);
  }
  return (tab[1][lo+1]-tab[1][lo])/(tab[0][lo+1]-tab[0][lo]);
}

inline void vecskew (const double v[3], double A[9])
{
  A[0] = 0; A[1] = v[2]; A[2] = -v[1];
  A[3] = -v[2]; A[4] = 0; A[5] = v[0];
  A[6] = v[1]; A[7] = -v[0]; A[8] = 0;
}

inline void tnmul (const double A[9], const double B[9], double C[9])
{
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++)
      C[3*c+r] = A[3*r]*B[3*c] + A[3*r+1]*B[3*c+1] + A[3*r+2]*B[3*c+2];
}

/* extremum_eigenvalue, macros.h:1538-1593: the largest eigenvalue of a
 * symmetric 3x3 matrix, or with `minimum` the smallest; a diagonal matrix
 * returns its largest diagonal entry either way, as the reference does */
double extremum_eigenvalue (const double A[9], bool minimum)
{
  double p1 = A[3]*A[3] + A[6]*A[6] + A[7]*A[7];
  if (p1 == 0.0)
  {
    double r = A[0] > A[4] ? A[0] : A[4];
    return r > A[8] ? r : A[8];
  }
  double q = (A[0]+A[4]+A[8])/3.0;
  double p2 = (A[0]-q)*(A[0]-q) + (A[4]-q)*(A[4]-q) + (A[8]-q)*(A[8]-q) + 2*p1;
  double p = std::sqrt (p2 / 6.0);
  double invp = 1.0 / p;
  double B[9] = {invp*(A[0]-q), invp*A[1], invp*A[2], invp*A[3], invp*(A[4]-q), invp*A[5],
                 invp*A[6], invp*A[7], invp*(A[8]-q)};
  double det = B[0]*(B[4]*B[8]-B[5]*B[7]) + B[3]*(B[2]*B[7]-B[1]*B[8]) + B[6]*(B[1]*B[5]-B[2]*B[4]);
  double r = 0.5 * det;
  double phi;
  if (r <= -1.0) phi = 3.14159265358979323846 / 3.0;
  else if (r >= 1.0) phi = 0.0;
  else phi = pmxt::acos (r) / 3.0;
  if (minimum) return q + 2.0 * p * pmxt::cos (phi + (2*3.14159265358979323846/3));
  return q + 2.0 * p * pmxt::cos (phi);
}

/* equivalent point mass along a direction, parmec.ispc:423-482 (point1 and
 * point2 are passed as the springs task passes them: reference point minus
 * the particle's reference position) */
double eqm (const orc_ctx &s, int part1, const double point1[3], int part2, const double point2[3], const double dir[3])
{
  size_t N = s.parnum;
  double Jiv[9], im, A[3], Ask[9], tmp[9], Miv[9];
  for (int k = 0; k < 9; k++) Jiv[k] = s.inverse[k*N + part1];
  im = s.invm[part1];
  A[0] = s.position[3*N + part1] - point1[0];
  A[1] = s.position[4*N + part1] - point1[1];
  A[2] = s.position[5*N + part1] - point1[2];
  vecskew (A, Ask);
  tnmul (Ask, Jiv, tmp);
  nnmul (tmp, Ask, Miv);
  Miv[0] += im; Miv[4] += im; Miv[8] += im;
  if (part2 >= 0)
  {
    for (int k = 0; k < 9; k++) Jiv[k] = s.inverse[k*N + part2];
    im = s.invm[part2];
    A[0] = s.position[3*N + part2] - point2[0];
    A[1] = s.position[4*N + part2] - point2[1];
    A[2] = s.position[5*N + part2] - point2[2];
    vecskew (A, Ask);
    tnmul (Ask, Jiv, tmp);
    nnmul (tmp, Ask, Jiv);
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

/* springs_task, forces.ispc:641-1412, run as one task: per spring force and
 * torque, accumulated locally per part1 run and per part2 run and flushed to
 * the particles where the run ends (forces.ispc:1203-1411) */
int springs (orc_ctx &s, double step)
{
  int S = s.sprnum;
  size_t N = s.parnum, SS = S;
  auto P = [&] (int p, int c) { return s.sprpnt.data () + (size_t) (6*p+c) * SS; };
  double *force = s.force.data (), *torque = s.torque.data ();
  double fj[3] = {0, 0, 0}, tj[3] = {0, 0, 0}, fk[3] = {0, 0, 0}, tk[3] = {0, 0, 0};
  int j0 = -1, k0 = -1;
  /* adaptive step: the run sums of forces.ispc:725-730, flushed with the
   * forces (1216-1241, 1288-1321, 1377-1411) */
  const bool adaptive = s.adaptive;
  double *adp = adaptive ? s.adp.data () : nullptr;
  double kact_j[3] = {0, 0, 0}, kact_k[3] = {0, 0, 0}, kmax_j = 0.0, kmax_k = 0.0, emax_j = 0.0, emax_k = 0.0;
  double krot_j[6] = {0, 0, 0, 0, 0, 0}, krot_k[6] = {0, 0, 0, 0, 0, 0};
  auto flush = [&] (int b, double ka[3], double &km, double &em, double kr[6]) {
    for (int c = 0; c < 3; c++) adp[c*N + b] += ka[c];
    adp[3*N + b] = km > adp[3*N + b] ? km : adp[3*N + b];
    adp[4*N + b] = em > adp[4*N + b] ? em : adp[4*N + b];
    for (int c = 0; c < 6; c++) adp[(5+c)*N + b] += kr[c];
  };
  auto reset = [] (double ka[3], double &km, double &em, double kr[6]) {
    ka[0] = ka[1] = ka[2] = 0.0;
    km = em = 0.0;
    for (int c = 0; c < 6; c++) kr[c] = 0.0;
  };
  auto addterms = [] (const double L[9], const double arm[3], const double z[3], double stf, double dmp,
                      double ka[3], double &km, double &em, double kr[6]) {
    for (int c = 0; c < 3; c++) ka[c] += std::fabs (z[c]);
    km = stf > km ? stf : km;
    em = dmp > em ? dmp : em;
    double A[3];
    tvmul (L, arm, A);
    double dot = A[0]*A[0] + A[1]*A[1] + A[2]*A[2];
    kr[0] += dot-A[0]*A[0];
    kr[1] += dot-A[1]*A[1];
    kr[2] += dot-A[2]*A[2];
    kr[3] += -A[0]*A[1];
    kr[4] += -A[0]*A[2];
    kr[5] += -A[1]*A[2];
  };
  double oj[3] = {}, vj[3] = {}, xj[3] = {}, Xj[3] = {}, Lj[9] = {}, ok[3] = {}, vk[3] = {}, xk[3] = {}, Xk[3] = {}, Lk[9] = {};
  const double *sx = s.spring.data (), *sy = s.spring.data () + s.sprcrv;
  const double *dx = s.dashpot.data (), *dy = s.dashpot.data () + s.dashcrv;
  const double *ux = s.unload.data (), *uy = s.unload.data () + s.uncrv;
  for (int i = 0; i < S; i++)
  {
    double len = 0.0, inv, velocity, spring_force = 0.0, dashpot_force = 0.0, total_force;
    double Yj[3], Yk[3] = {0, 0, 0}, p[3], q[3], z[3] = {0, 0, 0}, w[3], frc[3], trq[3], u[3], a[3], b[3];
    double stiffness = 0.0, damping = 0.0;
    int j = s.sprpart[i], k = s.sprpart[SS + i];
    if (j0 != j)
    {
      for (int c = 0; c < 3; c++)
      {
        oj[c] = s.angular[(3+c)*N + j]; vj[c] = s.linear[c*N + j];
        xj[c] = s.position[c*N + j]; Xj[c] = s.position[(3+c)*N + j];
      }
      for (int c = 0; c < 9; c++) Lj[c] = s.rotation[c*N + j];
    }
    for (int c = 0; c < 3; c++) Yj[c] = P(0, 3+c)[i] - Xj[c];
    p[0] = xj[0] + (Lj[0]*Yj[0] + Lj[3]*Yj[1] + Lj[6]*Yj[2]);
    p[1] = xj[1] + (Lj[1]*Yj[0] + Lj[4]*Yj[1] + Lj[7]*Yj[2]);
    p[2] = xj[2] + (Lj[2]*Yj[0] + Lj[5]*Yj[1] + Lj[8]*Yj[2]);
    for (int c = 0; c < 3; c++) P(0, c)[i] = p[c];
    for (int c = 0; c < 3; c++) u[c] = vj[c] * -1.0;
    for (int c = 0; c < 3; c++) a[c] = p[c] - xj[c];
    u[0] -= oj[1]*a[2] - oj[2]*a[1];
    u[1] -= oj[2]*a[0] - oj[0]*a[2];
    u[2] -= oj[0]*a[1] - oj[1]*a[0];
    int flg = s.sprflg[i];
    if (k >= 0)
    {
      if (k0 != k)
      {
        for (int c = 0; c < 3; c++)
        {
          ok[c] = s.angular[(3+c)*N + k]; vk[c] = s.linear[c*N + k];
          xk[c] = s.position[c*N + k]; Xk[c] = s.position[(3+c)*N + k];
        }
        for (int c = 0; c < 9; c++) Lk[c] = s.rotation[c*N + k];
      }
      if (flg & PMX_SPRDIR_PROJECT)
      {
        for (int c = 0; c < 3; c++) w[c] = s.sprdir[(3+c)*SS + i];
        nvmul (Lk, w, z);
        for (int c = 0; c < 3; c++) s.sprdir[c*SS + i] = z[c];
      }
      for (int c = 0; c < 3; c++) Yk[c] = P(1, 3+c)[i] - Xk[c];
      q[0] = xk[0] + (Lk[0]*Yk[0] + Lk[3]*Yk[1] + Lk[6]*Yk[2]);
      q[1] = xk[1] + (Lk[1]*Yk[0] + Lk[4]*Yk[1] + Lk[7]*Yk[2]);
      q[2] = xk[2] + (Lk[2]*Yk[0] + Lk[5]*Yk[1] + Lk[8]*Yk[2]);
      for (int c = 0; c < 3; c++) P(1, c)[i] = q[c];
      for (int c = 0; c < 3; c++) u[c] += vk[c];
      for (int c = 0; c < 3; c++) b[c] = q[c] - xk[c];
      u[0] += ok[1]*b[2] - ok[2]*b[1];
      u[1] += ok[2]*b[0] - ok[0]*b[2];
      u[2] += ok[0]*b[1] - ok[1]*b[0];
    }
    else
    {
      for (int c = 0; c < 3; c++) q[c] = P(1, c)[i];
      if (flg & PMX_SPRDIR_PROJECT) for (int c = 0; c < 3; c++) z[c] = s.sprdir[c*SS + i];
    }
    switch (flg & (PMX_SPRDIR_FOLLOWER|PMX_SPRDIR_CONSTANT|PMX_SPRDIR_PLANAR|PMX_SPRDIR_PROJECT))
    {
    case PMX_SPRDIR_FOLLOWER:
      for (int c = 0; c < 3; c++) w[c] = q[c] - p[c];
      len = std::sqrt (w[0]*w[0] + w[1]*w[1] + w[2]*w[2]);
      inv = len > 0.0 ? 1.0/len : 0.0;
      for (int c = 0; c < 3; c++) { z[c] = inv*w[c]; s.sprdir[c*SS + i] = z[c]; }
      break;
    case PMX_SPRDIR_CONSTANT:
      for (int c = 0; c < 3; c++) { w[c] = q[c] - p[c]; z[c] = s.sprdir[c*SS + i]; }
      len = w[0]*z[0] + w[1]*z[1] + w[2]*z[2];
      break;
    case PMX_SPRDIR_PLANAR:
      for (int c = 0; c < 3; c++) w[c] = q[c] - p[c];
      inv = w[0]*s.sprdir[3*SS + i] + w[1]*s.sprdir[4*SS + i] + w[2]*s.sprdir[5*SS + i];
      for (int c = 0; c < 3; c++) w[c] -= inv*s.sprdir[(3+c)*SS + i];
      len = std::sqrt (w[0]*w[0] + w[1]*w[1] + w[2]*w[2]);
      inv = len > 0.0 ? 1.0/len : 0.0;
      for (int c = 0; c < 3; c++) { z[c] = inv*w[c]; s.sprdir[c*SS + i] = z[c]; }
      break;
    case PMX_SPRDIR_PROJECT:
      for (int c = 0; c < 3; c++) w[c] = p[c] - q[c];
      len = w[0]*z[0] + w[1]*z[1] + w[2]*z[2];
      break;
    }
    /* the reference evaluates the relative normal velocity only on the
     * force-law branch; evaluated here for every spring (its friction uses it) */
    velocity = z[0]*u[0] + z[1]*u[1] + z[2]*u[2];
    /* forces.ispc:915-919: the stroke offset, a load curve of time */
    double offset = 0.0;
    int ilc = s.sproffset.empty () ? -1 : s.sproffset[i];
    if (ilc >= 0)
    {
      double slope;
      offset = look_up (s.lcurve.data (), s.lcurve.data () + s.lccrv, s.lcidx[ilc], s.lcidx[ilc+1], s.curtime, &slope);
    }
    int unspr = s.unspring[i];
    double *stroke = s.stroke.data (), *yld = s.yield.data ();
    if (unspr <= -2)
    {
      if (s.sprtype[i] == PMX_SPRING_NONLINEAR_ELASTIC)
      {
        stroke[i] = len - s.stroke0[i];
        spring_force = look_up (sx, sy, s.spridx[i], s.spridx[i+1], stroke[i] - offset, &stiffness);
      }
      else
      {
        if ((s.sprflg[i] & PMX_SPRING_YIELDED) == 0)
        {
          stroke[i] = len - s.stroke0[i];
          spring_force = look_up (sx, sy, s.spridx[i], s.spridx[i+1], stroke[i] - offset, &stiffness);
          if (spring_force < yld[i] || spring_force > yld[SS + i])
          {
            s.sprflg[i] |= PMX_SPRING_YIELDED;
            if (spring_force < 0) { stroke[SS + i] = stroke[i]; stroke[2*SS + i] = 0.0; yld[SS + i] = 0.0; }
            else { stroke[SS + i] = 0.0; stroke[2*SS + i] = stroke[i]; yld[i] = 0.0; }
          }
        }
        else
        {
          double s0 = stroke[i];
          double s1 = len - s.stroke0[i];
          double ds = s1 - s0;
          double f0 = s.sprfrc[SS + i];
          bool loading = false;
          stroke[i] = s1;
          if (f0 * ds > 0)
          {
            if (s.sprflg[i] & PMX_SPRING_UNLOADING)
            {
              double dfds = look_up_slope (1, ux, uy, s.unidx[i], s.unidx[i+1], f0, nullptr);
              spring_force = f0 + dfds*ds;
              stiffness = dfds;   /* the reference leaves it unset on this branch */
              if (spring_force < yld[i] || spring_force > yld[SS + i])
              {
                s.sprflg[i] &= ~PMX_SPRING_UNLOADING;
                loading = true;
              }
            }
            else loading = true;
          }
          else
          {
            double at;
            double dfds = look_up_slope (1, ux, uy, s.unidx[i], s.unidx[i+1], f0, &at);
            stiffness = dfds;
            spring_force = f0 + dfds*ds;
            if (spring_force * f0 < 0.0)
              spring_force = look_up (ux, uy, s.unidx[i], s.unidx[i+1], ds+at, &stiffness);
            if (spring_force < yld[i] || spring_force > yld[SS + i])
            {
              s.sprflg[i] &= ~PMX_SPRING_UNLOADING;
              loading = true;
            }
            else s.sprflg[i] |= PMX_SPRING_UNLOADING;
          }
          if (loading)
          {
            if (ds < 0) stroke[SS + i] += ds;
            else stroke[2*SS + i] += ds;
            spring_force = look_up (sx, sy, s.spridx[i], s.spridx[i+1], ds < 0 ? stroke[SS + i] : stroke[2*SS + i], &stiffness);
            if (ds < 0) yld[i] = spring_force;
            else yld[SS + i] = spring_force;
          }
        }
      }
      if (s.dashidx[i]+1 == s.dashidx[i+1])
      {
        double mass = eqm (s, j, Yj, k, Yk, z);
        double ratio = dx[s.dashidx[i]];
        damping = ratio * 2.0 * std::sqrt (std::fabs (stiffness) * mass);
        dashpot_force = velocity * damping;
      }
      else dashpot_force = look_up (dx, dy, s.dashidx[i], s.dashidx[i+1], velocity, &damping);
      total_force = spring_force + (spring_force == 0.0 ? 0.0 : dashpot_force);
    }
    else if (unspr == -1)
    {
      stroke[i] = len - s.stroke0[i];
      stiffness = 0.;
      damping = 0.;
      spring_force = dashpot_force = total_force = 0.;
    }
    else
    {
      /* forces.ispc:1080-1115: unloading along load curve unspr */
      double s0 = stroke[i];
      double s1 = len - s.stroke0[i];
      double ds = s1 - s0;
      double f0 = s.sprfrc[SS + i];
      double sgn = f0 > 0. ? 1. : -1.;
      stroke[i] = s1;
      double at;
      double dfds = look_up_slope (1, s.lcurve.data (), s.lcurve.data () + s.lccrv, s.lcidx[unspr], s.lcidx[unspr+1], f0, &at);
      stiffness = dfds;
      spring_force = f0 - sgn*std::fabs (dfds*ds);
      if (spring_force * f0 < 0.0)
      {
        s.unspring[i] = -1;
        spring_force = 0.;
      }
      damping = 0.;
      dashpot_force = 0.;
      total_force = spring_force;
    }

    for (int c = 0; c < 3; c++) frc[c] = z[c] * total_force;
    s.sprfrc[i] = total_force;
    s.sprfrc[SS + i] = spring_force;
    double fric = s.sprfric[i];
    if (fric > 0.0)
    {
      double us[3] = {u[0] - velocity*z[0], u[1] - velocity*z[1], u[2] - velocity*z[2]};
      double flim = fric*std::fabs (total_force);
      double fs[3] = {0.0, 0.0, 0.0};
      double kskn = s.sprkskn[i];
      double *sd = s.sprsdsp.data ();
      if (kskn == 0.0)
      {
        double ulen = std::sqrt (us[0]*us[0] + us[1]*us[1] + us[2]*us[2]);
        if (ulen > 0.0)
        {
          double coef = flim/ulen;
          for (int c = 0; c < 3; c++) fs[c] = coef*us[c];
        }
      }
      else
      {
        if (spring_force == 0.0) { sd[i] = 0.0; sd[SS + i] = 0.0; sd[2*SS + i] = 0.0; }
        double ds[3] = {sd[i] + us[0]*step, sd[SS + i] + us[1]*step, sd[2*SS + i] + us[2]*step};
        double c0 = kskn * std::fabs (stiffness);
        double c1 = spring_force == 0.0 ? 0.0 : kskn * std::fabs (damping);
        for (int c = 0; c < 3; c++) fs[c] = c0 * ds[c] + c1 * us[c];
        double flen = std::sqrt (fs[0]*fs[0] + fs[1]*fs[1] + fs[2]*fs[2]);
        if (flen > flim)
        {
          double coef = flim/flen;
          for (int c = 0; c < 3; c++) fs[c] *= coef;
          double dl0 = std::sqrt (ds[0]*ds[0] + ds[1]*ds[1] + ds[2]*ds[2]);
          coef = flim/(c0*dl0);
          for (int c = 0; c < 3; c++) sd[c*SS + i] = coef*ds[c];
        }
        else for (int c = 0; c < 3; c++) sd[c*SS + i] = ds[c];
      }
      s.sprfrc[2*SS + i] = std::sqrt (fs[0]*fs[0] + fs[1]*fs[1] + fs[2]*fs[2]);
      for (int c = 0; c < 3; c++) frc[c] += fs[c];
    }
    trq[0] = a[1]*frc[2] - a[2]*frc[1];
    trq[1] = a[2]*frc[0] - a[0]*frc[2];
    trq[2] = a[0]*frc[1] - a[1]*frc[0];
    if (j0 != j && j0 >= 0)
    {
      for (int c = 0; c < 3; c++) { force[c*N + j0] += fj[c]; torque[c*N + j0] += tj[c]; fj[c] = frc[c]; tj[c] = trq[c]; }
      if (adaptive) { flush (j0, kact_j, kmax_j, emax_j, krot_j); reset (kact_j, kmax_j, emax_j, krot_j); }
    }
    else
      for (int c = 0; c < 3; c++) { fj[c] += frc[c]; tj[c] += trq[c]; }
    if (adaptive && total_force != 0.0)
    {
      stiffness = std::fabs (stiffness);
      damping = std::fabs (damping);
      addterms (Lj, a, z, stiffness, damping, kact_j, kmax_j, emax_j, krot_j);
    }
    if (k >= 0)
    {
      trq[0] = b[1]*frc[2] - b[2]*frc[1];
      trq[1] = b[2]*frc[0] - b[0]*frc[2];
      trq[2] = b[0]*frc[1] - b[1]*frc[0];
      if (k0 != k && k0 >= 0)
      {
        for (int c = 0; c < 3; c++) { force[c*N + k0] -= fk[c]; torque[c*N + k0] -= tk[c]; fk[c] = frc[c]; tk[c] = trq[c]; }
        if (adaptive) { flush (k0, kact_k, kmax_k, emax_k, krot_k); reset (kact_k, kmax_k, emax_k, krot_k); }
      }
      else
        for (int c = 0; c < 3; c++) { fk[c] += frc[c]; tk[c] += trq[c]; }
      /* part2's arm rotated by part1's L, as forces.ispc:1343 does */
      if (adaptive && total_force != 0.0) addterms (Lj, b, z, stiffness, damping, kact_k, kmax_k, emax_k, krot_k);
      k0 = k;
    }
    j0 = j;
  }
  if (j0 >= 0)
  {
    for (int c = 0; c < 3; c++) { force[c*N + j0] += fj[c]; torque[c*N + j0] += tj[c]; }
    if (adaptive) flush (j0, kact_j, kmax_j, emax_j, krot_j);
  }
  if (k0 >= 0)
  {
    for (int c = 0; c < 3; c++) { force[c*N + k0] -= fk[c]; torque[c*N + k0] -= tk[c]; }
    if (adaptive) flush (k0, kact_k, kmax_k, emax_k, krot_k);
  }
  return 0;
}

/* --------------------------------------------------------- torsion springs */

/* trqspr_task, forces.ispc:1415-2055, as one task over all torsion springs
 * (sorted by part1, part2): curve torques and dampers per axis or on a cone,
 * local sums flushed where part1 or part2 changes */
void torsion_springs (orc_ctx &s, double step)
{
  const int T = s.trqsprnum;
  const size_t N = s.parnum, TT = T;
  double *torque = s.torque.data ();
  const double *kx = s.krpy.data (), *ky = kx + s.krpycrv, *dx = s.drpy.data (), *dy = dx + s.drpycrv;
  auto kidx = [&] (int o, int i) { return s.krpyidx[o*(TT+1) + i]; };
  auto didx = [&] (int o, int i) { return s.drpyidx[o*(TT+1) + i]; };
  double oj[3] = {0, 0, 0}, Lj[9], ok[3] = {0, 0, 0};   /* loaded at the first spring (j0, k0 start at -1) */
  int j0 = -1, k0 = -1;
  double tj[3] = {0, 0, 0}, tk[3] = {0, 0, 0};
  double krot_j[6] = {0, 0, 0, 0, 0, 0}, krot_k[6] = {0, 0, 0, 0, 0, 0};
  double *adp = s.adaptive ? s.adp.data () : nullptr;
  for (int i = 0; i < T; i++)
  {
    double zdir0[3], xdir0[3], ydir0[3], rot1j[9];
    int j = s.trqsprpart[i], k = s.trqsprpart[TT + i];
    if (j0 != j)
    {
      for (int c = 0; c < 3; c++) oj[c] = s.angular[(3+c)*N + j];
      for (int c = 0; c < 9; c++) Lj[c] = s.rotation[c*N + j];
    }
    for (int c = 0; c < 3; c++) { zdir0[c] = s.trqzdir0[c*TT + i]; xdir0[c] = s.trqxdir0[c*TT + i]; }
    ydir0[0] = zdir0[1]*xdir0[2] - zdir0[2]*xdir0[1];
    ydir0[1] = zdir0[2]*xdir0[0] - zdir0[0]*xdir0[2];
    ydir0[2] = zdir0[0]*xdir0[1] - zdir0[1]*xdir0[0];
    nvmul (Lj, xdir0, rot1j);
    nvmul (Lj, ydir0, rot1j+3);
    nvmul (Lj, zdir0, rot1j+6);
    for (int c = 0; c < 3; c++) { s.trqzdir1[c*TT + i] = rot1j[6+c]; s.trqxdir1[c*TT + i] = rot1j[c]; }
    if (k >= 0) { if (k0 != k) for (int c = 0; c < 3; c++) ok[c] = s.angular[(3+c)*N + k]; }
    else ok[0] = ok[1] = ok[2] = 0.0;
    double ojk[3] = {oj[0]-ok[0], oj[1]-ok[1], oj[2]-ok[2]};
    double rolldot = rot1j[0]*ojk[0] + rot1j[1]*ojk[1] + rot1j[2]*ojk[2];
    double pitchdot = rot1j[3]*ojk[0] + rot1j[4]*ojk[1] + rot1j[5]*ojk[2];
    double yawdot = rot1j[6]*ojk[0] + rot1j[7]*ojk[1] + rot1j[8]*ojk[2];
    double roll = s.trqrpy[i], pitch = s.trqrpy[TT + i], yaw = s.trqrpy[2*TT + i];
    roll += rolldot*step;
    pitch += pitchdot*step;
    yaw += yawdot*step;
    s.trqrpy[i] = roll; s.trqrpy[TT + i] = pitch; s.trqrpy[2*TT + i] = yaw;
    double I[9], tmp[9], Ixx = 0.0, Iyy = 0.0, Izz = 0.0;
    if (didx (0, i)+1 == didx (0, i+1) || didx (1, i)+1 == didx (1, i+1) || didx (2, i)+1 == didx (2, i+1))
    {
      for (int c = 0; c < 9; c++) I[c] = s.inverse[c*N + j];
      if (k >= 0) for (int c = 0; c < 9; c++) I[c] += s.inverse[c*N + k];
      ntmul (I, rot1j, tmp);
      nnmul (rot1j, tmp, I);
      invert (I, tmp);
      Ixx = tmp[0]; Iyy = tmp[4]; Izz = tmp[8];
    }
    double kroll = 0.0, kpitch = 0.0, kyaw = 0.0, droll, dpitch, dyaw;
    double trqroll = 0.0, trqpitch = 0.0, trqyaw = 0.0, trqrolldot = 0.0, trqpitchdot = 0.0, trqyawdot = 0.0;
    double trqrolltot = 0.0, trqpitchtot = 0.0, trqyawtot = 0.0;
    /* curve o: forces.ispc:1625-1637 */
    auto curve = [&] (int o, double arg, double *kk) -> double {
      if (kidx (o, i+1) > kidx (o, i)) return look_up (kx, ky, kidx (o, i), kidx (o, i+1), arg, kk);
      *kk = 0.0;
      return 0.0;
    };
    /* damper o, its ratio read from table r (forces.ispc:1641-1675) */
    auto damper = [&] (int o, int r, double kk, double Iq, double vel, double *d) -> double {
      if (didx (o, i)+1 == didx (o, i+1))
      {
        double ratio = dx[didx (r, i)];
        *d = - ratio * 2.0 * std::sqrt (std::fabs (kk) * Iq);
        return vel * *d;
      }
      return look_up (dx, dy, didx (o, i), didx (o, i+1), vel, d);
    };
    switch (s.trqcone[i])
    {
    case 0:
      trqroll = curve (0, roll, &kroll);
      trqpitch = curve (1, pitch, &kpitch);
      trqyaw = curve (2, yaw, &kyaw);
      trqrolldot = damper (0, 0, kroll, Ixx, rolldot, &droll);
      trqpitchdot = damper (1, 1, kpitch, Iyy, pitchdot, &dpitch);
      trqyawdot = damper (2, 2, kyaw, Izz, yawdot, &dyaw);
      trqrolltot = trqroll + (trqroll == 0.0 ? 0.0 : trqrolldot);
      trqpitchtot = trqpitch + (trqpitch == 0.0 ? 0.0 : trqpitchdot);
      trqyawtot = trqyaw + (trqyaw == 0.0 ? 0.0 : trqyawdot);
      break;
    case 3:
    {
      double angle = std::sqrt (roll*roll + pitch*pitch);
      double angledot = angle > 0.0 ? (rolldot*roll+pitchdot*pitch)/angle : 0.0;
      trqroll = curve (0, angle, &kroll);
      kpitch = kroll;
      trqrolldot = damper (0, 0, kroll, std::sqrt (Ixx*Ixx + Iyy*Iyy), angledot, &droll);
      trqyaw = curve (2, yaw, &kyaw);
      trqyawdot = damper (2, 2, kyaw, Izz, yawdot, &dyaw);
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
    case 5:
    {
      double angle = std::sqrt (roll*roll + yaw*yaw);
      double angledot = angle > 0.0 ? (rolldot*roll+yawdot*yaw)/angle : 0.0;
      trqroll = curve (0, angle, &kroll);
      kyaw = kroll;
      trqrolldot = damper (0, 0, kroll, std::sqrt (Ixx*Ixx + Izz*Izz), angledot, &droll);
      trqpitch = curve (1, pitch, &kpitch);
      trqpitchdot = damper (1, 1, kpitch, Iyy, pitchdot, &dpitch);
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
    case 6:
    {
      trqroll = curve (0, roll, &kroll);
      trqrolldot = damper (0, 0, kroll, Ixx, rolldot, &droll);
      double angle = std::sqrt (pitch*pitch + yaw*yaw);
      double angledot = angle > 0.0 ? (pitchdot*pitch+yawdot*yaw)/angle : 0.0;
      trqpitch = curve (1, angle, &kpitch);
      kyaw = kpitch;
      /* the ratio read from roll's table (forces.ispc:1840) */
      trqpitchdot = damper (1, 0, kpitch, std::sqrt (Iyy*Iyy + Izz*Izz), angledot, &dpitch);
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
    case 7:
    {
      double angle = std::sqrt (roll*roll + pitch*pitch + yaw*yaw);
      double angledot = angle > 0.0 ? (rolldot*roll+pitchdot*pitch+yawdot*yaw)/angle : 0.0;
      trqroll = curve (0, angle, &kroll);
      kpitch = kyaw = kroll;
      trqrolldot = damper (0, 0, kroll, std::sqrt (Ixx*Ixx + Iyy*Iyy + Izz*Izz), angledot, &droll);
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
    double trq[3];
    trq[0] = trqrolltot*rot1j[0] + trqpitchtot*rot1j[3] + trqyawtot*rot1j[6];
    trq[1] = trqrolltot*rot1j[1] + trqpitchtot*rot1j[4] + trqyawtot*rot1j[7];
    trq[2] = trqrolltot*rot1j[2] + trqpitchtot*rot1j[5] + trqyawtot*rot1j[8];
    s.trqrpytot[i] = trqrolltot; s.trqrpytot[TT + i] = trqpitchtot; s.trqrpytot[2*TT + i] = trqyawtot;
    s.trqrpyspr[i] = trqroll; s.trqrpyspr[TT + i] = trqpitch; s.trqrpyspr[2*TT + i] = trqyaw;
    if (j0 != j && j0 >= 0)
    {
      for (int c = 0; c < 3; c++) { torque[c*N + j0] += tj[c]; tj[c] = trq[c]; }
      if (adp) for (int c = 0; c < 6; c++) { adp[(5+c)*N + j0] += krot_j[c]; krot_j[c] = 0.0; }
    }
    else for (int c = 0; c < 3; c++) tj[c] += trq[c];
    if (adp)
    {
      krot_j[0] += std::fabs (kroll)*rot1j[0];
      krot_j[1] += std::fabs (kpitch)*rot1j[4];
      krot_j[2] += std::fabs (kyaw)*rot1j[8];
    }
    if (k >= 0)
    {
      if (k0 != k && k0 >= 0)
      {
        for (int c = 0; c < 3; c++) torque[c*N + k0] -= tk[c];
        if (adp) for (int c = 0; c < 6; c++) adp[(5+c)*N + k0] += krot_k[c];
        for (int c = 0; c < 3; c++) tk[c] = trq[c];
        if (adp) for (int c = 0; c < 6; c++) krot_k[c] = 0.0;
      }
      else for (int c = 0; c < 3; c++) tk[c] += trq[c];
      if (adp)
      {
        krot_k[0] += std::fabs (kroll)*rot1j[0];
        krot_k[1] += std::fabs (kpitch)*rot1j[4];
        krot_k[2] += std::fabs (kyaw)*rot1j[8];
      }
      k0 = k;
    }
    j0 = j;
  }
  if (j0 >= 0)
  {
    for (int c = 0; c < 3; c++) torque[c*N + j0] += tj[c];
    if (adp) for (int c = 0; c < 6; c++) adp[(5+c)*N + j0] += krot_j[c];
  }
  if (k0 >= 0)
  {
    for (int c = 0; c < 3; c++) torque[c*N + k0] -= tk[c];
    if (adp) for (int c = 0; c < 6; c++) adp[(5+c)*N + k0] += krot_k[c];
  }
}

/* unspring_task, forces.ispc:2058-2127 */
void unspring_test (orc_ctx &s)
{
  const size_t SS = s.sprnum;
  for (int i = 0; i < s.unsprnum; i++)
  {
    if (s.stepnum % s.nsteps[i]) continue;
    if (s.nfreq[i] == 0) continue;
    int op = s.unop[i];
    double result = op == PMX_OP_MIN ? 1.7976931348623157e308 : op == PMX_OP_MAX ? -1.7976931348623157e308 : 0.;
    for (int q = s.tspridx[i]; q < s.tspridx[i+1]; q++)
    {
      int j = s.tsprings[q];
      double value;
      switch (s.unent[i])
      {
      case PMX_HIS_STROKE: value = s.stroke[j]; break;
      case PMX_HIS_F: value = s.sprfrc[j]; break;
      case PMX_HIS_SF: value = s.sprfrc[SS + j]; break;
      default: value = s.sprfrc[2*SS + j]; break;
      }
      if (s.unabs[i]) value = std::fabs (value);
      switch (op)
      {
      case PMX_OP_MIN: result = value < result ? value : result; break;
      case PMX_OP_MAX: result = value > result ? value : result; break;
      default: result += value; break;
      }
    }
    if (result < s.unlim[i] || result > s.unlim[s.unsprnum + i])
    {
      s.nfreq[i] --;
      if (s.nfreq[i] == 0)
      {
        for (int q = s.mspridx[i]; q < s.mspridx[i+1]; q++) s.unspring[s.msprings[q]] = s.unaction[i];
        for (int q = s.actidx[i]; q < s.actidx[i+1]; q++) s.unspring[s.activate[q]] = -3;
      }
    }
  }
}

/* -------------------------------------------------------------- restraints */

/* SUBMUL (v, DOT (v, d), d, v) */
void unproject (double v[3], const double d[3])
{
  double dot = v[0]*d[0] + v[1]*d[1] + v[2]*d[2];
  for (int c = 0; c < 3; c++) v[c] = v[c] - dot*d[c];
}

/* restrain_velocities_task, restrain.ispc:30-89, entries in order */
void restrain_velocities (orc_ctx &s)
{
  const size_t N = s.parnum, R = s.rstnum;
  for (size_t i = 0; i < R; i++)
  {
    int j = s.rstpart[i];
    double v[3], O[3], o[3], L[9], l[9], a[9], A[9];
    for (int c = 0; c < 3; c++) { v[c] = s.linear[c*N + j]; O[c] = s.angular[c*N + j]; o[c] = s.angular[(3+c)*N + j]; }
    for (int c = 0; c < 9; c++) { L[c] = s.rotation[c*N + j]; l[c] = s.rstlin[c*R + i]; a[c] = s.rstang[c*R + i]; }
    unproject (v, l); unproject (v, l+3); unproject (v, l+6);
    unproject (o, a); unproject (o, a+3); unproject (o, a+6);
    tvmul (L, a, A); tvmul (L, a+3, A+3); tvmul (L, a+6, A+6);
    unproject (O, A); unproject (O, A+3); unproject (O, A+6);
    for (int c = 0; c < 3; c++) { s.linear[c*N + j] = v[c]; s.angular[c*N + j] = O[c]; s.angular[(3+c)*N + j] = o[c]; }
  }
}

/* restrain_forces_task, restrain.ispc:92-133 */
void restrain_forces (orc_ctx &s)
{
  const size_t N = s.parnum, R = s.rstnum;
  for (size_t i = 0; i < R; i++)
  {
    int j = s.rstpart[i];
    double f[3], t[3], l[9], a[9];
    for (int c = 0; c < 3; c++) { f[c] = s.force[c*N + j]; t[c] = s.torque[c*N + j]; }
    for (int c = 0; c < 9; c++) { l[c] = s.rstlin[c*R + i]; a[c] = s.rstang[c*R + i]; }
    unproject (f, l); unproject (f, l+3); unproject (f, l+6);
    unproject (t, a); unproject (t, a+3); unproject (t, a+6);
    for (int c = 0; c < 3; c++) { s.force[c*N + j] = f[c]; s.torque[c*N + j] = t[c]; }
  }
}

/* ------------------------------------------------------------ ball joints */

/* amgcl::reorder::cuthill_mckee<false>::get, cuthill_mckee.hpp:49-170, on a
 * CSR pattern (its level-set walk visits nodes while `node > 0`) */
void cuthill_mckee (int n, const std::vector<int> &ptr, const std::vector<int> &col, std::vector<int> &perm)
{
  perm.assign (n, 0);
  if (n == 0) return;
  std::vector<int> degree (n), levelSet (n, 0), nextSameDegree (n, -1);
  int maxDegree = 0;
  for (int i = 0; i < n; i++) { degree[i] = ptr[i+1] - ptr[i]; maxDegree = std::max (maxDegree, degree[i]); }
  std::vector<int> firstWithDegree (maxDegree + 1, -1), nFirstWithDegree (maxDegree + 1);
  int initialNode = 0;
  perm[0] = initialNode;
  int currentLevelSet = 1;
  levelSet[initialNode] = currentLevelSet;
  int maxDegreeInCurrentLevelSet = degree[initialNode];
  firstWithDegree[maxDegreeInCurrentLevelSet] = initialNode;
  for (int next = 1; next < n; )
  {
    int nMDICLS = 0;
    std::fill (nFirstWithDegree.begin (), nFirstWithDegree.end (), -1);
    bool empty = true;
    for (int soughtDegree = 0; soughtDegree != maxDegreeInCurrentLevelSet + 1; soughtDegree++)
    {
      int node = firstWithDegree[soughtDegree];
      while (node > 0)
      {
        for (int q = ptr[node]; q < ptr[node+1]; q++)
        {
          int c = col[q];
          if (levelSet[c] == 0)
          {
            levelSet[c] = currentLevelSet + 1;
            perm[next] = c;
            ++next;
            empty = false;
            nextSameDegree[c] = nFirstWithDegree[degree[c]];
            nFirstWithDegree[degree[c]] = c;
            nMDICLS = std::max (nMDICLS, degree[c]);
          }
        }
        node = nextSameDegree[node];
      }
    }
    ++currentLevelSet;
    maxDegreeInCurrentLevelSet = nMDICLS;
    for (int i = 0; i <= nMDICLS; ++i) firstWithDegree[i] = nFirstWithDegree[i];
    if (empty)
    {
      for (int i = 0; i < n; ++i)
        if (levelSet[i] == 0)
        {
          perm[next] = i;
          ++next;
          levelSet[i] = currentLevelSet;
          maxDegreeInCurrentLevelSet = degree[i];
          firstWithDegree[maxDegreeInCurrentLevelSet] = i;
          break;
        }
    }
  }
}

/* amgcl::solver::skyline_lu (skyline_lu.hpp:74-303): the matrix reordered,
 * stored on its skyline of non-zero entries, factorised by Crout's method
 * with the diagonal held as inverses, then solved; returns 1 on a zero
 * pivot (the reference's precondition fails there) */
int skyline_lu (int n, const std::vector<int> &aptr, const std::vector<int> &col, const std::vector<double> &val,
                const std::vector<double> &rhs, std::vector<double> &x)
{
  std::vector<int> perm, invperm (n), ptr (n + 1, 0);
  cuthill_mckee (n, aptr, col, perm);
  for (int i = 0; i < n; ++i) invperm[perm[i]] = i;
  for (int i = 0; i < n; ++i)
    for (int q = aptr[i]; q < aptr[i+1]; q++)
    {
      int j = col[q], newi = invperm[i], newj = invperm[j];
      if (val[q] != 0.0)
      {
        if (newi > newj) { if (ptr[newi] < newi - newj) ptr[newi] = newi - newj; }
        else if (newi < newj) { if (ptr[newj] < newj - newi) ptr[newj] = newj - newi; }
      }
    }
  {
    int last = 0;
    for (int i = 1; i <= n; ++i) { int tmp = ptr[i]; ptr[i] = ptr[i-1] + last; last = tmp; }
  }
  std::vector<double> L (ptr.back (), 0.0), U (ptr.back (), 0.0), D (n, 0.0), y (n);
  for (int i = 0; i < n; ++i)
    for (int q = aptr[i]; q < aptr[i+1]; q++)
    {
      int j = col[q], newi = invperm[i], newj = invperm[j];
      double v = val[q];
      if (v != 0.0)
      {
        if (newi < newj) U[ptr[newj + 1] + newi - newj] = v;
        else if (newi == newj) D[newi] = v;
        else L[ptr[newi + 1] + newj - newi] = v;
      }
    }
  if (D[0] == 0.0) return 1;
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
      int jBeginMult = std::max (iBeginCol, jBeginRow);
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
      int jBeginMult = std::max (jBeginCol, jBeginRow);
      int indexL = ptr[k+1] + jBeginMult - jBeginRow;
      int indexU = ptr[i] + jBeginMult - jBeginCol;
      for (int j = jBeginMult; j < i; ++j, ++indexL, ++indexU) sum -= L[indexL] * U[indexU];
      L[indexEntry] = sum;
    }
    double sum = D[k+1];
    for (int j = ptr[k+1]; j < ptr[k+2]; ++j) sum -= L[j] * U[j];
    if (sum == 0.0) return 1;
    D[k+1] = 1.0 / sum;
  }
  for (int i = 0; i < n; ++i)
  {
    double sum = rhs[perm[i]];
    for (int k = ptr[i], j = i - ptr[i+1] + k; k < ptr[i+1]; ++k, ++j) sum -= L[k] * y[j];
    y[i] = D[i] * sum;
  }
  for (int j = n - 1; j >= 0; --j)
    for (int k = ptr[j], i = j - ptr[j+1] + k; k < ptr[j+1]; ++k, ++i) y[i] -= U[k] * y[j];
  x.assign (n, 0.0);
  for (int i = 0; i < n; ++i) x[perm[i]] = y[i];
  return 0;
}

/* the recursive jmark of joints.cpp:67-80 */
void jmark (std::vector<std::set<int>> &jadj, std::set<int> &jall, int joint, std::vector<int> &iset)
{
  if (jall.find (joint) != jall.end ())
  {
    jall.erase (joint);
    iset.push_back (joint);
    for (int j : jadj[joint]) jmark (jadj, jall, j, iset);
  }
}

/* Rot * skew (X - P) */
void joint_h (const orc_ctx &s, int part, const double Rot[9], const double P[3], double H[9])
{
  size_t N = s.parnum;
  double A[3], Ask[9];
  for (int c = 0; c < 3; c++) A[c] = s.position[(3+c)*N + part] - P[c];
  Ask[0] = 0; Ask[1] = A[2]; Ask[2] = -A[1];
  Ask[3] = -A[2]; Ask[4] = 0; Ask[5] = A[0];
  Ask[6] = A[1]; Ask[7] = -A[0]; Ask[8] = 0;
  nnmul (Rot, Ask, H);
}

/* reset_joints_symbolic and solve_joints, joints.cpp:83-567; returns 1 on a
 * zero pivot */
int solve_joints (orc_ctx &s, const double damping[6], double step0, double step1)
{
  const int J = s.jnum;
  if (J == 0) return 0;
  const size_t N = s.parnum, JJ = J;
  const int *jp0 = s.jpart.data (), *jp1 = s.jpart.data () + J;
  auto JP = [&] (int k, int i) { return k ? jp1[i] : jp0[i]; };
  std::map<int, std::set<int>> partadj;
  for (int i = 0; i < J; i++) { partadj[jp0[i]].insert (i); if (jp1[i] >= 0) partadj[jp1[i]].insert (i); }
  std::vector<std::set<int>> jadj (J);
  for (auto &pa : partadj)
    for (int a : pa.second) for (int b : pa.second) if (a != b) jadj[a].insert (b);
  std::vector<int> jmap (J);
  std::vector<std::vector<int>> isets;
  std::set<int> jall;
  for (int i = 0; i < J; i++) jall.insert (i);
  while (!jall.empty ())
  {
    std::vector<int> iset;
    jmark (jadj, jall, *jall.begin (), iset);
    for (size_t i = 0; i < iset.size (); i++) jmap[iset[i]] = (int) i;
    isets.push_back (iset);
  }
  const double half = 0.5*step0, step = 0.5*(step0+step1);
  for (auto &iset : isets)
  {
    int n = 3 * (int) iset.size ();
    /* CSR of the dense pattern, columns ascending */
    std::vector<int> ptr (n + 1), col ((size_t) n * n);
    for (int i = 0; i <= n; i++) ptr[i] = i * n;
    for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) col[(size_t) i*n + j] = j;
    std::vector<double> val ((size_t) n * n, 0.0);
    auto V = [&] (int r, int c) -> double & { return val[(size_t) r * n + c]; };
    for (size_t i0 = 0; i0 < iset.size (); i0++)
    {
      int i = iset[i0];
      double Wii[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Pi[3] = {s.jpoint[i], s.jpoint[JJ + i], s.jpoint[2*JJ + i]};
      for (int k = 0; k < 2; k++)
      {
        int part = JP (k, i);
        if (part < 0) continue;
        double Rot[9], Jiv[9], Hi[9], Hj[9], C[9], D[9], Wij[9];
        for (int c = 0; c < 9; c++) { Rot[c] = s.rotation[c*N + part]; Jiv[c] = s.inverse[c*N + part]; }
        double im = s.invm[part];
        joint_h (s, part, Rot, Pi, Hi);
        ntmul (Jiv, Hi, C);
        nnmul (Hi, C, D);
        for (int c = 0; c < 9; c++) Wii[c] = Wii[c] + D[c];
        Wii[0] += im; Wii[4] += im; Wii[8] += im;
        for (int j : partadj[part])
        {
          if (i == j) continue;
          int j0 = jmap[j];
          double Pj[3] = {s.jpoint[j], s.jpoint[JJ + j], s.jpoint[2*JJ + j]};
          joint_h (s, part, Rot, Pj, Hj);
          ntmul (Jiv, Hj, C);
          nnmul (Hi, C, Wij);
          Wij[0] += im; Wij[4] += im; Wij[8] += im;
          if ((part == jp0[i] && part == jp1[j]) || (part == jp1[i] && part == jp0[j]))
            for (int c = 0; c < 9; c++) Wij[c] *= -1.0;
          for (int i1 = 0; i1 < 3; i1++)
            for (int j1 = 0; j1 < 3; j1++) V (3*(int) i0 + i1, 3*j0 + j1) += Wij[i1*3+j1];
        }
      }
      for (int i1 = 0; i1 < 3; i1++)
        for (int j1 = 0; j1 < 3; j1++) V (3*(int) i0 + i1, 3*(int) i0 + j1) += Wii[i1*3+j1];
    }
    std::vector<double> rhs (n, 0.0), x;
    for (size_t i0 = 0; i0 < iset.size (); i0++)
    {
      int i = iset[i0];
      double *B = &rhs[3*i0], Pi[3] = {s.jpoint[i], s.jpoint[JJ + i], s.jpoint[2*JJ + i]};
      for (int k = 0; k < 2; k++)
      {
        int part = JP (k, i);
        if (part < 0) continue;
        double Rot[9], Jiv[9], Jm[9], O[3], o[3], v[3], f[3], t[3], A[3], Vv[3], C[3], T[3], dRot[9], Hi[9];
        for (int c = 0; c < 9; c++)
        {
          Rot[c] = s.rotation[c*N + part]; Jiv[c] = s.inverse[c*N + part]; Jm[c] = s.inertia[c*N + part];
        }
        for (int c = 0; c < 3; c++)
        {
          O[c] = s.angular[c*N + part]; f[c] = s.force[c*N + part]; t[c] = s.torque[c*N + part]; v[c] = s.linear[c*N + part];
        }
        double im = s.invm[part];
        joint_h (s, part, Rot, Pi, Hi);
        nvmul (Hi, O, Vv);
        for (int c = 0; c < 3; c++) Vv[c] += v[c];
        if (damping[0] != 0.0 || damping[1] != 0.0 || damping[2] != 0.0)
        {
          double ma = s.mass[i];   /* joints.cpp:399 reads mass[joint] */
          for (int c = 0; c < 3; c++) f[c] -= ma * damping[c] * v[c];
        }
        if (damping[3] != 0.0 || damping[4] != 0.0 || damping[5] != 0.0)
        {
          for (int c = 0; c < 3; c++) o[c] = damping[3+c]*s.angular[(3+c)*N + part];
          tvmul (Rot, o, C);
          nvmul (Jm, C, T);
          nvmul (Rot, T, C);
          for (int c = 0; c < 3; c++) t[c] -= C[c];
        }
        expmap (-half*O[0], -half*O[1], -half*O[2], dRot);
        tvmul (Rot, t, T);
        nvmul (Jm, O, A);
        nvmul (dRot, A, C);
        for (int c = 0; c < 3; c++) C[c] = C[c] + half*T[c];
        nvmul (Jiv, C, A);
        nvmul (Jm, A, C);
        T[0] -= A[1]*C[2] - A[2]*C[1];
        T[1] -= A[2]*C[0] - A[0]*C[2];
        T[2] -= A[0]*C[1] - A[1]*C[0];
        for (int c = 0; c < 3; c++) T[c] *= step;
        double Ot[3];
        Ot[0] = O[0] + (Jiv[0]*T[0]+Jiv[3]*T[1]+Jiv[6]*T[2]);
        Ot[1] = O[1] + (Jiv[1]*T[0]+Jiv[4]*T[1]+Jiv[7]*T[2]);
        Ot[2] = O[2] + (Jiv[2]*T[0]+Jiv[5]*T[1]+Jiv[8]*T[2]);
        nvmul (Hi, Ot, C);
        for (int c = 0; c < 3; c++) C[c] += v[c];
        for (int c = 0; c < 3; c++) C[c] = C[c] + (step*im)*f[c];
        if (part == jp0[i]) for (int c = 0; c < 3; c++) B[c] -= Vv[c]+C[c];
        else for (int c = 0; c < 3; c++) B[c] += Vv[c]+C[c];
      }
    }
    if (skyline_lu (n, ptr, col, val, rhs, x)) return 1;
    double invstep = 1./step;
    for (size_t i0 = 0; i0 < iset.size (); i0++)
    {
      int i = iset[i0];
      double R[3] = {x[3*i0] * invstep, x[3*i0+1] * invstep, x[3*i0+2] * invstep};
      for (int c = 0; c < 3; c++) s.jreac[c*JJ + i] = R[c];
      for (int k = 0; k < 2; k++)
      {
        int part = JP (k, i);
        if (part < 0) continue;
        double Rot[9], A[3], a[3], t[3];
        for (int c = 0; c < 9; c++) Rot[c] = s.rotation[c*N + part];
        for (int c = 0; c < 3; c++) A[c] = s.position[(3+c)*N + part] - s.jpoint[c*JJ + i];
        nvmul (Rot, A, a);
        t[0] = R[1]*a[2] - R[2]*a[1];
        t[1] = R[2]*a[0] - R[0]*a[2];
        t[2] = R[0]*a[1] - R[1]*a[0];
        if (part == jp0[i]) for (int c = 0; c < 3; c++) { s.torque[c*N + part] += t[c]; s.force[c*N + part] += R[c]; }
        else for (int c = 0; c < 3; c++) { s.torque[c*N + part] -= t[c]; s.force[c*N + part] -= R[c]; }
      }
    }
  }
  return 0;
}

/* ------------------------------------------------- critical step estimates */

/* time_step_task, parmec.ispc:275-420, for one particle; min and max as
 * ISPC's (a < b ? a : b, a > b ? a : b); INVERT leaves the inverse as it
 * was where the determinant is zero (zeros here) */
double time_step (const double kact[3], double kmax, double emax, const double krot[6], double mxyz, const double J[9],
                  double *ocri, double *rcri)
{
  double nx = kact[0], ny = kact[1], nz = kact[2];
  double nxy = nx > ny ? nx : ny;
  double nxyz = nxy > nz ? nxy : nz;
  if (!(nxyz > 0.0)) return 1.7976931348623157E+308;
  double k = kmax;
  double omegaxyz = std::sqrt (nxyz*k/mxyz);
  double K[9], L[9], inv, invL[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, temp[9], A[9];
  K[0] = k*krot[0]; K[1] = k*krot[3]; K[2] = k*krot[4];
  K[3] = K[1]; K[4] = k*krot[1]; K[5] = k*krot[5];
  K[6] = K[2]; K[7] = K[5]; K[8] = k*krot[2];
  double mrot = extremum_eigenvalue (J, true);
  L[0] = std::sqrt (J[0]);
  inv = 1.0/L[0];
  L[1] = inv * J[1];
  L[2] = inv * J[2];
  L[3] = 0.0;
  L[4] = std::sqrt (J[4] - L[1]*L[1]);
  inv = 1.0/L[4];
  L[5] = inv * (J[5] - L[2]*L[1]);
  L[6] = 0.0;
  L[7] = 0.0;
  L[8] = std::sqrt (J[8] - (L[2]*L[2] + L[5]*L[5]));
  double det = L[0]*(L[4]*L[8]-L[5]*L[7]) + L[3]*(L[2]*L[7]-L[1]*L[8]) + (L[1]*L[5]-L[2]*L[4])*L[6];   /* INVERT */
  if (det != 0.0)
  {
    det = 1.0 / det;
    invL[0] = (L[4]*L[8]-L[5]*L[7])*det;
    invL[1] = (L[2]*L[7]-L[1]*L[8])*det;
    invL[2] = (L[1]*L[5]-L[2]*L[4])*det;
    invL[3] = (L[5]*L[6]-L[3]*L[8])*det;
    invL[4] = (L[0]*L[8]-L[2]*L[6])*det;
    invL[5] = (L[2]*L[3]-L[0]*L[5])*det;
    invL[6] = (L[3]*L[7]-L[4]*L[6])*det;
    invL[7] = (L[1]*L[6]-L[0]*L[7])*det;
    invL[8] = (L[0]*L[4]-L[1]*L[3])*det;
  }
  /* NTMUL (K, invL, temp): temp = K invL^T */
  for (int c = 0; c < 3; c++)
    for (int r = 0; r < 3; r++)
      temp[3*c+r] = K[r]*invL[c] + K[3+r]*invL[3+c] + K[6+r]*invL[6+c];
  nnmul (invL, temp, A);
  double lambda = extremum_eigenvalue (A, false);
  double omegarot = std::sqrt (lambda);
  double damping = emax;
  double ratioxyz = damping / (2.0*mxyz*omegaxyz);
  double ratiorot = damping / (2.0*mrot*omegarot);
  double hxyz = (2.0/omegaxyz)*(std::sqrt (1.0+ratioxyz*ratioxyz) - ratioxyz);
  double hrot = (2.0/omegarot)*(std::sqrt (1.0+ratiorot*ratiorot) - ratiorot);
  if (ocri) *ocri = omegaxyz > omegarot ? omegaxyz : omegarot;
  if (rcri) *rcri = ratioxyz > ratiorot ? ratioxyz : ratiorot;
  return hxyz < hrot ? hxyz : hrot;
}

/* adaptive_timestep, parmec.ispc:391-420 */
double adaptive_timestep (const orc_ctx &s, double curstep, double adaptive)
{
  size_t N = s.parnum;
  double hmin = 1.7976931348623157E+308;
  for (size_t i = 0; i < N; i++)
  {
    double ka[3], kr[6], J[9];
    for (int c = 0; c < 3; c++) ka[c] = s.adp[c*N + i];
    for (int c = 0; c < 6; c++) kr[c] = s.adp[(5+c)*N + i];
    for (int c = 0; c < 9; c++) J[c] = s.inertia[c*N + i];
    double h = time_step (ka, s.adp[3*N + i], s.adp[4*N + i], kr, s.mass[i], J, nullptr, nullptr);
    hmin = h < hmin ? h : hmin;
  }
  if (hmin < 1.7976931348623157E+308) return adaptive*hmin;
  return curstep;
}

/* the largest |slope| of a curve, parmec.ispc:510-515 */
double max_slope (const double *x, const double *y, int lo, int hi)
{
  double m = 0.0;
  for (int q = lo; q < hi-1; q++)
  {
    double k = (y[q+1]-y[q]) / (x[q+1]-x[q]);
    k = std::fabs (k);
    m = k > m ? k : m;
  }
  return m;
}

/* critical_perparticle, parmec.ispc:547-888: the contacts of the last step
 * (in the canonical order of the force sums), then the springs in spring
 * order, then time_step_task */
void critical_perparticle (orc_ctx &s, double *hcri, double *ocri, double *rcri)
{
  size_t N = s.parnum, S = s.sprnum;
  std::vector<double> adp (11 * N, 0.0);
  for (size_t i = 0; i < N; i++)   /* critical_perparticle_contacts, parmec.ispc:547-656 */
  {
    struct E { unsigned long long k0, k1; const double *p; double f[3]; const double *n; double kc, ec; };
    std::vector<E> es;
    for (const Contact &con : s.master[i])
      es.push_back ({side_key (con.slave1), side_key (con.master), con.point, {con.force[0], con.force[1], con.force[2]},
                     con.normal, con.kcur, con.ecur});
    for (const Slave &sv : s.slave[i])
      es.push_back ({side_key (sv.slave1 < 0 ? sv.slave1 : sv.master1), side_key (sv.slave1 < 0 ? sv.master1 : sv.slave1),
                     sv.point, {sv.force[0], sv.force[1], sv.force[2]}, nullptr, 0.0, 0.0});
    std::sort (es.begin (), es.end (), [] (const E &a, const E &b) { return a.k0 < b.k0 || (a.k0 == b.k0 && a.k1 < b.k1); });
    double L[9], po[3];
    for (int c = 0; c < 9; c++) L[c] = s.rotation[c*N + i];
    for (int c = 0; c < 3; c++) po[c] = s.position[c*N + i];
    double kact0[3] = {0.0, 0.0, 0.0}, kmax0 = 0.0, emax0 = 0.0, krot0[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (const E &e : es)
    {
      double a[3] = {e.p[0]-po[0], e.p[1]-po[1], e.p[2]-po[2]}, f[3];
      if (e.n) for (int c = 0; c < 3; c++) f[c] = e.n[c];
      else
      {
        for (int c = 0; c < 3; c++) f[c] = e.f[c];
        double len = std::sqrt (f[0]*f[0] + f[1]*f[1] + f[2]*f[2]);
        if (len != 0.0)
        {
          len = 1.0/len;
          f[0] *= len; f[1] *= len; f[2] *= len;
        }
        else f[0] = f[1] = f[2] = 0.0;
      }
      for (int c = 0; c < 3; c++) kact0[c] += std::fabs (f[c]);
      kmax0 = e.kc > kmax0 ? e.kc : kmax0;
      emax0 = e.ec > emax0 ? e.ec : emax0;
      double A[3];
      nvmul (L, a, A);
      double dot = A[0]*A[0] + A[1]*A[1] + A[2]*A[2];
      krot0[0] += dot - A[0]*A[0];
      krot0[1] += dot - A[1]*A[1];
      krot0[2] += dot - A[2]*A[2];
      krot0[3] += -A[0]*A[1];
      krot0[4] += -A[0]*A[2];
      krot0[5] += -A[1]*A[2];
    }
    double v[11] = {kact0[0], kact0[1], kact0[2], kmax0, emax0, krot0[0], krot0[1], krot0[2], krot0[3], krot0[4], krot0[5]};
    for (int c = 0; c < 11; c++) adp[c*N + i] = v[c];
  }
  /* critical_perparticle_springs, parmec.ispc:659-857, one spring after
   * another.  dashidx and sprflg are read at particle indices there (i, k);
   * where such an index lies beyond the spring arrays the spring's own
   * entry (dashidx) or no flag (sprflg) is taken */
  const double *sx = s.spring.data (), *sy = s.spring.data () + s.sprcrv;
  const double *dx = s.dashpot.data (), *dy = s.dashpot.data () + s.dashcrv;
  for (size_t j = 0; j < S; j++)
  {
    int i = s.sprpart[j], k = s.sprpart[S + j];
    double d[3] = {s.sprdir[j], s.sprdir[S + j], s.sprdir[2*S + j]};
    double kspr = max_slope (sx, sy, s.spridx[j], s.spridx[j+1]);
    double espr = 0.0;
    size_t di = (size_t) i < S ? (size_t) i : j;
    if (s.dashidx[di]+1 == s.dashidx[di+1])
    {
      double Pi[3], Pk[3];
      for (int c = 0; c < 3; c++) { Pi[c] = s.sprpnt[(3+c)*S + j]; Pk[c] = s.sprpnt[(9+c)*S + j]; }
      double mass = eqm (s, i, Pi, k, Pk, d);
      double ratio = dx[s.dashidx[di]];
      espr = ratio * 2.0 * std::sqrt (kspr * mass);
    }
    else espr = max_slope (dx, dy, s.dashidx[j], s.dashidx[j+1]);
    for (int side = 0; side < 2; side++)
    {
      int b = side ? k : i;
      if (b < 0) continue;
      int flg = (size_t) b < S ? s.sprflg[b] : 0;
      if (d[0]*d[0] + d[1]*d[1] + d[2]*d[2] > 0.0)
      {
        if (flg & PMX_SPRDIR_PLANAR)
        {
          double t1[3] = {d[0], d[1], d[2]}, t2[3];
          t1[0] += 0.123;
          t1[1] += 0.346;
          t1[2] += 0.789;
          t2[0] = d[1]*t1[2] - d[2]*t1[1];
          t2[1] = d[2]*t1[0] - d[0]*t1[2];
          t2[2] = d[0]*t1[1] - d[1]*t1[0];
          double l2 = std::sqrt (t2[0]*t2[0] + t2[1]*t2[1] + t2[2]*t2[2]);
          double i2 = 1.0/l2;
          t2[0] *= i2; t2[1] *= i2; t2[2] *= i2;
          t1[0] = t2[1]*d[2] - t2[2]*d[1];
          t1[1] = t2[2]*d[0] - t2[0]*d[2];
          t1[2] = t2[0]*d[1] - t2[1]*d[0];
          for (int c = 0; c < 3; c++) adp[c*N + b] += std::fabs (t1[c]) + std::fabs (t2[c]);
        }
        else for (int c = 0; c < 3; c++) adp[c*N + b] += std::fabs (d[c]);
      }
      else for (int c = 0; c < 3; c++) adp[c*N + b] += 1.0;
      double pnt[3], Lb[9], a[3], A[3];
      for (int c = 0; c < 3; c++) pnt[c] = s.sprpnt[(6*side+c)*S + j];
      for (int c = 0; c < 9; c++) Lb[c] = s.rotation[c*N + b];
      for (int c = 0; c < 3; c++) a[c] = pnt[c] - s.position[c*N + b];
      tvmul (Lb, a, A);
      double dot = A[0]*A[0] + A[1]*A[1] + A[2]*A[2];
      adp[5*N + b] += dot-A[0]*A[0];
      adp[6*N + b] += dot-A[1]*A[1];
      adp[7*N + b] += dot-A[2]*A[2];
      adp[8*N + b] += -A[0]*A[1];
      adp[9*N + b] += -A[0]*A[2];
      adp[10*N + b] += -A[1]*A[2];
      adp[3*N + b] = adp[3*N + b] > kspr ? adp[3*N + b] : kspr;
      adp[4*N + b] = adp[4*N + b] > espr ? adp[4*N + b] : espr;
    }
  }
  for (size_t i = 0; i < N; i++)   /* time_step_task */
  {
    double ka[3], kr[6], J[9], o = 0.0, r = 0.0;
    for (int c = 0; c < 3; c++) ka[c] = adp[c*N + i];
    for (int c = 0; c < 6; c++) kr[c] = adp[(5+c)*N + i];
    for (int c = 0; c < 9; c++) J[c] = s.inertia[c*N + i];
    hcri[i] = time_step (ka, adp[3*N + i], adp[4*N + i], kr, s.mass[i], J, &o, &r);
    if (ocri) ocri[i] = o;
    if (rcri) rcri[i] = r;
  }
}

/* critical_perspring, parmec.ispc:485-535 */
void critical_perspring (orc_ctx &s, double *hcri, double *ocri, double *rcri)
{
  size_t S = s.sprnum;
  const double *sx = s.spring.data (), *sy = s.spring.data () + s.sprcrv;
  const double *dx = s.dashpot.data (), *dy = s.dashpot.data () + s.dashcrv;
  for (size_t j = 0; j < S; j++)
  {
    int i = s.sprpart[j], k = s.sprpart[S + j];
    double pi[3], pk[3], d[3];
    for (int c = 0; c < 3; c++) { pi[c] = s.sprpnt[(3+c)*S + j]; pk[c] = s.sprpnt[(9+c)*S + j]; d[c] = s.sprdir[c*S + j]; }
    double kspr = max_slope (sx, sy, s.spridx[j], s.spridx[j+1]);
    double mspr = eqm (s, i, pi, k, pk, d);
    double omega = std::sqrt (kspr/mspr), ksi = 0.0;
    if (s.dashidx[j]+1 == s.dashidx[j+1]) ksi = dx[s.dashidx[j]];
    else
    {
      double espr = max_slope (dx, dy, s.dashidx[j], s.dashidx[j+1]);
      if (omega > 0.0) ksi = espr / (2*mspr*omega);
    }
    hcri[j] = omega > 0.0 ? (2.0/omega)*(std::sqrt (1.0+ksi*ksi) - ksi) : 1.0;
    if (ocri) ocri[j] = omega;
    if (rcri) rcri[j] = ksi;
  }
}

} /* namespace */

/* ------------------------------------------------------------------ API */

extern "C" {

const char *orc_last_error (void) { return g_err.c_str (); }

int orc_create (int nthreads, orc_ctx **ctx)
{
  if (!ctx) return fail (PMX_EINVAL, "null ctx");
  *ctx = new orc_ctx ();
#ifdef _OPENMP
  (*ctx)->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads ();
#else
  (*ctx)->nthreads = 1;
  (void) nthreads;
#endif
  return PMX_OK;
}

void orc_destroy (orc_ctx *ctx) { delete ctx; }

int orc_upload (orc_ctx *s, const pmx_model *m)
{
  if (!s || !m) return fail (PMX_EINVAL, "null argument");
  if (m->obsnum < 0) return fail (PMX_EINVAL, "invalid obstacle count");
  for (int i = 0; i < m->prsnum; i++)
    if (m->prspart[i] < 0 || m->prspart[i] >= m->parnum) return fail (PMX_EINVAL, "prescribed particle out of range");
  for (int t = 0; t < m->trinum; t++)
    if (m->triobs[t] <= -2 && -m->triobs[t]-2 >= m->obsnum) return fail (PMX_EINVAL, "triangle obstacle index out of range");
  for (int i = 0; i < m->obsnum; i++)
    if (m->trirng[2*i] < 0 || m->trirng[2*i] > m->trirng[2*i+1] || m->trirng[2*i+1] > m->trinum)
      return fail (PMX_EINVAL, "obstacle triangle range out of bounds");
  for (int i = 0; i < m->sprnum; i++)
  {
    if (i && (m->sprpart[i-1] > m->sprpart[i] ||
              (m->sprpart[i-1] == m->sprpart[i] && m->sprpart[m->sprnum + i-1] > m->sprpart[m->sprnum + i])))
      return fail (PMX_EINVAL, "springs must be sorted by (part1, part2)");
  }
  bool same = s->parnum == m->parnum && s->ellnum == m->ellnum && s->trinum == m->trinum;
  s->parnum = m->parnum; s->ellnum = m->ellnum; s->trinum = m->trinum;
  s->nodnum = m->nodnum; s->facnum = m->facnum; s->pairnum = m->pairnum;
  size_t np = m->parnum, ne = m->ellnum, nt = m->trinum;
  take (s->angular, m->angular, 6*np);
  take (s->linear, m->linear, 3*np);
  take (s->rotation, m->rotation, 9*np);
  take (s->position, m->position, 6*np);
  take (s->inertia, m->inertia, 9*np);
  take (s->inverse, m->inverse, 9*np);
  take (s->mass, m->mass, np);
  take (s->invm, m->invm, np);
  take (s->force, m->force, 3*np);
  take (s->torque, m->torque, 3*np);
  take (s->flags, m->flags, np);
  take (s->center, m->center, 6*ne);
  take (s->radii, m->radii, 3*ne);
  take (s->part, m->part, ne);
  take (s->ellcol, m->ellcol, ne);
  take (s->tri, m->tri, 9*nt);
  take (s->tricol, m->tricol, nt);
  take (s->triobs, m->triobs, nt);
  s->prsnum = m->prsnum;
  take (s->prspart, m->prspart, (size_t) m->prsnum);
  take (s->prskind, m->prskind, 2*(size_t) m->prsnum);
  s->prscol.resize (m->prsnum);
  s->prscols = 0;
  for (int i = 0; i < m->prsnum; i++)
  {
    s->prscol[i] = m->prscol ? m->prscol[i] : i;
    s->prscols = std::max (s->prscols, s->prscol[i] + 1);
  }
  s->obsnum = m->obsnum;
  take (s->trirng, m->trirng, 2*(size_t) m->obsnum);
  take (s->obspnt, m->obspnt, 3*(size_t) m->obsnum);
  take (s->obsang, m->obsang, 3*(size_t) m->obsnum);
  take (s->obslin, m->obslin, 3*(size_t) m->obsnum);
  take (s->nodes, m->nodes, 6*(size_t) m->nodnum);
  take (s->nodpart, m->nodpart, (size_t) m->nodnum);
  take (s->facnod, m->facnod, 3*(size_t) m->facnum);
  take (s->factri, m->factri, (size_t) m->facnum);
  take (s->pairs, m->pairs, 2*(size_t) m->pairnum);
  take (s->ikind, m->ikind, (size_t) m->pairnum);
  take (s->iparam, m->iparam, (size_t) PMX_NIPARAM * m->pairnum);
  size_t S = m->sprnum;
  s->sprnum = m->sprnum;
  s->sprcrv = m->sprcrv; s->dashcrv = m->dashcrv; s->uncrv = m->uncrv;
  take (s->sprtype, m->sprtype, S);
  take (s->sprpart, m->sprpart, 2*S);
  take (s->sprpnt, m->sprpnt, 12*S);
  take (s->spridx, m->spridx, S ? S+1 : 0);
  take (s->spring, m->spring, 2*(size_t) m->sprcrv);
  take (s->dashidx, m->dashidx, S ? S+1 : 0);
  take (s->dashpot, m->dashpot, 2*(size_t) m->dashcrv);
  take (s->unidx, m->unidx, S ? S+1 : 0);
  take (s->unload, m->unload, 2*(size_t) m->uncrv);
  take (s->yield, m->yield, 2*S);
  take (s->sprdir, m->sprdir, 6*S);
  take (s->sprflg, m->sprflg, S);
  take (s->sprfric, m->sprfric, S);
  take (s->sprkskn, m->sprkskn, S);
  take (s->sprsdsp, m->sprsdsp, 3*S);
  take (s->stroke0, m->stroke0, S);
  take (s->stroke, m->stroke, 3*S);
  take (s->sprfrc, m->sprfrc, 3*S);
  take (s->unspring, m->unspring, S);
  take (s->sproffset, m->sproffset, S);
  if (!m->sproffset) s->sproffset.assign (S, -1);
  s->lcnum = m->lcnum; s->lccrv = m->lccrv;
  take (s->lcidx, m->lcidx, m->lcnum > 0 ? (size_t) m->lcnum + 1 : 0);
  take (s->lcurve, m->lcurve, 2*(size_t) m->lccrv);
  size_t T = m->trqsprnum;
  s->trqsprnum = m->trqsprnum; s->krpycrv = m->krpycrv; s->drpycrv = m->drpycrv;
  take (s->trqsprpart, m->trqsprpart, 2*T);
  take (s->trqzdir0, m->trqzdir0, 3*T);
  take (s->trqxdir0, m->trqxdir0, 3*T);
  take (s->krpyidx, m->krpyidx, T ? 3*(T+1) : 0);
  take (s->krpy, m->krpy, 2*(size_t) m->krpycrv);
  take (s->drpyidx, m->drpyidx, T ? 3*(T+1) : 0);
  take (s->drpy, m->drpy, 2*(size_t) m->drpycrv);
  take (s->trqcone, m->trqcone, T);
  take (s->trqzdir1, m->trqzdir1, 3*T);
  take (s->trqxdir1, m->trqxdir1, 3*T);
  take (s->trqrpy, m->trqrpy, 3*T);
  take (s->trqrpytot, m->trqrpytot, 3*T);
  take (s->trqrpyspr, m->trqrpyspr, 3*T);
  size_t U = m->unsprnum;
  s->unsprnum = m->unsprnum;
  take (s->tspridx, m->tspridx, U ? U+1 : 0);
  take (s->tsprings, m->tsprings, U ? (size_t) m->tspridx[U] : 0);
  take (s->mspridx, m->mspridx, U ? U+1 : 0);
  take (s->msprings, m->msprings, U ? (size_t) m->mspridx[U] : 0);
  take (s->actidx, m->actidx, U ? U+1 : 0);
  take (s->activate, m->activate, U ? (size_t) m->actidx[U] : 0);
  take (s->unlim, m->unlim, 2*U);
  take (s->unent, m->unent, U);
  take (s->unop, m->unop, U);
  take (s->unabs, m->unabs, U);
  take (s->nsteps, m->nsteps, U);
  take (s->nfreq, m->nfreq, U);
  take (s->unaction, m->unaction, U);
  size_t R = m->rstnum;
  s->rstnum = m->rstnum;
  take (s->rstpart, m->rstpart, R);
  take (s->rstlin, m->rstlin, 9*R);
  take (s->rstang, m->rstang, 9*R);
  size_t JN = m->jnum;
  s->jnum = m->jnum;
  take (s->jpart, m->jpart, 2*JN);
  take (s->jpoint, m->jpoint, 3*JN);
  take (s->jreac, m->jreac, 3*JN);
  s->curtime = m->curtime; s->curstep = m->curstep; s->stepnum = m->stepnum;
  if (!same)
  {
    s->master.assign (np, {});
    s->slave.assign (np, {});
  }
  return PMX_OK;
}

int orc_download (orc_ctx *s, pmx_model *m)
{
  if (!s || !m) return fail (PMX_EINVAL, "null argument");
  if (m->parnum != s->parnum || m->ellnum != s->ellnum || m->trinum != s->trinum)
    return fail (PMX_EINVAL, "model counts differ from the uploaded state");
  auto put = [] (double *dst, const std::vector<double> &src) { if (dst) std::memcpy (dst, src.data (), src.size ()*sizeof (double)); };
  put (m->angular, s->angular); put (m->linear, s->linear); put (m->rotation, s->rotation);
  put (m->position, s->position); put (m->inverse, s->inverse); put (m->invm, s->invm);
  put (m->force, s->force); put (m->torque, s->torque); put (m->center, s->center);
  put (m->tri, s->tri); put (m->nodes, s->nodes);
  if (m->obsnum == s->obsnum) { put (m->obspnt, s->obspnt); put (m->obsang, s->obsang); put (m->obslin, s->obslin); }
  put (m->sprpnt, s->sprpnt); put (m->sprdir, s->sprdir); put (m->sprsdsp, s->sprsdsp); put (m->stroke, s->stroke);
  put (m->sprfrc, s->sprfrc); put (m->yield, s->yield);
  if (m->sprflg) std::memcpy (m->sprflg, s->sprflg.data (), s->sprflg.size () * sizeof (int));
  if (m->unspring) std::memcpy (m->unspring, s->unspring.data (), s->unspring.size () * sizeof (int));
  put (m->trqzdir1, s->trqzdir1); put (m->trqxdir1, s->trqxdir1); put (m->trqrpy, s->trqrpy);
  put (m->trqrpytot, s->trqrpytot); put (m->trqrpyspr, s->trqrpyspr); put (m->jreac, s->jreac);
  if (m->nfreq && s->unsprnum) std::memcpy (m->nfreq, s->nfreq.data (), s->nfreq.size () * sizeof (int));
  m->curtime = s->curtime; m->curstep = s->curstep; m->stepnum = s->stepnum;
  return PMX_OK;
}

int orc_step_times (double duration, double step, double curtime, double *times, int cap, int *nsteps);

/* dem(), parmec.cpp:2315-2530; adaptive in (0, 1] takes each step's
 * length from adaptive_timestep (2444-2447) and the host inputs from
 * fc->step_fn step by step */
static int dem_run (orc_ctx *s, double duration, double step, double adaptive, const pmx_forcing *fc, double *elapsed)
{
  if (!s) return fail (PMX_EINVAL, "null ctx");
  if (!(duration > 0.0) || !(step > 0.0)) return fail (PMX_EINVAL, "duration and step must be positive");
  auto t0 = std::chrono::steady_clock::now ();
  const bool adapt = adaptive > 0.0 && adaptive <= 1.0;
  s->adaptive = adapt;
  if (adapt) s->adp.assign (11 * (size_t) s->parnum, 0.0);
  std::vector<double> obsv (6 * (size_t) std::max (s->obsnum, 1)), prsv (6 * (size_t) std::max (s->prscols, 1));
  double step0, step1;
  if (s->curtime == 0.0)
  {
    step0 = step;
    euler (*s, 0.5*step0);
    shapes (*s);
    obstaclev (*s, fc ? fc->obstab0 : nullptr);   /* velocities at step0/2 */
    obstacles (*s, step0);
  }
  else step0 = s->curstep;
  step1 = step0;
  invert_inertia (*s);
  restrain_velocities (*s);
  Tree T;
  tree_create (s->ellnum, s->cen(0), s->cen(1), s->cen(2), T);
  long long k = 0, rebuilds = 1;
  for (double time = 0.0; time < duration; time += 0.5*(step0+step1), s->curtime += 0.5*(step0+step1), step0 = step1, k++)
  {
    if (tree_store (T, s->ellnum, s->cen(0), s->cen(1), s->cen(2)) > 0)
    {
      tree_create (s->ellnum, s->cen(0), s->cen(1), s->cen(2), T);
      rebuilds++;
      if (tree_store (T, s->ellnum, s->cen(0), s->cen(1), s->cen(2)) != 0) return fail (PMX_EREPART, "Repartitioning failed");
    }
    int bad = 0;
    if (s->ellnum > 0)
    {
      bad += update_existing (*s);
      bad += test_ellipsoids (*s, T);
      bad += test_triangles (*s, T);
    }
    if (bad) return fail (PMX_EUNSUPPORTED, "ellipsoid (non-sphere) shapes are not supported");
    double g[3], damp[6];
    const double *prow, *orow;
    pmx_step_inputs in;
    in.output = 0;
    if (adapt)
    {
      /* the step's inputs at its curtime (read_gravity_and_damping, prescribe,
       * obstaclev at curtime+step0) */
      for (int d = 0; d < 3; d++) in.gravity[d] = fc ? fc->gravity[d] : 0.0;
      for (int d = 0; d < 6; d++) in.damping[d] = fc ? fc->damping[d] : 0.0;
      std::fill (obsv.begin (), obsv.end (), 0.0);
      std::fill (prsv.begin (), prsv.end (), 0.0);
      in.obsvel = obsv.data ();
      in.prs = prsv.data ();
      in.hisrow = -1;
      if (fc && fc->step_fn && fc->step_fn (fc->step_user, (int) k, s->curtime, step0, &in))
        return fail (PMX_EINVAL, "step_fn failed");
      for (int d = 0; d < 3; d++) g[d] = in.gravity[d];
      for (int d = 0; d < 6; d++) damp[d] = in.damping[d];
      prow = prsv.data ();
      orow = obsv.data ();
    }
    else
    {
      for (int d = 0; d < 3; d++) g[d] = fc && fc->gtab ? fc->gtab[3*k+d] : (fc ? fc->gravity[d] : 0.0);
      for (int d = 0; d < 6; d++) damp[d] = fc && fc->dtab ? fc->dtab[6*k+d] : (fc ? fc->damping[d] : 0.0);
      prow = fc && fc->prstab ? fc->prstab + 6*(size_t) k * (fc->prscols > 0 ? fc->prscols : s->prscols) : nullptr;
      orow = fc && fc->obstab ? fc->obstab + 6*(size_t) k * s->obsnum : nullptr;
    }
    if (forces (*s, step0, g)) return fail (PMX_EUNSUPPORTED, "unsupported contact kind");
    if (s->sprnum && springs (*s, step0)) return fail (PMX_EUNSUPPORTED, "unsupported spring kind");
    if (s->trqsprnum) torsion_springs (*s, step0);
    if (s->unsprnum) unspring_test (*s);
    step1 = adapt ? adaptive_timestep (*s, step0, adaptive) : step0;
    if (!(step1 > 0.0)) return fail (PMX_EINVAL, "adaptive step is not positive");
    if (solve_joints (*s, damp, step0, step1))
      return fail (PMX_EINVAL, "zero pivot in the ball joint system (skyline LU, as amgcl's precondition)");
    restrain_forces (*s);
    prescribe_acceleration (*s, prow);
    dynamics (*s, damp, step0, step1);
    prescribe_velocity (*s, prow);
    shapes (*s);
    obstaclev (*s, orow);   /* at curtime+step0 */
    obstacles (*s, 0.5*(step0+step1));
    /* output points (parmec.cpp:2502-2507): the caller reads the state */
    if (adapt ? in.output != 0 : false)
    {
      if (fc && fc->out_fn) fc->out_fn (fc->out_user, (int) k);
    }
    else if (fc && fc->out_fn && fc->outstep)
      for (int q = 0; q < fc->noutstep; q++)
        if (fc->outstep[q] == k) fc->out_fn (fc->out_user, (int) k);
  }
  s->adaptive = false;
  s->curstep = step1;
  s->stepnum++;
  size_t pairs = 0;
  for (auto &ml : s->master) pairs += ml.size ();
  s->stats.steps = k;
  s->stats.rebuilds = rebuilds;
  s->stats.pairs = (long long) pairs;
  s->stats.leaves = (int) T.leaf.size ();
  double dt = std::chrono::duration<double> (std::chrono::steady_clock::now () - t0).count ();
  s->stats.seconds = dt;
  if (elapsed) *elapsed = dt;
  return PMX_OK;
}

int orc_dem (orc_ctx *s, double duration, double step, const pmx_forcing *fc, double *elapsed)
{
  return dem_run (s, duration, step, 0.0, fc, elapsed);
}

int orc_dem_adaptive (orc_ctx *s, double duration, double step, double adaptive, const pmx_forcing *fc, double *elapsed)
{
  if (!(adaptive > 0.0 && adaptive <= 1.0)) return fail (PMX_EINVAL, "adaptive must lie in (0, 1]");
  return dem_run (s, duration, step, adaptive, fc, elapsed);
}

int orc_critical (orc_ctx *s, int which, double *h, double *o, double *r)
{
  if (!s || !h) return fail (PMX_EINVAL, "null argument");
  if (which == PMX_CRIT_SPRING) critical_perspring (*s, h, o, r);
  else if (which == PMX_CRIT_PARTICLE) critical_perparticle (*s, h, o, r);
  else return fail (PMX_EINVAL, "unknown estimate");
  return PMX_OK;
}

void orc_trig (int fn, const double *x, double *y, int n)
{
  for (int i = 0; i < n; i++)
    y[i] = fn == 0 ? pmxt::sin (x[i]) : fn == 1 ? pmxt::cos (x[i]) : pmxt::acos (x[i]);
}

double orc_time_step (const double kact[3], double kmax, double emax, const double krot[6], double mass,
                      const double inertia[9], double *o, double *r)
{
  return time_step (kact, kmax, emax, krot, mass, inertia, o, r);
}

double orc_extremum_eigenvalue (const double A[9], int minimum) { return extremum_eigenvalue (A, minimum != 0); }

int orc_step_times (double duration, double step, double curtime, double *times, int cap, int *nsteps)
{
  int k = 0;
  for (double time = 0.0; time < duration; time += 0.5*(step+step), curtime += 0.5*(step+step), k++)
    if (times && k < cap) times[k] = curtime;
  if (nsteps) *nsteps = k;
  return PMX_OK;
}

int orc_contacts (orc_ctx *s, pmx_pair *out, size_t cap, size_t *count)
{
  if (!s) return fail (PMX_EINVAL, "null ctx");
  std::vector<pmx_pair> all;
  for (auto &ml : s->master)
    for (const Contact &c : ml)
    {
      pmx_pair p;
      p.master = c.master;
      p.slave = c.slave1;
      for (int d = 0; d < 3; d++) p.state[d] = c.state[d];
      all.push_back (p);
    }
  std::sort (all.begin (), all.end (), [] (const pmx_pair &a, const pmx_pair &b) {
    return a.master < b.master || (a.master == b.master && side_key (a.slave) < side_key (b.slave)); });
  if (count) *count = all.size ();
  if (out) std::memcpy (out, all.data (), std::min (cap, all.size ()) * sizeof (pmx_pair));
  return PMX_OK;
}

int orc_get_stats (orc_ctx *s, pmx_stats *st)
{
  if (!s || !st) return fail (PMX_EINVAL, "null argument");
  *st = s->stats;
  return PMX_OK;
}

void orc_expmap (const double omega[3], double lambda[9]) { expmap (omega[0], omega[1], omega[2], lambda); }

int orc_skyline_lu (int n, const double *A, const double *b, double *x)
{
  std::vector<int> ptr (n + 1), col ((size_t) n * n);
  for (int i = 0; i <= n; i++) ptr[i] = i * n;
  for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) col[(size_t) i*n + j] = j;
  std::vector<double> val (A, A + (size_t) n * n), rhs (b, b + n), y;
  if (skyline_lu (n, ptr, col, val, rhs, y)) return 1;
  std::copy (y.begin (), y.end (), x);
  return 0;
}

void orc_cuthill_mckee (int n, const int *ptr, const int *col, int *perm)
{
  std::vector<int> p (ptr, ptr + n + 1), c (col, col + ptr[n]), out;
  cuthill_mckee (n, p, c, out);
  std::copy (out.begin (), out.end (), perm);
}

double orc_triangle_sphere (const double a[3], const double b[3], const double c[3],
                            const double p[3], double r, double point[3], double normal[3])
{
  return triangle_sphere (a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], p[0], p[1], p[2], r, point, normal);
}

unsigned int orc_morton (const double p[3], const double extents[6]) { return morton (p[0], p[1], p[2], extents); }

long long orc_tree_pairs (int n, const double *center, const double *radius, int *out, long long cap)
{
  Tree T;
  const double *cx = center, *cy = center + n, *cz = center + 2*(size_t) n;
  tree_create (n, cx, cy, cz, T);
  if (tree_store (T, n, cx, cy, cz) > 0) return -1;
  std::vector<std::pair<int,int>> pr;
  for (int i = 0; i < n; i++)
  {
    double p[3] = {cx[i], cy[i], cz[i]}, rx = radius[i];
    double lo[3] = {p[0]-rx, p[1]-rx, p[2]-rx}, hi[3] = {p[0]+rx, p[1]+rx, p[2]+rx};
    std::vector<int> leaves;
    tree_reach (T, 0, lo, hi, leaves);
    for (int L : leaves)
      for (int j : T.leaf[L])
      {
        double q[3] = {p[0]-cx[j], p[1]-cy[j], p[2]-cz[j]};
        double len = std::sqrt (q[0]*q[0] + q[1]*q[1] + q[2]*q[2]);
        if (rx + radius[j] - len > 0.0 && i != j) pr.push_back ({std::min (i, j), std::max (i, j)});
      }
  }
  std::sort (pr.begin (), pr.end ());
  pr.erase (std::unique (pr.begin (), pr.end ()), pr.end ());
  long long m = (long long) pr.size ();
  for (long long k = 0; k < m && k < cap; k++) { out[2*k] = pr[k].first; out[2*k+1] = pr[k].second; }
  return m;
}

/* the leaf each centre descends to in a fresh tree (partitioning_create +
 * drop_ellipsoid, partition.ispc:600-660, 663-746), leaves numbered in the
 * left-first depth-first order of partitioning_tree_create (525-569),
 * whether or not the leaf overflows; returns the number of leaves and, in
 * *overflow, the count tree_store reports */
int orc_tree_leaves (int n, const double *center, int *leafof, int *overflow)
{
  Tree T;
  const double *cx = center, *cy = center + n, *cz = center + 2*(size_t) n;
  tree_create (n, cx, cy, cz, T);
  *overflow = tree_store (T, n, cx, cy, cz);
  const double *c[3] = {cx, cy, cz};
  for (int i = 0; i < n; i++)
  {
    int node = 0;
    while (T.node[node].dimension >= 0)
    {
      int d = T.node[node].dimension;
      node = c[d][i] < T.node[node].coord ? T.node[node].left : T.node[node].right;
    }
    leafof[i] = T.node[node].leaf;
  }
  return (int) T.leaf.size ();
}

/* the leaves of a fresh tree over the centres that each box [lo, hi]
 * (box[6*b .. 6*b+5] = lo xyz, hi xyz) reaches by the traversal rule of
 * condet.ispc:104-113 (left when lo[d] <= coord, right when hi[d] >= coord),
 * in visiting order; off[b] .. off[b+1] index `leaves` (at most cap written);
 * returns the total count */
long long orc_tree_reach (int n, const double *center, int nbox, const double *box, long long *off, int *leaves,
                          long long cap)
{
  Tree T;
  const double *cx = center, *cy = center + n, *cz = center + 2*(size_t) n;
  tree_create (n, cx, cy, cz, T);
  long long k = 0;
  for (int b = 0; b < nbox; b++)
  {
    off[b] = k;
    std::vector<int> out;
    tree_reach (T, 0, box + 6*(size_t) b, box + 6*(size_t) b + 3, out);
    for (int L : out) { if (k < cap) leaves[k] = L; k++; }
  }
  off[nbox] = k;
  return k;
}

/* broad-phase statistics of a fresh tree: [0] leaves, [1] mean leaves reached
 * per sphere AABB, [2] mean leaf tests per sphere, [3] mean candidate leaves
 * of a leaf's union box, [4] max candidates, [5] mean spheres per leaf */
void orc_tree_stats (int n, const double *center, const double *radius, double *out)
{
  Tree T;
  const double *cx = center, *cy = center + n, *cz = center + 2*(size_t) n;
  tree_create (n, cx, cy, cz, T);
  tree_store (T, n, cx, cy, cz);
  double reach = 0, tests = 0;
  size_t L = T.leaf.size ();
  std::vector<double> blo (3*L, 1e300), bhi (3*L, -1e300);
  std::vector<int> lof (n);
  for (size_t l = 0; l < L; l++) for (int j : T.leaf[l]) lof[j] = (int) l;
  for (int i = 0; i < n; i++)
  {
    double p[3] = {cx[i], cy[i], cz[i]}, rx = radius[i];
    double lo[3] = {p[0]-rx, p[1]-rx, p[2]-rx}, hi[3] = {p[0]+rx, p[1]+rx, p[2]+rx};
    std::vector<int> leaves;
    tree_reach (T, 0, lo, hi, leaves);
    reach += leaves.size ();
    for (int l : leaves) tests += T.leaf[l].size ();
    int l = lof[i];
    for (int k = 0; k < 3; k++) { blo[3*l+k] = std::min (blo[3*l+k], lo[k]); bhi[3*l+k] = std::max (bhi[3*l+k], hi[k]); }
  }
  double cand = 0, cmax = 0;
  for (size_t l = 0; l < L; l++)
  {
    if (T.leaf[l].empty ()) continue;
    std::vector<int> leaves;
    tree_reach (T, 0, &blo[3*l], &bhi[3*l], leaves);
    cand += leaves.size ();
    cmax = std::max (cmax, (double) leaves.size ());
  }
  out[0] = (double) L; out[1] = reach / n; out[2] = tests / n; out[3] = cand / L; out[4] = cmax; out[5] = (double) n / L;
}

} /* extern "C" */
