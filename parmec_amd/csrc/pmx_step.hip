/*
 * pmx_step.hip -- per-step kernels of the FP64 DEM step (the body of
 * dem()'s loop, parmec.cpp:2418-2517, for spheres and static triangles).
 *
 * Contact bookkeeping.  The reference keeps per-particle master lists that
 * persist between steps and slave copies rebuilt every step (condet.h,
 * forces.ispc:195-407).  Here every sphere owns a row: its contacts sorted by
 * the partner's identity (partner ellipsoid, or 0x80000000|triangle), stored
 * slot-major (slot k of sphere e at k*E+e) in ping-pong buffers.  Both
 * spheres of a pair hold the entry; the tangential state lives in the master's
 * row (master = sphere of the smaller particle index, as condet.ispc:187).
 * Each step:
 *   bin     : centre descends the frozen tree to its leaf (partition.ispc:600-646)
 *   scatter : leaf-ordered copy of centres/radii
 *   detect  : sphere AABB reaches leaves (condet.ispc:98-113); hits with positive
 *             depth are new-pair candidates.  A hit whose partner's AABB does not
 *             reach this sphere's leaf is also posted to the partner (the
 *             reference finds such a pair from one side only).  Triangles come
 *             from per-leaf lists of triangles whose AABB reaches the leaf
 *             (condet.ispc:412-426).
 *   rows    : merge old row with hits; recompute geometry (update_existing,
 *             condet.ispc:627-700); drop pairs with depth < 0 (forces.ispc:329-364);
 *             granular law per pair (forces.ispc:125-192, evaluated identically
 *             by both spheres of a pair from the master's side); per-sphere sum
 *             in row order plus m*g (forces.ispc:453-596).
 *   dynamics, shapes (dynamics.ispc:30-169, shapes.ispc:29-98).
 */
#include "pmx_internal.h"
#include "pmx_device.h"
#include <algorithm>
#include <cfloat>

using namespace pmxd;

namespace {

inline dim3 grid (int n) { return dim3 ((n + PMX_TPB - 1) / PMX_TPB); }

__global__ void k_zero_i (int *a, int n)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

/* the step completed: advance the step counter (unless it was aborted) and
 * the sequence number of enqueued steps */
__global__ void k_zero_u64 (unsigned long long *a, int n)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ void k_step_end (int *ctl)
{
  if (threadIdx.x != 0) return;
  if (!ctl[CTL_ABORT]) ctl[CTL_KSTEP] += 1;
  ctl[CTL_SEQ] += 1;
  /* the fused binning for the next step overflowed a leaf: this step is
   * complete; the next is aborted, to be repeated after the re-partition */
  if (ctl[CTL_REBIN])
  {
    ctl[CTL_REBIN] = 0;
    if (!ctl[CTL_ABORT]) ctl[CTL_ABORT] = ctl[CTL_SEQ] + 1;
  }
}

/* the leaf a centre descends to ("c[d] < coord -> left", partition.ispc:
 * 600-646) from the instance's root; or the hint leaf h (the sphere's leaf
 * in the previous step, h < 0: none) when the centre lies in h's reach
 * cell, lower <= c < upper along every axis, which is where the descent
 * ends */
__device__ __forceinline__ int bin_leaf (const double c[3], int h, int root, const int *left, const int *right,
                                         const int *dim, const double *coord, const double *cell, int nleaves)
{
  if (h >= 0)
  {
    bool in = true;
#pragma unroll
    for (int k = 0; k < 3; k++)
      in = in && cell[(size_t) k*nleaves + h] <= c[k] && c[k] < cell[(size_t) (3+k)*nleaves + h];
    if (in) return h;
  }
  int ref = root;
  while (ref >= 0)
  {
    int d = dim[ref];
    ref = c[d] < coord[ref] ? left[ref] : right[ref];
  }
  return -ref-1;
}

/* slot = leaf count before this sphere: the wave's spheres grouped by
 * leaf, then one atomic per group, all groups' atomics in one instruction.
 * Called by every active lane of the wave; valid: the lane bins a sphere
 * into leaf L */
__device__ __forceinline__ int bin_slot (bool valid, int L, int *lcount)
{
  int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot (valid);
  unsigned long long lt = lane ? ((1ull << lane) - 1ull) : 0ull;
  int leader = lane;
  unsigned long long grp = 0;
  while (todo)
  {
    int ld = __ffsll ((long long) todo) - 1;
    int Lw = __shfl (L, ld, 64);
    unsigned long long g = __ballot (valid && L == Lw) & todo;
    if ((g >> lane) & 1ull) { leader = ld; grp = g; }
    todo &= ~g;
  }
  int base = 0;
  if (valid && leader == lane) base = atomicAdd (&lcount[L], __popcll (grp));
  base = __shfl (base, leader, 64);
  return base + __popcll (grp & lt);
}

/* a sphere's record in slot s of leaf L (the slotted layout) */
__device__ __forceinline__ void bin_record (size_t pos, int e, int pe, const double c[3], double rad, int *lmem,
                                            int *lpart, double *lpos)
{
  lmem[pos] = e;
  lpart[pos] = pe;
  double2 *rec = (double2*) (lpos + 4*pos);
  rec[0] = make_double2 (c[0], c[1]);
  rec[1] = make_double2 (c[2], rad);
}

/* partitioning_store: descend to the leaf, take a slot (partition.ispc:600-646) */
/* the leaf of each sphere and its slot there.  With lstride (the slotted
 * layout: PMX_LSIZE slots per leaf) the sphere's record, member and
 * particle go straight to slot s of its leaf and its reverse-hit count is
 * cleared, the work k_scatter does for the packed layout */
__global__ void k_bin (int E, const double *center, const int *roots, const int *ell_inst, const int *left,
                       const int *right, const int *dim, const double *coord, const double *cell, int nleaves, int *ctl,
                       int *leafof, int *slot, int *lcount, int *flags, int *ovf, const int *stop, int lstride,
                       const double *radii, const int *part, int *lmem, double *lpos, int *lpart, int *xcnt)
{
  /* an abort of an earlier step: nothing to do.  An abort raised by this
   * launch (another sphere's leaf overflowed) does not stop the binning:
   * every instance with an overflowing leaf must be flagged, so that one
   * re-partition serves them all, as each would re-partition alone */
  int seq = ctl[CTL_SEQ], ab = ctl[CTL_ABORT];
  if (ab && ab != seq + 1) return;
  int hint = ctl[CTL_HINT];
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  bool valid = e < E && !frozen (stop, ell_inst[e < E ? e : 0], ctl);
  int L = -1, hprev = -1;
  double c[3] = {0.0, 0.0, 0.0}, rad = 0.0;
  int pe = 0;
  if (valid)
  {
    c[0] = center[e]; c[1] = center[(size_t) E + e]; c[2] = center[2*(size_t) E + e];
    if (lstride) { rad = radii[e]; pe = part[e]; }   /* in flight with the centre */
    if (hint) hprev = leafof[e];
    L = bin_leaf (c, hprev, roots[ell_inst[e]], left, right, dim, coord, cell, nleaves);
  }
  int s = bin_slot (valid, L, lcount);
  if (!valid) return;
  if (L != hprev) leafof[e] = L;   /* unchanged where the hint held */
  if (s >= PMX_LSIZE)
  {
    atomicAdd (&flags[FLAG_LEAF_OVERFLOW], 1);
    atomicExch (&ovf[ell_inst[e]], 1);   /* the instance whose tree is rebuilt */
    atomicExch (&ctl[CTL_ABORT], seq + 1);   /* re-partition before this step goes on */
    return;
  }
  if (!lstride) { slot[e] = s; return; }
  bin_record ((size_t) L * lstride + s, e, pe, c, rad, lmem, lpart, lpos);
}

__global__ void k_scatter (int *ctl, int E, const double *center, const double *radii, const int *leafof, const int *slot,
                           const int *loff, int *lmem, double *lpos, const int *stop, const int *ell_inst, int *xcnt,
                           const int *part, int *lpart)
{
  if (ctl[CTL_ABORT]) return;
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E || frozen (stop, ell_inst[e], ctl)) return;
  int pos = loff[leafof[e]] + slot[e];
  lmem[pos] = e;
  lpart[pos] = part[e];
  double2 *rec = (double2*) (lpos + 4*(size_t) pos);
  rec[0] = make_double2 (center[e], center[(size_t) E + e]);
  rec[1] = make_double2 (center[2*(size_t) E + e], radii[e]);
}

/* FP32 pre-test of a (sphere, record) pair on offsets from an origin O,
 * x32 = (float) (x - O), r32 = (float) r, with |c_a - O| <= B per axis for
 * the block sphere a.  With u = 2^-24: a pair that the FP64 test accepts has
 * |c_a - c_b| < (r_a + r_b)(1 + 1e-15) and so |c_b - O| <= B + |c_a - c_b|;
 * the rounding of the offsets (1.01u relative), of the difference, the
 * squares and sums (u per operation) and of s = r32_a + r32_b then keep
 * sqrt(d2) <= s (1 + 5.7u) + 3.6u B, below thr = s (1 + 16u) + 8u B + 1e-18
 * even after thr and thr^2 are rounded: such a pair is never rejected.  B >
 * 1e15 (FP32 offsets could overflow) turns the pre-test off (thr = inf);
 * NaN and flushed denormals only keep pairs. */
__device__ __forceinline__ bool near32 (float4 a, float4 b, float k1, float k2)
{
  float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  float d2 = dx*dx + dy*dy + dz*dz;
  float thr = (a.w + b.w) * k1 + k2;
  return !(d2 > thr * thr);
}

/* Workgroups are dealt to the 8 XCDs round robin (block b to XCD b % 8),
 * each XCD with its own L2.  Block b takes item xcd_block (b, n) so that
 * each XCD walks one contiguous stretch of the items: neighbouring leaves
 * (near in Morton order) then meet in the same L2.  A bijection of [0, n). */
__device__ __forceinline__ int xcd_block (int b, int n)
{
  int x = b & 7, k = b >> 3, q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + k;
}

/* bits of x set in the lanes below this one (mbcnt: no lane mask held in
 * registers) */
__device__ __forceinline__ int below (unsigned long long x)
{
  return (int) __builtin_amdgcn_mbcnt_hi ((unsigned) (x >> 32), __builtin_amdgcn_mbcnt_lo ((unsigned) x, 0u));
}

__device__ __forceinline__ unsigned long long readlane_u64 (unsigned long long x, int k)
{
  unsigned lo = (unsigned) __builtin_amdgcn_readlane ((int) (unsigned) x, k);
  unsigned hi = (unsigned) __builtin_amdgcn_readlane ((int) (unsigned) (x >> 32), k);
  return ((unsigned long long) hi << 32) | lo;
}

/* FP32 bounds below / above any double v.  (float) v rounds to nearest,
 * within half an ulp of the result (a quarter at a binade's bottom);
 * stepping out by |f| 2^-23 (rounded to nearest: at least one ulp) and by
 * 2^-120 (above every subnormal rounding) stays on the safe side, and the
 * overflowed cases (inf - inf = NaN, or v beyond the float range) end at
 * the largest finite float, which is still on the safe side of v */
__device__ __forceinline__ float f32_below (double v)
{
  float f = (float) v;
  return fminf (f - fabsf (f) * 0x1p-23f - 0x1p-120f, FLT_MAX);
}

__device__ __forceinline__ float f32_above (double v)
{
  float f = (float) v;
  return fmaxf (f + fabsf (f) * 0x1p-23f + 0x1p-120f, -FLT_MAX);
}

/* a sphere's outward-rounded AABB (lo, hi) against a cell packed as
 * (lo x, lo y), (lo z, hi x), (hi y, hi z) */
__device__ __forceinline__ bool reach32 (const float *lo, const float *hi, float2 a, float2 b, float2 e)
{
  /* bitwise: one straight line of compares, no branches */
  return (lo[0] <= b.y) & (lo[1] <= e.x) & (lo[2] <= e.y) & (hi[0] >= a.x) & (hi[1] >= a.y) & (hi[2] >= b.x);
}

/* inv20[n] = ceil(2^20 / n), n = 1..PMX_LSIZE: for q n < 2^20 (here q <= 64,
 * n <= 96), q / n = (q inv20[n]) >> 20 exactly; a scalar load and 24-bit
 * multiplies in place of an integer division */
struct Inv20
{
  unsigned v[PMX_LSIZE + 1];
  constexpr Inv20 () : v ()
  {
    for (int n = 1; n <= PMX_LSIZE; n++) v[n] = ((1u << 20) + (unsigned) n - 1u) / (unsigned) n;
  }
};
__constant__ Inv20 inv20 = Inv20 ();

/* (q / n, q mod n) for 0 <= q <= 64, 1 <= n <= PMX_LSIZE */
__device__ __forceinline__ int2 divmod_small (int q, int n)
{
  int d = (int) (__umul24 ((unsigned) q, inv20.v[n]) >> 20);
  return make_int2 (d, q - (int) __umul24 ((unsigned) d, (unsigned) n));
}

/* k_detect runs NW waves per block (1, 2, 4 or 8; the block's spheres are
 * spread over its threads); NW is picked per workload by the autotuner */

/* one sphere-sphere test of the reference leaf test (condet.ispc:128-144):
 * depth = r_a + r_b - |c_a - c_b| > 0, with the correctly rounded sqrt.
 * d2 >= sr^2 (1 + 1e-10) implies sqrt(d2) rounds to >= sr, hence depth <= 0;
 * d2 < sr^2 (1 - 1e-10) (sr^2 finite) implies sqrt(d2) < sr (1 - 4.9e-11),
 * which rounds below sr, hence depth > 0: only pairs in between take the
 * sqrt (the roundings of sr^2 and of the bounds are below 3e-16) */
__device__ __forceinline__ bool overlap (double ax, double ay, double az, double ar, double bx, double by, double bz,
                                         double br)
{
  double qx = ax-bx, qy = ay-by, qz = az-bz;
  double d2 = qx*qx + qy*qy + qz*qz;
  double sr = ar + br;
  double s2 = sr * sr;
  if (d2 >= s2 * 1.0000000001) return false;
  if (d2 < s2 * 0.9999999999 && s2 <= DBL_MAX) return true;
  asm volatile ("");   /* a real branch: the sqrt is not computed for every lane */
  double len = sqrt (d2);
  return sr - len > 0.0;
}

/* Detection (test_ellipsoids + test_triangles, condet.ispc:98-516), one block
 * per leaf Q.  A sphere reaches leaf L through the reference descent iff its
 * AABB overlaps L's cell.  The block's spheres sit in LDS; each wave takes
 * candidate leaves (Q's neighbour list) in turn: it forms the list of the
 * block's spheres whose AABB reaches the candidate (two ballots), stages the
 * candidate's records in its own LDS area and spreads the (sphere, record)
 * tests over its 64 lanes.  Waves never wait for each other inside the
 * candidate loop.  Hit slots per sphere come from LDS counters; the hit
 * order is irrelevant (rows sort their keys). */
/* the box of each leaf's records in this step, one wave per leaf: the AABB
 * of its spheres (c -/+ r per axis) widened by a relative margin, a pair
 * that the FP64 test accepts having |c_a - c_b| < (r_a + r_b)(1 + 2e-10)
 * along every axis, lbox [6][nleaves] (lo x y z, hi x y z).  k_detect takes
 * its own leaf's box from here, and skips a reached neighbour whose box
 * misses it: each of that neighbour's records then misses the own box by
 * more than the neighbour's margin, which bounds the margin of the record
 * filter (1e-9 (|c| + r) <= 1e-9 (|lo| + |hi| + (hi - lo))), so the filter
 * would have kept none of them.  An empty leaf gets zeros (no sphere reads
 * it: an empty candidate is skipped by its count). */
/* a DPP lane move of a double: the value of the lane the control selects,
 * or the lane's own where there is none */
template <int CTRL, int ROWS, int BANKS>
__device__ __forceinline__ double dpp_f64 (double v)
{
  int lo = __double2loint (v), hi = __double2hiint (v);
  lo = __builtin_amdgcn_update_dpp (lo, lo, CTRL, ROWS, BANKS, false);
  hi = __builtin_amdgcn_update_dpp (hi, hi, CTRL, ROWS, BANKS, false);
  return __hiloint2double (hi, lo);
}

/* the minimum over the wave's 64 lanes, in lane 63 (the other lanes hold
 * partial minima): row_shr 1, 2, 3, 4, 8 within rows of 16, then
 * row_bcast 15 and 31 across rows.  min is exact, so the order of the
 * folds does not change the result */
__device__ __forceinline__ double wave_min63 (double v)
{
  v = fmin (v, dpp_f64<0x111, 0xf, 0xf> (v));   /* row_shr:1 */
  v = fmin (v, dpp_f64<0x112, 0xf, 0xf> (v));   /* row_shr:2 */
  v = fmin (v, dpp_f64<0x113, 0xf, 0xf> (v));   /* row_shr:3 */
  v = fmin (v, dpp_f64<0x114, 0xf, 0xe> (v));   /* row_shr:4 */
  v = fmin (v, dpp_f64<0x118, 0xf, 0xc> (v));   /* row_shr:8 */
  v = fmin (v, dpp_f64<0x142, 0xa, 0xf> (v));   /* row_bcast:15 */
  v = fmin (v, dpp_f64<0x143, 0xc, 0xf> (v));   /* row_bcast:31 */
  return v;
}

__global__ __launch_bounds__(256)
void k_leafbox (int nleaves, const int *lcount, const int *loff, const double *lpos, int lstride, double *lbox,
                const int *ctl)
{
  if (ctl[CTL_ABORT]) return;
  int L = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (L >= nleaves) return;
  /* slotted records: the first 64 slots are read together with the count
   * (every slot of a leaf is in bounds; the ones past the count are not
   * used), one memory round trip instead of two */
  double2 pa = make_double2 (0.0, 0.0), pb = pa;
  if (lstride)
  {
    const double2 *rp = (const double2*) (lpos + 4*((size_t) L * lstride + lane));
    pa = rp[0];
    pb = rp[1];
  }
  int cnt = lcount[L], off = lstride ? L * lstride : loff[L];
  double blo[3] = {0.0, 0.0, 0.0}, bhi[3] = {0.0, 0.0, 0.0};
  if (cnt > 0)
  {
#pragma unroll
    for (int d = 0; d < 3; d++) { blo[d] = DBL_MAX; bhi[d] = -DBL_MAX; }
    for (int k = lane; k < cnt; k += 64)
    {
      double2 a = pa, b = pb;
      if (!lstride || k != lane)
      {
        const double2 *rp = (const double2*) (lpos + 4*(size_t) (off + k));
        a = rp[0];
        b = rp[1];
      }
      double c[3] = {a.x, a.y, b.x}, r = b.y;
#pragma unroll
      for (int d = 0; d < 3; d++)
      {
        blo[d] = fmin (blo[d], c[d] - r);
        bhi[d] = fmax (bhi[d], c[d] + r);
      }
    }
    /* the wave's min and max into lane 63 by DPP row shifts and
     * broadcasts (plain lane moves, no LDS crossbar) */
#pragma unroll
    for (int d = 0; d < 3; d++)
    {
      blo[d] = wave_min63 (blo[d]);
      bhi[d] = -wave_min63 (-bhi[d]);
    }
#pragma unroll
    for (int d = 0; d < 3; d++)
    {
      double tol = 1e-9 * (fabs (blo[d]) + fabs (bhi[d]) + (bhi[d] - blo[d]));
      blo[d] -= tol;
      bhi[d] += tol;
    }
  }
  if (lane == 63)
#pragma unroll
    for (int d = 0; d < 3; d++)
    {
      lbox[(size_t) d * nleaves + L] = blo[d];
      lbox[(size_t) (3+d) * nleaves + L] = bhi[d];
    }
}

/* 5 waves per SIMD: the register budget (96 VGPRs, a 16-byte spill) and
 * the LDS per one-wave block (7.3 KB) keep 20 blocks resident per CU, up
 * from 16 at the unbounded register count; detection is bound by the
 * latency of its dependent loads, so residency pays */
template <int NW>
__global__ __launch_bounds__(NW*64, 5)
void k_detect (int E, const double *center, const double *radii, const int *part, const double *cell, int nleaves,
               const int *lcount, const int *loff, const int *lmem, const double *lpos, const int *lpart,
               const double *tri, int T, const int *tricnt, const int *trioff, const int *trilist,
               const int *mtcnt, const int *mtlist, int mcap, const int *nbrcnt, const int *nbroff, const int *nbrlist,
               int nbrstride,
               uint32_t *hits, int *hitcnt, int hcap, int *xbuf, int *xcnt, int xcap, int *flags, int *ctl,
               const int *stop, const int *leafinst, int lstride, const double *lbox)
{
  if (ctl[CTL_ABORT]) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl[CTL_HINT] = 1;   /* this step's bin completed: leafof is valid */
  int Q = xcd_block (blockIdx.x, gridDim.x);
  if (frozen (stop, leafinst[Q], ctl)) return;   /* uniform over the block */
  /* LDS holds what the loops read on every pair; the block spheres'
   * members and particles are read from the leaf's slots (cache-resident)
   * on the rare paths that need them, so that more blocks stay resident */
  __shared__ double ssph[4*PMX_LSIZE];   /* block spheres: x, y, z, rx */
  __shared__ int shc[PMX_LSIZE];
  __shared__ float4 ssph32[PMX_LSIZE];       /* block spheres: FP32 offsets from O, radius */
  __shared__ float4 wrec32[NW][PMX_LSIZE];   /* per wave: the candidate's kept records, the same */
  __shared__ int wmem[NW][PMX_LSIZE], wlist[NW][PMX_LSIZE];
  int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int qoff = lstride ? Q * lstride : loff[Q], qcnt = lcount[Q];
  int nc = nbrcnt[Q];
  size_t c0 = nbrstride ? (size_t) Q * nbrstride : (size_t) nbroff[Q];
  int ncw = nc > w ? (nc - w + NW - 1) / NW : 0;   /* candidates of this wave: w, w + NW, ... */
  /* the leaf's triangle lists (static from the build, moving from this step) */
  int nsq = T > 0 ? tricnt[Q] : 0, toff = T > 0 ? trioff[Q] : 0;
  int nmq = mtcnt ? mtcnt[Q] : 0;
  nmq = nmq < mcap ? nmq : mcap;
  int ntq = nsq + nmq;
  /* the wave's first (usually only) batch of candidate leaves, one per
   * lane: leaf, count, offset and cell, loaded before the block's spheres
   * so that both dependent loads are in flight together */
  /* the AABB of the block's spheres, widened by a relative margin
   * (k_leafbox, from the same records; zeros for an empty leaf): scalar
   * loads, ahead of the candidates that are tested against it */
  double blo[3], bhi[3];
#pragma unroll
  for (int d = 0; d < 3; d++)
  {
    blo[d] = lbox[(size_t) d * nleaves + Q];
    bhi[d] = lbox[(size_t) (3+d) * nleaves + Q];
  }
  int bL = 0, bcnt = 0, boff = 0;
  bool bfar = false;   /* a neighbour whose records' box misses the block's: nothing to stage */
  double Ll[3] = {0.0, 0.0, 0.0}, Lu[3] = {0.0, 0.0, 0.0};
  auto batch = [&] (int p0) {
    int np = ncw - p0 < 64 ? ncw - p0 : 64;
    bL = 0; bcnt = 0; boff = 0; bfar = false;
    if (lane < np)
    {
      /* the lane index taken afresh (mbcnt): fewer registers held through
       * the candidate loop than with the first batch's address kept */
      int ln = (int) __builtin_amdgcn_mbcnt_hi (~0u, __builtin_amdgcn_mbcnt_lo (~0u, 0u));
      const int *nb = nbrlist + c0;
      bL = nb[w + (p0 + ln) * NW];
      bcnt = lcount[bL];
      boff = lstride ? bL * lstride : loff[bL];
#pragma unroll
      for (int d = 0; d < 3; d++)
      {
        Ll[d] = cell[(size_t) d*nleaves + bL];
        Lu[d] = cell[(size_t) (3+d)*nleaves + bL];
      }
      if (bL != Q)
#pragma unroll
        for (int d = 0; d < 3; d++)
          bfar = bfar || lbox[(size_t) d * nleaves + bL] > bhi[d] || lbox[(size_t) (3+d) * nleaves + bL] < blo[d];
    }
  };
  batch (0);
  /* the block's spheres from the leaf-ordered records (x, y, z, r0: a
   * sphere's r1, r2 are negative, so r0 is its AABB half-width) */
  for (int s = t; s < qcnt; s += NW*64)
  {
    const double2 *rp = (const double2*) (lpos + 4*(size_t) (qoff + s));
    double2 a = rp[0], b = rp[1];
    ssph[4*s] = a.x; ssph[4*s+1] = a.y; ssph[4*s+2] = b.x; ssph[4*s+3] = b.y;
    shc[s] = 0;
  }
  double cl[3] = {cell[Q], cell[(size_t) nleaves + Q], cell[2*(size_t) nleaves + Q]};
  double cu[3] = {cell[3*(size_t) nleaves + Q], cell[4*(size_t) nleaves + Q], cell[5*(size_t) nleaves + Q]};
  __syncthreads ();

  /* origin O of the FP32 offsets: the box centre; B bounds |c - O| per axis
   * for every block sphere */
  double O[3], B = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++)
  {
    O[d] = 0.5 * (blo[d] + bhi[d]);
    B = fmax (B, fmax (bhi[d] - O[d], O[d] - blo[d]));
  }
  const float k1 = 1.0f + 0x1p-20f;   /* 1 + 16u, u = 2^-24 */
  const float k2 = B > 1e15 ? INFINITY : (float) (0x1p-21 * B * (1.0 + 1e-6)) + 1e-18f;   /* 8u B + 1e-18 */
  for (int s = t; s < qcnt; s += NW*64)
    ssph32[s] = make_float4 ((float) (ssph[4*s] - O[0]), (float) (ssph[4*s+1] - O[1]), (float) (ssph[4*s+2] - O[2]),
                             (float) ssph[4*s+3]);
  __syncthreads ();

  float4 *rec32 = wrec32[w];
  int *mem = wmem[w], *lst = wlist[w];
  for (int p0 = 0; p0 < ncw; p0 += 64)
  {
    /* up to 64 candidates at once, one per lane, and which block spheres'
     * own AABB reaches each one's cell (condet.ispc:104-113) */
    int np = ncw - p0 < 64 ? ncw - p0 : 64;
    if (p0 > 0) batch (p0);
    /* which block spheres' own AABB reaches each candidate's cell
     * (condet.ispc:104-113), tested in FP32 with every bound rounded
     * outward: a mask holds every sphere that the FP64 test accepts and, at
     * a bound within FP32 rounding, rarely one more (its hits are confirmed
     * against the FP64 cell below).  The candidates' cells pass through the
     * wave's staging area (free until the candidates are processed) and are
     * read as LDS broadcasts; lane s tests block spheres s and s + 64 */
    float *cf = (float*) rec32;   /* 64 cells of 6 floats: lo x, y, z, hi x, y, z */
    __builtin_amdgcn_wave_barrier ();
    if (lane < np)
    {
#pragma unroll
      for (int d = 0; d < 3; d++)
      {
        cf[6*lane+d] = f32_below (Ll[d]);
        cf[6*lane+3+d] = f32_above (Lu[d]);
      }
    }
    __builtin_amdgcn_wave_barrier ();
    /* sphere AABB bounds as the FP64 test forms them (c -/+ r), rounded
     * outward; NaN (no comparison holds) past the block's spheres */
    float slo0[3], shi0[3], slo1[3], shi1[3];
#pragma unroll
    for (int k = 0; k < 2; k++)
    {
      int s = lane + 64*k;
      float *lo = k ? slo1 : slo0, *hi = k ? shi1 : shi0;
#pragma unroll
      for (int d = 0; d < 3; d++)
      {
        lo[d] = __builtin_nanf ("");
        hi[d] = __builtin_nanf ("");
      }
      if (s < qcnt)
      {
        double r = ssph[4*s+3];
#pragma unroll
        for (int d = 0; d < 3; d++)
        {
          lo[d] = f32_below (ssph[4*s+d] - r);
          hi[d] = f32_above (ssph[4*s+d] + r);
        }
      }
    }
    /* the candidate count of the wave, uniform: the loops below run with
     * the full wave and keep their counters scalar */
    /* the masks of the candidates with records whose box meets the
     * block's (one bit per lane, uniform), the others left empty */
    unsigned long long km = __ballot (lane < np && bcnt > 0 && !bfar);
    unsigned long long bm0 = 0ull, bm1 = 0ull;
    if (qcnt <= 64)
      for (unsigned long long k = km; k; k &= k - 1ull)
      {
        int c = __ffsll ((long long) k) - 1;
        const float2 *cp = (const float2*) (cf + 6*c);
        unsigned long long r0 = __ballot (reach32 (slo0, shi0, cp[0], cp[1], cp[2]));
        bm0 = lane == c ? r0 : bm0;
      }
    else
      for (unsigned long long k = km; k; k &= k - 1ull)
      {
        int c = __ffsll ((long long) k) - 1;
        const float2 *cp = (const float2*) (cf + 6*c);
        float2 a = cp[0], b = cp[1], e = cp[2];
        unsigned long long r0 = __ballot (reach32 (slo0, shi0, a, b, e));
        unsigned long long r1 = __ballot (reach32 (slo1, shi1, a, b, e));
        bm0 = lane == c ? r0 : bm0;
        bm1 = lane == c ? r1 : bm1;
      }
    __builtin_amdgcn_wave_barrier ();
    /* the candidates that some block sphere reaches, in turn */
    unsigned long long todo = __ballot ((bm0 | bm1) != 0);
    while (todo)
    {
      int k = __ffsll ((long long) todo) - 1;
      todo &= todo - 1ull;
      int L = __builtin_amdgcn_readlane (bL, k);
      int cnt = __builtin_amdgcn_readlane (bcnt, k);
      int off = __builtin_amdgcn_readlane (boff, k);
      unsigned long long m0 = readlane_u64 (bm0, k), m1 = readlane_u64 (bm1, k);
      int n0 = __popcll (m0), nsl = n0 + __popcll (m1);
      bool own = L == Q;
      /* stage: the reaching spheres' slots and the candidate's records */
      __builtin_amdgcn_wave_barrier ();
      if ((m0 >> lane) & 1ull) lst[below (m0)] = lane;
      if ((m1 >> lane) & 1ull) lst[n0 + below (m1)] = lane + 64;
      if (!own)
      {
        /* only the records whose AABB meets the block spheres' box, packed */
        int kept = 0;
        for (int b0 = 0; b0 < cnt; b0 += 64)
        {
          int kk = b0 + lane;
          bool ok = false;
          double2 x = make_double2 (0.0, 0.0), y = x;
          int mj = 0;
          if (kk < cnt)
          {
            const double2 *rp = (const double2*) (lpos + 4*(size_t) (off + kk));
            x = rp[0]; y = rp[1];
            mj = lmem[off + kk];
            double c[3] = {x.x, x.y, y.x}, r = y.y;
            ok = true;
#pragma unroll
            for (int d = 0; d < 3; d++)
            {
              double tb = 1e-9 * (fabs (c[d]) + r);
              ok = ok && (c[d] - r) - bhi[d] <= tb && blo[d] - (c[d] + r) <= tb;
            }
          }
          unsigned long long mk = __ballot (ok);
          if (ok)
          {
            int p = kept + below (mk);
            rec32[p] = make_float4 ((float) (x.x - O[0]), (float) (x.y - O[1]), (float) (y.x - O[2]), (float) y.y);
            mem[p] = mj;
          }
          kept += __popcll (mk);
        }
        cnt = kept;
      }
      __builtin_amdgcn_wave_barrier ();
      if (cnt == 0) continue;
      if (own)
      {
        /* own leaf: every sphere reaches it and the test is symmetric
         * (depth r_a + r_b - |c_a - c_b|), so each unordered pair is tested
         * once: pairs (a, (a+k) mod n), k = 1..(n-1)/2, plus k = n/2, a < n/2 */
        int n = cnt, halfn = (n - 1) / 2;
        /* the FP64 test and the hits of a pair that passed the FP32 one */
        auto pair = [&] (int a, int bb) {
          if (!overlap (ssph[4*a], ssph[4*a+1], ssph[4*a+2], ssph[4*a+3],
                        ssph[4*bb], ssph[4*bb+1], ssph[4*bb+2], ssph[4*bb+3]))
            return;
          if (lpart[qoff + a] == lpart[qoff + bb]) return;
          int ea = lmem[qoff + a], eb = lmem[qoff + bb];
          int sa = atomicAdd (&shc[a], 1), sb = atomicAdd (&shc[bb], 1);
          if (sa < hcap) hits[(size_t) sa * E + ea] = (uint32_t) eb; else { atomicAdd (&flags[FLAG_HIT_OVERFLOW], 1); abort_step (ctl); }
          if (sb < hcap) hits[(size_t) sb * E + eb] = (uint32_t) ea; else { atomicAdd (&flags[FLAG_HIT_OVERFLOW], 1); abort_step (ctl); }
        };
        /* rx = max (r0, r1, r2) is r0 for a sphere (r1, r2 < 0) */
        if (n <= 64)
        {
          /* lane a keeps sphere a and takes its partners (a + k) mod n two
           * at a time (two independent LDS reads in flight), k = 1..(n-1)/2,
           * then k = n/2 on the lanes a < n/2; idle lanes read slot 0 */
          int kmax = halfn + ((n & 1) ? 0 : 1);
          bool mine = lane < n;
          float4 A = ssph32[mine ? lane : 0];
          int bb = mine ? lane : 0;
          for (int k = 1; k <= kmax; k += 2)
          {
            int b1 = bb + 1 == n ? 0 : bb + 1;
            int b2 = b1 + 1 == n ? 0 : b1 + 1;
            bb = b2;
            bool t1 = lane < (k <= halfn ? n : n / 2);
            bool t2 = k < kmax && lane < (k + 1 <= halfn ? n : n / 2);
            float4 B1 = ssph32[b1], B2 = ssph32[b2];
            bool p1 = t1 & near32 (A, B1, k1, k2);
            bool p2 = t2 & near32 (A, B2, k1, k2);
            if (p1) pair (lane, b1);
            if (p2) pair (lane, b2);
          }
        }
        else
        {
          int work = n * halfn + ((n & 1) ? 0 : n / 2);
          /* pair q = kk n + a, carried from pass to pass: q advances by
           * 64 = D n + R, so a by R and kk by D, with one wrap at most */
          int2 st = divmod_small (64, n), l0 = divmod_small (lane, n);
          int D = st.x, R = st.y;
          int kk = l0.x, an = l0.y;
          for (int q = lane; q < work; q += 64)
          {
            int a = an, bb = a + kk + 1;
            bb = bb >= n ? bb - n : bb;
            an += R;
            kk += D;
            if (an >= n) { an -= n; kk++; }
            if (!near32 (ssph32[a], ssph32[bb], k1, k2)) continue;
            pair (a, bb);
          }
        }
      }
      else
      {
        int work = nsl * cnt;
        /* pair q = a cnt + m (sphere a of the list, record m), carried from
         * pass to pass like the own leaf's */
        int2 st = divmod_small (64, cnt), l0 = divmod_small (lane, cnt);
        int D = st.x, R = st.y;
        int an = l0.x, mn = l0.y;
        for (int q = lane; q < work; q += 64)
        {
          int a = an, m = mn;
          mn += R;
          an += D;
          if (mn >= cnt) { mn -= cnt; an++; }
          int ti = lst[a];
          if (!near32 (ssph32[ti], rec32[m], k1, k2)) continue;
          /* the partner's FP64 values (the ones its record holds) */
          int j = mem[m];
          double cx = center[j], cy = center[(size_t) E + j], cz = center[2*(size_t) E + j], rj = radii[j];
          if (!overlap (ssph[4*ti], ssph[4*ti+1], ssph[4*ti+2], ssph[4*ti+3], cx, cy, cz, rj)) continue;
          /* the reach of sphere ti to L, confirmed in FP64 (the masks are
           * a superset of it) */
          {
            double x = ssph[4*ti], y = ssph[4*ti+1], z = ssph[4*ti+2], r = ssph[4*ti+3];
            if (!(x-r <= cell[3*(size_t) nleaves + L] && y-r <= cell[4*(size_t) nleaves + L] &&
                  z-r <= cell[5*(size_t) nleaves + L] && x+r >= cell[L] && y+r >= cell[(size_t) nleaves + L] &&
                  z+r >= cell[2*(size_t) nleaves + L]))
              continue;
          }
          if (lpart[qoff + ti] == part[j]) continue;
          int e = lmem[qoff + ti];
          int slot = atomicAdd (&shc[ti], 1);
          if (slot < hcap) hits[(size_t) slot * E + e] = (uint32_t) j;
          else { atomicAdd (&flags[FLAG_HIT_OVERFLOW], 1); abort_step (ctl); }
          /* does the partner's AABB reach this sphere's leaf Q? */
          bool back = cx-rj <= cu[0] && cy-rj <= cu[1] && cz-rj <= cu[2] &&
                      cx+rj >= cl[0] && cy+rj >= cl[1] && cz+rj >= cl[2];
          if (!back)
          {
            int sx = atomicAdd (&xcnt[j], 1);
            if (sx < xcap) xbuf[(size_t) sx * E + j] = e;
            else { atomicAdd (&flags[FLAG_EXTRA_OVERFLOW], 1); abort_step (ctl); }
          }
        }
      }
    }
  }
  __syncthreads ();

  /* triangles whose AABB reaches Q (test_triangles, condet.ispc:412-516):
   * static ones from the rebuild-time lists, moving ones from this step's */
  if (ntq > 0 && qcnt > 0)
  {
    /* the leaf's triangles staged once in the (now free) candidate staging
     * area when they fit (21 per wave of the block), so the pairs read
     * their vertices from LDS instead of each gathering them from the
     * triangle arrays */
    double *stri = (double*) &wrec32[0][0];
    bool staged = ntq * 9 * (int) sizeof (double) <= (int) sizeof (wrec32);
    if (staged)
    {
      for (int q = t; q < 9 * ntq; q += NW*64)
      {
        int m = q / 9, k = q - 9*m;
        int tt = m < nsq ? trilist[toff + m] : mtlist[(size_t) Q*mcap + (m - nsq)];
        stri[q] = tri[(size_t) k*T + tt];
      }
      __syncthreads ();
    }
    int work = qcnt * ntq;
    for (int q = t; q < work; q += NW*64)
    {
      int ti = q / ntq, m = q - ti * ntq;
      int tt = -1;
      double v[9];
      if (staged)
      {
#pragma unroll
        for (int k = 0; k < 9; k++) v[k] = stri[9*m + k];
      }
      else
      {
        tt = m < nsq ? trilist[toff + m] : mtlist[(size_t) Q*mcap + (m - nsq)];
#pragma unroll
        for (int k = 0; k < 9; k++) v[k] = tri[(size_t) k*T + tt];
      }
      double px = ssph[4*ti], py = ssph[4*ti+1], pz = ssph[4*ti+2], rr = ssph[4*ti+3];   /* the record's r: radii[e] */
      /* the returned point lies on the triangle, so its distance is at least
       * the distance to the triangle's box; skip when that exceeds r by a
       * margin far above the rounding of these few operations */
      double gx = fmax (fmax (fmin (fmin (v[0], v[3]), v[6]) - px, px - fmax (fmax (v[0], v[3]), v[6])), 0.0);
      double gy = fmax (fmax (fmin (fmin (v[1], v[4]), v[7]) - py, py - fmax (fmax (v[1], v[4]), v[7])), 0.0);
      double gz = fmax (fmax (fmin (fmin (v[2], v[5]), v[8]) - pz, pz - fmax (fmax (v[2], v[5]), v[8])), 0.0);
      if (gx*gx + gy*gy + gz*gz > rr*rr*1.000001) continue;
      double pt[3], nr[3];
      double depth = triangle_sphere (v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], px, py, pz, rr, pt, nr);
      if (depth > 0.0)
      {
        if (tt < 0) tt = m < nsq ? trilist[toff + m] : mtlist[(size_t) Q*mcap + (m - nsq)];
        int slot = atomicAdd (&shc[ti], 1);
        if (slot < hcap) hits[(size_t) slot * E + lmem[qoff + ti]] = PMX_TRI_FLAG | (uint32_t) tt;
        else { atomicAdd (&flags[FLAG_HIT_OVERFLOW], 1); abort_step (ctl); }
      }
    }
  }
  __syncthreads ();
  /* hit counts of the spheres with hits (every other sphere's count is
   * zero already: k_rows clears the counts it consumes) */
  for (int s = t; s < qcnt; s += NW*64)
    if (shc[s] > 0) hitcnt[lmem[qoff + s]] = shc[s] < hcap ? shc[s] : hcap;
}

__global__ void k_force_init (const int *ctl, int N, const int *list, int nlist, const double *mass, const double *gtab,
                              int ninst, const int *par_inst, double *force, double *torque, const int *stop)
{
  if (ctl[CTL_ABORT]) return;
  int k = ctl[CTL_KSTEP];
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nlist) return;
  int i = list[q];
  if (frozen (stop, par_inst[i], ctl)) return;
  double ma = mass[i];
  const double *g = gtab + 3*((size_t) k*ninst + (ninst > 1 ? par_inst[i] : 0));
  force[i] = 0.0 + ma * g[0];
  force[(size_t) N + i] = 0.0 + ma * g[1];
  force[2*(size_t) N + i] = 0.0 + ma * g[2];
  torque[i] = 0.0;
  torque[(size_t) N + i] = 0.0;
  torque[2*(size_t) N + i] = 0.0;
}


/* the reactions on meshed particles: the slave copies of the contacts of
 * spheres with their triangles (forces.ispc:383-405), summed as
 * contacts_acc_task sums slave entries (453-596).  Per particle, its
 * triangles in ascending number and each triangle's contacts by ascending
 * sphere (the reference's order follows its task interleaving); force =
 * sum + m g, torque = sum of (point - x) x f.  With adp, the adaptive sums of
 * the same entries (unit force directions, kcur = ecur = 0 on the slave
 * side, arms L (point - x)); crit: CRITICAL's contact part only (sums from
 * the contacts the last step left in the slots, current x and L) */
__global__ void k_react (int nmp, const int *mp_part, const int *mp_toff, const int *mp_tlist, int T,
                         const int *tcnt, const int *trec_e, const double *trec, int tcap, int N,
                         const double *position, const double *rotation, const double *mass, const double *gtab,
                         int ninst, const int *par_inst, double *force, double *torque, double *adp,
                         const int *ctl, const int *stop, int crit)
{
  if (ctl && ctl[CTL_ABORT]) return;
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nmp) return;
  int p = mp_part[q];
  if (!crit && frozen (stop, par_inst[p], ctl)) return;
  size_t NN = N, TT = T;
  double x[3] = {position[p], position[NN + p], position[2*NN + p]}, L[9];
  if (adp)
    for (int c = 0; c < 9; c++) L[c] = rotation[c*NN + p];
  double fs[3] = {0.0, 0.0, 0.0}, ts[3] = {0.0, 0.0, 0.0}, ka[3] = {0.0, 0.0, 0.0};
  double kr[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int k = mp_toff[q]; k < mp_toff[q+1]; k++)
  {
    int t = mp_tlist[k];
    int n = tcnt[t];
    n = n < tcap ? n : tcap;
    int last = -1;
    for (int r = 0; r < n; r++)
    {
      int best = 0, be = 0x7FFFFFFF;   /* the next sphere in ascending order */
      for (int sl = 0; sl < n; sl++)
      {
        int es = trec_e[(size_t) sl * TT + t];
        if (es > last && es < be) { be = es; best = sl; }
      }
      last = be;
      double f[3], a[3];
      for (int c = 0; c < 3; c++)
      {
        f[c] = trec[(size_t) (c*tcap + best) * TT + t];
        a[c] = trec[(size_t) ((3+c)*tcap + best) * TT + t] - x[c];
      }
      fs[0] += f[0]; fs[1] += f[1]; fs[2] += f[2];
      ts[0] += a[1]*f[2] - a[2]*f[1];
      ts[1] += a[2]*f[0] - a[0]*f[2];
      ts[2] += a[0]*f[1] - a[1]*f[0];
      if (adp)
      {
        double d[3] = {f[0], f[1], f[2]};
        double len = sqrt (d[0]*d[0] + d[1]*d[1] + d[2]*d[2]);
        if (len != 0.0)
        {
          len = 1.0/len;
          d[0] *= len; d[1] *= len; d[2] *= len;
        }
        else d[0] = d[1] = d[2] = 0.0;
        for (int c = 0; c < 3; c++) ka[c] += fabs (d[c]);
        double A[3];
        nvmul (L, a, A);
        double dot = A[0]*A[0] + A[1]*A[1] + A[2]*A[2];
        kr[0] += dot - A[0]*A[0];
        kr[1] += dot - A[1]*A[1];
        kr[2] += dot - A[2]*A[2];
        kr[3] += -A[0]*A[1];
        kr[4] += -A[0]*A[2];
        kr[5] += -A[1]*A[2];
      }
    }
  }
  if (!crit)
  {
    const double *g = gtab + 3*((size_t) ctl[CTL_KSTEP]*ninst + (ninst > 1 ? par_inst[p] : 0));
    double ma = mass[p];
    force[p] = fs[0] + ma * g[0];
    force[NN + p] = fs[1] + ma * g[1];
    force[2*NN + p] = fs[2] + ma * g[2];
    torque[p] = ts[0];
    torque[NN + p] = ts[1];
    torque[2*NN + p] = ts[2];
  }
  if (adp)
  {
    double v[11] = {ka[0], ka[1], ka[2], 0.0, 0.0, kr[0], kr[1], kr[2], kr[3], kr[4], kr[5]};
    for (int c = 0; c < 11; c++) adp[c*NN + p] = v[c];
  }
}

#define ROWS_SORT 8   /* hit keys a thread sorts in LDS (more: sorted in the hit array) */

/* The rows of a step in two launches.  k_rows_idle (TPB threads per block,
 * few registers) takes every sphere with neither a contact in the previous
 * step's row nor a hit in this step: its row is empty and its sums are
 * zero, so the merge would give force 0 + m g and torque 0, written here
 * directly.  The block's other (busy) spheres are listed, in order, in the
 * block's segment of the busy list (no atomics).  k_rows then merges the
 * busy spheres' rows with one 64-thread block per list slice, so the long
 * per-sphere merges do not hold the idle spheres' blocks resident. */
template <int TPB>
__global__ __launch_bounds__(TPB)
void k_rows_idle (int E, int N, const int *part, const double *mass, int ninst, const int *ell_inst, int *cnt0,
                  int *cnt1, int cur0, const int *hitcnt, const int *xcnt, const double *gtab, const int *ctl,
                  double *force, double *torque, const int *stop, int *lcount, int nleaves, int *busy, int *nbusy,
                  const signed char *flags)
{
  __shared__ int wbase[TPB / 64];  /* busy spheres of the waves before */
  if (ctl[CTL_ABORT]) return;
  /* the leaf counts were last read by k_detect: cleared for the next bin
   * (nleaves <= E: every leaf held a sphere when the tree was built) */
  int e = blockIdx.x * TPB + threadIdx.x;
  if (e < nleaves) lcount[e] = 0;
  int kstep = ctl[CTL_KSTEP];
  bool odd = ((cur0 + kstep) & 1) != 0;   /* rows of the previous step */
  const int *ocnt = odd ? cnt1 : cnt0;
  int *ncnt = odd ? cnt0 : cnt1;
  bool live = e < E && !frozen (stop, ell_inst[e], ctl);
  bool bs = live && (hitcnt[e] | xcnt[e] | ocnt[e]) != 0;
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long m = __ballot (bs);
  if (lane == 0) wbase[w] = __popcll (m);
  __syncthreads ();
  int base = 0, tot = 0;
  for (int k = 0; k < TPB / 64; k++) { base += k < w ? wbase[k] : 0; tot += wbase[k]; }
  if (bs) busy[(size_t) blockIdx.x * TPB + base + __popcll (m & ((1ull << lane) - 1ull))] = e;
  if (threadIdx.x == 0) nbusy[blockIdx.x] = tot;
  if (!live || bs) return;
  ncnt[e] = 0;
  int pe = part[e];
  if (flags[pe] & 4) return;   /* k_dynamics forms 0 + m g and 0 itself */
  const double *g = gtab + 3*((size_t) kstep*ninst + (ninst > 1 ? ell_inst[e] : 0));
  double ma = mass[pe];
  force[pe] = 0.0 + ma * g[0];
  force[(size_t) N + pe] = 0.0 + ma * g[1];
  force[2*(size_t) N + pe] = 0.0 + ma * g[2];
  torque[pe] = 0.0;
  torque[(size_t) N + pe] = 0.0;
  torque[2*(size_t) N + pe] = 0.0;
}

/* row merge + contact forces + per-sphere accumulation of the busy spheres.
 * R is the arithmetic of the contact geometry and force law: double (the
 * FP64 path) or float (mixed precision: differences of FP64 positions are
 * formed in FP64 and then narrowed, the depth that decides whether a contact
 * persists stays FP64, the tangential state is stored in FP32 and the
 * per-sphere sums are Kahan-compensated FP32).  Block j takes slice
 * j % slices of list j / slices (lists of LIST entries at most). */
/* ADP: the instantiation of adaptive, critical-step and sphere-on-mesh
 * runs, which also forms per particle the sums the reference's
 * contacts_acc_task forms when adaptive stepping is on (forces.ispc:505-533,
 * 542-577): the summed |direction| (the normal on the master side, the unit
 * force on the slave side), the largest contact stiffness and damping
 * coefficient (kn and en = 2 damper sqrt(kn m) of the master side; the
 * reference never writes them on the slave side, taken as 0 here) and the
 * rotational stiffness sums of the arms L (point - x), into adp [11][N]
 * (when not null); stores each row entry's point, |direction|, kn and en
 * into crec [8][cap][E] (when not null) for the critical step estimate; and
 * puts the reaction of each contact with a meshed particle's triangle (the
 * slave copy of forces.ispc:383-405) into that triangle's slots (when tcnt
 * is not null) */
template <class R, bool ADP = false>
__global__ __launch_bounds__(64)
void k_rows (int E, int N, const double *center, const double *radii, const int *part, const int *ellcol,
             const double *tri, int T, const int *tricol, const int *triobs,
             const double *position, const double *linear, const double *angular, const double *mass,
             const double *invm, int pairnum, const int *pairs, const int *ikind, const double *iparam,
             int ninst, const int *ell_inst, const int *pair_off,
             uint32_t *key0, R *st0, int *cnt0, uint32_t *key1, R *st1, int *cnt1, int cur0, int cap,
             uint32_t *hits, int *hitcnt, int hcap, const int *xbuf, int *xcnt, int xcap,
             const double *gtab, int *ctl, double h, double *force, double *torque, int *flags,
             const double *obspnt, const double *obsang, const double *obslin,
             const int *busy, const int *nbusy, int list, int slices,
             const double *rotation, double *adp, double *crec, int *tcnt, int *trec_e, double *trec, int tcap)
{
  constexpr int TPB = 64;
  __shared__ uint32_t sh[ROWS_SORT * TPB];
  if (ctl[CTL_ABORT]) return;
  int kstep = ctl[CTL_KSTEP];
  bool odd = ((cur0 + kstep) & 1) != 0;   /* rows of the previous step */
  constexpr bool mixed = sizeof (R) == 4;
  const uint32_t *okey = odd ? key1 : key0;
  const R *ost = odd ? st1 : st0;
  const int *ocnt = odd ? cnt1 : cnt0;
  uint32_t *nkey = odd ? key0 : key1;
  R *nst = odd ? st0 : st1;
  int *ncnt = odd ? cnt0 : cnt1;
  int lb = blockIdx.x / slices, sl = blockIdx.x - lb * slices;
  int nb = nbusy[lb];
  int mx = 0;   /* longest row of the thread's spheres */
  for (int q = sl * TPB + threadIdx.x; q < nb; q += TPB * slices)
  {
  int cnt = 0;
  int e = busy[(size_t) lb * list + q];
  /* the step's hit keys (own hits, then reverse hits), sorted: in a small
   * LDS buffer, or in place in the sphere's column of the hit array when
   * there are more (both orders are the same unique sorted set) */
  int nh = hitcnt[e];
  int nx = xcnt[e];
  /* the next step's hit counts start at zero (only busy spheres have any) */
  if (nh) hitcnt[e] = 0;
  if (nx) xcnt[e] = 0;
  nx = nx < xcap ? nx : xcap;
  bool small = nh + nx <= ROWS_SORT;
  uint32_t *H = small ? sh + threadIdx.x : hits + e;
  size_t hs = small ? (size_t) TPB : (size_t) E;   /* stride of H */
  if (small) for (int q = 0; q < nh; q++) H[q*hs] = hits[(size_t) q * E + e];
  else if (nh + nx > hcap) { atomicAdd (&flags[FLAG_HIT_OVERFLOW], 1); abort_step (ctl); nx = hcap - nh; }
  for (int q = 0; q < nx; q++) H[(nh + q)*hs] = (uint32_t) xbuf[(size_t) q * E + e];
  nh += nx;
  for (int a = 1; a < nh; a++)   /* insertion sort of the hit keys */
  {
    uint32_t v = H[a*hs];
    int b = a - 1;
    while (b >= 0 && H[b*hs] > v) { H[(b+1)*hs] = H[b*hs]; b--; }
    H[(b+1)*hs] = v;
  }

  int pe = part[e];
  double xe[3] = {position[pe], position[(size_t) N + pe], position[2*(size_t) N + pe]};
  double pc[3] = {center[e], center[(size_t) E + e], center[2*(size_t) E + e]};
  double re = radii[e];
  int oc = ocnt[e];
  int inst = ninst > 1 ? ell_inst[e] : 0;
  int po = 0, pn = pairnum;   /* the instance's pairing table */
  if (ninst > 1) { po = pair_off[inst]; pn = pair_off[inst+1] - po; }
  int io = 0, ih = 0;
  R fs[3] = {0.0, 0.0, 0.0}, ts[3] = {0.0, 0.0, 0.0};
  R cf[3] = {0.0, 0.0, 0.0}, ct[3] = {0.0, 0.0, 0.0};   /* Kahan compensation (mixed) */
  double ka[3] = {0.0, 0.0, 0.0}, km = 0.0, em = 0.0, kr[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, Le[9];
  if constexpr (ADP)
    if (adp)
      for (int c = 0; c < 9; c++) Le[c] = rotation[(size_t) c*N + pe];

  while (io < oc || ih < nh)
  {
    uint32_t ko = io < oc ? okey[(size_t) io * E + e] : 0xFFFFFFFFu;
    uint32_t kh = ih < nh ? H[ih*hs] : 0xFFFFFFFFu;
    uint32_t key;
    bool old;
    int kold = -1;
    if (io < oc && (ih >= nh || ko <= kh))
    {
      key = ko; old = true; kold = io; io++;
      if (ih < nh && kh == ko) ih++;
    }
    else { key = kh; old = false; ih++; }

    double point[3], depth;
    R normal[3], st[3] = {0.0, 0.0, 0.0};
    bool master;
    int i, j, col0, col1, tk = -1;
    if (key & PMX_TRI_FLAG)
    {
      int t = (int) (key & ~PMX_TRI_FLAG);
      tk = t;
      double nd[3];   /* closest-point regions and depth in FP64: they decide the pair */
      depth = triangle_sphere (tri[t], tri[(size_t) T + t], tri[2*(size_t) T + t],
                               tri[3*(size_t) T + t], tri[4*(size_t) T + t], tri[5*(size_t) T + t],
                               tri[6*(size_t) T + t], tri[7*(size_t) T + t], tri[8*(size_t) T + t],
                               pc[0], pc[1], pc[2], re, point, nd);
      normal[0] = (R) nd[0]; normal[1] = (R) nd[1]; normal[2] = (R) nd[2];
      if (old && depth < 0.0) continue;
      master = true;
      if (old)
      {
        st[0] = ost[(size_t) (0*cap + kold) * E + e];
        st[1] = ost[(size_t) (1*cap + kold) * E + e];
        st[2] = ost[(size_t) (2*cap + kold) * E + e];
      }
      i = pe;
      j = triobs[t];
      col0 = ellcol[e];
      col1 = tricol[t];
    }
    else
    {
      int e2 = (int) key;
      int p2 = part[e2];
      master = pe < p2;
      int em = master ? e : e2, es = master ? e2 : e;
      double pm[3] = {center[em], center[(size_t) E + em], center[2*(size_t) E + em]};
      double ps[3] = {center[es], center[(size_t) E + es], center[2*(size_t) E + es]};
      double qx = pm[0]-ps[0], qy = pm[1]-ps[1], qz = pm[2]-ps[2];
      double len = sqrt (qx*qx + qy*qy + qz*qz);
      point[0] = 0.5*(pm[0]+ps[0]);
      point[1] = 0.5*(pm[1]+ps[1]);
      point[2] = 0.5*(pm[2]+ps[2]);
      if (mixed)
      {
        R lf = (R) len;
        R il = old ? R (1.0)/lf : (lf > R (0.0) ? R (1.0)/lf : R (1.0));
        normal[0] = il*(R) qx;
        normal[1] = il*(R) qy;
        normal[2] = il*(R) qz;
      }
      else
      {
        double ilen = old ? 1.0/len : (len > 0.0 ? 1.0/len : 1.0);
        normal[0] = (R) (ilen*qx);
        normal[1] = (R) (ilen*qy);
        normal[2] = (R) (ilen*qz);
      }
      depth = radii[em] + radii[es] - len;
      if (old && depth < 0.0) continue;
      if (old)
      {
        if (master)
        {
          st[0] = ost[(size_t) (0*cap + kold) * E + e];
          st[1] = ost[(size_t) (1*cap + kold) * E + e];
          st[2] = ost[(size_t) (2*cap + kold) * E + e];
        }
        else
        {
          int lo = 0, hi = ocnt[e2] - 1, at = -1;
          while (lo <= hi)
          {
            int mid = (lo + hi) >> 1;
            uint32_t km = okey[(size_t) mid * E + e2];
            if (km == (uint32_t) e) { at = mid; break; }
            if (km < (uint32_t) e) lo = mid + 1; else hi = mid - 1;
          }
          if (at < 0) atomicAdd (&flags[FLAG_BAD_KIND], 1);
          else
          {
            st[0] = ost[(size_t) (0*cap + at) * E + e2];
            st[1] = ost[(size_t) (1*cap + at) * E + e2];
            st[2] = ost[(size_t) (2*cap + at) * E + e2];
          }
        }
      }
      i = part[em];
      j = part[es];
      col0 = ellcol[em];
      col1 = ellcol[es];
    }

    /* contacts_task velocities, forces.ispc:219-287: v = omega x z + v at
     * the contact point for both bodies, then v_j - v_i */
    double xi[3] = {position[i], position[(size_t) N + i], position[2*(size_t) N + i]};
    R oi[3] = {(R) angular[3*(size_t) N + i], (R) angular[4*(size_t) N + i], (R) angular[5*(size_t) N + i]};
    R ri[3], rj[3] = {0.0, 0.0, 0.0}, z[3];   /* omega x z of body i and of its partner */
    double li[3] = {linear[i], linear[(size_t) N + i], linear[2*(size_t) N + i]}, lj[3] = {0.0, 0.0, 0.0};
    z[0] = (R) (point[0]-xi[0]);
    z[1] = (R) (point[1]-xi[1]);
    z[2] = (R) (point[2]-xi[2]);
    ri[0] = oi[1]*z[2]-oi[2]*z[1];
    ri[1] = oi[2]*z[0]-oi[0]*z[2];
    ri[2] = oi[0]*z[1]-oi[1]*z[0];
    if (j >= 0)
    {
      z[0] = (R) (point[0]-position[j]);
      z[1] = (R) (point[1]-position[(size_t) N + j]);
      z[2] = (R) (point[2]-position[2*(size_t) N + j]);
      R oj[3] = {(R) angular[3*(size_t) N + j], (R) angular[4*(size_t) N + j], (R) angular[5*(size_t) N + j]};
      rj[0] = oj[1]*z[2]-oj[2]*z[1];
      rj[1] = oj[2]*z[0]-oj[0]*z[2];
      rj[2] = oj[0]*z[1]-oj[1]*z[0];
      lj[0] = linear[j]; lj[1] = linear[(size_t) N + j]; lj[2] = linear[2*(size_t) N + j];
    }
    else if (j <= -2)   /* moving obstacle, forces.ispc:265-280 */
    {
      int l = -j-2;
      z[0] = (R) (point[0]-obspnt[3*l]);
      z[1] = (R) (point[1]-obspnt[3*l+1]);
      z[2] = (R) (point[2]-obspnt[3*l+2]);
      R oj[3] = {(R) obsang[3*l], (R) obsang[3*l+1], (R) obsang[3*l+2]};
      rj[0] = oj[1]*z[2]-oj[2]*z[1];
      rj[1] = oj[2]*z[0]-oj[0]*z[2];
      rj[2] = oj[0]*z[1]-oj[1]*z[0];
      lj[0] = obslin[3*l]; lj[1] = obslin[3*l+1]; lj[2] = obslin[3*l+2];
    }
    R vij[3];
    if constexpr (mixed)
    {
      /* the translational parts differenced in FP64 before narrowing: two
       * bodies moving together (a pack falling or shaken as a whole) would
       * otherwise lose the digits of their small relative velocity to
       * cancellation in FP32 */
      for (int k = 0; k < 3; k++) vij[k] = (R) (lj[k] - li[k]) + (rj[k] - ri[k]);
    }
    else
    {
      /* the reference's operation order: (omega x z + v) per body, then the
       * difference (a static partner's velocity is +0) */
      R vi[3], vj[3];
      for (int k = 0; k < 3; k++) vi[k] = ri[k] + (R) li[k];
      if (j == -1) { vj[0] = vj[1] = vj[2] = 0.0; }
      else for (int k = 0; k < 3; k++) vj[k] = rj[k] + (R) lj[k];
      for (int k = 0; k < 3; k++) vij[k] = vj[k] - vi[k];
    }
    int ij = po + pairing (pn, pairs + 2*(size_t) po, col0, col1);
    if (ikind[ij] != PMX_GRANULAR_FORCE) atomicAdd (&flags[FLAG_BAD_KIND], 1);
    double ma = j >= 0 ? 1.0 / (invm[i] + invm[j]) : mass[i];
    R f[3];
    granular<R> (normal, vij, (R) depth, (R) ma, (R) iparam[PMX_SPRING*pairnum + ij], (R) iparam[PMX_DAMPER*pairnum + ij],
                 (R) iparam[PMX_KSKN*pairnum + ij], (R) iparam[PMX_FRISTAT*pairnum + ij],
                 (R) iparam[PMX_FRIDYN*pairnum + ij], st, (R) h, f);

    if (cnt < cap)
    {
      nkey[(size_t) cnt * E + e] = key;
      if (master)
      {
        nst[(size_t) (0*cap + cnt) * E + e] = st[0];
        nst[(size_t) (1*cap + cnt) * E + e] = st[1];
        nst[(size_t) (2*cap + cnt) * E + e] = st[2];
      }
    }
    else { atomicAdd (&flags[FLAG_ROW_OVERFLOW], 1); abort_step (ctl); }
    cnt++;

    if constexpr (ADP)
      if (tcnt && tk >= 0 && j >= 0)
      {
        /* a meshed particle's triangle: its reaction -f at the point */
        int sl = atomicAdd (&tcnt[tk], 1);
        if (sl < tcap)
        {
          trec_e[(size_t) sl * T + tk] = e;
          double rv[6] = {-(double) f[0], -(double) f[1], -(double) f[2], point[0], point[1], point[2]};
          for (int c = 0; c < 6; c++) trec[(size_t) (c*tcap + sl) * T + tk] = rv[c];
        }
        else { atomicAdd (&flags[FLAG_TREV_OVERFLOW], 1); abort_step (ctl); }
      }
    /* contacts_acc_task, forces.ispc:492-533: slaves receive -f */
    if (!master) { f[0] = -f[0]; f[1] = -f[1]; f[2] = -f[2]; }
    if constexpr (ADP)
    {
      double dir[3] = {0.0, 0.0, 0.0}, kc = 0.0, ec = 0.0;
      if (master)
      {
        for (int c = 0; c < 3; c++) dir[c] = fabs ((double) normal[c]);
        kc = iparam[PMX_SPRING*pairnum + ij];
        ec = iparam[PMX_DAMPER*pairnum + ij] * 2.0 * sqrt (kc*ma);
      }
      else
      {
        double g[3] = {(double) f[0], (double) f[1], (double) f[2]};
        double fl = sqrt (g[0]*g[0] + g[1]*g[1] + g[2]*g[2]);
        if (fl != 0.0)
        {
          fl = 1.0/fl;
          for (int c = 0; c < 3; c++) dir[c] = fabs (g[c]*fl);
        }
      }
      if (crec && cnt - 1 < cap)
      {
        double rec[8] = {point[0], point[1], point[2], dir[0], dir[1], dir[2], kc, ec};
        for (int c = 0; c < 8; c++) crec[(size_t) (c*cap + cnt - 1) * E + e] = rec[c];
      }
      if (adp)
      {
        for (int c = 0; c < 3; c++) ka[c] += dir[c];
        km = kc > km ? kc : km;
        em = ec > em ? ec : em;
        double av[3] = {point[0]-xe[0], point[1]-xe[1], point[2]-xe[2]}, A[3];
        nvmul (Le, av, A);
        double dot = A[0]*A[0] + A[1]*A[1] + A[2]*A[2];
        kr[0] += dot - A[0]*A[0];
        kr[1] += dot - A[1]*A[1];
        kr[2] += dot - A[2]*A[2];
        kr[3] += -A[0]*A[1];
        kr[4] += -A[0]*A[2];
        kr[5] += -A[1]*A[2];
      }
    }
    R a[3] = {(R) (point[0]-xe[0]), (R) (point[1]-xe[1]), (R) (point[2]-xe[2])};
    if constexpr (mixed)
    {
      kahan_add (fs[0], cf[0], f[0]); kahan_add (fs[1], cf[1], f[1]); kahan_add (fs[2], cf[2], f[2]);
      kahan_add (ts[0], ct[0], a[1]*f[2] - a[2]*f[1]);
      kahan_add (ts[1], ct[1], a[2]*f[0] - a[0]*f[2]);
      kahan_add (ts[2], ct[2], a[0]*f[1] - a[1]*f[0]);
    }
    else
    {
      fs[0] += f[0]; fs[1] += f[1]; fs[2] += f[2];
      ts[0] += a[1]*f[2] - a[2]*f[1];
      ts[1] += a[2]*f[0] - a[0]*f[2];
      ts[2] += a[0]*f[1] - a[1]*f[0];
    }
  }
  (void) cf; (void) ct;
  ncnt[e] = cnt < cap ? cnt : cap;
  const double *g = gtab + 3*((size_t) kstep*ninst + inst);
  double ma = mass[pe];
  force[pe] = (double) fs[0] + ma * g[0];
  force[(size_t) N + pe] = (double) fs[1] + ma * g[1];
  force[2*(size_t) N + pe] = (double) fs[2] + ma * g[2];
  torque[pe] = (double) ts[0];
  torque[(size_t) N + pe] = (double) ts[1];
  torque[2*(size_t) N + pe] = (double) ts[2];
  if constexpr (ADP)
    if (adp)
    {
      double v[11] = {ka[0], ka[1], ka[2], km, em, kr[0], kr[1], kr[2], kr[3], kr[4], kr[5]};
      for (int c = 0; c < 11; c++) adp[(size_t) c*N + pe] = v[c];
    }
  mx = cnt > mx ? cnt : mx;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
  {
    int o = __shfl_xor (mx, off, 64);
    mx = o > mx ? o : mx;
  }
  if ((threadIdx.x & 63) == 0 && mx > 0) atomicMax (&flags[FLAG_MAXROW], mx);
}

/* the binning for the next step fused into k_dynamics (fused: lstride >
 * 0, the slotted leaf records): partitioning_store of the next step
 * (partition.ispc:600-646) from the centre k_dynamics has just formed,
 * which is the centre the next step's store would read.  until: the
 * call's step count (the last step bins nothing) */
struct BinArgs
{
  const int *roots, *left, *right, *dim;
  const double *coord, *cell, *radii;
  int nleaves, lstride, until;
  int *leafof, *lcount, *flags, *ovf, *lmem, *lpart;
  double *lpos;
};

/* dynamics_task, dynamics.ispc:30-169 */
/* ISO: the launch over particles whose inertia is diag(j, j, j) with exact
 * zeros elsewhere (inverse diag(v, v, v), INVERT macros.h:1033-1051): J x
 * and I x are the scalar products j x and v x, equal in value to the
 * matrix products; the other launch takes the remaining particles */
template <bool ISO, bool FUSE>
__global__ __launch_bounds__(PMX_TPB)   /* the launch size: no register cap below what the kernel takes (no spill) */
void k_dynamics (int N, double *angular, double *linear, double *rotation, double *position,
                            const double *inertia, const double *inverse, const double *mass, const double *invm,
                            const double *dtab, const int *ctl, double *force, double *torque, double step0, double step1,
                            const int *parell, int E, double *center, const signed char *iso, int ninst,
                            const int *par_inst, const int *stop, const int *cnt0, const int *cnt1, int cur0,
                            const double *gtab, const int *list, int nlist, BinArgs bn, int *ctlw)
{
  if (ctl[CTL_ABORT]) return;
  int kstep = ctl[CTL_KSTEP];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlist) return;
  if (list) i = list[i];
  int flag = iso[i];
  if ((flag & 1) != (int) ISO || frozen (stop, par_inst[i], ctl)) return;
  /* dtab null: no damping anywhere in the call (no loads, no terms) */
  const double *damping = dtab ? dtab + 6*((size_t) kstep*ninst + (ninst > 1 ? par_inst[i] : 0)) : nullptr;
  double half = 0.5*step0;
  double step = 0.5*(step0+step1);
  double O[3], o[3], v[3], L1[9], im, f[3], t[3], T[3], DL[9], A[3], B[3];
  double J[ISO ? 1 : 9], I[ISO ? 1 : 9];
#pragma unroll
  for (int k = 0; k < 3; k++) { O[k] = angular[(size_t) k*N + i]; v[k] = linear[(size_t) k*N + i]; }
#pragma unroll
  for (int k = 0; k < 9; k++) L1[k] = rotation[(size_t) k*N + i];
  if constexpr (ISO)
  {
    J[0] = inertia[i];
    I[0] = inverse[i];
  }
  else
  {
#pragma unroll
    for (int k = 0; k < 9; k++)
    {
      J[k] = inertia[(size_t) k*N + i];
      I[k] = inverse[(size_t) k*N + i];
    }
  }
  auto Jx = [&] (const double x[3], double y[3]) {
    if constexpr (ISO) { y[0] = J[0]*x[0]; y[1] = J[0]*x[1]; y[2] = J[0]*x[2]; } else nvmul (J, x, y); };
  auto Ix = [&] (const double x[3], double y[3]) {
    if constexpr (ISO) { y[0] = I[0]*x[0]; y[1] = I[0]*x[1]; y[2] = I[0]*x[2]; } else nvmul (I, x, y); };
  im = invm[i];
  int e = parell ? parell[i] : -1;
  /* fused binning: the sphere's radius, its previous leaf (the hint) and
   * that leaf's reach cell are fetched now, in flight while the motion is
   * integrated (indices clamped: every load is in bounds) */
  int hprev = -1;
  double rad = 0.0, hc[6];
  if constexpr (FUSE)
  {
    int es = e < 0 ? 0 : e;
    rad = bn.radii[es];
    int h = bn.leafof[es];
    int hs = (unsigned) h < (unsigned) bn.nleaves ? h : 0;
#pragma unroll
    for (int k = 0; k < 6; k++) hc[k] = bn.cell[(size_t) k*bn.nleaves + hs];
    hprev = ctl[CTL_HINT] && e >= 0 ? h : -1;
  }
  /* a particle loaded by contacts and gravity only whose row of this step
   * is empty: k_rows would have stored force 0 + m g and torque 0 */
  const int *ncnt = ((cur0 + kstep) & 1) ? cnt0 : cnt1;
  if ((flag & 4) && e >= 0 && ncnt[e] == 0)
  {
    double ma = mass[i];
    const double *g = gtab + 3*((size_t) kstep*ninst + (ninst > 1 ? par_inst[i] : 0));
    f[0] = 0.0 + ma * g[0]; f[1] = 0.0 + ma * g[1]; f[2] = 0.0 + ma * g[2];
    t[0] = 0.0; t[1] = 0.0; t[2] = 0.0;
  }
  else
  {
    f[0] = force[i]; f[1] = force[(size_t) N + i]; f[2] = force[2*(size_t) N + i];
    t[0] = torque[i]; t[1] = torque[(size_t) N + i]; t[2] = torque[2*(size_t) N + i];
  }
  if (damping && (damping[0] != 0.0 || damping[1] != 0.0 || damping[2] != 0.0))
  {
    double ma = mass[i];
    f[0] -= ma * damping[0] * v[0];
    f[1] -= ma * damping[1] * v[1];
    f[2] -= ma * damping[2] * v[2];
    force[i] = f[0]; force[(size_t) N + i] = f[1]; force[2*(size_t) N + i] = f[2];
  }
  if (damping && (damping[3] != 0.0 || damping[4] != 0.0 || damping[5] != 0.0))
  {
    double w[3], tt[3], TT[3];
    w[0] = damping[3]*angular[3*(size_t) N + i];
    w[1] = damping[4]*angular[4*(size_t) N + i];
    w[2] = damping[5]*angular[5*(size_t) N + i];
    tvmul (L1, w, tt);
    Jx (tt, TT);
    nvmul (L1, TT, tt);
    t[0] -= tt[0]; t[1] -= tt[1]; t[2] -= tt[2];
    torque[i] = t[0]; torque[(size_t) N + i] = t[1]; torque[2*(size_t) N + i] = t[2];
  }
  tvmul (L1, t, T);
  expmap (-half*O[0], -half*O[1], -half*O[2], DL);
  Jx (O, A);
  nvmul (DL, A, B);
  B[0] = B[0] + half*T[0];
  B[1] = B[1] + half*T[1];
  B[2] = B[2] + half*T[2];
  Ix (B, A);
  Jx (A, B);
  T[0] -= A[1]*B[2] - A[2]*B[1];
  T[1] -= A[2]*B[0] - A[0]*B[2];
  T[2] -= A[0]*B[1] - A[1]*B[0];
  T[0] *= step0; T[1] *= step0; T[2] *= step0;
  Ix (T, A);
  O[0] = O[0] + A[0];
  O[1] = O[1] + A[1];
  O[2] = O[2] + A[2];
  im *= step0;
  v[0] = v[0] + im*f[0];
  v[1] = v[1] + im*f[1];
  v[2] = v[2] + im*f[2];
  expmap (step*O[0], step*O[1], step*O[2], DL);
  double xn[3];
  xn[0] = position[i] + step*v[0];
  xn[1] = position[(size_t) N + i] + step*v[1];
  xn[2] = position[2*(size_t) N + i] + step*v[2];
  /* the particle's sphere (ellipsoids_task, shapes.ispc:29-67) */
  double C[3] = {0.0, 0.0, 0.0};   /* kept when the offset is +0 (flag bit 1) */
  if (e >= 0 && !(flag & 2))
  {
#pragma unroll
    for (int k = 0; k < 3; k++) C[k] = center[(size_t) (3+k)*E + e] - position[(size_t) (3+k)*N + i];
  }
  /* L2 = L1 DL (NNMUL) one column at a time: each column is stored at once
   * and folded into omega = L2 Omega and L2 C with the same left-to-right
   * sums as the whole-matrix products */
  double cc[3];
#pragma unroll
  for (int c = 0; c < 3; c++)
  {
#pragma unroll
    for (int r = 0; r < 3; r++)
    {
      double l = L1[r]*DL[3*c] + L1[3+r]*DL[3*c+1] + L1[6+r]*DL[3*c+2];
      __builtin_nontemporal_store (l, &rotation[(size_t) (3*c+r)*N + i]);
      o[r] = c ? o[r] + l*O[c] : l*O[c];
      cc[r] = c ? cc[r] + l*C[c] : l*C[c];
    }
  }
  __builtin_nontemporal_store (xn[0], &position[i]);
  __builtin_nontemporal_store (xn[1], &position[(size_t) N + i]);
  __builtin_nontemporal_store (xn[2], &position[2*(size_t) N + i]);
#pragma unroll
  for (int k = 0; k < 3; k++)
  {
    __builtin_nontemporal_store (O[k], &angular[(size_t) k*N + i]);
    __builtin_nontemporal_store (o[k], &angular[(size_t) (3+k)*N + i]);
    __builtin_nontemporal_store (v[k], &linear[(size_t) k*N + i]);
  }
  double nc[3] = {xn[0] + cc[0], xn[1] + cc[1], xn[2] + cc[2]};
  if (e >= 0)
  {
    __builtin_nontemporal_store (nc[0], &center[e]);
    __builtin_nontemporal_store (nc[1], &center[(size_t) E + e]);
    __builtin_nontemporal_store (nc[2], &center[2*(size_t) E + e]);
  }
  if constexpr (FUSE)
  {
    /* the next step's binning (every active lane takes part in the slot
     * grouping); an instance's last step bins nothing */
    int inst = ninst > 1 ? par_inst[i] : 0;
    bool bv = e >= 0 && kstep + 1 < bn.until && !(stop && kstep + 1 >= stop[inst]);
    int L = -1;
    if (bv)
    {
      bool in = hprev >= 0;
#pragma unroll
      for (int k = 0; k < 3; k++) in = in && hc[k] <= nc[k] && nc[k] < hc[3+k];
      L = in ? hprev : bin_leaf (nc, -1, bn.roots[inst], bn.left, bn.right, bn.dim, bn.coord, bn.cell, bn.nleaves);
    }
    int s = bin_slot (bv, L, bn.lcount);
    if (!bv) return;
    if (L != hprev) bn.leafof[e] = L;
    if (s >= PMX_LSIZE)
    {
      /* this step stands; the next one is aborted (k_step_end) and
       * repeated once the instance is re-partitioned */
      atomicAdd (&bn.flags[FLAG_LEAF_OVERFLOW], 1);
      atomicExch (&bn.ovf[inst], 1);
      atomicExch (&ctlw[CTL_REBIN], 1);
      return;
    }
    bin_record ((size_t) L * bn.lstride + s, e, i, nc, rad, bn.lmem, bn.lpart, bn.lpos);
  }

}

/* force and torque of the flag-4 particles (loaded by contacts and gravity
 * only) of the instances finishing now, as k_rows would have stored them at
 * their last step k = fin[instance] when their row came out empty; where
 * damping was on at that step, k_dynamics stored them itself */
__global__ void k_materialize (int N, const signed char *flags, const int *parell, const int *cnt0, const int *cnt1,
                               int cur0, int ninst, const int *par_inst, const int *fin, const double *gtab,
                               const double *dtab, const double *mass, double *force, double *torque)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || !(flags[i] & 4)) return;
  int inst = ninst > 1 ? par_inst[i] : 0;
  int k = fin[inst];
  if (k < 0) return;
  const int *ncnt = ((cur0 + k) & 1) ? cnt0 : cnt1;
  if (ncnt[parell[i]] != 0) return;
  const double *d = dtab + 6*((size_t) k*ninst + inst);
  if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0)
  {
    double ma = mass[i];
    const double *g = gtab + 3*((size_t) k*ninst + inst);
    force[i] = 0.0 + ma * g[0];
    force[(size_t) N + i] = 0.0 + ma * g[1];
    force[2*(size_t) N + i] = 0.0 + ma * g[2];
  }
  if (d[3] == 0.0 && d[4] == 0.0 && d[5] == 0.0)
  {
    torque[i] = 0.0;
    torque[(size_t) N + i] = 0.0;
    torque[2*(size_t) N + i] = 0.0;
  }
}

/* ellipsoids_task centre update, shapes.ispc:29-67 */
/* prescribe_acceleration, input.cpp:4247-4341, one thread per prescribed
 * particle: lf = 2*column + kind of its last linear entry (kind 1:
 * f = m a, kind 0: f = 0), af the same for the angular part (t = J alpha
 * with the referential inertia), values from row ctl[KSTEP] of tab */
__global__ void k_prs_acc (int n, int N, const int *part, const int *lf, const int *af, const double *tab, int P,
                           const double *mass, const double *inertia, double *force, double *torque, const int *ctl,
                           const int *stop, const int *par_inst)
{
  if (ctl[CTL_ABORT]) return;
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int j = part[q];
  if (frozen (stop, par_inst[j], ctl)) return;
  const double *row = tab + 6*(size_t) ctl[CTL_KSTEP] * P;
  int l = lf[q], a = af[q];
  if (l >= 0)
  {
    const double *v = row + 6*(l >> 1);
    double ma = mass[j];
    for (int d = 0; d < 3; d++) force[(size_t) d*N + j] = (l & 1) ? ma * v[d] : 0.0;
  }
  if (a >= 0)
  {
    const double *v = row + 6*(a >> 1) + 3;
    double J[9], t[3] = {0.0, 0.0, 0.0};
    if (a & 1)
    {
      for (int k = 0; k < 9; k++) J[k] = inertia[(size_t) k*N + j];
      nvmul (J, v, t);
    }
    for (int d = 0; d < 3; d++) torque[(size_t) d*N + j] = t[d];
  }
}

/* prescribe_velocity, input.cpp:4344-4421: lv / av the column of the
 * particle's last velocity entry; body Omega = L^T omega with the rotation
 * dynamics just updated */
__global__ void k_prs_vel (int n, int N, const int *part, const int *lv, const int *av, const double *tab, int P,
                           const double *rotation, double *linear, double *angular, const int *ctl, const int *stop,
                           const int *par_inst)
{
  if (ctl[CTL_ABORT]) return;
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int j = part[q];
  if (frozen (stop, par_inst[j], ctl)) return;
  const double *row = tab + 6*(size_t) ctl[CTL_KSTEP] * P;
  if (lv[q] >= 0)
    for (int d = 0; d < 3; d++) linear[(size_t) d*N + j] = row[6*lv[q] + d];
  if (av[q] >= 0)
  {
    const double *o = row + 6*av[q] + 3;
    double L[9], O[3];
    for (int k = 0; k < 9; k++) L[k] = rotation[(size_t) k*N + j];
    tvmul (L, o, O);
    for (int d = 0; d < 3; d++) { angular[(size_t) d*N + j] = O[d]; angular[(size_t) (3+d)*N + j] = o[d]; }
  }
}

/* obstacles, obstacles.ispc:28-78, one thread per moving triangle: the
 * rigid motion of its obstacle over the step, rotation expmap(h*omega) about
 * the obstacle point x and translation h*v, with the velocities the
 * obstaclev call of this step set (row `row` of tab [rows][O][6], or row
 * ctl[KSTEP] when row < 0).  Canonical stride 3 for the velocities. */
__global__ void k_obstacles (int nm, const int *mtri, const int *mtobs, int T, double *tri, const double *obspnt,
                             const double *tab, int O, int row, double h, const int *ctl, const int *stop,
                             const int *tri_inst)
{
  if (ctl[CTL_ABORT]) return;
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nm) return;
  int t = mtri[q], i = mtobs[q];
  if (i < 0 || frozen (stop, tri_inst[t], ctl)) return;   /* i < 0: a meshed particle's triangle */
  const double *r = tab + 6*((size_t) (row >= 0 ? row : ctl[CTL_KSTEP]) * O + i);
  double x[3] = {obspnt[3*i], obspnt[3*i+1], obspnt[3*i+2]};
  double y[3], DL[9];
  y[0] = x[0] + h*r[3];
  y[1] = x[1] + h*r[4];
  y[2] = x[2] + h*r[5];
  expmap (h*r[0], h*r[1], h*r[2], DL);
#pragma unroll
  for (int a = 0; a < 3; a++)
  {
    double z[3] = {tri[(size_t) (3*a)*T + t] - x[0], tri[(size_t) (3*a+1)*T + t] - x[1], tri[(size_t) (3*a+2)*T + t] - x[2]};
    tri[(size_t) (3*a)*T + t] = DL[0]*z[0]+DL[3]*z[1]+DL[6]*z[2] + y[0];
    tri[(size_t) (3*a+1)*T + t] = DL[1]*z[0]+DL[4]*z[1]+DL[7]*z[2] + y[1];
    tri[(size_t) (3*a+2)*T + t] = DL[2]*z[0]+DL[5]*z[1]+DL[8]*z[2] + y[2];
  }
}

/* then per obstacle: x += h*v and the velocities of the step become the
 * current ones (read by the next step's contact law) */
__global__ void k_obspnt (int O, double *obspnt, double *obsang, double *obslin, const double *tab, int row, double h,
                          const int *ctl, const int *stop, const int *obs_inst)
{
  if (ctl[CTL_ABORT]) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= O || frozen (stop, obs_inst[i], ctl)) return;
  const double *r = tab + 6*((size_t) (row >= 0 ? row : ctl[CTL_KSTEP]) * O + i);
#pragma unroll
  for (int d = 0; d < 3; d++)
  {
    obspnt[3*i+d] = obspnt[3*i+d] + h*r[3+d];
    obsang[3*i+d] = r[d];
    obslin[3*i+d] = r[3+d];
  }
}

/* sel (per instance, or null for all): the instances the launch updates */
__global__ void k_shapes (int E, int N, const int *part, double *center, const double *rotation, const double *position,
                          const int *sel, const int *ell_inst)
{
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E || (sel && !sel[ell_inst[e]])) return;
  int j = part[e];
  double L[9], C[3], c[3];
#pragma unroll
  for (int k = 0; k < 9; k++) L[k] = rotation[(size_t) k*N + j];
#pragma unroll
  for (int k = 0; k < 3; k++) C[k] = center[(size_t) (3+k)*E + e] - position[(size_t) (3+k)*N + j];
  c[0] = position[j] + (L[0]*C[0] + L[3]*C[1] + L[6]*C[2]);
  c[1] = position[(size_t) N + j] + (L[1]*C[0] + L[4]*C[1] + L[7]*C[2]);
  c[2] = position[2*(size_t) N + j] + (L[2]*C[0] + L[5]*C[1] + L[8]*C[2]);
  center[e] = c[0];
  center[(size_t) E + e] = c[1];
  center[2*(size_t) E + e] = c[2];
}

/* euler_task, parmec.ispc:224-266 */
__global__ void k_euler (int N, double *angular, const double *linear, double *rotation, double *position, double step,
                         const int *sel, const int *par_inst)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || (sel && !sel[par_inst[i]])) return;
  double O[3], L1[9], DL[9], L2[9], o[3];
  O[0] = step*angular[i];
  O[1] = step*angular[(size_t) N + i];
  O[2] = step*angular[2*(size_t) N + i];
  expmap (O[0], O[1], O[2], DL);
#pragma unroll
  for (int k = 0; k < 9; k++) L1[k] = rotation[(size_t) k*N + i];
  nnmul (L1, DL, L2);
#pragma unroll
  for (int k = 0; k < 9; k++) rotation[(size_t) k*N + i] = L2[k];
  nvmul (L2, O, o);
  angular[3*(size_t) N + i] = o[0];
  angular[4*(size_t) N + i] = o[1];
  angular[5*(size_t) N + i] = o[2];
  position[i] += step * linear[i];
  position[(size_t) N + i] += step * linear[(size_t) N + i];
  position[2*(size_t) N + i] += step * linear[2*(size_t) N + i];
}

/* invert, parmec.ispc:62-87 (INVERT, macros.h:1033-1051) */
__global__ void k_invert (int N, const double *inertia, double *inverse, const double *mass, double *invm)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double F[9], det;
#pragma unroll
  for (int k = 0; k < 9; k++) F[k] = inertia[(size_t) k*N + i];
  det = F[0]*(F[4]*F[8]-F[5]*F[7]) + F[3]*(F[2]*F[7]-F[1]*F[8]) + (F[1]*F[5]-F[2]*F[4])*F[6];
  if (det != 0.0)
  {
    det = 1.0 / det;
    inverse[i] = (F[4]*F[8]-F[5]*F[7])*det;
    inverse[(size_t) 1*N + i] = (F[2]*F[7]-F[1]*F[8])*det;
    inverse[(size_t) 2*N + i] = (F[1]*F[5]-F[2]*F[4])*det;
    inverse[(size_t) 3*N + i] = (F[5]*F[6]-F[3]*F[8])*det;
    inverse[(size_t) 4*N + i] = (F[0]*F[8]-F[2]*F[6])*det;
    inverse[(size_t) 5*N + i] = (F[2]*F[3]-F[0]*F[5])*det;
    inverse[(size_t) 6*N + i] = (F[3]*F[7]-F[4]*F[6])*det;
    inverse[(size_t) 7*N + i] = (F[1]*F[6]-F[0]*F[7])*det;
    inverse[(size_t) 8*N + i] = (F[0]*F[4]-F[1]*F[3])*det;
  }
  invm[i] = 1.0 / mass[i];
}

} /* namespace */

/* zero leaf counts, whatever the control block says (after a build, or
 * after runs outside a step) */
int clear_leaf_counts (pmx_ctx &c)
{
  int L = std::max (c.tree.nleaves, 1);
  hipLaunchKernelGGL (k_zero_i, grid (L), dim3 (PMX_TPB), 0, c.stream, c.lcount.p, L);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

int step_prepare (pmx_ctx &c)
{
  int L = std::max (c.tree.nleaves, 1);
  if (c.lcount.alloc (L) || c.loff.alloc (L) || c.lbox.alloc (6 * (size_t) L))
  { pmx_set_error ("out of device memory (bins)"); return PMX_ENOMEM; }
  int rc = clear_leaf_counts (c);
  if (rc) return rc;
  /* leaf records: PMX_LSIZE slots per leaf (k_bin writes each record into
   * its slot; no scan, no k_scatter) while that takes at most 4 slots per
   * sphere and 2^31 slots in all, else packed at the scanned counts (also
   * when asked for: the environment variable PMX_LEAF_RECORDS=packed) */
  size_t slots = (size_t) L * PMX_LSIZE, E = std::max (c.E, 1);
  c.lstride = !c.rec_packed && slots <= 4 * E && slots < ((size_t) 1 << 31) ? PMX_LSIZE : 0;
  size_t R = c.lstride ? slots : E;
  if (c.lmem.alloc (R) || c.lpart.alloc (R) || c.lpos.alloc (4 * R))
  { pmx_set_error ("out of device memory (leaf records)"); return PMX_ENOMEM; }
  /* size the scan scratch for the per-step leaf scan now: nothing may
   * allocate while a step is being captured */
  size_t need = (size_t) L / (PMX_TPB*4) + 64;
  if (c.tree.scanbuf.alloc (need) != hipSuccess) { pmx_set_error ("out of device memory (scan)"); return PMX_ENOMEM; }
  if (c.nmtri > 0 && (c.mtcnt.alloc (L) || c.mtlist.alloc ((size_t) L * c.mcap)))
  { pmx_set_error ("out of device memory (obstacle lists)"); return PMX_ENOMEM; }
  return PMX_OK;
}

int step_bin (pmx_ctx &c)
{
  TreeDev &t = c.tree;
  /* the leaf counts are zero here: cleared by step_prepare after a build,
   * by k_rows in every step and by the autotuner after its runs */
  {
    /* per sphere: centre 24, leaf 4 (read for the hint; written where it
     * changes), slot 4; slotted: no slot, but radius 8, particle 4 and the
     * record 32, member 4, particle 4 */
    ProfScope ps (c, K_BIN, (c.lstride ? 80.0 : 32.0) * c.E);
    hipLaunchKernelGGL (k_bin, grid (c.E), dim3 (PMX_TPB), 0, c.stream, c.E, (const double*) c.center.p,
                        (const int*) t.roots.p, (const int*) c.ell_inst.p, (const int*) t.left.p, (const int*) t.right.p, (const int*) t.dim.p, (const double*) t.coord.p,
                        (const double*) t.cell.p, t.nleaves, c.ctl.p, c.leafof.p, c.slot.p, c.lcount.p, c.dflags.p,
                        c.ovf.p, c.staggered ? (const int*) c.istop.p : (const int*) nullptr, c.lstride,
                        (const double*) c.radii.p, (const int*) c.part.p, c.lmem.p, c.lpos.p, c.lpart.p, c.xcnt.p);
  }
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

/* packed layout: leaf offsets, leaf-ordered records and cleared
 * reverse-hit counts, the inputs of k_detect (k_bin has written them in
 * the slotted layout) */
int step_scatter (pmx_ctx &c)
{
  if (c.lstride) return PMX_OK;   /* slotted layout: k_bin wrote the records */
  TreeDev &t = c.tree;
  int L = std::max (t.nleaves, 1);
  int rc = scan_exclusive_to (c, (const int*) c.lcount.p, c.loff.p, L, c.ctl.p);
  if (rc) return rc;
  {
    ProfScope ps (c, K_SCATTER, 80.0 * c.E);   /* centre 24, radius 8, leaf + slot 8, particle 4; member 4, record 32, particle 4 */
    hipLaunchKernelGGL (k_scatter, grid (c.E), dim3 (PMX_TPB), 0, c.stream, c.ctl.p, c.E, (const double*) c.center.p,
                        (const double*) c.radii.p, (const int*) c.leafof.p, (const int*) c.slot.p, (const int*) c.loff.p,
                        c.lmem.p, c.lpos.p, c.staggered ? (const int*) c.istop.p : (const int*) nullptr,
                        (const int*) c.ell_inst.p, c.xcnt.p, (const int*) c.part.p, c.lpart.p);
  }
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

template <int NW>
void detect_launch (pmx_ctx &c)
{
  TreeDev &t = c.tree;
  hipLaunchKernelGGL (k_detect<NW>, dim3 (t.nleaves), dim3 (NW*64), 0, c.stream, c.E, (const double*) c.center.p,
                      (const double*) c.radii.p, (const int*) c.part.p, (const double*) t.cell.p,
                      t.nleaves, (const int*) c.lcount.p, (const int*) c.loff.p,
                      (const int*) c.lmem.p, (const double*) c.lpos.p, (const int*) c.lpart.p, (const double*) c.tri.p, c.T,
                      (const int*) t.tricnt.p, (const int*) t.trioff.p, (const int*) t.trilist.p,
                      c.nmtri > 0 ? (const int*) c.mtcnt.p : (const int*) nullptr, (const int*) c.mtlist.p, c.mcap,
                      (const int*) t.nbrcnt.p, (const int*) t.nbroff.p, (const int*) t.nbrlist.p, t.nbrstride,
                      c.hits.p, c.hitcnt.p, c.hcap, c.xbuf.p, c.xcnt.p, c.xcap, c.dflags.p, c.ctl.p,
                      c.staggered ? (const int*) c.istop.p : (const int*) nullptr, (const int*) t.leafinst.p,
                      c.lstride, (const double*) c.lbox.p);
}

/* k_detect with nw waves per block (1, 2, 4 or 8) */
int launch_detect (pmx_ctx &c, int nw)
{
  /* per sphere: its leaf record 32, member 4, particle 4, hit count 4 */
  ProfScope ps (c, K_DETECT, 44.0 * c.E);
  hipLaunchKernelGGL (k_leafbox, dim3 ((c.tree.nleaves + 3) / 4), dim3 (256), 0, c.stream, c.tree.nleaves,
                      (const int*) c.lcount.p, (const int*) c.loff.p, (const double*) c.lpos.p, c.lstride, c.lbox.p,
                      (const int*) c.ctl.p);
  if (nw == 1) detect_launch<1> (c);
  else if (nw == 2) detect_launch<2> (c);
  else if (nw == 8) detect_launch<8> (c);
  else detect_launch<4> (c);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

int step_detect (pmx_ctx &c)
{
  int rc = step_scatter (c);
  if (rc) return rc;
  return launch_detect (c, c.det_nw);
}

/* k_rows over the busy lists of nb idle blocks of tpb threads: one block
 * per list, or up to 8 per list (slices) when there are fewer than 4096
 * lists, so that small models spread their busy spheres over more waves */
template <class R, bool ADP>
void rows_launch_t (pmx_ctx &c, R *st0, R *st1, double h, int tpb, int nb)
{
  int slices = std::max (1, std::min (8, 4096 / std::max (nb, 1)));
  hipLaunchKernelGGL ((k_rows<R, ADP>), dim3 (nb * slices), dim3 (64), 0, c.stream, c.E, c.N,
                      (const double*) c.center.p, (const double*) c.radii.p, (const int*) c.part.p,
                      (const int*) c.ellcol.p, (const double*) c.tri.p, c.T, (const int*) c.tricol.p,
                      (const int*) c.triobs.p, (const double*) c.position.p, (const double*) c.linear.p,
                      (const double*) c.angular.p, (const double*) c.mass.p, (const double*) c.invm.p, c.P,
                      (const int*) c.pairs.p, (const int*) c.ikind.p, (const double*) c.iparam.p,
                      c.I, (const int*) c.ell_inst.p, (const int*) c.pair_off.p,
                      c.rowkey[0].p, st0, c.rowcnt[0].p, c.rowkey[1].p, st1, c.rowcnt[1].p,
                      c.cur0, c.cap, c.hits.p, c.hitcnt.p, c.hcap,
                      (const int*) c.xbuf.p, c.xcnt.p, c.xcap, (const double*) c.gtab.p,
                      c.ctl.p, h, c.force.p, c.torque.p, c.dflags.p,
                      (const double*) c.obspnt.p, (const double*) c.obsang.p, (const double*) c.obslin.p,
                      (const int*) c.rbusy.p, (const int*) c.rnbusy.p, tpb, slices,
                      (const double*) c.rotation.p, ADP && c.adp_on ? c.adp.p : (double*) nullptr,
                      ADP && c.opt.critical ? c.crec.p : (double*) nullptr,
                      ADP && c.mesh_contact ? c.tcnt.p : (int*) nullptr, c.trec_e.p, c.trec.p, c.tcap);
}

template <class R>
void rows_launch (pmx_ctx &c, R *st0, R *st1, double h, int tpb, int nb)
{
  if (c.adp_on || c.opt.critical || c.mesh_contact) rows_launch_t<R, true> (c, st0, st1, h, tpb, nb);
  else rows_launch_t<R, false> (c, st0, st1, h, tpb, nb);
}

template <int TPB>
void rows_idle_launch (pmx_ctx &c)
{
  hipLaunchKernelGGL ((k_rows_idle<TPB>), dim3 ((c.E + TPB - 1) / TPB), dim3 (TPB), 0, c.stream, c.E, c.N,
                      (const int*) c.part.p, (const double*) c.mass.p, c.I, (const int*) c.ell_inst.p,
                      c.rowcnt[0].p, c.rowcnt[1].p, c.cur0, (const int*) c.hitcnt.p, (const int*) c.xcnt.p,
                      (const double*) c.gtab.p, (const int*) c.ctl.p, c.force.p, c.torque.p,
                      c.staggered ? (const int*) c.istop.p : (const int*) nullptr, c.lcount.p, c.tree.nleaves,
                      c.rbusy.p, c.rnbusy.p, (const signed char*) c.iso.p);
}

/* the rows of a step: k_rows_idle with tpb threads per block (64 to 1024),
 * then k_rows over the busy lists */
int launch_rows (pmx_ctx &c, int tpb, double h)
{
  if (c.E == 0) return PMX_OK;
  ProfScope ps (c, K_ROWS, 136.0 * c.E);
  int nb = (c.E + tpb - 1) / tpb;
  if (tpb == 64) rows_idle_launch<64> (c);
  else if (tpb == 128) rows_idle_launch<128> (c);
  else if (tpb == 256) rows_idle_launch<256> (c);
  else if (tpb == 512) rows_idle_launch<512> (c);
  else rows_idle_launch<1024> (c);
  if (c.opt.precision == PMX_MIXED) rows_launch<float> (c, c.rowst32[0].p, c.rowst32[1].p, h, tpb, nb);
  else rows_launch<double> (c, c.rowst[0].p, c.rowst[1].p, h, tpb, nb);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

int step_rows (pmx_ctx &c, double h)
{
  if (c.nnosph > 0)
  {
    ProfScope ps (c, K_FORCE_INIT, 60.0 * c.nnosph);
    hipLaunchKernelGGL (k_force_init, grid (c.nnosph), dim3 (PMX_TPB), 0, c.stream, (const int*) c.ctl.p, c.N,
                        (const int*) c.nosph.p, c.nnosph, (const double*) c.mass.p, (const double*) c.gtab.p,
                        c.I, (const int*) c.par_inst.p, c.force.p, c.torque.p,
                        c.staggered ? (const int*) c.istop.p : (const int*) nullptr);
  }
  if (c.mesh_contact)   /* this step's reactions on meshed particles start from empty slots */
    hipLaunchKernelGGL (k_zero_i, grid (c.T), dim3 (PMX_TPB), 0, c.stream, c.tcnt.p, c.T);
  int rc = launch_rows (c, c.rows_tpb, h);
  if (rc) return rc;
  if (c.mesh_contact && (rc = step_react (c, false))) return rc;
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

int step_react (pmx_ctx &c, bool critical)
{
  if (!c.mesh_contact || c.nmp == 0) return PMX_OK;
  double *adp = critical || c.adp_on ? c.adp.p : nullptr;
  hipLaunchKernelGGL (k_react, grid (c.nmp), dim3 (PMX_TPB), 0, c.stream, c.nmp, (const int*) c.mp_part.p,
                      (const int*) c.mp_toff.p, (const int*) c.mp_tlist.p, c.T, (const int*) c.tcnt.p,
                      (const int*) c.trec_e.p, (const double*) c.trec.p, c.tcap, c.N, (const double*) c.position.p,
                      (const double*) c.rotation.p, (const double*) c.mass.p, (const double*) c.gtab.p, c.I,
                      (const int*) c.par_inst.p, c.force.p, c.torque.p, adp,
                      critical ? (const int*) nullptr : (const int*) c.ctl.p,
                      c.staggered ? (const int*) c.istop.p : (const int*) nullptr, critical ? 1 : 0);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

int step_dynamics (pmx_ctx &c, double step0, double step1)
{
  /* algorithmic bytes as each particle's branch reads and writes them, with
   * every contact row empty and no damping: per particle the flag 1, state
   * (Omega, v, L, x) 144 read and 168 written, 1/m 8, inertia and inverse 16
   * (isotropic) or 144, its sphere's index 4, its instance 4 (batches); per
   * sphere the centre written 24, the reference offset 48 unless it is +0;
   * force and torque 48 read, except for a particle loaded by contacts and
   * gravity only, which reads its row count 4 and mass 8 instead (each such
   * particle whose row is not empty reads 40 B more than counted here) */
  double dyn = 321.0 * c.N + 16.0 * c.niso + 144.0 * (c.N - c.niso) + (c.E > 0 ? 4.0 * c.N : 0.0) +
               (c.I > 1 ? 4.0 * c.N : 0.0) + 24.0 * c.E + 48.0 * (c.E - c.nzc) + 12.0 * c.nfree +
               48.0 * (c.N - c.nfree);
  /* fused binning (slotted records): per sphere its radius 8 and leaf 4
   * read, the record 32, member 4 and particle 4 written (the centre is in
   * registers) */
  BinArgs B {};
  if (c.fuse_bin)
  {
    TreeDev &t = c.tree;
    dyn += 52.0 * c.E;
    B = BinArgs {(const int*) t.roots.p, (const int*) t.left.p, (const int*) t.right.p, (const int*) t.dim.p,
                 (const double*) t.coord.p, (const double*) t.cell.p, (const double*) c.radii.p, t.nleaves,
                 c.lstride, c.bin_until, c.leafof.p, c.lcount.p, c.dflags.p, c.ovf.p, c.lmem.p, c.lpart.p,
                 c.lpos.p};
  }
  ProfScope ps (c, K_DYNAMICS, dyn);
#define DYN_ARGS c.N, c.angular.p, c.linear.p, c.rotation.p, \
                 c.position.p, (const double*) c.inertia.p, (const double*) c.inverse.p, (const double*) c.mass.p, \
                 (const double*) c.invm.p, c.damped ? (const double*) c.dtab.p : nullptr, (const int*) c.ctl.p, \
                 c.force.p, c.torque.p, \
                 step0, step1, (const int*) c.parell.p, c.E, c.center.p, (const signed char*) c.iso.p, c.I, \
                 (const int*) c.par_inst.p, c.staggered ? (const int*) c.istop.p : (const int*) nullptr, \
                 (const int*) c.rowcnt[0].p, (const int*) c.rowcnt[1].p, c.cur0, (const double*) c.gtab.p
  /* one launch per inertia form, each over its own particles (dynlist) */
  const int *dl = c.niso > 0 && c.niso < c.N ? (const int*) c.dynlist.p : (const int*) nullptr;
  if (c.niso > 0)
  {
    if (c.fuse_bin)
      hipLaunchKernelGGL ((k_dynamics<true, true>), grid (c.niso), dim3 (PMX_TPB), 0, c.stream, DYN_ARGS, dl, c.niso, B,
                          c.ctl.p);
    else
      hipLaunchKernelGGL ((k_dynamics<true, false>), grid (c.niso), dim3 (PMX_TPB), 0, c.stream, DYN_ARGS, dl, c.niso,
                          B, c.ctl.p);
  }
  if (c.niso < c.N)
  {
    if (c.fuse_bin)
      hipLaunchKernelGGL ((k_dynamics<false, true>), grid (c.N - c.niso), dim3 (PMX_TPB), 0, c.stream, DYN_ARGS,
                          dl ? dl + c.niso : dl, c.N - c.niso, B, c.ctl.p);
    else
      hipLaunchKernelGGL ((k_dynamics<false, false>), grid (c.N - c.niso), dim3 (PMX_TPB), 0, c.stream, DYN_ARGS,
                          dl ? dl + c.niso : dl, c.N - c.niso, B, c.ctl.p);
  }
#undef DYN_ARGS
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

/* entries of every sphere's current row and spheres with a non-empty row,
 * each instance read from the row buffer holding its rows (icur) */
__global__ void k_row_totals (int E, const int *ell_inst, const int *icur, const int *cnt0, const int *cnt1,
                              unsigned long long *out)
{
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n = 0, b = 0;
  if (e < E)
  {
    int buf = icur[ell_inst ? ell_inst[e] : 0];
    int k = buf ? cnt1[e] : cnt0[e];
    n = (unsigned long long) k;
    b = k > 0;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
  {
    n += __shfl_xor (n, off, 64);
    b += __shfl_xor (b, off, 64);
  }
  if ((threadIdx.x & 63) == 0 && (n || b))
  {
    atomicAdd (&out[0], n);
    atomicAdd (&out[1], b);
  }
}

int row_totals (pmx_ctx &c, long long out[2])
{
  out[0] = out[1] = 0;
  if (c.E == 0) return PMX_OK;
  if (c.fin.alloc (std::max (c.I, 4)) || c.hmin.alloc (2)) { pmx_set_error ("out of device memory"); return PMX_ENOMEM; }
  PMX_HIP (hipMemcpyAsync (c.fin.p, c.icur.data (), sizeof (int) * c.I, hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL (k_zero_u64, dim3 (1), dim3 (64), 0, c.stream, c.hmin.p, 2);
  hipLaunchKernelGGL (k_row_totals, grid (c.E), dim3 (PMX_TPB), 0, c.stream, c.E,
                      c.I > 1 ? (const int*) c.ell_inst.p : (const int*) nullptr, (const int*) c.fin.p,
                      (const int*) c.rowcnt[0].p, (const int*) c.rowcnt[1].p, c.hmin.p);
  PMX_HIP (hipGetLastError ());
  unsigned long long h[2];
  PMX_HIP (hipMemcpyAsync (h, c.hmin.p, sizeof (h), hipMemcpyDeviceToHost, c.stream));
  PMX_HIP (hipStreamSynchronize (c.stream));
  out[0] = (long long) h[0];
  out[1] = (long long) h[1];
  return PMX_OK;
}

/* fin[i]: the last step (in this call) of each instance finishing now, or -1 */
int step_materialize (pmx_ctx &c, const std::vector<int> &fin)
{
  if (c.N == 0 || c.E == 0) return PMX_OK;
  if (c.fin.alloc (std::max (c.I, 1))) { pmx_set_error ("out of device memory"); return PMX_ENOMEM; }
  PMX_HIP (hipMemcpyAsync (c.fin.p, fin.data (), sizeof (int) * c.I, hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL (k_materialize, grid (c.N), dim3 (PMX_TPB), 0, c.stream, c.N, (const signed char*) c.iso.p,
                      (const int*) c.parell.p, (const int*) c.rowcnt[0].p, (const int*) c.rowcnt[1].p, c.cur0, c.I,
                      (const int*) c.par_inst.p, (const int*) c.fin.p, (const double*) c.gtab.p,
                      (const double*) c.dtab.p, (const double*) c.mass.p, c.force.p, c.torque.p);
  PMX_HIP (hipGetLastError ());
  /* fin lives on the caller's stack: the copy must be done before it returns */
  PMX_HIP (hipStreamSynchronize (c.stream));
  return PMX_OK;
}

int step_end (pmx_ctx &c)
{
  hipLaunchKernelGGL (k_step_end, dim3 (1), dim3 (64), 0, c.stream, c.ctl.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

/* obstaclev + obstacles at the end of a step (parmec.cpp:2498-2500); tab
 * holds the velocities per step (row < 0: the step's own row) */
int step_obstacles (pmx_ctx &c, const double *tab, int row, double h)
{
  if (c.O == 0) return PMX_OK;
  if (c.nmtri > 0)
    hipLaunchKernelGGL (k_obstacles, grid (c.nmtri), dim3 (PMX_TPB), 0, c.stream, c.nmtri, (const int*) c.mtri.p,
                        (const int*) c.mtobs.p, c.T, c.tri.p, (const double*) c.obspnt.p, tab, c.O, row, h,
                        (const int*) c.ctl.p, c.staggered ? (const int*) c.istop.p : (const int*) nullptr,
                        (const int*) c.tri_inst.p);
  hipLaunchKernelGGL (k_obspnt, grid (c.O), dim3 (PMX_TPB), 0, c.stream, c.O, c.obspnt.p, c.obsang.p, c.obslin.p, tab,
                      row, h, (const int*) c.ctl.p, c.staggered ? (const int*) c.istop.p : (const int*) nullptr,
                      (const int*) c.obs_inst.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

int step_prescribe_acc (pmx_ctx &c)
{
  if (c.nprs == 0) return PMX_OK;
  hipLaunchKernelGGL (k_prs_acc, grid (c.nprs), dim3 (PMX_TPB), 0, c.stream, c.nprs, c.N, (const int*) c.prs_part.p,
                      (const int*) c.prs_lf.p, (const int*) c.prs_af.p, (const double*) c.prstab.p, c.PRSC,
                      (const double*) c.mass.p, (const double*) c.inertia.p, c.force.p, c.torque.p, (const int*) c.ctl.p,
                      c.staggered ? (const int*) c.istop.p : (const int*) nullptr, (const int*) c.par_inst.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

int step_prescribe_vel (pmx_ctx &c)
{
  if (c.nprs == 0) return PMX_OK;
  hipLaunchKernelGGL (k_prs_vel, grid (c.nprs), dim3 (PMX_TPB), 0, c.stream, c.nprs, c.N, (const int*) c.prs_part.p,
                      (const int*) c.prs_lv.p, (const int*) c.prs_av.p, (const double*) c.prstab.p, c.PRSC,
                      (const double*) c.rotation.p, c.linear.p, c.angular.p, (const int*) c.ctl.p,
                      c.staggered ? (const int*) c.istop.p : (const int*) nullptr, (const int*) c.par_inst.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

/* a step up to its forces (partitioning_store .. forces, prescribed
 * accelerations; with the adaptive step on, the per-particle estimates and
 * their minimum): everything that runs before the step's length step1 is
 * known.  h is step0, the length of the tangential contact update */
int enqueue_step_forces (pmx_ctx &c, double h)
{
  int rc;
  /* with fused binning the previous step's dynamics (or the call's own
   * k_bin) has binned this step's centres */
  if (c.E > 0 && ((!c.fuse_bin && (rc = step_bin (c))) || (rc = step_mtri (c)) || (rc = step_detect (c)))) return rc;
  if (c.adp_on && (rc = adp_begin (c))) return rc;
  /* mesh nodes and faces (shapes.ispc:101-211) are updated per step only
   * where spheres can touch them (enqueue_step_motion): elsewhere nothing on
   * this path reads them (springs carry their own points), so they are formed
   * from the final particle state when an instance finishes, with the same
   * arithmetic and the same result */
  /* forces(): contacts, springs, torsion springs, UNSPRING tests
   * (forces.ispc:2153-2217) */
  if ((rc = step_rows (c, h)) || (rc = step_springs (c, h)) || (rc = step_torsion (c, h)) || (rc = step_unspring (c)))
    return rc;
  if (c.adp_on && (rc = adp_time_step (c, nullptr, nullptr, nullptr))) return rc;
  return PMX_OK;
}

/* the rest of a step: ball joints and restraints, prescribed
 * accelerations, dynamics with step0 and step1, prescribed velocities,
 * obstacles moved by (step0+step1)/2, the history sample and the step
 * counter */
int enqueue_step_motion (pmx_ctx &c, double step0, double step1)
{
  int rc;
  /* solve_joints, restrain_forces, prescribe_acceleration, dynamics,
   * prescribe_velocity (parmec.cpp:2453-2466) */
  if ((rc = step_joints (c, step0, step1)) || (rc = step_prescribe_acc (c)) || (rc = step_dynamics (c, step0, step1)) ||
      (rc = step_prescribe_vel (c)))
    return rc;
  /* spheres touch meshed particles: their nodes and triangles follow every
   * step (shapes.ispc:101-211), as the next step's detection reads them */
  if (c.mesh_contact && (rc = step_mesh_shapes (c, true))) return rc;
  if ((rc = step_obstacles (c, c.obstab.p, -1, 0.5*(step0+step1))) || (rc = step_history (c)) || (rc = step_end (c)))
    return rc;
  return PMX_OK;
}

int enqueue_step (pmx_ctx &c, double h)
{
  int rc = enqueue_step_forces (c, h);
  return rc ? rc : enqueue_step_motion (c, h, h);
}

int step_shapes (pmx_ctx &c)
{
  if (c.E == 0) return PMX_OK;
  ProfScope ps (c, K_SHAPES, 168.0 * c.E);
  hipLaunchKernelGGL (k_shapes, grid (c.E), dim3 (PMX_TPB), 0, c.stream, c.E, c.N, (const int*) c.part.p, c.center.p,
                      (const double*) c.rotation.p, (const double*) c.position.p,
                      c.first_sel ? (const int*) c.isel.p : (const int*) nullptr, (const int*) c.ell_inst.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

int step_euler (pmx_ctx &c, double h)
{
  hipLaunchKernelGGL (k_euler, grid (c.N), dim3 (PMX_TPB), 0, c.stream, c.N, c.angular.p, (const double*) c.linear.p,
                      c.rotation.p, c.position.p, h, c.first_sel ? (const int*) c.isel.p : (const int*) nullptr,
                      (const int*) c.par_inst.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

int step_invert (pmx_ctx &c)
{
  hipLaunchKernelGGL (k_invert, grid (c.N), dim3 (PMX_TPB), 0, c.stream, c.N, (const double*) c.inertia.p,
                      c.inverse.p, (const double*) c.mass.p, c.invm.p);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

/* Autotuner: pick the waves per block of k_detect for this workload by
 * timing each variant on the current state (detection only reads its inputs
 * and rewrites its outputs, so runs repeat exactly; block shapes never change
 * a result).  The block size of k_rows is not tuned: its cost depends on the
 * share of spheres in contact, which the early steps where tuning runs do
 * not show, and its largest block (most spheres per compaction) is best or
 * level on the workloads measured.  Called at a step boundary with the
 * control block clear; the caller resets the control block afterwards (the
 * next step redoes everything the timings wrote).  Counts a pass in c.tuned
 * when it completes. */
int autotune (pmx_ctx &c, double h)
{
  struct NoProf   /* the timings are not profiled steps */
  {
    Prof &p; int was;
    NoProf (Prof &p_) : p (p_), was (p_.mode) { p.mode = 0; }
    ~NoProf () { p.mode = was; }
  } np (c.prof);
  int rc;
  if ((rc = step_bin (c)) || (rc = step_mtri (c)) || (rc = step_scatter (c))) return rc;
  int ctl[CTL_COUNT];
  PMX_HIP (hipMemcpyAsync (ctl, c.ctl.p, sizeof (ctl), hipMemcpyDeviceToHost, c.stream));
  PMX_HIP (hipStreamSynchronize (c.stream));
  if (ctl[CTL_ABORT])
  {
    /* a leaf (or moving-triangle list) overflows: tune at a later call; the
     * instances this binning flagged are not re-partitioned for it */
    PMX_HIP (hipMemsetAsync (c.ovf.p, 0, sizeof (int) * std::max (c.I, 1), c.stream));
    return clear_leaf_counts (c);
  }
  hipEvent_t e0, e1;
  PMX_HIP (hipEventCreate (&e0));
  PMX_HIP (hipEventCreate (&e1));
  /* prep runs before the timed region, run inside it */
  auto best_of = [&] (const int *opts, int nopt, auto prep, auto run) -> int {
    int best = opts[0];
    float bt = 1e30f;
    for (int k = 0; k < nopt; k++)
      for (int rep = 0; rep < 3; rep++)
      {
        if (prep ()) return -1;
        (void) hipEventRecord (e0, c.stream);
        if (run (opts[k])) return -1;
        (void) hipEventRecord (e1, c.stream);
        float ms = 0.0f;
        if (hipEventSynchronize (e1) != hipSuccess || hipEventElapsedTime (&ms, e0, e1) != hipSuccess) return -1;
        if (ms < bt) { bt = ms; best = opts[k]; }
      }
    return best;
  };
  const int nws[4] = {1, 2, 4, 8};
  /* the reverse-hit counts start at zero for each detection, and so does the
   * abort word (a run whose hits overflow their capacity raises it; the
   * counts stay bounded by the capacities, the hits are discarded) */
  auto clear = [&] () {
    hipLaunchKernelGGL (k_zero_i, grid (c.E), dim3 (PMX_TPB), 0, c.stream, c.xcnt.p, c.E);
    hipLaunchKernelGGL (k_zero_i, dim3 (1), dim3 (64), 0, c.stream, c.ctl.p + CTL_ABORT, 1);
    return hipGetLastError () == hipSuccess ? 0 : -1; };
  int nw = best_of (nws, 4, clear, [&] (int v) { return launch_detect (c, v); });
  (void) hipEventDestroy (e0);
  (void) hipEventDestroy (e1);
  if (nw < 0) { pmx_set_error ("autotuning failed"); return PMX_EDEVICE; }
  c.det_nw = nw;
  c.tuned++;
  /* hit and reverse-hit counts of the runs back to zero */
  hipLaunchKernelGGL (k_zero_i, grid (c.E), dim3 (PMX_TPB), 0, c.stream, c.hitcnt.p, c.E);
  if (clear ()) { pmx_set_error ("autotuning failed"); return PMX_EDEVICE; }
  if ((rc = clear_leaf_counts (c))) return rc;   /* the next bin starts from zero */
  return PMX_OK;
}
