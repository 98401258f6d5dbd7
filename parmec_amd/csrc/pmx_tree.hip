/*
 * pmx_tree.hip -- device build of Parmec's frozen partitioning tree
 * (partitioning_create, partition.ispc:663-746).
 *
 *   extents -> 30-bit Morton codes (partition.ispc:58-94)
 *           -> stable LSD radix sort, 6-bit digits, per-wave ballot ranking
 *              staged in LDS (replaces the task quick sort + tournament merge,
 *              partition.ispc:97-243)
 *           -> Karras internal nodes, CUTOFF terminals, split dimension
 *              (dnode-2)%3 and split coordinate = min coordinate of the right
 *              child's points (partition.ispc:268-394)
 *           -> compact leaf numbering of the children of branch nodes
 *              (partitioning_tree_create, partition.ispc:525-569)
 *           -> per-leaf reach cell: an AABB [lo,hi] reaches a leaf through the
 *              descent rule "lo[d] <= coord -> left, hi[d] >= coord -> right"
 *              (condet.ispc:104-113) iff lo <= U and hi >= L component-wise,
 *              U = min coord over ancestors left of the path, L = max over
 *              ancestors right of it.
 */
#include "pmx_internal.h"
#include <cfloat>
#include <cmath>

namespace {

__device__ __forceinline__ unsigned expandbits (unsigned v)
{
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__device__ __forceinline__ int lzc (unsigned x) { return x ? __clz (x) : 32; }

/* generalised common prefix length within one instance's sorted range
 * [lo, hi] (partition.ispc:272-281); equal codes tie-break on the positions
 * relative to the range, as in a run of that instance alone */
__device__ __forceinline__ int delta (int i, unsigned codei, int j, int lo, int hi, const uint64_t *key)
{
  if (j < lo || j > hi) return -1;
  unsigned codej = (unsigned) (key[j] & 0x3FFFFFFFull);
  if (codei == codej) return 32 + lzc ((unsigned) ((i - lo) ^ (j - lo)));
  return lzc (codei ^ codej);
}

/* ---------------------------------------------------------------- scan */

__global__ void k_scan_block (const int *src, int *data, int n, int *bsums, const int *ctl)
{
  if (ctl && ctl[CTL_ABORT]) return;
  __shared__ int wsum[PMX_TPB/64];
  int base = blockIdx.x * (PMX_TPB*4) + threadIdx.x * 4;
  int v[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) { v[k] = base+k < n ? src[base+k] : 0; s += v[k]; }
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = s;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1)
  {
    int y = __shfl_up (x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads ();
  int wpre = 0;
  for (int i = 0; i < w; i++) wpre += wsum[i];
  int run = wpre + x - s;
#pragma unroll
  for (int k = 0; k < 4; k++) { if (base+k < n) data[base+k] = run; run += v[k]; }
  if (bsums && threadIdx.x == PMX_TPB-1) bsums[blockIdx.x] = wpre + x;
}

__global__ void k_scan_add (int *data, int n, const int *boffs, const int *ctl)
{
  if (ctl && ctl[CTL_ABORT]) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) data[i] += boffs[i / (PMX_TPB*4)];
}

/* exclusive scan of src into data (src may be data) */
int scan_rec (pmx_ctx &c, const int *src, int *data, int n, int *scratch, size_t scratch_n, const int *ctl)
{
  const int per = PMX_TPB*4;
  int nb = (n + per - 1) / per;
  if (nb <= 1)
  {
    hipLaunchKernelGGL (k_scan_block, dim3 (1), dim3 (PMX_TPB), 0, c.stream, src, data, n, (int*) nullptr, ctl);
    PMX_HIP (hipGetLastError ());
    return PMX_OK;
  }
  if ((size_t) nb > scratch_n) { pmx_set_error ("scan scratch too small"); return PMX_ENOMEM; }
  hipLaunchKernelGGL (k_scan_block, dim3 (nb), dim3 (PMX_TPB), 0, c.stream, src, data, n, scratch, ctl);
  PMX_HIP (hipGetLastError ());
  int rc = scan_rec (c, scratch, scratch, nb, scratch + nb, scratch_n - nb, ctl);
  if (rc) return rc;
  hipLaunchKernelGGL (k_scan_add, dim3 ((n + PMX_TPB - 1)/PMX_TPB), dim3 (PMX_TPB), 0, c.stream, data, n, (const int*) scratch, ctl);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}

/* ------------------------------------------------------------- extents */

__device__ __forceinline__ void block_minmax6 (double e[6], double *sh /* [6][PMX_TPB/64] */)
{
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
  {
#pragma unroll
    for (int k = 0; k < 3; k++)
    {
      double a = __shfl_xor (e[k], off, 64);
      double b = __shfl_xor (e[3+k], off, 64);
      e[k] = a < e[k] ? a : e[k];
      e[3+k] = b > e[3+k] ? b : e[3+k];
    }
  }
  if (lane == 0)
    for (int k = 0; k < 6; k++) sh[k*(PMX_TPB/64) + w] = e[k];
  __syncthreads ();
  if (threadIdx.x == 0)
  {
    for (int ww = 1; ww < PMX_TPB/64; ww++)
      for (int k = 0; k < 3; k++)
      {
        double a = sh[k*(PMX_TPB/64) + ww], b = sh[(3+k)*(PMX_TPB/64) + ww];
        e[k] = a < e[k] ? a : e[k];
        e[3+k] = b > e[3+k] ? b : e[3+k];
      }
  }
}

__global__ void k_extents_partial (int n, const double *cx, const double *cy, const double *cz, double *part)
{
  __shared__ double sh[6*(PMX_TPB/64)];
  double e[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
  {
    double x = cx[i], y = cy[i], z = cz[i];
    if (x < e[0]) e[0] = x;
    if (y < e[1]) e[1] = y;
    if (z < e[2]) e[2] = z;
    if (x > e[3]) e[3] = x;
    if (y > e[4]) e[4] = y;
    if (z > e[5]) e[5] = z;
  }
  block_minmax6 (e, sh);
  if (threadIdx.x == 0) for (int k = 0; k < 6; k++) part[6*blockIdx.x + k] = e[k];
}

__global__ void k_extents_final (int nb, double *ext /* ext[0..5] out, partials from ext+6 */)
{
  __shared__ double sh[6*(PMX_TPB/64)];
  double e[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
  const double *part = ext + 6;
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int k = 0; k < 3; k++)
    {
      double a = part[6*b + k], c = part[6*b + 3 + k];
      e[k] = a < e[k] ? a : e[k];
      e[3+k] = c > e[3+k] ? c : e[3+k];
    }
  block_minmax6 (e, sh);
  if (threadIdx.x == 0) for (int k = 0; k < 6; k++) ext[k] = e[k];
}

/* per-instance extents, one block per instance (only the instances with
 * a nonzero flag when flag is not null) */
__global__ void k_extents_seg (const int *off, const double *cx, const double *cy, const double *cz, double *ext,
                               const int *flag)
{
  __shared__ double sh[6*(PMX_TPB/64)];
  int inst = blockIdx.x;
  if (flag && !flag[inst]) return;   /* uniform over the block */
  double e[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int i = off[inst] + threadIdx.x; i < off[inst+1]; i += blockDim.x)
  {
    double x = cx[i], y = cy[i], z = cz[i];
    if (x < e[0]) e[0] = x;
    if (y < e[1]) e[1] = y;
    if (z < e[2]) e[2] = z;
    if (x > e[3]) e[3] = x;
    if (y > e[4]) e[4] = y;
    if (z > e[5]) e[5] = z;
  }
  block_minmax6 (e, sh);
  if (threadIdx.x == 0) for (int k = 0; k < 6; k++) ext[6*inst + k] = e[k];
}

__global__ void k_set_i (int *a, int n, int v)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

/* the same inside a step (a no-op once the step is aborted) */
__global__ void k_set_ctl (const int *ctl, int *a, int n, int v)
{
  if (ctl[CTL_ABORT]) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

/* copy the current centres of the selected instances (all when sel is null) */
__global__ void k_snap (int n, const double *center, double *bcen, const int *inst, const int *sel)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (sel && !sel[inst[i]]) return;
  bcen[i] = center[i];
  bcen[(size_t) n + i] = center[(size_t) n + i];
  bcen[2*(size_t) n + i] = center[2*(size_t) n + i];
}

/* -------------------------------------------------------------- morton */

/* code and sphere of the p-th sphere of the build: sphere sel[p], or p
 * when sel is null */
__global__ void k_morton (int n, const double *cx, const double *cy, const double *cz, const double *ext_all,
                          const int *inst, uint64_t *code, int *order, const int *sel)
{
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n) return;
  int i = sel ? sel[w] : w;
  int ins = inst[i];
  const double *ext = ext_all + 6*ins;
  double p[3] = {cx[i], cy[i], cz[i]};
  unsigned q[3];
#pragma unroll
  for (int k = 0; k < 3; k++)
  {
    double w = ext[3+k] - ext[k];
    double pk = w > 0.0 ? (p[k]-ext[k])/w : 0.0;   /* zero extent -> cell 0 */
    double s = pk * 1024.0;
    s = s > 0.0 ? s : 0.0;
    s = s < 1023.0 ? s : 1023.0;
    q[k] = expandbits ((unsigned) (int) s);
  }
  code[w] = ((uint64_t) ins << 30) | (uint64_t) (q[0]*4 + q[1]*2 + q[2]);
  order[w] = i;
}

/* a partial build sorts the spheres of whole instances, listed in sphere
 * order: the instance segments of the sorted list are those of sel, so the
 * p-th sorted entry goes to position sel[p] of the instance-ordered arrays */
__global__ void k_place (int n, const int *sel, const uint64_t *key, const int *val, uint64_t *key2, int *val2)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int i = sel[p];
  key2[i] = key[p];
  val2[i] = val[p];
}

/* cm keys of the selected nodes to the maximum (the identity of min) */
__global__ void k_cm_init (int n, int stride, const int *sel, unsigned long long *cm)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int i = sel[p];
  for (int d = 0; d < 3; d++) cm[(size_t) d*stride + i] = ~0ull;
}

/* ---------------------------------------------------------- radix sort */

#define RADIX_BITS 6
#define RADIX 64

__global__ void k_radix_hist (int n, const uint64_t *key, int shift, int *hist, int nb)
{
  __shared__ int h[RADIX];
  if (threadIdx.x < RADIX) h[threadIdx.x] = 0;
  __syncthreads ();
  int i = blockIdx.x * PMX_TPB + threadIdx.x;
  if (i < n) atomicAdd (&h[(key[i] >> shift) & (RADIX-1)], 1);
  __syncthreads ();
  if (threadIdx.x < RADIX) hist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

/* stable scatter: rank inside the block = same-digit elements of lower
 * thread index, from per-wave match masks (6 ballots) and per-wave digit
 * counts staged in LDS */
__global__ void k_radix_scatter (int n, const uint64_t *key, const int *val, int shift, const int *hist, int nb,
                                 uint64_t *key2, int *val2)
{
  __shared__ int wcnt[PMX_TPB/64][RADIX];
  int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int k = tid; k < (PMX_TPB/64)*RADIX; k += PMX_TPB) (&wcnt[0][0])[k] = 0;
  __syncthreads ();
  int i = blockIdx.x * PMX_TPB + tid;
  bool valid = i < n;
  uint64_t k = valid ? key[i] : 0ull;
  int digit = (int) ((k >> shift) & (RADIX-1));
  unsigned long long m = __ballot (valid);
#pragma unroll
  for (int b = 0; b < RADIX_BITS; b++)
  {
    int bit = (digit >> b) & 1;
    unsigned long long bb = __ballot (bit);
    m &= bit ? bb : ~bb;
  }
  unsigned long long lt = lane ? ((1ull << lane) - 1ull) : 0ull;
  int rank = __popcll (m & lt);
  if (valid && (m >> lane) == 1ull) wcnt[w][digit] = __popcll (m);
  __syncthreads ();
  if (!valid) return;
  int pre = 0;
  for (int ww = 0; ww < w; ww++) pre += wcnt[ww][digit];
  int pos = hist[digit * nb + blockIdx.x] + pre + rank;
  key2[pos] = k;
  val2[pos] = val[i];
}

/* ----------------------------------------------------------- karras */

/* node sel[p] (node p when sel is null) of the instance's Karras tree */
__global__ void k_karras (int n, const uint64_t *key, const int *off, int *first, int *size, int *split, int *dim,
                          const int *sel)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (sel) i = sel[i];
  int ins = (int) (key[i] >> 30);
  int lo = off[ins], hi = off[ins+1] - 1;
  if (i >= hi) { dim[i] = -1; size[i] = 0; first[i] = i; split[i] = 0; return; }   /* not an internal node */
  unsigned codei = (unsigned) (key[i] & 0x3FFFFFFFull);
  int dd = delta (i, codei, i+1, lo, hi, key) - delta (i, codei, i-1, lo, hi, key);
  int d = (dd > 0) - (dd < 0);
  int dmin = delta (i, codei, i-d, lo, hi, key);
  int lmax = 2;
  while (delta (i, codei, i+lmax*d, lo, hi, key) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax/2; t >= 1; t /= 2)
    if (delta (i, codei, i+(l+t)*d, lo, hi, key) > dmin) l = l+t;
  int j = i+l*d;
  first[i] = d > 0 ? i : j;
  size[i] = l+1;
  if (l+1 <= PMX_CUTOFF)
  {
    split[i] = 0;
    dim[i] = -1;
  }
  else
  {
    int dnode = delta (i, codei, j, lo, hi, key);
    int s = 0;
    for (int t = (l+1)/2;; t = (t+1)/2)
    {
      if (delta (i, codei, i+(s+t)*d, lo, hi, key) > dnode) s = s+t;
      if (t == 1) break;
    }
    split[i] = i+s*d+min (d, 0);
    dim[i] = (dnode-2)%3;
  }
}

/* leaf children per branch node (plus the root itself when an instance's
 * whole range is one terminal node); branch flags for the coordinate pass */
__global__ void k_children (int n, const uint64_t *key, const int *off, const int *first, const int *size,
                            const int *split, const int *dim, int *nleafkids, int *parent, const int *sel)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (sel) i = sel[i];
  int ins = (int) (key[i] >> 30);
  int lo = off[ins];
  int k = 0;
  if (i == lo)
  {
    parent[i] = -1;
    if (dim[i] < 0) k = 1;            /* root leaf */
  }
  if (dim[i] < 0) { nleafkids[i] = k; return; }
  int j = split[i];
  if (first[i] == j || dim[j] < 0) k++; else parent[j] = 2*i;
  if (first[i]+size[i]-1 == j+1 || dim[j+1] < 0) k++; else parent[j+1] = 2*i+1;
  nleafkids[i] = k;
}

__global__ void k_assign (int n, const uint64_t *key, const int *off, const int *first, const int *size,
                          const int *split, const int *dim, const int *leafbase, int *left, int *right, int *leafpar,
                          int *leafinst, int *roots)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int ins = (int) (key[i] >> 30);
  int L = leafbase[i];
  if (i == off[ins])
  {
    if (dim[i] < 0) { roots[ins] = -(L+1); leafpar[L] = -1; leafinst[L] = ins; L++; }
    else roots[ins] = i;
  }
  if (dim[i] < 0) return;
  int j = split[i];
  if (first[i] == j || dim[j] < 0) { left[i] = -(L+1); leafpar[L] = 2*i; leafinst[L] = ins; L++; }
  else left[i] = j;
  if (first[i]+size[i]-1 == j+1 || dim[j+1] < 0) { right[i] = -(L+1); leafpar[L] = 2*i+1; leafinst[L] = ins; }
  else right[i] = j+1;
}

/* split coordinate = min of the right child's points along the split
 * dimension (mincoord, partition.ispc:297-308), for every branch node at once.
 * Each terminal child range (at most CUTOFF points, or one point) takes its
 * minima directly and folds them into every ancestor with 64-bit atomic
 * minima of order-preserving keys (exact; relaxed atomics suffice since the
 * results are read by the next kernel); then a node's split coordinate is
 * its right child's minimum along the split dimension.  cm[d][node] holds
 * the key of the minimum of the node's whole range. */
__device__ __forceinline__ unsigned long long dkey (double v)
{
  unsigned long long b = (unsigned long long) __double_as_longlong (v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double dval (unsigned long long k)
{
  unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  return __longlong_as_double ((long long) b);
}

__device__ __forceinline__ void range_min (int a, int b, const int *order, const double *cx, const double *cy,
                                           const double *cz, double m[3])
{
  m[0] = m[1] = m[2] = DBL_MAX;
  for (int k = a; k < b; k++)
  {
    int q = order[k];
    double x = cx[q], y = cy[q], z = cz[q];
    m[0] = x < m[0] ? x : m[0];
    m[1] = y < m[1] ? y : m[1];
    m[2] = z < m[2] ? z : m[2];
  }
}

/* one thread per (branch node, side) whose child is a terminal range */
__global__ void k_minfold (int cnt, int n, const int *first, const int *size, const int *split, const int *dim,
                           const int *parent, const int *order, const double *cx, const double *cy, const double *cz,
                           unsigned long long *cm, double *rmin, const int *sel)
{
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 2*cnt) return;
  int i = sel ? sel[q >> 1] : q >> 1, side = q & 1;
  if (dim[i] < 0) return;
  int j = split[i] + side;
  int a, b;
  if (side == 0 && first[i] == j) { a = j; b = j+1; }
  else if (side == 1 && first[i] + size[i] - 1 == j) { a = j; b = j+1; }
  else if (dim[j] < 0) { a = first[j]; b = first[j] + size[j]; }
  else return;                       /* a branch child: folded by its own ranges */
  double m[3];
  range_min (a, b, order, cx, cy, cz, m);
  if (side == 1) for (int d = 0; d < 3; d++) rmin[(size_t) d*n + i] = m[d];   /* terminal right child */
  unsigned long long k[3] = {dkey (m[0]), dkey (m[1]), dkey (m[2])};
  for (int node = i, guard = 0; node >= 0 && guard < 4096; guard++)
  {
    for (int d = 0; d < 3; d++) atomicMin (&cm[(size_t) d*n + node], k[d]);
    int p = parent[node];
    node = p < 0 ? -1 : (p >> 1);
  }
}

__global__ void k_coord_set (int m, int n, const int *first, const int *size, const int *split, const int *dim,
                             const unsigned long long *cm, const double *rmin, double *coord, const int *sel)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  if (sel) i = sel[i];
  if (dim[i] < 0) return;
  int d = dim[i], j = split[i] + 1;
  bool term = first[i] + size[i] - 1 == j || dim[j] < 0;
  coord[i] = term ? rmin[(size_t) d*n + i] : dval (cm[(size_t) d*n + j]);
}

__global__ void k_cells (int nleaves, const int *leafpar, const int *parent, const int *dim, const double *coord,
                         double *cell)
{
  int L = blockIdx.x * blockDim.x + threadIdx.x;
  if (L >= nleaves) return;
  double lo[3] = {-INFINITY, -INFINITY, -INFINITY}, hi[3] = {INFINITY, INFINITY, INFINITY};
  int p = leafpar[L];
  int guard = 0;
  while (p >= 0 && guard < 4096)
  {
    int node = p >> 1, side = p & 1;
    int d = dim[node];
    double c = coord[node];
    if (side == 0) hi[d] = c < hi[d] ? c : hi[d];
    else lo[d] = c > lo[d] ? c : lo[d];
    p = parent[node];
    guard++;
  }
  for (int k = 0; k < 3; k++) { cell[k*nleaves + L] = lo[k]; cell[(3+k)*nleaves + L] = hi[k]; }
}

/* leaves reached by a box, in the reference's descent order */
__device__ __forceinline__ int reach_leaves_count (int root, const int *left, const int *right, const int *dim,
                                                    const double *coord, const double lo[3], const double hi[3],
                                                    int *tricnt, int t, const int *trioff, int *trilist, int *fill,
                                                    int *flag)
{
  int stack[64];
  int sp = 0, ref = root, visited = 0;
  for (;;)
  {
    if (ref >= 0)
    {
      int d = dim[ref];
      double c = coord[ref];
      bool gl = lo[d] <= c, gr = hi[d] >= c;
      if (gl && gr)
      {
        if (sp < 64) stack[sp++] = right[ref]; else atomicAdd (flag, 1);
        ref = left[ref];
        continue;
      }
      if (gl) { ref = left[ref]; continue; }
      if (gr) { ref = right[ref]; continue; }
    }
    else
    {
      int L = -ref-1;
      if (trilist) { int s = atomicAdd (&fill[L], 1); trilist[trioff[L] + s] = t; }
      else atomicAdd (&tricnt[L], 1);
      visited++;
    }
    if (sp == 0) break;
    ref = stack[--sp];
  }
  return visited;
}

/* Static triangles of instances with at most TRI_LEAF_MAX triangles are
 * listed leaf by leaf (k_tri_leaf), the others triangle by triangle through
 * the tree (k_tri_reach).  Both give the same leaves: the descent (left when
 * lo <= coord, right when hi >= coord) reaches leaf L exactly when, on every
 * axis, lo <= the least coord of L's left-turn ancestors and hi >= the
 * greatest coord of its right-turn ancestors, which are L's cell bounds. */
#define TRI_LEAF_MAX 64

__global__ void k_tri_reach (int T, const double *tri, const int *tmov, const int *roots, const int *tri_inst,
                             const int *tri_off, const int *left, const int *right, const int *dim,
                             const double *coord, int *tricnt, const int *trioff, int *trilist, int *fill, int *flag)
{
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  if (tmov[t]) return;                 /* moving obstacle triangles: reach lists per step */
  int inst = tri_inst[t];
  if (tri_off[inst+1] - tri_off[inst] <= TRI_LEAF_MAX) return;   /* listed by k_tri_leaf */
  int root = roots[inst];
  if (root == PMX_NO_ROOT) return;     /* the triangle's instance has no spheres */
  double v[3][3];
  for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) v[a][b] = tri[(size_t) (3*a+b)*T + t];
  double lo[3], hi[3];
  for (int b = 0; b < 3; b++)
  {
    double mn = v[0][b] < v[1][b] ? v[0][b] : v[1][b];
    mn = mn < v[2][b] ? mn : v[2][b];
    double mx = v[0][b] > v[1][b] ? v[0][b] : v[1][b];
    mx = mx > v[2][b] ? mx : v[2][b];
    lo[b] = mn; hi[b] = mx;
  }
  reach_leaves_count (root, left, right, dim, coord, lo, hi, tricnt, t, trioff, trilist, fill, flag);
}

/* the static triangles of leaf L's instance whose box reaches L's cell, in
 * triangle order: counted into tricnt[L] (trilist null) or listed from
 * trioff[L]; only for instances with at most TRI_LEAF_MAX triangles */
__global__ void k_tri_leaf (int nleaves, const int *leafinst, const int *tri_off, const int *tmov, const double *tri,
                            int T, const double *cell, int *tricnt, const int *trioff, int *trilist)
{
  int L = blockIdx.x * blockDim.x + threadIdx.x;
  if (L >= nleaves) return;
  int inst = leafinst[L];
  int t0 = tri_off[inst], t1 = tri_off[inst+1];
  if (t1 - t0 > TRI_LEAF_MAX) return;
  double clo[3], chi[3];
  for (int d = 0; d < 3; d++) { clo[d] = cell[(size_t) d*nleaves + L]; chi[d] = cell[(size_t) (3+d)*nleaves + L]; }
  int n = 0, o = trilist ? trioff[L] : 0;
  for (int t = t0; t < t1; t++)
  {
    if (tmov[t]) continue;
    bool in = true;
    for (int d = 0; d < 3; d++)
    {
      double x0 = tri[(size_t) d*T + t], x1 = tri[(size_t) (3+d)*T + t], x2 = tri[(size_t) (6+d)*T + t];
      double mn = x0 < x1 ? x0 : x1, mx = x0 > x1 ? x0 : x1;
      mn = mn < x2 ? mn : x2;
      mx = mx > x2 ? mx : x2;
      in = in && mn <= chi[d] && mx >= clo[d];
    }
    if (!in) continue;
    if (trilist) trilist[o + n] = t;
    n++;
  }
  if (!trilist) tricnt[L] = n;
}

/* leaves reached by each moving obstacle triangle's current box (the
 * traversal of test_triangles, condet.ispc:412-426), appended to per-leaf
 * lists of capacity mcap; run every step since the triangles move */
__global__ void k_mtri_reach (int nm, const int *mtri, int T, const double *tri, const int *roots, const int *tri_inst,
                              const int *left, const int *right, const int *dim, const double *coord, int *mtcnt,
                              int *mtlist, int mcap, int *flag, int *ctl, const int *stop)
{
  if (ctl[CTL_ABORT]) return;
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nm) return;
  int t = mtri[q];
  if (frozen (stop, tri_inst[t], ctl)) return;
  int root = roots[tri_inst[t]];
  if (root == PMX_NO_ROOT) return;
  double lo[3], hi[3];
  for (int b = 0; b < 3; b++)
  {
    double x0 = tri[(size_t) b*T + t], x1 = tri[(size_t) (3+b)*T + t], x2 = tri[(size_t) (6+b)*T + t];
    double mn = x0 < x1 ? x0 : x1, mx = x0 > x1 ? x0 : x1;
    lo[b] = mn < x2 ? mn : x2;
    hi[b] = mx > x2 ? mx : x2;
  }
  int stack[64];
  int sp = 0, ref = root;
  for (;;)
  {
    if (ref >= 0)
    {
      int d = dim[ref];
      double cc = coord[ref];
      bool gl = lo[d] <= cc, gr = hi[d] >= cc;
      if (gl && gr)
      {
        if (sp < 64) stack[sp++] = right[ref];
        else { atomicAdd (flag + FLAG_TRILIST_OVERFLOW, 1); abort_step (ctl); return; }
        ref = left[ref];
      }
      else if (gl) ref = left[ref];
      else if (gr) ref = right[ref];
      else { if (sp == 0) break; ref = stack[--sp]; }
      continue;
    }
    int L = -ref-1;
    int s = atomicAdd (&mtcnt[L], 1);
    if (s < mcap) mtlist[(size_t) L*mcap + s] = t;
    else { atomicAdd (flag + FLAG_MTRI_OVERFLOW, 1); abort_step (ctl); }
    if (sp == 0) break;
    ref = stack[--sp];
  }
}

/* leaves whose closed cell overlaps leaf Q's cell inflated by rmax: every leaf
 * a sphere binned in Q can reach (its centre lies in Q's cell and its AABB
 * half-width is at most rmax; rounding is monotone) */
__device__ int leaf_nbr_one (int Q, int nleaves, const double *cell, double rmax, const int *roots,
                             const int *leafinst, const int *left, const int *right, const int *dim,
                             const double *coord, const int *off, int *list, int stride, int *flag)
{
  double lo[3], hi[3];
  for (int k = 0; k < 3; k++)
  {
    lo[k] = cell[(size_t) k*nleaves + Q] - rmax;
    hi[k] = cell[(size_t) (3+k)*nleaves + Q] + rmax;
  }
  int stack[64];
  int sp = 0, ref = roots[leafinst[Q]], n = 0;
  for (;;)
  {
    if (ref >= 0)
    {
      int d = dim[ref];
      double c = coord[ref];
      bool gl = lo[d] <= c, gr = hi[d] >= c;
      if (gl && gr)
      {
        if (sp < 64) stack[sp++] = right[ref]; else atomicAdd (flag, 1);
        ref = left[ref];
        continue;
      }
      if (gl) { ref = left[ref]; continue; }
      if (gr) { ref = right[ref]; continue; }
    }
    else
    {
      if (list && (!stride || n < stride)) list[stride ? (size_t) Q * stride + n : (size_t) off[Q] + n] = -ref-1;
      n++;
    }
    if (sp == 0) break;
    ref = stack[--sp];
  }
  return n;
}

/* leaves whose cell meets leaf Q's cell inflated by rmax: counted (cnt,
 * the longest count folded into cmax) and/or listed (at Q * stride, entries
 * past the stride dropped, or at off[Q]); only the leaves of instances with
 * a nonzero iflag when iflag is not null */
__global__ void k_leaf_nbr (int nleaves, const double *cell, double rmax, const int *roots, const int *leafinst,
                            const int *left, const int *right, const int *dim, const double *coord, int *cnt,
                            const int *off, int *list, int stride, int *flag, const int *iflag, int *cmax)
{
  int Q = blockIdx.x * blockDim.x + threadIdx.x;
  int n = 0;
  if (Q < nleaves && !(iflag && !iflag[leafinst[Q]]))
    n = leaf_nbr_one (Q, nleaves, cell, rmax, roots, leafinst, left, right, dim, coord, off, list, stride, flag);
  if (Q < nleaves && cmax && !(iflag && !iflag[leafinst[Q]])) cnt[Q] = n;
  if (cmax)
  {
    for (int o = 32; o >= 1; o >>= 1) n = max (n, __shfl_xor (n, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax (cmax, n);
  }
}

/* the previous build's lists of the instances a partial build left
 * unchanged: leaf Q' of instance i was leaf Q' - (nbase[i] - obase[i]),
 * and so was each leaf of its list (same stride); the flagged instances'
 * leaves are listed afresh */
__global__ void k_nbr_shift (int nleaves, const int *leafinst, const int *iflag, const int *obase, const int *nbase,
                             int stride, const int *cnt0, const int *list0, int *cnt, int *list)
{
  /* one thread per list slot (leaf Q, entry k), so a wave moves whole rows */
  size_t q = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (size_t) nleaves * stride) return;
  int Q = (int) (q / (unsigned) stride), k = (int) (q - (size_t) Q * stride);
  int i = leafinst[Q];
  if (iflag[i]) return;
  int delta = nbase[i] - obase[i];
  int Qo = Q - delta;
  int c = cnt0[Qo];
  if (k == 0) cnt[Q] = c;
  if (k < c) list[(size_t) Q * stride + k] = list0[(size_t) Qo * stride + k] + delta;
}

/* first leaf of each instance (base[I] = nleaves): leaves are numbered in
 * node order and an instance's nodes are its sphere range */
__global__ void k_inst_base (int I, int n, const int *off, const int *lfirst, int nleaves, int *base)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > I) return;
  int o = i < I ? off[i] : n;
  base[i] = o < n ? lfirst[o] : nleaves;
}


__global__ void k_fill_int (int *a, int n, int v)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

inline dim3 grid (int n) { return dim3 ((n + PMX_TPB - 1) / PMX_TPB); }

} /* namespace */

int step_mtri (pmx_ctx &c)
{
  if (c.nmtri == 0 || c.E == 0) return PMX_OK;
  TreeDev &t = c.tree;
  int L = std::max (t.nleaves, 1);
  hipLaunchKernelGGL (k_set_ctl, grid (L), dim3 (PMX_TPB), 0, c.stream, (const int*) c.ctl.p, c.mtcnt.p, L, 0);
  hipLaunchKernelGGL (k_mtri_reach, grid (c.nmtri), dim3 (PMX_TPB), 0, c.stream, c.nmtri, (const int*) c.mtri.p, c.T,
                      (const double*) c.tri.p, (const int*) t.roots.p, (const int*) c.tri_inst.p, (const int*) t.left.p,
                      (const int*) t.right.p, (const int*) t.dim.p, (const double*) t.coord.p, c.mtcnt.p, c.mtlist.p,
                      c.mcap, c.dflags.p, c.ctl.p,
                      c.staggered ? (const int*) c.istop.p : (const int*) nullptr);
  PMX_HIP (hipGetLastError ());
  return PMX_OK;
}


int scan_exclusive (pmx_ctx &c, int *data, int n, int *total_host, const int *ctl)
{
  if (n <= 0) { if (total_host) *total_host = 0; return PMX_OK; }
  size_t need = (size_t) n / (PMX_TPB*4) + 64;
  if (need > c.tree.scanbuf.n && ctl) { pmx_set_error ("scan scratch not prepared"); return PMX_ENOMEM; }
  if (c.tree.scanbuf.alloc (need) != hipSuccess) { pmx_set_error ("out of device memory (scan)"); return PMX_ENOMEM; }
  int last = 0;
  if (total_host)
    PMX_HIP (hipMemcpyAsync (&last, data + n - 1, sizeof (int), hipMemcpyDeviceToHost, c.stream));
  int rc = scan_rec (c, data, data, n, c.tree.scanbuf.p, c.tree.scanbuf.n, ctl);
  if (rc) return rc;
  if (total_host)
  {
    int lastex = 0;
    PMX_HIP (hipMemcpyAsync (&lastex, data + n - 1, sizeof (int), hipMemcpyDeviceToHost, c.stream));
    PMX_HIP (hipStreamSynchronize (c.stream));
    *total_host = lastex + last;
  }
  return PMX_OK;
}

/* exclusive scan of src into dst inside a step (scratch sized by
 * step_prepare; no allocation, so it can be captured) */
int scan_exclusive_to (pmx_ctx &c, const int *src, int *dst, int n, const int *ctl)
{
  if (n <= 0) return PMX_OK;
  size_t need = (size_t) n / (PMX_TPB*4) + 64;
  if (need > c.tree.scanbuf.n) { pmx_set_error ("scan scratch not prepared"); return PMX_ENOMEM; }
  return scan_rec (c, src, dst, n, c.tree.scanbuf.p, c.tree.scanbuf.n, ctl);
}

int tree_build (pmx_ctx &c, bool all)
{
  TreeDev &t = c.tree;
  int n = c.E, I = c.I;
  t.n = n;
  c.bin_hint = false;
  if (n == 0) { t.nleaves = 0; t.root = -1; return PMX_OK; }
  /* the tree of an instance is a function of the centres it was last built
   * from, so a rebuild forced by one instance leaves the others unchanged */
  /* all: every instance, except resumed ones, which brought their snapshots
   * (snap_sel: isel flags the instances to take) */
  if (!c.snap_sel && c.bcen.alloc (3*(size_t) n) != hipSuccess) { pmx_set_error ("out of device memory (tree)"); return PMX_ENOMEM; }
  hipLaunchKernelGGL (k_snap, grid (n), dim3 (PMX_TPB), 0, c.stream, n, (const double*) c.center.p, c.bcen.p,
                      (const int*) c.ell_inst.p,
                      all ? (c.snap_sel ? (const int*) c.isel.p : (const int*) nullptr) : (const int*) c.ovf.p);
  PMX_HIP (hipGetLastError ());
  const double *cx = c.bcen.p, *cy = c.bcen.p + n, *cz = c.bcen.p + 2*(size_t) n;
  int nbext = std::min (1024, (n + PMX_TPB - 1) / PMX_TPB);
  int nbsort = (n + PMX_TPB - 1) / PMX_TPB;
  if (t.ext.alloc (std::max (6*(size_t) I, 6 + 6*(size_t) nbext)) || t.code.alloc (n) || t.code2.alloc (n) ||
      t.order.alloc (n) || t.order2.alloc (n) || t.hist.alloc (std::max ((size_t) RADIX * nbsort + 1, 2*(size_t) n + 2)) ||
      t.first.alloc (n) || t.size.alloc (n) || t.split.alloc (n) || t.dim.alloc (n) || t.parent.alloc (n) ||
      t.coord.alloc (n) || t.left.alloc (n) || t.right.alloc (n) || t.nleafkids.alloc (n + 1) ||
      t.lfirst.alloc (n + 1) || t.roots.alloc (I) || t.cmin.alloc (3*(size_t) n) || t.cmkey.alloc (3*(size_t) n))
  { pmx_set_error ("out of device memory (tree)"); return PMX_ENOMEM; }

  /* A rebuild forced by some instances (ovf) re-sorts only their spheres
   * and rebuilds only their nodes: every other instance's sorted codes,
   * nodes, leaf-children counts and split coordinates are functions of its
   * unchanged snapshot and stay as the last build left them (node and
   * sorted-sphere ranges are the instance's sphere range).  The leaf
   * numbering, cells and per-leaf lists below are redone for all.  A full
   * build when most of the spheres are flagged. */
  int m = n;
  const int *sel = nullptr;
  std::vector<int> list;
  t.last_partial = false;
  if (!all && t.ready && I > 1)
  {
    std::vector<int> flag (I);
    PMX_HIP (hipMemcpyAsync (flag.data (), c.ovf.p, sizeof (int) * I, hipMemcpyDeviceToHost, c.stream));
    PMX_HIP (hipStreamSynchronize (c.stream));
    size_t ns = 0;
    for (int i = 0; i < I; i++) if (flag[i]) ns += (size_t) (c.off[i+1].ell - c.off[i].ell);
    if (4 * ns <= 3 * (size_t) n)
    {
      list.reserve (ns);
      for (int i = 0; i < I; i++)
        if (flag[i]) for (int e = c.off[i].ell; e < c.off[i+1].ell; e++) list.push_back (e);
      m = (int) ns;
      if (t.sel.alloc (std::max (m, 1)) || t.scode.alloc (std::max (m, 1)) || t.sorder.alloc (std::max (m, 1)))
      { pmx_set_error ("out of device memory (tree)"); return PMX_ENOMEM; }
      if (m > 0) PMX_HIP (hipMemcpyAsync (t.sel.p, list.data (), sizeof (int) * m, hipMemcpyHostToDevice, c.stream));
      PMX_HIP (hipStreamSynchronize (c.stream));   /* list is freed when this function returns */
      sel = t.sel.p;
      t.last_partial = true;
    }
  }

  if (m > 0)
  {
    /* extents per instance (partition.ispc:29-54, 669-684) */
    if (I == 1)
    {
      hipLaunchKernelGGL (k_extents_partial, dim3 (nbext), dim3 (PMX_TPB), 0, c.stream, n, cx, cy, cz, t.ext.p + 6);
      hipLaunchKernelGGL (k_extents_final, dim3 (1), dim3 (PMX_TPB), 0, c.stream, nbext, t.ext.p);
    }
    else
      hipLaunchKernelGGL (k_extents_seg, dim3 (I), dim3 (PMX_TPB), 0, c.stream, (const int*) c.ell_off.p, cx, cy, cz,
                          t.ext.p, sel ? (const int*) c.ovf.p : (const int*) nullptr);
    /* codes of the build's spheres: in place (full build) or packed in
     * sel's order (partial build), sorted, then placed at sel */
    uint64_t *k0 = sel ? t.code2.p : t.code.p, *k1 = sel ? t.scode.p : t.code2.p;
    int *v0 = sel ? t.order2.p : t.order.p, *v1 = sel ? t.sorder.p : t.order2.p;
    hipLaunchKernelGGL (k_morton, grid (m), dim3 (PMX_TPB), 0, c.stream, m, cx, cy, cz, (const double*) t.ext.p,
                        (const int*) c.ell_inst.p, k0, v0, sel);
    PMX_HIP (hipGetLastError ());

    /* stable LSD radix sort of (instance, code) */
    int bits = 30;
    while ((1ll << (bits - 30)) < I) bits++;
    int nbs = (m + PMX_TPB - 1) / PMX_TPB;
    for (int shift = 0; shift < bits; shift += RADIX_BITS)
    {
      hipLaunchKernelGGL (k_radix_hist, dim3 (nbs), dim3 (PMX_TPB), 0, c.stream, m, (const uint64_t*) k0, shift,
                          t.hist.p, nbs);
      PMX_HIP (hipGetLastError ());
      int rc = scan_exclusive (c, t.hist.p, RADIX * nbs, nullptr);
      if (rc) return rc;
      hipLaunchKernelGGL (k_radix_scatter, dim3 (nbs), dim3 (PMX_TPB), 0, c.stream, m, (const uint64_t*) k0,
                          (const int*) v0, shift, (const int*) t.hist.p, nbs, k1, v1);
      PMX_HIP (hipGetLastError ());
      std::swap (k0, k1);
      std::swap (v0, v1);
    }
    if (sel)
      hipLaunchKernelGGL (k_place, grid (m), dim3 (PMX_TPB), 0, c.stream, m, sel, (const uint64_t*) k0,
                          (const int*) v0, t.code.p, t.order.p);
    else if (k0 != t.code.p)
    {
      PMX_HIP (hipMemcpyAsync (t.code.p, k0, sizeof (uint64_t) * n, hipMemcpyDeviceToDevice, c.stream));
      PMX_HIP (hipMemcpyAsync (t.order.p, v0, sizeof (int) * n, hipMemcpyDeviceToDevice, c.stream));
    }
    PMX_HIP (hipGetLastError ());
    const uint64_t *key = t.code.p;
    const int *order = t.order.p;

    /* Karras nodes per instance range (partition.ispc:311-394) */
    hipLaunchKernelGGL (k_karras, grid (m), dim3 (PMX_TPB), 0, c.stream, m, key, (const int*) c.ell_off.p, t.first.p,
                        t.size.p, t.split.p, t.dim.p, sel);
    PMX_HIP (hipGetLastError ());
    hipLaunchKernelGGL (k_children, grid (m), dim3 (PMX_TPB), 0, c.stream, m, key, (const int*) c.ell_off.p,
                        (const int*) t.first.p, (const int*) t.size.p, (const int*) t.split.p, (const int*) t.dim.p,
                        t.nleafkids.p, t.parent.p, sel);
    PMX_HIP (hipGetLastError ());

    /* split coordinates of the branch nodes (mincoord, partition.ispc:297-308) */
    if (sel)
      hipLaunchKernelGGL (k_cm_init, grid (m), dim3 (PMX_TPB), 0, c.stream, m, n, sel, t.cmkey.p);
    else
      PMX_HIP (hipMemsetAsync (t.cmkey.p, 0xFF, sizeof (unsigned long long) * 3 * (size_t) n, c.stream));
    hipLaunchKernelGGL (k_minfold, grid (2*m), dim3 (PMX_TPB), 0, c.stream, m, n, (const int*) t.first.p,
                        (const int*) t.size.p, (const int*) t.split.p, (const int*) t.dim.p, (const int*) t.parent.p,
                        order, cx, cy, cz, t.cmkey.p, t.cmin.p, sel);
    hipLaunchKernelGGL (k_coord_set, grid (m), dim3 (PMX_TPB), 0, c.stream, m, n, (const int*) t.first.p,
                        (const int*) t.size.p, (const int*) t.split.p, (const int*) t.dim.p,
                        (const unsigned long long*) t.cmkey.p, (const double*) t.cmin.p, t.coord.p, sel);
    PMX_HIP (hipGetLastError ());
  }
  const uint64_t *key = t.code.p;
  /* an instance without spheres keeps PMX_NO_ROOT */
  hipLaunchKernelGGL (k_set_i, grid (I), dim3 (PMX_TPB), 0, c.stream, t.roots.p, I, PMX_NO_ROOT);

  /* leaf numbering and reach cells */
  int nleaves = 0;
  if (t.ibase.alloc (I + 1) || t.ibase2.alloc (I + 1)) { pmx_set_error ("out of device memory (tree)"); return PMX_ENOMEM; }
  if (sel)   /* each instance's first leaf before this build (its lists are shifted by the change) */
    hipLaunchKernelGGL (k_inst_base, grid (I + 1), dim3 (PMX_TPB), 0, c.stream, I, n, (const int*) c.ell_off.p,
                        (const int*) t.lfirst.p, t.nleaves, t.ibase2.p);
  PMX_HIP (hipMemcpyAsync (t.lfirst.p, t.nleafkids.p, sizeof (int) * n, hipMemcpyDeviceToDevice, c.stream));
  int rc = scan_exclusive (c, t.lfirst.p, n, &nleaves);
  if (rc) return rc;
  hipLaunchKernelGGL (k_inst_base, grid (I + 1), dim3 (PMX_TPB), 0, c.stream, I, n, (const int*) c.ell_off.p,
                      (const int*) t.lfirst.p, nleaves, t.ibase.p);
  t.nleaves = nleaves;
  if (t.leafpar.alloc (nleaves) || t.leafinst.alloc (nleaves) || t.cell.alloc (6*(size_t) nleaves))
  { pmx_set_error ("out of device memory (tree)"); return PMX_ENOMEM; }
  hipLaunchKernelGGL (k_assign, grid (n), dim3 (PMX_TPB), 0, c.stream, n, key, (const int*) c.ell_off.p,
                      (const int*) t.first.p, (const int*) t.size.p, (const int*) t.split.p, (const int*) t.dim.p,
                      (const int*) t.lfirst.p, t.left.p, t.right.p, t.leafpar.p, t.leafinst.p, t.roots.p);
  hipLaunchKernelGGL (k_cells, grid (nleaves), dim3 (PMX_TPB), 0, c.stream, nleaves, (const int*) t.leafpar.p,
                      (const int*) t.parent.p, (const int*) t.dim.p, (const double*) t.coord.p, t.cell.p);
  PMX_HIP (hipGetLastError ());
  PMX_HIP (hipMemcpyAsync (&t.root, t.roots.p, sizeof (int), hipMemcpyDeviceToHost, c.stream));
  PMX_HIP (hipStreamSynchronize (c.stream));
  t.ready = true;
  return PMX_OK;
}

int tree_nbrlists (pmx_ctx &c)
{
  TreeDev &t = c.tree;
  int L = std::max (t.nleaves, 1);
  bool shift = t.last_partial && t.nbr_ready && t.nbrstride > 0 && !c.nbr_csr;
  t.nbr_ready = false;
  if (shift)
  {
    std::swap (t.nbrcnt.p, t.nbrcnt2.p); std::swap (t.nbrcnt.n, t.nbrcnt2.n);
    std::swap (t.nbrlist.p, t.nbrlist2.p); std::swap (t.nbrlist.n, t.nbrlist2.n);
  }
  if (t.nbrcnt.alloc (L) || t.nbroff.alloc (L + 1) || t.dmax.alloc (1))
  { pmx_set_error ("out of device memory (neighbour lists)"); return PMX_ENOMEM; }
  t.nnbr = 0;
  if (c.E == 0 || t.nleaves == 0) return PMX_OK;
  int *flag = c.dflags.p + FLAG_TRILIST_OVERFLOW;
  auto count = [&] (const int *iflag, int *mx) -> int {
    hipLaunchKernelGGL (k_fill_int, dim3 (1), dim3 (64), 0, c.stream, t.dmax.p, 1, 0);
    hipLaunchKernelGGL (k_leaf_nbr, grid (t.nleaves), dim3 (PMX_TPB), 0, c.stream, t.nleaves, (const double*) t.cell.p,
                        c.rmax, (const int*) t.roots.p, (const int*) t.leafinst.p, (const int*) t.left.p,
                        (const int*) t.right.p, (const int*) t.dim.p, (const double*) t.coord.p, t.nbrcnt.p,
                        (const int*) nullptr, (int*) nullptr, 0, flag, iflag, t.dmax.p);
    PMX_HIP (hipGetLastError ());
    PMX_HIP (hipMemcpyAsync (mx, t.dmax.p, sizeof (int), hipMemcpyDeviceToHost, c.stream));
    PMX_HIP (hipStreamSynchronize (c.stream));
    return PMX_OK;
  };
  auto fill = [&] (const int *iflag) -> int {
    hipLaunchKernelGGL (k_leaf_nbr, grid (t.nleaves), dim3 (PMX_TPB), 0, c.stream, t.nleaves, (const double*) t.cell.p,
                        c.rmax, (const int*) t.roots.p, (const int*) t.leafinst.p, (const int*) t.left.p,
                        (const int*) t.right.p, (const int*) t.dim.p, (const double*) t.coord.p, t.nbrcnt.p,
                        (const int*) t.nbroff.p, t.nbrlist.p, t.nbrstride, flag, iflag, (int*) nullptr);
    PMX_HIP (hipGetLastError ());
    return PMX_OK;
  };
  int rc, mx = 0;
  if (shift && (size_t) t.nleaves * t.nbrstride * sizeof (int) <= ((size_t) 2 << 30))
  {
    /* a partial build: the flagged instances' leaves are counted and listed
     * at the previous stride in one traversal, and the unchanged instances'
     * lists move with their leaves; a list longer than the stride sends the
     * build to the full path below */
    t.nnbr = t.nleaves * t.nbrstride;
    if (t.nbrlist.alloc (std::max (t.nnbr, 1))) { pmx_set_error ("out of device memory (neighbour lists)"); return PMX_ENOMEM; }
    hipLaunchKernelGGL (k_fill_int, dim3 (1), dim3 (64), 0, c.stream, t.dmax.p, 1, 0);
    hipLaunchKernelGGL (k_leaf_nbr, grid (t.nleaves), dim3 (PMX_TPB), 0, c.stream, t.nleaves, (const double*) t.cell.p,
                        c.rmax, (const int*) t.roots.p, (const int*) t.leafinst.p, (const int*) t.left.p,
                        (const int*) t.right.p, (const int*) t.dim.p, (const double*) t.coord.p, t.nbrcnt.p,
                        (const int*) t.nbroff.p, t.nbrlist.p, t.nbrstride, flag, (const int*) c.ovf.p, t.dmax.p);
    PMX_HIP (hipGetLastError ());
    PMX_HIP (hipMemcpyAsync (&mx, t.dmax.p, sizeof (int), hipMemcpyDeviceToHost, c.stream));
    PMX_HIP (hipStreamSynchronize (c.stream));
    if (mx <= t.nbrstride)
    {
      hipLaunchKernelGGL (k_nbr_shift, grid (t.nnbr), dim3 (PMX_TPB), 0, c.stream, t.nleaves, (const int*) t.leafinst.p,
                          (const int*) c.ovf.p, (const int*) t.ibase2.p, (const int*) t.ibase.p, t.nbrstride,
                          (const int*) t.nbrcnt2.p, (const int*) t.nbrlist2.p, t.nbrcnt.p, t.nbrlist.p);
      PMX_HIP (hipGetLastError ());
      t.nbr_ready = true;
      return PMX_OK;
    }
  }
  /* all leaves: leaf Q's list at Q * stride (stride = the longest list), so
   * that a detection block finds its list without first reading an offset;
   * CSR when that layout would take more than 2 GiB (or when asked for: the
   * environment variable PMX_NEIGHBOUR_LISTS=csr) */
  if ((rc = count (nullptr, &mx))) return rc;
  mx = std::max (mx, 1);
  t.nbrstride = !c.nbr_csr && (size_t) t.nleaves * mx * sizeof (int) <= ((size_t) 2 << 30) ? mx : 0;
  if (t.nbrstride)
  {
    t.nnbr = t.nleaves * mx;
    hipLaunchKernelGGL (k_fill_int, grid (L + 1), dim3 (PMX_TPB), 0, c.stream, t.nbroff.p, L + 1, 0);
  }
  else
  {
    PMX_HIP (hipMemcpyAsync (t.nbroff.p, t.nbrcnt.p, sizeof (int) * t.nleaves, hipMemcpyDeviceToDevice, c.stream));
    int total = 0;
    if ((rc = scan_exclusive (c, t.nbroff.p, t.nleaves, &total))) return rc;
    t.nnbr = total;
  }
  if (t.nbrlist.alloc (std::max (t.nnbr, 1))) { pmx_set_error ("out of device memory (neighbour lists)"); return PMX_ENOMEM; }
  if ((rc = fill (nullptr))) return rc;
  t.nbr_ready = true;
  return PMX_OK;
}

int tree_trilists (pmx_ctx &c)
{
  TreeDev &t = c.tree;
  int L = std::max (t.nleaves, 1);
  if (t.tricnt.alloc (L) || t.trioff.alloc (L + 1)) { pmx_set_error ("out of device memory (trilists)"); return PMX_ENOMEM; }
  hipLaunchKernelGGL (k_fill_int, grid (L), dim3 (PMX_TPB), 0, c.stream, t.tricnt.p, L, 0);
  PMX_HIP (hipGetLastError ());
  t.ntrilist = 0;
  if (c.T == 0 || c.E == 0 || t.nleaves == 0)
  {
    hipLaunchKernelGGL (k_fill_int, grid (L), dim3 (PMX_TPB), 0, c.stream, t.trioff.p, L, 0);
    PMX_HIP (hipGetLastError ());
    return PMX_OK;
  }
  int *flag = c.dflags.p + FLAG_TRILIST_OVERFLOW;
  hipLaunchKernelGGL (k_tri_leaf, grid (t.nleaves), dim3 (PMX_TPB), 0, c.stream, t.nleaves, (const int*) t.leafinst.p,
                      (const int*) c.tri_off.p, (const int*) c.tmov.p, (const double*) c.tri.p, c.T,
                      (const double*) t.cell.p, t.tricnt.p, (const int*) nullptr, (int*) nullptr);
  hipLaunchKernelGGL (k_tri_reach, grid (c.T), dim3 (PMX_TPB), 0, c.stream, c.T, (const double*) c.tri.p,
                      (const int*) c.tmov.p, (const int*) t.roots.p, (const int*) c.tri_inst.p, (const int*) c.tri_off.p,
                      (const int*) t.left.p, (const int*) t.right.p,
                      (const int*) t.dim.p, (const double*) t.coord.p, t.tricnt.p, (const int*) nullptr, (int*) nullptr,
                      (int*) nullptr, flag);
  PMX_HIP (hipGetLastError ());
  PMX_HIP (hipMemcpyAsync (t.trioff.p, t.tricnt.p, sizeof (int) * t.nleaves, hipMemcpyDeviceToDevice, c.stream));
  int total = 0;
  int rc = scan_exclusive (c, t.trioff.p, t.nleaves, &total);
  if (rc) return rc;
  t.ntrilist = total;
  if (t.trilist.alloc (std::max (total, 1))) { pmx_set_error ("out of device memory (trilists)"); return PMX_ENOMEM; }
  dvec<int> fill;
  if (fill.alloc (L)) { pmx_set_error ("out of device memory (trilists)"); return PMX_ENOMEM; }
  hipLaunchKernelGGL (k_fill_int, grid (L), dim3 (PMX_TPB), 0, c.stream, fill.p, L, 0);
  hipLaunchKernelGGL (k_tri_leaf, grid (t.nleaves), dim3 (PMX_TPB), 0, c.stream, t.nleaves, (const int*) t.leafinst.p,
                      (const int*) c.tri_off.p, (const int*) c.tmov.p, (const double*) c.tri.p, c.T,
                      (const double*) t.cell.p, (int*) nullptr, (const int*) t.trioff.p, t.trilist.p);
  hipLaunchKernelGGL (k_tri_reach, grid (c.T), dim3 (PMX_TPB), 0, c.stream, c.T, (const double*) c.tri.p,
                      (const int*) c.tmov.p, (const int*) t.roots.p, (const int*) c.tri_inst.p, (const int*) c.tri_off.p,
                      (const int*) t.left.p, (const int*) t.right.p,
                      (const int*) t.dim.p, (const double*) t.coord.p, (int*) nullptr, (const int*) t.trioff.p,
                      t.trilist.p, fill.p, flag);
  PMX_HIP (hipGetLastError ());
  PMX_HIP (hipStreamSynchronize (c.stream));
  fill.release ();
  return PMX_OK;
}
