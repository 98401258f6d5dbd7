/*
 * pmx_internal.h -- device buffers and launcher declarations shared by the
 * translation units of libpmx.so.  Not part of the public ABI.
 */
#ifndef PMX_INTERNAL_H
#define PMX_INTERNAL_H

#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pmx.h"

#define PMX_TPB 256            /* threads per block of the elementwise kernels */
#define PMX_RETUNE_STEPS 256
#define PMX_LOOKAHEAD 2      /* groups of steps enqueued ahead of the one the host waits for */
#define PMX_PROF_EVERY 8     /* profiling mode 1 times every 8th step (timing events slow the kernels) */
#define PMX_GRAPH_STEPS 8    /* steps per graph launch where steps are short (a launch costs ~20 us) */
#define PMX_RING 4           /* slots of the pinned flag ring (> PMX_LOOKAHEAD) */   /* steps of a context before the second autotuner pass */
#define PMX_TRI_FLAG 0x80000000u  /* row key flag of a triangle partner */
#define PMX_NO_ROOT INT_MIN       /* tree root of an instance without spheres */

/* device control block of a dem call: abort word (a leaf or contact
 * capacity overflow in the current step; every later kernel of the step and
 * of the steps enqueued behind it is a no-op; it holds the aborted step's
 * sequence number + 1), completed steps (indexes the per-step tables and the
 * contact-row ping-pong), bin hint validity (leafof holds leaves of the
 * current tree), step sequence number (counts every enqueued step, aborted
 * or not: k_bin tells an abort raised by its own launch, where it keeps
 * binning so that every overflowing instance is flagged, from one raised by
 * an earlier step) */
enum { CTL_ABORT = 0, CTL_KSTEP, CTL_HINT, CTL_SEQ, CTL_REBIN, CTL_COUNT = 8 };
/* CTL_REBIN: the binning for the next step, done at the end of this step's
 * dynamics (k_dynamics with fused binning), overflowed a leaf: this step
 * completes, and k_step_end aborts the next one, which the host repeats
 * after re-partitioning */

/* device error/overflow flags, one int each.  Leaf, row, hit, reverse-hit
 * and moving-triangle list overflows abort the step (the host re-partitions
 * or grows the capacity and repeats it); a traversal stack overflow in a
 * tree build (FLAG_TRILIST_OVERFLOW) is an error */
enum { FLAG_LEAF_OVERFLOW = 0, FLAG_ROW_OVERFLOW, FLAG_HIT_OVERFLOW, FLAG_EXTRA_OVERFLOW,
       FLAG_TRILIST_OVERFLOW, FLAG_MAXROW, FLAG_BAD_KIND, FLAG_MTRI_OVERFLOW, FLAG_TREV_OVERFLOW,
       FLAG_JOINT_PIVOT, FLAG_COUNT };

/* abort the current step: the step is repeated after the host has grown the
 * capacity that overflowed (or re-partitioned), from unchanged state */
__device__ __forceinline__ void abort_step (int *ctl)
{
  atomicExch (&ctl[CTL_ABORT], ctl[CTL_SEQ] + 1);
}

void pmx_set_error (const std::string &msg);

#define PMX_HIP(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) { \
  pmx_set_error (std::string ("HIP error: ") + hipGetErrorString (e_) + " at " + __FILE__ + ":" + std::to_string (__LINE__)); \
  return PMX_EDEVICE; } } while (0)

/* device array owned by its holder (freed with it) */
template <class T> struct dvec
{
  T *p = nullptr;
  size_t n = 0;
  dvec () = default;
  dvec (const dvec &) = delete;
  dvec &operator= (const dvec &) = delete;
  ~dvec () { release (); }
  hipError_t alloc (size_t count)
  {
    if (count <= n && p) return hipSuccess;
    if (p) { (void) hipFree (p); p = nullptr; n = 0; }
    if (count == 0) return hipSuccess;
    hipError_t e = hipMalloc ((void**) &p, count * sizeof (T));
    if (e == hipSuccess) n = count;
    return e;
  }
  void release () { if (p) (void) hipFree (p); p = nullptr; n = 0; }
};

/* partitioning tree on the device (the frozen k-d tree of partition.ispc) */
struct TreeDev
{
  int n = 0;                 /* points the tree was built over */
  int nleaves = 0;
  int root = -1;             /* node reference of the root of instance 0 */
  dvec<int> roots;           /* [I] node reference of each instance's root */
  dvec<int> leafinst;        /* [nleaves] instance of each leaf */
  /* scratch of the build */
  dvec<double> ext;          /* [6][I] extents (single instance: partials follow) */
  dvec<uint64_t> code, code2; /* (instance << 30) | Morton code */
  dvec<int> order, order2;
  dvec<int> hist;            /* radix histogram / scan scratch */
  dvec<int> scanbuf;         /* block sums for scans */
  /* internal nodes (n-1) */
  dvec<int> first, size, split, dim, parent;
  dvec<double> coord;
  dvec<int> left, right;     /* child refs: >= 0 branch node, < 0 leaf -(L+1) */
  dvec<int> nleafkids;       /* per internal node: leaf children */
  dvec<int> lfirst;          /* per internal node: its first leaf id (scan of nleafkids) */
  dvec<int> sel;             /* spheres (and nodes) of the instances a partial build redoes */
  dvec<uint64_t> scode;      /* sort scratch of a partial build */
  dvec<int> sorder;
  bool ready = false;        /* a full build has been made: partial builds may follow */
  bool last_partial = false; /* the last build redid only the overflowing instances */
  dvec<double> cmin;         /* [3][n] minima of a node's terminal right child */
  dvec<unsigned long long> cmkey;  /* [3][n] ordered keys of the minima of each node's range */
  /* leaves */
  dvec<int> leafpar;         /* [nleaves] parent branch node*2 + side (or -1 for a root leaf) */
  dvec<double> cell;         /* [6][nleaves]: lower bounds 0-2, upper bounds 3-5 */
  dvec<int> tricnt, trioff, trilist;  /* leaf -> triangles whose AABB reaches it */
  int ntrilist = 0;
  dvec<int> nbrcnt, nbroff, nbrlist;  /* leaf -> leaves overlapping its cell inflated by rmax */
  int nbrstride = 0;                  /* list Q at Q * nbrstride, or CSR by nbroff when 0 */
  dvec<int> nbrcnt2, nbrlist2;        /* the previous build's lists while a partial build shifts them */
  dvec<int> ibase, ibase2;            /* [I+1] first leaf of each instance: this build, the previous one */
  dvec<int> dmax;                     /* device maximum of the list lengths */
  bool nbr_ready = false;             /* nbrcnt/nbrlist hold the lists of the current leaves */
  int nnbr = 0;
};

/* offsets of one instance's items in the concatenated arrays of a batch */
struct InstOff
{
  int par = 0, ell = 0, tri = 0, nod = 0, fac = 0, spr = 0, scrv = 0, dcrv = 0, ucrv = 0, pair = 0, obs = 0, prs = 0, pcol = 0;
};

/* kernels timed by the profiler */
enum { K_BIN = 0, K_SCATTER, K_DETECT, K_FORCE_INIT, K_ROWS, K_DYNAMICS, K_SHAPES, K_TREE, K_TRILIST, K_SPRINGS,
       K_SPRING_ACC, K_COUNT };

/* HIP-event timing of the launches on the context stream; events are pooled
 * and resolved at the per-step synchronisation */
struct Prof
{
  int mode = 0;                /* 0 off; 1 every PMX_PROF_EVERY-th step replays a copy of the step graph
                                  with an event pair per kernel as record nodes (fresh events swapped in
                                  before each launch), the others the plain graph; 2 every step enqueued
                                  eagerly with an event pair per kernel */
  bool capturing = false;      /* a step is being captured: pairs become event record nodes */
  hipError_t err = hipSuccess; /* a capture-time query failed */
  int step = -1;               /* step being enqueued (-1: launches outside the step loop) */
  int win0 = -1, win1 = -1;    /* pmx_set_window: only these steps of the call are timed (-1: all) */
  bool in_window (int s) const { return win0 < 0 || (s >= win0 && s <= win1); }
  struct Pend { int kid; hipEvent_t a, b; double bytes; int step; };
  /* a profiled kernel scope of the captured step: the chain's tail node
   * before and after it, and the event record nodes put around it */
  struct Cap { int kid; double bytes; hipGraphNode_t before, after, na, nb; hipEvent_t a, b; };
  bool eager = false;          /* mode 1 could not place its nodes: steps run as in mode 2 */
  std::vector<hipEvent_t> pool;
  std::vector<Pend> pend;
  std::vector<Cap> cap;
  double ms[K_COUNT] = {0}, bytes[K_COUNT] = {0};
  long long n[K_COUNT] = {0};
  hipEvent_t get ()
  {
    if (!pool.empty ()) { hipEvent_t e = pool.back (); pool.pop_back (); return e; }
    /* no system-scope fence: the event marks the kernel's completion on the
     * device timeline without writing the caches back for the host (a
     * writeback after each timed kernel would be counted in its time) */
    hipEvent_t e = nullptr;
    (void) hipEventCreateWithFlags (&e, hipEventDisableSystemFence);
    return e;
  }
};

struct pmx_ctx
{
  pmx_options opt {};
  hipStream_t stream = nullptr;
  int device = 0;
  int N = 0, E = 0, T = 0, P = 0, NOD = 0, FAC = 0;
  /* instances stepped together (1 for a single model); each owns a
   * contiguous range of particles, spheres and triangles and its own tree */
  int I = 1;
  std::vector<InstOff> off;    /* [I+1] prefix offsets of each item family */
  dvec<int> par_inst, ell_inst, tri_inst, ell_off, pair_off, tri_off;
  dvec<double> bcen;           /* [3][E] centres the current tree was built from */
  dvec<int> ovf;               /* [I] leaf overflow of the instance in the aborted step */
  /* per-instance step counts of a dem call whose instances run different
   * durations: instance i is frozen (every kernel skips its items) once
   * ctl[KSTEP] reaches istop[i]; null pointer when all run the same count */
  dvec<int> istop, obs_inst;
  bool staggered = false;      /* the last dem call left instances at different times */
  std::vector<int> icur;       /* row buffer holding each instance's rows */
  std::vector<double> icurtime, icurstep;
  std::vector<int> istepnum;

  /* particles */
  dvec<double> angular, linear, rotation, position, inertia, inverse, mass, invm, force, torque;
  dvec<int> flags;
  /* ellipsoids */
  dvec<double> center, radii;
  dvec<int> part, ellcol;
  /* triangles */
  dvec<double> tri;
  dvec<int> tricol, triobs;
  /* moving obstacles: point, angular and linear velocity [3*O]; the
   * triangles they move (mtri) with their obstacle (mtobs); per-step reach
   * lists leaf -> moving triangles (mtlist [nleaves][mcap]) */
  int O = 0, nmtri = 0, mcap = 32;
  dvec<double> obspnt, obsang, obslin, obstab;
  dvec<int> mtri, mtobs, tmov, mtcnt, mtlist;
  /* prescribed motion: one device entry per particle with the table columns
   * (of [step][PRS][6]) its force and velocity come from */
  int PRS = 0, PRSC = 0, nprs = 0;   /* entries, history columns, device entries */
  dvec<int> prs_part, prs_lf, prs_lv, prs_af, prs_av;
  dvec<double> prstab;
  /* pairing */
  dvec<int> pairs, ikind;
  dvec<double> iparam;
  /* per-step tables */
  dvec<double> gtab, dtab;
  /* meshed particles: nodes and faces */
  dvec<double> nodes;
  dvec<int> nodpart, facnod, factri;
  /* springs */
  int S = 0, sprcrv = 0, dashcrv = 0, uncrv = 0;
  dvec<int> sp_type, sp_part, sp_spridx, sp_dashidx, sp_unidx, sp_unspring, sp_flg;
  dvec<double> sp_pnt, sp_dir, sp_yield, sp_sdsp, sp_stroke, sp_frc, sp_stroke0, sp_fric, sp_kskn;
  dvec<double> sp_curve, sp_dash, sp_unload, sp_out;
  dvec<int> sp_evoff, sp_evrun, sp_runoff, sp_runspr;
  dvec<int> sp_offset;         /* [S] load curve of each spring's stroke offset, or -1 */
  bool sp_offsets = false;     /* some spring has a stroke offset: k_springs reads ttab */
  /* load curves (spring stroke offsets, UNSPRING unloading) */
  int LC = 0, lccrv = 0;
  dvec<int> lcidx;
  dvec<double> lcurve;         /* [2][lccrv] */
  dvec<double> ttab;           /* [step] curtime of each step of the call (stroke offsets) */
  /* torsion springs (sorted by part1, part2) */
  int TS = 0, krpycrv = 0, drpycrv = 0;
  dvec<int> ts_part, ts_kidx, ts_didx, ts_cone;
  dvec<double> ts_z0, ts_x0, ts_krpy, ts_drpy, ts_z1, ts_x1, ts_rpy, ts_tot, ts_spr;
  dvec<double> ts_out;         /* [6][TS] torque, adaptive terms |k| R(d,d) of each spring */
  dvec<int> ts_evoff, ts_evrun, ts_runoff, ts_runspr;   /* flush order, as the springs' */
  /* UNSPRING conditions */
  int US = 0;
  dvec<int> un_tidx, un_t, un_midx, un_m, un_aidx, un_a, un_ent, un_op, un_abs, un_nsteps, un_nfreq, un_action;
  dvec<double> un_lim;
  /* restraints: runs of one particle's entries in entry order */
  int RS = 0, RR = 0;
  dvec<int> rs_part, rs_runoff;
  dvec<double> rs_lin, rs_ang;
  /* ball joints: independent sets solved per step (joints.cpp:83-567) */
  int J = 0, JS = 0, JP = 0;   /* joints, independent sets, particles with joints */
  dvec<int> jt_part, jt_set, jt_pos, jt_adjoff, jt_adj, jt_soff, jt_sk, jt_ppart, jt_poff, jt_plist, jt_iw;
  dvec<double> jt_point, jt_reac, jt_mat, jt_rhs, jt_work;
  dvec<long long> jt_moff, jt_woff;   /* per set: its dense matrix, its skyline work space */

  /* broad phase per step */
  TreeDev tree;
  dvec<int> leafof, slot, lcount, loff, lmem, lpart;   /* lmem, lpart: leaf-ordered sphere, its particle */
  dvec<double> lpos;           /* [4][records] leaf-ordered x, y, z, r */
  dvec<double> lbox;           /* [6][nleaves] this step's box of each leaf's records (k_leafbox) */
  int lstride = 0;             /* leaf L's records at L * lstride (PMX_LSIZE slots per leaf, written by
                                  k_bin), or 0: packed at loff[L] (scan of the counts, then k_scatter) */
  /* contact rows (ping-pong) */
  int cap = 32, hcap = 32, xcap = 16;
  int cur = 0;
  dvec<uint32_t> rowkey[2];    /* [cap][E] */
  dvec<double> rowst[2];       /* [3][cap][E] tangential state (FP64 path) */
  dvec<float> rowst32[2];      /* [3][cap][E] tangential state (mixed precision) */
  dvec<int> rowcnt[2];         /* [E] */
  dvec<uint32_t> hits;         /* [hcap][E] own hits */
  dvec<int> hitcnt;            /* [E] */
  dvec<int> xbuf;              /* [xcap][E] reverse hits */
  dvec<int> xcnt;              /* [E] */
  dvec<int> rbusy, rnbusy;     /* per block of k_rows_idle: its busy spheres, their count */
  dvec<int> fin;               /* [I] last step of each instance finishing at a report, or -1 */
  dvec<int> dflags;            /* [FLAG_COUNT] */
  dvec<int> ctl;               /* [CTL_COUNT] */
  int cur0 = 0;                /* row buffer holding the rows at step 0 of the call */
  hipGraph_t graph = nullptr;  /* the captured step */
  hipGraphExec_t gexec = nullptr;
  hipGraph_t pgraph = nullptr; /* profiling mode 1: the step with event record nodes (kept: they name
                                  the nodes of pgexec), replayed every PMX_PROF_EVERY steps */
  hipGraphExec_t pgexec = nullptr;
  hipGraphExec_t gexec_multi = nullptr;   /* PMX_GRAPH_STEPS steps in one graph (groups of that many
                                             steps or more, profiling off): one launch per 8 steps */
  bool graph_valid = false;
  /* time histories (single model): specs, the sampled row of each step of
   * the current call, the sampled values [row][history] */
  int nhis = 0;
  bool his_active = false;     /* the current call samples (k_history is in the step) */
  dvec<int> his_ent, his_off, his_list, hisrow;
  dvec<signed char> his_pt;
  dvec<double> his_pts, hisbuf;
  dvec<double> energy;         /* pmx_energy: per-block partial sums and the result */
  /* adaptive step and critical estimates (single model): per particle
   * kact[3], kmax, emax, krot[6] as [11][N]; per spring the adaptive terms
   * [17][S]; with opt.critical, the contact records of the last step taken
   * (point[3], |direction|[3], kcur, ecur as [8][cap][E]); the estimates
   * h, omega, ratio [3][max(N, S)]; the adaptive minimum as an ordered key */
  /* spheres in contact with triangles of meshed particles: each contact's
   * reaction (the sphere, -f, the point) in its triangle's slots
   * (trec_e [tcap][T], trec [6][tcap][T], tcnt [T]), summed per meshed
   * particle over its triangles (mp_toff/mp_tlist, ascending) by k_react */
  bool mesh_contact = false;
  int tcap = 8, nmp = 0;
  int nmnod = -1, nmfac = -1;  /* nodes and faces of the meshed particles in mp_part (mnod, mfac), or -1:
                                  all of them */
  dvec<int> mnod, mfac;
  dvec<int> tcnt, trec_e, mp_part, mp_toff, mp_tlist;
  dvec<double> trec;
  bool adp_on = false;         /* the step forms the adaptive sums (k_rows<R, true>) */
  bool crec_valid = false;     /* crec holds the contacts of the last step taken */
  bool crec_stale = false;     /* steps were taken since the upload without recording their contacts */
  dvec<double> adp, sp_adp, crec, crit;
  dvec<unsigned long long> hmin;
  dvec<int> sp_pofs, sp_plist; /* particle -> its springs (spring << 1 | side) in spring order */
  int *hring = nullptr;        /* pinned: per group of steps in flight, the control block and the flags */
  unsigned long long *hkey = nullptr;   /* pinned: the adaptive minimum of the step read back */
  int win[2] = {-1, -1};       /* pmx_set_window: steps timed on the device in the next dem call */
  hipEvent_t wev[2] = {nullptr, nullptr};
  hipEvent_t rev[PMX_RING] = {};   /* recorded after each group's copy into hring */
  int det_nw = 4;              /* waves per k_detect block (autotuned) */
  int rows_tpb = 1024;         /* threads per k_rows_idle block (spheres per busy list) */
  int tuned = 0;                /* autotuner passes done: at the first call, then once settled */
  long long stepped = 0;        /* steps taken by this context over all calls */
  dvec<int> parell;            /* [N] the particle's sphere, or -1 */
  dvec<int> nosph;             /* particles without a sphere */
  dvec<signed char> iso;       /* [N] bit 0: inertia is diag(j, j, j) with exact zeros; bit 1: the
                                  sphere's body offset is +0; bit 2: loaded by contacts and gravity
                                  only (no spring, no prescribed motion) */
  int niso = 0;
  dvec<int> dynlist;           /* with both inertia forms: the niso scalar-inertia particles, then the rest */
  int nzc = 0;                 /* particles whose sphere offset is +0 */
  int nfree = 0;               /* particles with flag bit 2 (loaded by contacts and gravity only) */
  bool nbr_csr = false;        /* neighbour lists always in CSR layout (PMX_NEIGHBOUR_LISTS=csr) */
  bool rec_packed = false;     /* leaf records always packed (PMX_LEAF_RECORDS=packed) */
  int nnosph = 0;
  bool bin_hint = false;       /* leafof holds leaves of the current tree */
  /* fused binning: k_dynamics bins the new centres for the next step (the
   * slotted leaf records of dem_core's steps); a separate k_bin runs only
   * where the records are not in place (need_bin: after a build, an aborted
   * step or a tuning run).  bin_until: the call's step count (the last
   * step bins nothing) */
  bool damped = true;          /* the call's damping table holds a nonzero entry (else k_dynamics skips it) */
  bool fuse_bin = false, need_bin = false;
  int bin_until = 0;
  /* a call whose instances did not all start at time 0 (instances resumed
   * from an earlier context beside new ones): isel [I] flags the instances
   * at time 0, the only ones the first-call stages (euler, shapes) update
   * when first_sel is set, and the only ones whose tree snapshot the call's
   * first build takes when snap_sel is set (the others bring theirs) */
  dvec<int> isel;
  bool first_sel = false, snap_sel = false;
  double rmax = 0.0;           /* largest sphere AABB half-width */

  double curtime = 0.0, curstep = 0.0;
  int stepnum = 0;
  bool uploaded = false;
  pmx_stats stats {};
  Prof prof;
};

/* the item's instance has completed its steps of the current dem call */
__device__ __forceinline__ bool frozen (const int *stop, int inst, const int *ctl)
{
  return stop && ctl[CTL_KSTEP] >= stop[inst];
}

/* bracket a launch with events when profiling; bytes = algorithmic HBM bytes */
struct ProfScope
{
  pmx_ctx &c;
  int kid;
  double bytes;
  hipEvent_t a = nullptr;
  hipGraphNode_t before = nullptr;
  bool on;
  ProfScope (pmx_ctx &c_, int kid_, double bytes_) : c (c_), kid (kid_), bytes (bytes_)
  {
    on = c.prof.mode && (c.prof.capturing || c.prof.in_window (c.prof.step));
    if (!on) return;
    if (c.prof.capturing) before = tail ();
    else { a = c.prof.get (); (void) hipEventRecord (a, c.stream); }
  }
  ~ProfScope ()
  {
    if (!on) return;
    if (c.prof.capturing)
    {
      hipGraphNode_t after = tail ();
      if (after && after != before) c.prof.cap.push_back ({kid, bytes, before, after, nullptr, nullptr, nullptr, nullptr});
      return;
    }
    if (!a) return;
    hipEvent_t b = c.prof.get ();
    (void) hipEventRecord (b, c.stream);
    c.prof.pend.push_back ({kid, a, b, bytes, c.prof.step});
  }
  /* the last node captured on the stream (a single-stream capture is a chain) */
  hipGraphNode_t tail ()
  {
    hipStreamCaptureStatus st;
    const hipGraphNode_t *deps = nullptr;
    size_t nd = 0;
    if (hipStreamGetCaptureInfo_v2 (c.stream, &st, nullptr, nullptr, &deps, &nd) != hipSuccess || nd > 1)
    { c.prof.err = hipErrorUnknown; return nullptr; }
    return nd ? deps[0] : nullptr;
  }
};

/* resolve the pending event pairs of steps before `upto` (all pairs when
 * upto is INT_MAX; their events must have completed); launches of steps at
 * or after `complete` (an aborted step and the steps enqueued behind it,
 * whose kernels returned at once) are not counted */
void prof_resolve (pmx_ctx &c, int complete = 0x7FFFFFFF, int upto = 0x7FFFFFFF);

/* ---- upload (pmx_api.hip) ---- */
int validate_model (const pmx_model *m);
InstOff model_counts (const pmx_model &m);
int upload_core (pmx_ctx *c, const pmx_model *m, const std::vector<InstOff> &off);

/* what a call may take beyond dem(): steps, each instance's step count
 * (in place of durations; instances may then be at different times, each
 * at time 0 or resumed at the call's step); stop_fn, asked after each group
 * of steps whether the call should pause there (paused_at: the steps taken,
 * the unfinished instances keep their time and take no stepnum) */
struct DemExtra
{
  const int *steps = nullptr;
  int (*stop_fn) (void*, int) = nullptr;
  void *stop_user = nullptr;
  int paused_at = -1;
};

int dem_core (pmx_ctx *c, const double *durations, double duration, double step, const pmx_forcing *fc, int nfc,
              void (*done_fn) (void*, int), void *user, double *elapsed, DemExtra *x = nullptr);
int alloc_rows (pmx_ctx &c);

/* an instance's contact rows and tree snapshot, carried from one context to
 * another (the rows in the instance's own numbering) */
struct ResumeRows
{
  double curtime = 0.0, curstep = 0.0;   /* the instance's time, step and dem() count */
  int stepnum = 0;
  int cap = 0, ne = 0, nt = 0;
  std::vector<uint32_t> key;        /* [cap][ne] */
  std::vector<unsigned char> st;    /* [3][cap][ne] of the context's state type */
  std::vector<int> cnt;             /* [ne] */
  std::vector<double> bcen;         /* [3][ne] */
};
int export_rows (pmx_ctx &c, int inst, ResumeRows &r);
/* into the row buffer of the context's next call, after pmx_upload_batch of
 * models at time 0: rs[i] sets instance i's time and rows; rs[i] null:
 * instance i starts at time 0 with no contact */
int import_rows (pmx_ctx &c, const std::vector<const ResumeRows*> &rs);

/* ---- instances (pmx_batch.hip) ---- */
/* host copy of a (possibly concatenated) model: owns the arrays the view points to */
struct Merged
{
  pmx_model m {};
  std::vector<std::vector<double>> dv;
  std::vector<std::vector<int>> iv;
  double *D (size_t n) { dv.emplace_back (std::max (n, (size_t) 1), 0.0); return dv.back ().data (); }
  int *Iv (size_t n) { iv.emplace_back (std::max (n, (size_t) 1), 0); return iv.back ().data (); }
};
/* concatenate models (validated one by one, index fields shifted) */
int merge_models (const pmx_model *ms, int count, Merged &g, std::vector<InstOff> &off);
/* append instance i's contact pairs (global numbering, or the instance's own
 * when local) from the row buffer holding its rows */
int contacts_range (pmx_ctx &c, int inst, bool local, std::vector<pmx_pair> &out);

/* ---- tree (pmx_tree.hip) ---- */
int tree_build (pmx_ctx &c, bool all);       /* all: snapshot every instance's centres; else only
                                                the instances flagged in c.ovf */
int tree_trilists (pmx_ctx &c);              /* leaf -> triangle lists */
int tree_nbrlists (pmx_ctx &c);              /* leaf -> neighbour leaf lists */
int step_mtri (pmx_ctx &c);                  /* leaf -> moving obstacle triangles, per step */

/* ---- scans (pmx_tree.hip) ---- */
int scan_exclusive (pmx_ctx &c, int *data, int n, int *total_host, const int *ctl = nullptr);
int scan_exclusive_to (pmx_ctx &c, const int *src, int *dst, int n, const int *ctl);

/* ---- step kernels (pmx_step.hip); step_* only enqueue work ---- */
int step_prepare (pmx_ctx &c);               /* per-tree buffers, after a (re)build */
int clear_leaf_counts (pmx_ctx &c);          /* zero the per-leaf counts of the binning */
int step_bin (pmx_ctx &c);
int step_detect (pmx_ctx &c);
int step_scatter (pmx_ctx &c);
int launch_detect (pmx_ctx &c, int nw);
int launch_rows (pmx_ctx &c, int tpb, double h);
int autotune (pmx_ctx &c, double h);          /* block shape of detection */
int step_rows (pmx_ctx &c, double h);
int step_dynamics (pmx_ctx &c, double step0, double step1);
int step_end (pmx_ctx &c);
int enqueue_step (pmx_ctx &c, double h);     /* one whole time step */
int enqueue_step_forces (pmx_ctx &c, double h);                  /* up to the forces (and the adaptive minimum) */
int enqueue_step_motion (pmx_ctx &c, double step0, double step1); /* dynamics to the end of the step */
/* ---- time histories (pmx_history.hip) ---- */
int history_set (pmx_ctx &c, const pmx_history *h);
int history_now (pmx_ctx &c, double *out);
int history_scratch_init (int device);        /* pmx_history_sample_noalloc's scratch address on a device */
int step_history (pmx_ctx &c);
int step_obstacles (pmx_ctx &c, const double *tab, int row, double h);
int step_prescribe_acc (pmx_ctx &c);           /* before dynamics */
int step_prescribe_vel (pmx_ctx &c);           /* after dynamics */
int step_materialize (pmx_ctx &c, const std::vector<int> &fin);   /* forces of finishing instances */
int row_totals (pmx_ctx &c, long long out[2]);   /* row entries and busy spheres of the current rows */
int step_shapes (pmx_ctx &c);
int step_euler (pmx_ctx &c, double h);
int step_invert (pmx_ctx &c);

/* ---- springs and meshes (pmx_spring.hip) ---- */
int springs_setup (pmx_ctx &c, const pmx_model *m);   /* per-instance ranges from c.off */
int springs_download (pmx_ctx &c, pmx_model *m);
int step_springs (pmx_ctx &c, double h);
int crit_springs (pmx_ctx &c);                 /* CRITICAL's spring part of the per-particle sums */
int crit_perspring (pmx_ctx &c, double *h, double *o, double *r);
int crit_contacts (pmx_ctx &c);                /* CRITICAL's contact part (from the recorded contacts) */
int step_react (pmx_ctx &c, bool critical);    /* meshed particles' contact reactions (forces, or CRITICAL's sums) */
int adp_begin (pmx_ctx &c);                    /* zero the per-particle sums; reset the adaptive minimum */
int adp_time_step (pmx_ctx &c, double *h, double *o, double *r);   /* per particle; minimum when h is null */
int step_mesh_shapes (pmx_ctx &c, bool gated = true);
int torsion_setup (pmx_ctx &c, const pmx_model *m);   /* torsion springs and UNSPRING conditions */
int torsion_download (pmx_ctx &c, pmx_model *m);
int step_torsion (pmx_ctx &c, double h);        /* torsion springs (after the springs) */
int step_unspring (pmx_ctx &c);                 /* UNSPRING tests (after the torsion springs) */

/* ---- restraints and ball joints (pmx_joint.hip) ---- */
int jt_validate (const pmx_model *m);          /* torsion springs, UNSPRING, load curves, restraints, joints */
int jt_setup (pmx_ctx &c, const pmx_model *m);
int jt_download (pmx_ctx &c, pmx_model *m);
int step_restrain_vel (pmx_ctx &c);            /* once per dem call, before the steps */
int step_joints (pmx_ctx &c, double step0, double step1);   /* ball joints, then restrained forces */
int step_times_table (pmx_ctx &c, const double *times, int n);   /* curtime of each step (stroke offsets) */

#endif
