/*
 * pmx_batch.hip -- many independent models stepped together.
 *
 * A batch is the concatenation of its instances' SoA arrays: every item
 * family (particles, spheres, triangles, nodes, faces, springs, curve points,
 * pairings) is laid out instance after instance, and every index an instance
 * stores (ellipsoid -> particle, triangle -> particle, node -> particle,
 * face -> node and triangle, spring -> particle and curve point) is shifted by
 * the offset of the family it points into.  The step kernels then run once
 * over the whole batch; what keeps the instances apart is
 *   - a partitioning tree per instance (segmented build), so no contact is
 *     ever formed across instances and each instance's tree is exactly the
 *     tree it would have alone (rebuilt only when that instance overflows);
 *   - per-instance pairing tables, gravity and damping tables;
 *   - per-instance spring flush sequences.
 * Contact keys of an instance are global ellipsoid/triangle numbers, shifted
 * by the same offsets, so their order is the order the instance has alone and
 * every per-body sum adds the same terms in the same order: each instance's
 * result is bitwise the result of stepping it alone.
 */
#include "pmx_internal.h"
#include <algorithm>
#include <cstring>

namespace {

int fail (int code, const std::string &msg) { pmx_set_error (msg); return code; }

/* component-major k-component arrays: dst[c*total + off_i + j] = src_i[c*n_i + j] (+ shift) */
template <class T, class Get, class Cnt, class Off, class Map>
void cat (T *dst, int k, int count, size_t total, Get get, Cnt cnt, Off off, Map map)
{
  for (int i = 0; i < count; i++)
  {
    const T *src = get (i);
    size_t n = cnt (i), o = off (i);
    if (!src || n == 0) continue;
    for (int c = 0; c < k; c++)
      for (size_t j = 0; j < n; j++) dst[c*total + o + j] = map (src[c*n + j], i);
  }
}

template <class T, class Get, class Cnt, class Off>
void split (const T *src, int k, int count, size_t total, Get get, Cnt cnt, Off off)
{
  for (int i = 0; i < count; i++)
  {
    T *dst = get (i);
    size_t n = cnt (i), o = off (i);
    if (!dst || n == 0) continue;
    for (int c = 0; c < k; c++)
      std::memcpy (dst + c*n, src + c*total + o, n * sizeof (T));
  }
}

} /* namespace */

int contacts_range (pmx_ctx &c, int inst, bool local, std::vector<pmx_pair> &out)
{
  int e0 = c.off[inst].ell, ne = c.off[inst+1].ell - e0, t0 = c.off[inst].tri;
  if (ne == 0) return PMX_OK;
  size_t E = c.E, cap = c.cap;
  int par = c.icur[inst];
  std::vector<uint32_t> key (cap * ne);
  std::vector<double> st (3 * cap * ne);
  std::vector<int> cnt (ne), part (ne);
  /* every copy on the context's stream: a copy on the legacy stream would
   * have to wait for the streams of other contexts, which is refused while
   * one of them captures a step */
  PMX_HIP (hipMemcpy2DAsync (key.data (), ne * sizeof (uint32_t), c.rowkey[par].p + e0, E * sizeof (uint32_t),
                             ne * sizeof (uint32_t), cap, hipMemcpyDeviceToHost, c.stream));
  if (c.opt.precision == PMX_MIXED)
  {
    std::vector<float> s32 (st.size ());
    PMX_HIP (hipMemcpy2DAsync (s32.data (), ne * sizeof (float), c.rowst32[par].p + e0, E * sizeof (float),
                               ne * sizeof (float), 3 * cap, hipMemcpyDeviceToHost, c.stream));
    PMX_HIP (hipStreamSynchronize (c.stream));
    for (size_t q = 0; q < st.size (); q++) st[q] = s32[q];
  }
  else
    PMX_HIP (hipMemcpy2DAsync (st.data (), ne * sizeof (double), c.rowst[par].p + e0, E * sizeof (double),
                               ne * sizeof (double), 3 * cap, hipMemcpyDeviceToHost, c.stream));
  PMX_HIP (hipMemcpyAsync (cnt.data (), c.rowcnt[par].p + e0, ne * sizeof (int), hipMemcpyDeviceToHost, c.stream));
  PMX_HIP (hipMemcpyAsync (part.data (), c.part.p + e0, ne * sizeof (int), hipMemcpyDeviceToHost, c.stream));
  PMX_HIP (hipStreamSynchronize (c.stream));
  std::vector<int> ppart;
  for (int e = 0; e < ne; e++)
    for (int k = 0; k < cnt[e]; k++)
    {
      uint32_t kk = key[(size_t) k * ne + e];
      bool tri = (kk & PMX_TRI_FLAG) != 0;
      if (!tri)
      {
        /* a sphere pair is listed once, by its master (smaller particle) */
        int p2 = part[(int) kk - e0];
        if (!(part[e] < p2)) continue;
      }
      pmx_pair p;
      p.master = e + (local ? 0 : e0);
      int t = (int) (kk & ~PMX_TRI_FLAG);
      p.slave = tri ? -(t - (local ? t0 : 0)) - 1 : (int) kk - (local ? e0 : 0);
      for (int d = 0; d < 3; d++) p.state[d] = st[((size_t) d * cap + k) * ne + e];
      out.push_back (p);
    }
  return PMX_OK;
}



/* an instance's rows (the buffer holding them, every slot of its row
 * capacity: a slot past the count is never read) and its tree snapshot,
 * partners renumbered to the instance's own spheres and triangles */
int export_rows (pmx_ctx &c, int inst, ResumeRows &r)
{
  int e0 = c.off[inst].ell, ne = c.off[inst+1].ell - e0, t0 = c.off[inst].tri;
  r.cap = c.cap; r.ne = ne; r.nt = c.off[inst+1].tri - t0;
  size_t cap = c.cap, E = c.E;
  size_t esz = c.opt.precision == PMX_MIXED ? sizeof (float) : sizeof (double);
  r.key.assign (cap * ne, 0u);
  r.st.assign (3 * cap * ne * esz, 0);
  r.cnt.assign (ne, 0);
  r.bcen.assign (3 * (size_t) ne, 0.0);
  if (ne == 0) return PMX_OK;
  int par = c.icur[inst];
  const void *st = c.opt.precision == PMX_MIXED ? (const void*) c.rowst32[par].p : (const void*) c.rowst[par].p;
  PMX_HIP (hipMemcpy2DAsync (r.key.data (), ne * sizeof (uint32_t), c.rowkey[par].p + e0, E * sizeof (uint32_t),
                             ne * sizeof (uint32_t), cap, hipMemcpyDeviceToHost, c.stream));
  PMX_HIP (hipMemcpy2DAsync (r.st.data (), ne * esz, (const char*) st + e0 * esz, E * esz, ne * esz, 3 * cap,
                             hipMemcpyDeviceToHost, c.stream));
  PMX_HIP (hipMemcpyAsync (r.cnt.data (), c.rowcnt[par].p + e0, ne * sizeof (int), hipMemcpyDeviceToHost, c.stream));
  PMX_HIP (hipMemcpy2DAsync (r.bcen.data (), ne * sizeof (double), c.bcen.p + e0, E * sizeof (double),
                             ne * sizeof (double), 3, hipMemcpyDeviceToHost, c.stream));
  PMX_HIP (hipStreamSynchronize (c.stream));
  for (int e = 0; e < ne; e++)
    for (int k = 0; k < r.cnt[e]; k++)
    {
      uint32_t &kk = r.key[(size_t) k * ne + e];
      kk = (kk & PMX_TRI_FLAG) ? (PMX_TRI_FLAG | ((kk & ~PMX_TRI_FLAG) - (uint32_t) t0)) : kk - (uint32_t) e0;
    }
  return PMX_OK;
}

int import_rows (pmx_ctx &c, const std::vector<const ResumeRows*> &rs)
{
  if ((int) rs.size () != c.I) { pmx_set_error ("resume states: one entry per instance"); return PMX_EINVAL; }
  int cap = c.cap;
  bool any = false;
  for (int i = 0; i < c.I; i++)
    if (rs[i])
    {
      if (rs[i]->ne != c.off[i+1].ell - c.off[i].ell || rs[i]->nt != c.off[i+1].tri - c.off[i].tri)
      { pmx_set_error ("resume state of instance " + std::to_string (i) + ": counts differ"); return PMX_EINVAL; }
      if (!(rs[i]->curtime > 0.0)) { pmx_set_error ("a resumed instance must carry its time"); return PMX_EINVAL; }
      cap = std::max (cap, rs[i]->cap);
      any = true;
    }
  (void) hipSetDevice (c.device);
  if (cap > c.cap)
  {
    /* rows as wide as the widest resumed row (grown in its earlier context) */
    c.cap = cap;
    int rc = alloc_rows (c);
    if (rc) return rc;
  }
  for (int i = 0; i < c.I; i++)
  {
    /* the fresh instances: no contact history, whatever rows an upload of
     * the same layout kept */
    c.icur[i] = c.cur;
    if (rs[i]) continue;
    int e0 = c.off[i].ell, ne = c.off[i+1].ell - e0;
    if (ne > 0) PMX_HIP (hipMemsetAsync (c.rowcnt[c.cur].p + e0, 0, ne * sizeof (int), c.stream));
  }
  if (!any) return PMX_OK;
  size_t E = c.E, esz = c.opt.precision == PMX_MIXED ? sizeof (float) : sizeof (double);
  if (c.bcen.alloc (3 * std::max (E, (size_t) 1)) != hipSuccess)
  { pmx_set_error ("out of device memory (tree)"); return PMX_ENOMEM; }
  int par = c.cur;
  void *st = c.opt.precision == PMX_MIXED ? (void*) c.rowst32[par].p : (void*) c.rowst[par].p;
  for (int i = 0; i < c.I; i++)
  {
    const ResumeRows *r = rs[i];
    if (!r) continue;
    c.icurtime[i] = r->curtime;
    c.icurstep[i] = r->curstep;
    c.istepnum[i] = r->stepnum;
    if (r->ne == 0) continue;
    int e0 = c.off[i].ell, ne = r->ne, t0 = c.off[i].tri;
    std::vector<uint32_t> key (r->key);
    for (int e = 0; e < ne; e++)
      for (int k = 0; k < r->cnt[e]; k++)
      {
        uint32_t &kk = key[(size_t) k * ne + e];
        kk = (kk & PMX_TRI_FLAG) ? (PMX_TRI_FLAG | ((kk & ~PMX_TRI_FLAG) + (uint32_t) t0)) : kk + (uint32_t) e0;
      }
    PMX_HIP (hipMemcpy2DAsync (c.rowkey[par].p + e0, E * sizeof (uint32_t), key.data (), ne * sizeof (uint32_t),
                               ne * sizeof (uint32_t), r->cap, hipMemcpyHostToDevice, c.stream));
    for (int d = 0; d < 3; d++)
      PMX_HIP (hipMemcpy2DAsync ((char*) st + ((size_t) d * c.cap * E + e0) * esz, E * esz,
                                 r->st.data () + (size_t) d * r->cap * ne * esz, ne * esz, ne * esz, r->cap,
                                 hipMemcpyHostToDevice, c.stream));
    PMX_HIP (hipMemcpyAsync (c.rowcnt[par].p + e0, r->cnt.data (), ne * sizeof (int), hipMemcpyHostToDevice, c.stream));
    PMX_HIP (hipMemcpy2DAsync (c.bcen.p + e0, E * sizeof (double), r->bcen.data (), ne * sizeof (double),
                               ne * sizeof (double), 3, hipMemcpyHostToDevice, c.stream));
    /* the host copies are the caller's: done before it may free them */
    PMX_HIP (hipStreamSynchronize (c.stream));
  }
  c.snap_sel = true;
  return PMX_OK;
}

int merge_models (const pmx_model *ms, int count, Merged &g, std::vector<InstOff> &off)
{
  if (!ms || count < 1) return fail (PMX_EINVAL, "null argument or empty batch");
  for (int i = 0; i < count; i++)
  {
    int rc = validate_model (&ms[i]);
    if (rc) return fail (rc, "instance " + std::to_string (i) + ": " + pmx_last_error ());
    if (ms[i].curtime != ms[0].curtime || ms[i].curstep != ms[0].curstep || ms[i].stepnum != ms[0].stepnum)
      return fail (PMX_EINVAL, "instances of a batch must share curtime, curstep and stepnum");
    bool offsets = false;
    for (int q = 0; ms[i].sproffset && q < ms[i].sprnum; q++) offsets = offsets || ms[i].sproffset[q] >= 0;
    if (ms[i].trqsprnum || ms[i].unsprnum || ms[i].rstnum || ms[i].jnum || ms[i].lcnum || offsets)
      return fail (PMX_EUNSUPPORTED, "torsion springs, UNSPRING, RESTRAIN, BALL_JOINT and load curves run in "
                                     "single-model contexts (pmx_upload)");
  }
  off.assign (count + 1, InstOff ());
  for (int i = 0; i < count; i++)
  {
    InstOff n = model_counts (ms[i]), &a = off[i], &b = off[i+1];
    b.par = a.par + n.par; b.ell = a.ell + n.ell; b.tri = a.tri + n.tri; b.nod = a.nod + n.nod;
    b.fac = a.fac + n.fac; b.spr = a.spr + n.spr; b.scrv = a.scrv + n.scrv; b.dcrv = a.dcrv + n.dcrv;
    b.ucrv = a.ucrv + n.ucrv; b.pair = a.pair + n.pair; b.obs = a.obs + n.obs; b.prs = a.prs + n.prs;
    b.pcol = a.pcol + n.pcol;
    if ((long long) a.ell + n.ell >= (long long) PMX_TRI_FLAG || (long long) a.tri + n.tri >= (long long) PMX_TRI_FLAG)
      return fail (PMX_EINVAL, "batch too large");
  }
  const InstOff &tot = off[count];
  pmx_model &m = g.m;
  auto id = [] (auto v, int) { return v; };
  auto P = [&] (int i) { return (size_t) ms[i].parnum; };
  auto Po = [&] (int i) { return (size_t) off[i].par; };
  auto E = [&] (int i) { return (size_t) ms[i].ellnum; };
  auto Eo = [&] (int i) { return (size_t) off[i].ell; };
  auto T = [&] (int i) { return (size_t) ms[i].trinum; };
  auto To = [&] (int i) { return (size_t) off[i].tri; };
  auto Nd = [&] (int i) { return (size_t) ms[i].nodnum; };
  auto Ndo = [&] (int i) { return (size_t) off[i].nod; };
  auto F = [&] (int i) { return (size_t) ms[i].facnum; };
  auto Fo = [&] (int i) { return (size_t) off[i].fac; };
  auto S = [&] (int i) { return (size_t) ms[i].sprnum; };
  auto So = [&] (int i) { return (size_t) off[i].spr; };
  auto Q = [&] (int i) { return (size_t) ms[i].pairnum; };
  auto Qo = [&] (int i) { return (size_t) off[i].pair; };
  auto shp = [&] (int v, int i) { return v + off[i].par; };
  auto shp1 = [&] (int v, int i) { return v >= 0 ? v + off[i].par : v; };
  size_t NP = tot.par, NE = tot.ell, NT = tot.tri, NN = tot.nod, NF = tot.fac, NS = tot.spr, NQ = tot.pair;

#define CATD(field, k, cnt, o, tot_) { m.field = g.D ((k) * (tot_)); \
    cat<double> (m.field, k, count, tot_, [&] (int i) { return (const double*) ms[i].field; }, cnt, o, id); }
#define CATI(field, k, cnt, o, tot_, map) { m.field = g.Iv ((k) * (tot_)); \
    cat<int> (m.field, k, count, tot_, [&] (int i) { return (const int*) ms[i].field; }, cnt, o, map); }

  m.parnum = (int) NP;
  CATD (angular, 6, P, Po, NP); CATD (linear, 3, P, Po, NP); CATD (rotation, 9, P, Po, NP);
  CATD (position, 6, P, Po, NP); CATD (inertia, 9, P, Po, NP); CATD (inverse, 9, P, Po, NP);
  CATD (mass, 1, P, Po, NP); CATD (invm, 1, P, Po, NP);
  CATI (flags, 1, P, Po, NP, id);
  bool haveforce = true;
  for (int i = 0; i < count; i++) haveforce = haveforce && (ms[i].parnum == 0 || (ms[i].force && ms[i].torque));
  if (haveforce) { CATD (force, 3, P, Po, NP); CATD (torque, 3, P, Po, NP); }
  m.ellnum = (int) NE;
  CATD (center, 6, E, Eo, NE); CATD (radii, 3, E, Eo, NE);
  CATI (part, 1, E, Eo, NE, shp); CATI (ellcol, 1, E, Eo, NE, id);
  m.trinum = (int) NT;
  CATD (tri, 9, T, To, NT); CATI (tricol, 1, T, To, NT, id);
  /* triobs: >= 0 a particle, -1 static, <= -2 moving obstacle -(l+2) */
  CATI (triobs, 1, T, To, NT, [&] (int v, int i) { return v >= 0 ? v + off[i].par : (v <= -2 ? v - off[i].obs : v); });
  m.obsnum = tot.obs;
  m.trirng = g.Iv (2*(size_t) tot.obs);
  m.obspnt = g.D (3*(size_t) tot.obs); m.obsang = g.D (3*(size_t) tot.obs); m.obslin = g.D (3*(size_t) tot.obs);
  for (int i = 0; i < count; i++)
    for (int o = 0; o < ms[i].obsnum; o++)
    {
      int q = off[i].obs + o;
      m.trirng[2*q] = ms[i].trirng[2*o] + off[i].tri;
      m.trirng[2*q+1] = ms[i].trirng[2*o+1] + off[i].tri;
      for (int d = 0; d < 3; d++)
      {
        m.obspnt[3*q+d] = ms[i].obspnt[3*o+d];
        m.obsang[3*q+d] = ms[i].obsang ? ms[i].obsang[3*o+d] : 0.0;
        m.obslin[3*q+d] = ms[i].obslin ? ms[i].obslin[3*o+d] : 0.0;
      }
    }
  m.nodnum = (int) NN;
  CATD (nodes, 6, Nd, Ndo, NN); CATI (nodpart, 1, Nd, Ndo, NN, shp);
  m.facnum = (int) NF;
  CATI (facnod, 3, F, Fo, NF, [&] (int v, int i) { return v + off[i].nod; });
  CATI (factri, 1, F, Fo, NF, [&] (int v, int i) { return v + off[i].tri; });
  m.pairnum = (int) NQ;
  /* colour pairs are (i, j) couples, row-major */
  m.pairs = g.Iv (2*NQ);
  for (int i = 0; i < count; i++)
    if (ms[i].pairnum) std::memcpy (m.pairs + 2*off[i].pair, ms[i].pairs, 2*sizeof (int)*ms[i].pairnum);
  CATI (ikind, 1, Q, Qo, NQ, id); CATD (iparam, PMX_NIPARAM, Q, Qo, NQ);
  m.prsnum = tot.prs;
  m.prspart = g.Iv ((size_t) tot.prs);
  m.prskind = g.Iv (2 * (size_t) tot.prs);
  m.prscol = g.Iv ((size_t) tot.prs);
  for (int i = 0; i < count; i++)
    for (int q = 0; q < ms[i].prsnum; q++)
    {
      m.prspart[off[i].prs + q] = ms[i].prspart[q] + off[i].par;
      m.prskind[off[i].prs + q] = ms[i].prskind[q];
      m.prskind[tot.prs + off[i].prs + q] = ms[i].prskind[ms[i].prsnum + q];
      m.prscol[off[i].prs + q] = (ms[i].prscol ? ms[i].prscol[q] : q) + off[i].pcol;
    }
  m.sprnum = (int) NS;
  if (NS > 0)
  {
    CATI (sprtype, 1, S, So, NS, id); CATI (sprpart, 2, S, So, NS, shp1);
    CATD (sprpnt, 12, S, So, NS); CATD (yield, 2, S, So, NS); CATD (sprdir, 6, S, So, NS);
    CATI (sprflg, 1, S, So, NS, id); CATD (sprfric, 1, S, So, NS); CATD (sprkskn, 1, S, So, NS);
    CATD (sprsdsp, 3, S, So, NS); CATD (stroke0, 1, S, So, NS); CATD (stroke, 3, S, So, NS);
    CATD (sprfrc, 3, S, So, NS); CATI (unspring, 1, S, So, NS, id);
    m.sproffset = nullptr;
    /* curve index ranges [sprnum+1] and the curve tables [2][points] */
    m.spridx = g.Iv (NS + 1); m.dashidx = g.Iv (NS + 1); m.unidx = g.Iv (NS + 1);
    for (int i = 0; i < count; i++)
      for (int s = 0; s < ms[i].sprnum; s++)
      {
        m.spridx[off[i].spr + s] = ms[i].spridx[s] + off[i].scrv;
        m.dashidx[off[i].spr + s] = ms[i].dashidx[s] + off[i].dcrv;
        m.unidx[off[i].spr + s] = ms[i].unidx[s] + off[i].ucrv;
      }
    m.spridx[NS] = tot.scrv; m.dashidx[NS] = tot.dcrv; m.unidx[NS] = tot.ucrv;
    m.sprcrv = tot.scrv; m.dashcrv = tot.dcrv; m.uncrv = tot.ucrv;
    auto SC = [&] (int i) { return (size_t) ms[i].sprcrv; };
    auto SCo = [&] (int i) { return (size_t) off[i].scrv; };
    auto DC = [&] (int i) { return (size_t) ms[i].dashcrv; };
    auto DCo = [&] (int i) { return (size_t) off[i].dcrv; };
    auto UC = [&] (int i) { return (size_t) ms[i].uncrv; };
    auto UCo = [&] (int i) { return (size_t) off[i].ucrv; };
    CATD (spring, 2, SC, SCo, (size_t) tot.scrv);
    CATD (dashpot, 2, DC, DCo, (size_t) tot.dcrv);
    CATD (unload, 2, UC, UCo, (size_t) tot.ucrv);
  }
#undef CATD
#undef CATI
  m.curtime = ms[0].curtime; m.curstep = ms[0].curstep; m.stepnum = ms[0].stepnum;
  return PMX_OK;
}

extern "C" {

int pmx_upload_batch (pmx_ctx *c, const pmx_model *ms, int count)
{
  if (!c) return fail (PMX_EINVAL, "null ctx");
  Merged g;
  std::vector<InstOff> off;
  int rc = merge_models (ms, count, g, off);
  if (rc) return rc;
  return upload_core (c, &g.m, off);
}

int pmx_download_batch (pmx_ctx *c, pmx_model *ms, int count)
{
  if (!c || !ms) return fail (PMX_EINVAL, "null argument");
  if (count != c->I) return fail (PMX_EINVAL, "instance count differs from the uploaded batch");
  const std::vector<InstOff> &off = c->off;
  for (int i = 0; i < count; i++)
  {
    InstOff n = model_counts (ms[i]);
    if (n.par != off[i+1].par - off[i].par || n.ell != off[i+1].ell - off[i].ell || n.tri != off[i+1].tri - off[i].tri ||
        n.nod != off[i+1].nod - off[i].nod || n.spr != off[i+1].spr - off[i].spr)
      return fail (PMX_EINVAL, "instance " + std::to_string (i) + ": counts differ from the uploaded state");
  }
  const InstOff &tot = off[count];
  size_t NP = tot.par, NE = tot.ell, NT = tot.tri, NN = tot.nod, NS = tot.spr;
  Merged g;
  pmx_model &m = g.m;
  m.parnum = (int) NP; m.ellnum = (int) NE; m.trinum = (int) NT; m.nodnum = (int) NN; m.sprnum = (int) NS;
  m.angular = g.D (6*NP); m.linear = g.D (3*NP); m.rotation = g.D (9*NP); m.position = g.D (6*NP);
  m.inverse = g.D (9*NP); m.invm = g.D (NP); m.force = g.D (3*NP); m.torque = g.D (3*NP);
  m.center = g.D (6*NE); m.tri = g.D (9*NT); m.nodes = g.D (6*NN);
  size_t NO = tot.obs;
  m.obsnum = (int) NO;
  m.obspnt = g.D (3*NO); m.obsang = g.D (3*NO); m.obslin = g.D (3*NO);
  if (NS > 0)
  {
    m.sprpnt = g.D (12*NS); m.sprdir = g.D (6*NS); m.yield = g.D (2*NS); m.sprsdsp = g.D (3*NS);
    m.stroke = g.D (3*NS); m.sprfrc = g.D (3*NS); m.sprflg = g.Iv (NS);
  }
  int rc = pmx_download (c, &m);
  if (rc) return rc;
  auto P = [&] (int i) { return (size_t) ms[i].parnum; };
  auto Po = [&] (int i) { return (size_t) off[i].par; };
  auto E = [&] (int i) { return (size_t) ms[i].ellnum; };
  auto Eo = [&] (int i) { return (size_t) off[i].ell; };
  auto T = [&] (int i) { return (size_t) ms[i].trinum; };
  auto To = [&] (int i) { return (size_t) off[i].tri; };
  auto Nd = [&] (int i) { return (size_t) ms[i].nodnum; };
  auto Ndo = [&] (int i) { return (size_t) off[i].nod; };
  auto S = [&] (int i) { return (size_t) ms[i].sprnum; };
  auto So = [&] (int i) { return (size_t) off[i].spr; };
#define SPD(field, k, cnt, o, tot_) split<double> (m.field, k, count, tot_, [&] (int i) { return ms[i].field; }, cnt, o)
  SPD (angular, 6, P, Po, NP); SPD (linear, 3, P, Po, NP); SPD (rotation, 9, P, Po, NP); SPD (position, 6, P, Po, NP);
  SPD (inverse, 9, P, Po, NP); SPD (invm, 1, P, Po, NP); SPD (force, 3, P, Po, NP); SPD (torque, 3, P, Po, NP);
  SPD (center, 6, E, Eo, NE); SPD (tri, 9, T, To, NT); SPD (nodes, 6, Nd, Ndo, NN);
  if (NS > 0)
  {
    SPD (sprpnt, 12, S, So, NS); SPD (sprdir, 6, S, So, NS); SPD (yield, 2, S, So, NS); SPD (sprsdsp, 3, S, So, NS);
    SPD (stroke, 3, S, So, NS); SPD (sprfrc, 3, S, So, NS);
    split<int> (m.sprflg, 1, count, NS, [&] (int i) { return ms[i].sprflg; }, S, So);
  }
#undef SPD
  for (int i = 0; i < count; i++)
  {
    ms[i].curtime = c->icurtime[i]; ms[i].curstep = c->icurstep[i]; ms[i].stepnum = c->istepnum[i];
    size_t o0 = off[i].obs, no = ms[i].obsnum;
    if (ms[i].obspnt) std::memcpy (ms[i].obspnt, m.obspnt + 3*o0, 3*no*sizeof (double));
    if (ms[i].obsang) std::memcpy (ms[i].obsang, m.obsang + 3*o0, 3*no*sizeof (double));
    if (ms[i].obslin) std::memcpy (ms[i].obslin, m.obslin + 3*o0, 3*no*sizeof (double));
  }
  return PMX_OK;
}

/* the contact pairs of one instance, in the instance's own numbering */
int pmx_contacts_instance (pmx_ctx *c, int inst, pmx_pair *out, size_t cap, size_t *count)
{
  if (!c) return fail (PMX_EINVAL, "null ctx");
  if (inst < 0 || inst >= c->I) return fail (PMX_EINVAL, "instance out of range");
  (void) hipSetDevice (c->device);
  std::vector<pmx_pair> v;
  int rc = contacts_range (*c, inst, true, v);
  if (rc) return rc;
  if (count) *count = v.size ();
  if (out) std::memcpy (out, v.data (), std::min (cap, v.size ()) * sizeof (pmx_pair));
  return PMX_OK;
}

/* one instance's state into its own view (its ranges of every array) */
int pmx_download_instance (pmx_ctx *c, int inst, pmx_model *m)
{
  if (!c || !m) return fail (PMX_EINVAL, "null argument");
  if (inst < 0 || inst >= c->I) return fail (PMX_EINVAL, "instance out of range");
  const InstOff &a = c->off[inst], &b = c->off[inst+1];
  InstOff n = model_counts (*m);
  if (n.par != b.par - a.par || n.ell != b.ell - a.ell || n.tri != b.tri - a.tri || n.nod != b.nod - a.nod ||
      n.spr != b.spr - a.spr || n.obs != b.obs - a.obs)
    return fail (PMX_EINVAL, "model counts differ from the instance");
  (void) hipSetDevice (c->device);
  const InstOff &t = c->off[c->I];
  /* k components of n items at offset o, out of total items per component */
  auto dn = [&] (void *dst, const void *src, size_t elem, int k, size_t tot, size_t o, size_t cnt) -> int {
    if (!dst || cnt == 0 || k == 0) return PMX_OK;
    PMX_HIP (hipMemcpy2DAsync (dst, cnt * elem, (const char*) src + o * elem, tot * elem, cnt * elem, k,
                               hipMemcpyDeviceToHost, c->stream));
    return PMX_OK;
  };
  size_t np = n.par, ne = n.ell, nt = n.tri, nn = n.nod, ns = n.spr, no = n.obs;
  int rc;
  if ((rc = dn (m->angular, c->angular.p, 8, 6, t.par, a.par, np)) || (rc = dn (m->linear, c->linear.p, 8, 3, t.par, a.par, np)) ||
      (rc = dn (m->rotation, c->rotation.p, 8, 9, t.par, a.par, np)) || (rc = dn (m->position, c->position.p, 8, 6, t.par, a.par, np)) ||
      (rc = dn (m->inverse, c->inverse.p, 8, 9, t.par, a.par, np)) || (rc = dn (m->invm, c->invm.p, 8, 1, t.par, a.par, np)) ||
      (rc = dn (m->force, c->force.p, 8, 3, t.par, a.par, np)) || (rc = dn (m->torque, c->torque.p, 8, 3, t.par, a.par, np)) ||
      (rc = dn (m->center, c->center.p, 8, 6, t.ell, a.ell, ne)) || (rc = dn (m->tri, c->tri.p, 8, 9, t.tri, a.tri, nt)) ||
      (rc = dn (m->nodes, c->nodes.p, 8, 6, t.nod, a.nod, nn)) ||
      (rc = dn (m->obspnt, c->obspnt.p, 8, 1, 3*(size_t) t.obs, 3*(size_t) a.obs, 3*no)) ||
      (rc = dn (m->obsang, c->obsang.p, 8, 1, 3*(size_t) t.obs, 3*(size_t) a.obs, 3*no)) ||
      (rc = dn (m->obslin, c->obslin.p, 8, 1, 3*(size_t) t.obs, 3*(size_t) a.obs, 3*no)))
    return rc;
  if (ns > 0 &&
      ((rc = dn (m->sprpnt, c->sp_pnt.p, 8, 12, t.spr, a.spr, ns)) || (rc = dn (m->sprdir, c->sp_dir.p, 8, 6, t.spr, a.spr, ns)) ||
       (rc = dn (m->yield, c->sp_yield.p, 8, 2, t.spr, a.spr, ns)) || (rc = dn (m->sprsdsp, c->sp_sdsp.p, 8, 3, t.spr, a.spr, ns)) ||
       (rc = dn (m->stroke, c->sp_stroke.p, 8, 3, t.spr, a.spr, ns)) || (rc = dn (m->sprfrc, c->sp_frc.p, 8, 3, t.spr, a.spr, ns)) ||
       (rc = dn (m->sprflg, c->sp_flg.p, 4, 1, t.spr, a.spr, ns))))
    return rc;
  PMX_HIP (hipStreamSynchronize (c->stream));
  m->curtime = c->icurtime[inst]; m->curstep = c->icurstep[inst]; m->stepnum = c->istepnum[inst];
  return PMX_OK;
}

int pmx_dem_batch_durations (pmx_ctx *c, const double *durations, double step, const pmx_forcing *forcing,
                             pmx_done_fn done, void *user, double *elapsed)
{
  if (!c || !durations) return fail (PMX_EINVAL, "null argument");
  return dem_core (c, durations, 0.0, step, forcing, c->I, done, user, elapsed);
}

} /* extern "C" */
