/*
 * pmx_api.hip -- C ABI of libpmx.so and the host step loop.
 *
 * pmx_dem() follows dem() (parmec.cpp:2315-2530) for the supported subset:
 * first call (curtime == 0): euler(h/2) and a full shapes update
 * (parmec.cpp:2386-2400); every call: invert_inertia (2411) and a fresh
 * partitioning tree (2415); per step: partitioning_store with rebuild on leaf
 * overflow (2420-2427), condet (2429), gravity (2432), forces (2434), step1 =
 * step0 (2450), dynamics (2461), shapes (2485-2496); loop control
 * `time < duration` with time += (step0+step1)/2 (2418); after the loop
 * curstep = step1 and stepnum++ (2521-2523).
 */
#include "pmx_internal.h"
#include <cstdlib>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>

static thread_local std::string g_last_error;

void pmx_set_error (const std::string &msg) { g_last_error = msg; }

namespace {

int fail (int code, const std::string &msg) { pmx_set_error (msg); return code; }

template <class T> int up (pmx_ctx &c, dvec<T> &d, const T *src, size_t n, const char *what)
{
  if (d.alloc (std::max (n, (size_t) 1)) != hipSuccess) return fail (PMX_ENOMEM, std::string ("out of device memory: ") + what);
  if (n == 0) return PMX_OK;
  if (!src) return fail (PMX_EINVAL, std::string ("missing array: ") + what);
  /* the source is pageable host memory that may be freed when this returns:
   * the copy completes first */
  PMX_HIP (hipMemcpyAsync (d.p, src, n * sizeof (T), hipMemcpyHostToDevice, c.stream));
  PMX_HIP (hipStreamSynchronize (c.stream));
  return PMX_OK;
}

template <class T> int down (pmx_ctx &c, T *dst, const dvec<T> &d, size_t n)
{
  if (!dst || n == 0) return PMX_OK;
  PMX_HIP (hipMemcpyAsync (dst, d.p, n * sizeof (T), hipMemcpyDeviceToHost, c.stream));
  return PMX_OK;
}

} /* namespace */

int alloc_rows (pmx_ctx &c)
{
  size_t E = std::max (c.E, 1);
  bool mixed = c.opt.precision == PMX_MIXED;
  for (int b = 0; b < 2; b++)
    if (c.rowkey[b].alloc ((size_t) c.cap * E) || c.rowcnt[b].alloc (E) ||
        (mixed ? c.rowst32[b].alloc (3 * (size_t) c.cap * E) : c.rowst[b].alloc (3 * (size_t) c.cap * E)))
      return fail (PMX_ENOMEM, "out of device memory (contact rows)");
  if (c.hits.alloc ((size_t) c.hcap * E) || c.hitcnt.alloc (E) || c.xbuf.alloc ((size_t) c.xcap * E) ||
      c.xcnt.alloc (E) || c.rbusy.alloc (E + 1024) || c.rnbusy.alloc (E / 64 + 2) || c.leafof.alloc (E) ||
      c.slot.alloc (E) || c.lmem.alloc (E) || c.lpart.alloc (E) || c.lpos.alloc (4 * E))
    return fail (PMX_ENOMEM, "out of device memory (broad phase)");
  if (c.opt.critical && c.crec.alloc (8 * (size_t) c.cap * E))
    return fail (PMX_ENOMEM, "out of device memory (contact records)");
  PMX_HIP (hipMemsetAsync (c.rowcnt[0].p, 0, E * sizeof (int), c.stream));
  PMX_HIP (hipMemsetAsync (c.rowcnt[1].p, 0, E * sizeof (int), c.stream));
  /* hit and reverse-hit counts: zero between steps (k_rows clears the ones it consumes) */
  PMX_HIP (hipMemsetAsync (c.xcnt.p, 0, E * sizeof (int), c.stream));
  PMX_HIP (hipMemsetAsync (c.hitcnt.p, 0, E * sizeof (int), c.stream));
  c.cur = 0;
  return PMX_OK;
}

/* checks a single model (one instance) for what the device path supports */
int validate_model (const pmx_model *m)
{
  if (m->parnum < 0 || m->ellnum < 0 || m->trinum < 0 || m->pairnum < 1)
    return fail (PMX_EINVAL, "invalid counts (pairnum must include the default pairing)");
  if (m->obsnum < 0) return fail (PMX_EINVAL, "invalid obstacle count");
  if (m->prsnum < 0) return fail (PMX_EINVAL, "invalid prescribed-motion count");
  if (int rc = jt_validate (m)) return rc;
  for (int i = 0; i < m->prsnum; i++)
  {
    if (m->prspart[i] < 0 || m->prspart[i] >= m->parnum) return fail (PMX_EINVAL, "prescribed particle out of range");
    for (int k = 0; k < 2; k++)
      if (m->prskind[(size_t) k*m->prsnum + i] < -1 || m->prskind[(size_t) k*m->prsnum + i] > 1)
        return fail (PMX_EINVAL, "prescribed motion kind must be -1, 0 or 1");
    if (m->prscol && m->prscol[i] < 0) return fail (PMX_EINVAL, "prescribed history column out of range");
  }
  {
    std::vector<char> moved (std::max (m->trinum, 1), 0);
    for (int i = 0; i < m->obsnum; i++)
    {
      if (m->trirng[2*i] < 0 || m->trirng[2*i] > m->trirng[2*i+1] || m->trirng[2*i+1] > m->trinum)
        return fail (PMX_EINVAL, "obstacle triangle range out of bounds");
      for (int t = m->trirng[2*i]; t < m->trirng[2*i+1]; t++)
      {
        if (moved[t]) return fail (PMX_EINVAL, "obstacle triangle ranges overlap");
        moved[t] = 1;
      }
    }
  }
  for (int i = 0; i < m->sprnum; i++)
  {
    if (m->sprpart[i] < 0 || m->sprpart[i] >= m->parnum || m->sprpart[m->sprnum + i] < -1 ||
        m->sprpart[m->sprnum + i] >= m->parnum) return fail (PMX_EINVAL, "spring particle index out of range");
    if (i && (m->sprpart[i-1] > m->sprpart[i] ||
              (m->sprpart[i-1] == m->sprpart[i] && m->sprpart[m->sprnum + i-1] > m->sprpart[m->sprnum + i])))
      return fail (PMX_EINVAL, "springs must be sorted by (part1, part2)");
  }
  for (int e = 0; e < m->ellnum; e++)
  {
    if (!(m->radii[(size_t) m->ellnum + e] < 0.0) || !(m->radii[2*(size_t) m->ellnum + e] < 0.0))
      return fail (PMX_EUNSUPPORTED, "only spherical ellipsoids are supported");
    if (m->part[e] < 0 || m->part[e] >= m->parnum) return fail (PMX_EINVAL, "ellipsoid particle index out of range");
  }
  for (int t = 0; t < m->trinum; t++)
  {
    if (m->triobs[t] <= -2 && -m->triobs[t]-2 >= m->obsnum) return fail (PMX_EINVAL, "triangle obstacle index out of range");
    if (m->triobs[t] >= m->parnum) return fail (PMX_EINVAL, "triangle particle index out of range");
  }
  for (int q = 0; q < m->nodnum; q++)
    if (m->nodpart[q] < 0 || m->nodpart[q] >= m->parnum) return fail (PMX_EINVAL, "node particle index out of range");
  for (int f = 0; f < m->facnum; f++)
  {
    if (m->factri[f] < 0 || m->factri[f] >= m->trinum) return fail (PMX_EINVAL, "face triangle index out of range");
    for (int k = 0; k < 3; k++)
      if (m->facnod[(size_t) k*m->facnum + f] < 0 || m->facnod[(size_t) k*m->facnum + f] >= m->nodnum)
        return fail (PMX_EINVAL, "face node index out of range");
  }
  std::vector<char> seen (std::max (m->parnum, 1), 0);
  for (int e = 0; e < m->ellnum; e++)
  {
    if (seen[m->part[e]]) return fail (PMX_EUNSUPPORTED, "a particle may own at most one sphere on this path");
    seen[m->part[e]] = 1;
  }
  return PMX_OK;
}

/* item counts of one model as offsets past its end */
InstOff model_counts (const pmx_model &m)
{
  InstOff o;
  o.par = m.parnum; o.ell = m.ellnum; o.tri = m.trinum; o.nod = m.nodnum; o.fac = m.facnum; o.spr = m.sprnum;
  o.scrv = m.sprcrv; o.dcrv = m.dashcrv; o.ucrv = m.uncrv; o.pair = m.pairnum; o.obs = m.obsnum; o.prs = m.prsnum;
  o.pcol = m.prsnum;
  if (m.prscol)
  {
    o.pcol = 0;
    for (int i = 0; i < m.prsnum; i++) o.pcol = std::max (o.pcol, m.prscol[i] + 1);
  }
  return o;
}

/* copy a (possibly concatenated) model to the device; off[i] are the
 * instance offsets, off[I] the totals */
int upload_core (pmx_ctx *c, const pmx_model *m, const std::vector<InstOff> &off)
{
  int I = (int) off.size () - 1;
  std::vector<int> parell (std::max (m->parnum, 1), -1), nosph;
  double rmax = 0.0;
  for (int e = 0; e < m->ellnum; e++)
    for (int k = 0; k < 3; k++) rmax = std::max (rmax, m->radii[(size_t) k*m->ellnum + e]);
  for (int e = 0; e < m->ellnum; e++) parell[m->part[e]] = e;
  for (int i = 0; i < m->parnum; i++) if (parell[i] < 0) nosph.push_back (i);
  std::vector<signed char> iso (std::max (m->parnum, 1), 0);
  int niso = 0, nzc = 0, nfree = 0;
  /* particles whose force another stage reads, adds to or sets (springs,
   * torsion springs, prescribed motion, restraints, ball joints): their force
   * always goes through the force array */
  std::vector<char> engaged (std::max (m->parnum, 1), 0);
  for (int q = 0; q < m->sprnum; q++)
  {
    engaged[m->sprpart[q]] = 1;
    if (m->sprpart[m->sprnum + q] >= 0) engaged[m->sprpart[m->sprnum + q]] = 1;
  }
  for (int q = 0; q < m->prsnum; q++) engaged[m->prspart[q]] = 1;
  for (int q = 0; q < m->trqsprnum; q++)
  {
    engaged[m->trqsprpart[q]] = 1;
    if (m->trqsprpart[m->trqsprnum + q] >= 0) engaged[m->trqsprpart[m->trqsprnum + q]] = 1;
  }
  for (int q = 0; q < m->rstnum; q++) engaged[m->rstpart[q]] = 1;
  for (int q = 0; q < m->jnum; q++)
  {
    engaged[m->jpart[q]] = 1;
    if (m->jpart[m->jnum + q] >= 0) engaged[m->jpart[m->jnum + q]] = 1;
  }
  for (int i = 0; i < m->parnum; i++)
  {
    const double *J = m->inertia;
    size_t N = m->parnum;
    bool d = J[i] > 0.0 && J[i] == J[4*N + i] && J[i] == J[8*N + i];
    for (int k : {1, 2, 3, 5, 6, 7}) d = d && J[k*N + i] == 0.0 && !std::signbit (J[k*N + i]);
    iso[i] = d ? 1 : 0;
    niso += d;
    /* bit 1: the sphere's body offset (reference centre - reference
     * position, as the device forms it) is +0 in every component */
    int e = parell[i];
    bool zc = e >= 0;
    for (int k = 0; zc && k < 3; k++)
    {
      double C = m->center[(size_t) (3+k)*m->ellnum + e] - m->position[(size_t) (3+k)*N + i];
      zc = C == 0.0 && !std::signbit (C);
    }
    if (zc) { iso[i] |= 2; nzc++; }
    /* bit 2: a sphere particle that nothing but its contacts and gravity
     * loads: when its row ends up empty its force is 0 + m g and its torque
     * 0, formed where they are used instead of stored */
    if (e >= 0 && !engaged[i]) { iso[i] |= 4; nfree++; }
  }
  std::vector<int> pinst (std::max (m->parnum, 1)), einst (std::max (m->ellnum, 1)), tinst (std::max (m->trinum, 1));
  std::vector<int> eoff (I + 1), poff (I + 1), toff (I + 1);
  for (int i = 0; i < I; i++)
  {
    for (int q = off[i].par; q < off[i+1].par; q++) pinst[q] = i;
    for (int q = off[i].ell; q < off[i+1].ell; q++) einst[q] = i;
    for (int q = off[i].tri; q < off[i+1].tri; q++) tinst[q] = i;
  }
  for (int i = 0; i <= I; i++) { eoff[i] = off[i].ell; poff[i] = off[i].pair; toff[i] = off[i].tri; }
  (void) hipSetDevice (c->device);
  bool same = c->uploaded && c->N == m->parnum && c->E == m->ellnum && c->T == m->trinum && c->I == I;
  for (int i = 0; same && i <= I; i++) same = c->off[i].ell == off[i].ell && c->off[i].tri == off[i].tri;
  /* contact history is kept across uploads only while every instance's rows
   * sit in the same row buffer */
  for (int i = 0; same && i < I; i++) same = (int) c->icur.size () == I && c->icur[i] == c->cur;
  c->I = I;
  c->off = off;
  c->N = m->parnum; c->E = m->ellnum; c->T = m->trinum; c->P = m->pairnum;
  c->NOD = m->nodnum; c->FAC = m->facnum;
  size_t N = c->N, E = c->E, T = c->T, P = c->P;
  int rc = 0;
  if ((rc = up (*c, c->angular, m->angular, 6*N, "angular")) || (rc = up (*c, c->linear, m->linear, 3*N, "linear")) ||
      (rc = up (*c, c->rotation, m->rotation, 9*N, "rotation")) || (rc = up (*c, c->position, m->position, 6*N, "position")) ||
      (rc = up (*c, c->inertia, m->inertia, 9*N, "inertia")) || (rc = up (*c, c->inverse, m->inverse, 9*N, "inverse")) ||
      (rc = up (*c, c->mass, m->mass, N, "mass")) || (rc = up (*c, c->invm, m->invm, N, "invm")) ||
      (rc = up (*c, c->center, m->center, 6*E, "center")) || (rc = up (*c, c->radii, m->radii, 3*E, "radii")) ||
      (rc = up (*c, c->part, m->part, E, "part")) || (rc = up (*c, c->ellcol, m->ellcol, E, "ellcol")) ||
      (rc = up (*c, c->tri, m->tri, 9*T, "tri")) || (rc = up (*c, c->tricol, m->tricol, T, "tricol")) ||
      (rc = up (*c, c->triobs, m->triobs, T, "triobs")) || (rc = up (*c, c->pairs, m->pairs, 2*P, "pairs")) ||
      (rc = up (*c, c->ikind, m->ikind, P, "ikind")) || (rc = up (*c, c->iparam, m->iparam, (size_t) PMX_NIPARAM*P, "iparam")))
    return rc;
  if ((rc = up (*c, c->par_inst, pinst.data (), N, "instances")) || (rc = up (*c, c->ell_inst, einst.data (), E, "instances")) ||
      (rc = up (*c, c->tri_inst, tinst.data (), T, "instances")) ||
      (rc = up (*c, c->ell_off, eoff.data (), (size_t) I + 1, "instances")) ||
      (rc = up (*c, c->pair_off, poff.data (), (size_t) I + 1, "instances")) ||
      (rc = up (*c, c->tri_off, toff.data (), (size_t) I + 1, "instances")))
    return rc;
  if (c->force.alloc (3*std::max (N, (size_t) 1)) || c->torque.alloc (3*std::max (N, (size_t) 1)) ||
      c->ovf.alloc (std::max (I, 1)))
    return fail (PMX_ENOMEM, "out of device memory");
  PMX_HIP (hipMemsetAsync (c->ovf.p, 0, sizeof (int) * std::max (I, 1), c->stream));
  c->nnosph = (int) nosph.size ();
  c->niso = niso;
  c->nzc = nzc;
  c->nfree = nfree;
  if ((rc = up (*c, c->iso, iso.data (), N, "iso"))) return rc;
  if (niso > 0 && niso < (int) N)
  {
    /* both inertia forms: k_dynamics runs once per form over its own
     * particles, the scalar ones first (ascending within each) */
    std::vector<int> dl;
    dl.reserve (N);
    for (int i = 0; i < (int) N; i++) if (iso[i] & 1) dl.push_back (i);
    for (int i = 0; i < (int) N; i++) if (!(iso[i] & 1)) dl.push_back (i);
    if ((rc = up (*c, c->dynlist, dl.data (), N, "dynamics lists"))) return rc;
  }
  if ((rc = up (*c, c->parell, parell.data (), N, "parell")) ||
      (rc = up (*c, c->nosph, nosph.data (), nosph.size (), "nosph")))
    return rc;
  c->bin_hint = false;
  c->rmax = rmax;
  if (m->force) PMX_HIP (hipMemcpyAsync (c->force.p, m->force, 3*N*sizeof (double), hipMemcpyHostToDevice, c->stream));
  if (m->torque) PMX_HIP (hipMemcpyAsync (c->torque.p, m->torque, 3*N*sizeof (double), hipMemcpyHostToDevice, c->stream));
  if (!same)
  {
    if ((rc = alloc_rows (*c))) return rc;
  }
  /* prescribed motion: entries in order, folded to one per particle (the
   * force comes from the particle's last linear/angular entry, the velocity
   * from its last velocity entry, as the reference's sequential loops) */
  {
    c->PRS = m->prsnum;
    c->PRSC = off[I].pcol;
    std::map<int, int> slot;
    std::vector<int> pp, lf, lv, af, av;
    for (int i = 0; i < m->prsnum; i++)
    {
      int j = m->prspart[i], lk = m->prskind[i], ak = m->prskind[(size_t) m->prsnum + i];
      int col = m->prscol ? m->prscol[i] : i;
      auto it = slot.find (j);
      int q;
      if (it == slot.end ()) { q = (int) pp.size (); slot[j] = q; pp.push_back (j); lf.push_back (-1); lv.push_back (-1); af.push_back (-1); av.push_back (-1); }
      else q = it->second;
      if (lk >= 0) { lf[q] = 2*col + lk; if (lk == 0) lv[q] = col; }
      if (ak >= 0) { af[q] = 2*col + ak; if (ak == 0) av[q] = col; }
    }
    c->nprs = (int) pp.size ();
    size_t n = pp.size ();
    if ((rc = up (*c, c->prs_part, pp.data (), n, "prescribed")) || (rc = up (*c, c->prs_lf, lf.data (), n, "prescribed")) ||
        (rc = up (*c, c->prs_lv, lv.data (), n, "prescribed")) || (rc = up (*c, c->prs_af, af.data (), n, "prescribed")) ||
        (rc = up (*c, c->prs_av, av.data (), n, "prescribed")))
      return rc;
  }
  /* moving obstacles; and, where spheres can touch them, the triangles of
   * meshed particles, which move every step as well: both reach the leaves
   * through per-step lists (obstacle index -1 for a meshed particle's
   * triangle, moved by its particle's faces instead) */
  {
    c->O = m->obsnum;
    std::vector<int> mtri, mtobs, tmov (std::max (m->trinum, 1), 0);
    for (int i = 0; i < m->obsnum; i++)
      for (int t = m->trirng[2*i]; t < m->trirng[2*i+1]; t++) { mtri.push_back (t); mtobs.push_back (i); tmov[t] = 1; }
    std::vector<int> mp_part, mp_toff (1, 0), mp_tlist;
    /* a meshed particle's triangles can meet spheres only in an instance
     * that has spheres: only those move every step (the others' nodes and
     * faces are formed where the state is read) */
    std::vector<std::vector<int>> bypart (std::max (m->parnum, 1));
    for (int i = 0; i < I; i++)
      if (off[i+1].ell > off[i].ell)
        for (int t = off[i].tri; t < off[i+1].tri; t++)
          if (m->triobs[t] >= 0) { mtri.push_back (t); mtobs.push_back (-1); tmov[t] = 1; bypart[m->triobs[t]].push_back (t); }
    std::vector<char> contact (std::max (m->parnum, 1), 0);
    for (int p = 0; p < m->parnum; p++)
      if (!bypart[p].empty ())
      {
        mp_part.push_back (p);
        contact[p] = 1;
        mp_tlist.insert (mp_tlist.end (), bypart[p].begin (), bypart[p].end ());
        mp_toff.push_back ((int) mp_tlist.size ());
      }
    c->nmtri = (int) mtri.size ();
    c->nmp = (int) mp_part.size ();
    c->mesh_contact = c->nmp > 0;
    /* the nodes and faces moved every step, where they are not all of them */
    std::vector<int> mnod, mfac;
    if (c->mesh_contact)
    {
      for (int q = 0; q < m->nodnum; q++) if (contact[m->nodpart[q]]) mnod.push_back (q);
      for (int q = 0; q < m->facnum; q++) if (tmov[m->factri[q]]) mfac.push_back (q);
    }
    c->nmnod = (int) mnod.size () < m->nodnum ? (int) mnod.size () : -1;
    c->nmfac = (int) mfac.size () < m->facnum ? (int) mfac.size () : -1;
    if ((c->nmnod >= 0 && (rc = up (*c, c->mnod, mnod.data (), mnod.size (), "mesh lists"))) ||
        (c->nmfac >= 0 && (rc = up (*c, c->mfac, mfac.data (), mfac.size (), "mesh lists"))))
      return rc;
    size_t O = c->O;
    if ((rc = up (*c, c->obspnt, m->obspnt, 3*O, "obspnt")) || (rc = up (*c, c->obsang, m->obsang, 3*O, "obsang")) ||
        (rc = up (*c, c->obslin, m->obslin, 3*O, "obslin")) || (rc = up (*c, c->mtri, mtri.data (), mtri.size (), "obstacles")) ||
        (rc = up (*c, c->mtobs, mtobs.data (), mtobs.size (), "obstacles")) ||
        (rc = up (*c, c->tmov, tmov.data (), T, "obstacles")) ||
        (rc = up (*c, c->mp_part, mp_part.data (), mp_part.size (), "meshed particles")) ||
        (rc = up (*c, c->mp_toff, mp_toff.data (), mp_toff.size (), "meshed particles")) ||
        (rc = up (*c, c->mp_tlist, mp_tlist.data (), mp_tlist.size (), "meshed particles")))
      return rc;
    if (c->mesh_contact)
    {
      if (c->tcnt.alloc (std::max (T, (size_t) 1)) || c->trec_e.alloc ((size_t) c->tcap * T) ||
          c->trec.alloc (6 * (size_t) c->tcap * T))
        return fail (PMX_ENOMEM, "out of device memory (mesh reactions)");
      PMX_HIP (hipMemsetAsync (c->tcnt.p, 0, sizeof (int) * T, c->stream));
    }
  }
  if ((rc = up (*c, c->nodes, m->nodes, 6*(size_t) c->NOD, "nodes")) ||
      (rc = up (*c, c->nodpart, m->nodpart, (size_t) c->NOD, "nodpart")) ||
      (rc = up (*c, c->facnod, m->facnod, 3*(size_t) c->FAC, "facnod")) ||
      (rc = up (*c, c->factri, m->factri, (size_t) c->FAC, "factri")) ||
      (rc = springs_setup (*c, m)) || (rc = jt_setup (*c, m)))
    return rc;
  c->curtime = m->curtime; c->curstep = m->curstep; c->stepnum = m->stepnum;
  c->icurtime.assign (I, m->curtime);
  c->icurstep.assign (I, m->curstep);
  c->istepnum.assign (I, m->stepnum);
  if (!same) c->icur.assign (I, c->cur);
  c->staggered = false;
  {
    std::vector<int> oi (std::max (off[I].obs, 1), 0);
    for (int i = 0; i < I; i++) for (int q = off[i].obs; q < off[i+1].obs; q++) oi[q] = i;
    if ((rc = up (*c, c->obs_inst, oi.data (), (size_t) off[I].obs, "obstacles"))) return rc;
  }
  PMX_HIP (hipStreamSynchronize (c->stream));
  c->uploaded = true;
  c->crec_valid = c->crec_stale = false;
  c->nhis = 0;   /* histories name particles of the previous model */
  c->his_active = false;
  return PMX_OK;
}

void prof_resolve (pmx_ctx &c, int complete, int upto)
{
  size_t keep = 0;
  for (size_t k = 0; k < c.prof.pend.size (); k++)
  {
    Prof::Pend &p = c.prof.pend[k];
    if (upto != 0x7FFFFFFF && (p.step < 0 || p.step >= upto)) { c.prof.pend[keep++] = p; continue; }
    float ms = 0.0f;
    if (p.step < complete && hipEventElapsedTime (&ms, p.a, p.b) == hipSuccess)
    {
      c.prof.ms[p.kid] += ms;
      c.prof.n[p.kid] += 1;
      c.prof.bytes[p.kid] += p.bytes;
    }
    c.prof.pool.push_back (p.a);
    c.prof.pool.push_back (p.b);
  }
  c.prof.pend.resize (keep);
}

/* the captured step graph and the events its record nodes were captured with */
static void drop_graph (pmx_ctx *c)
{
  if (c->gexec) { (void) hipGraphExecDestroy (c->gexec); c->gexec = nullptr; }
  if (c->graph) { (void) hipGraphDestroy (c->graph); c->graph = nullptr; }
  if (c->pgexec) { (void) hipGraphExecDestroy (c->pgexec); c->pgexec = nullptr; }
  if (c->gexec_multi) { (void) hipGraphExecDestroy (c->gexec_multi); c->gexec_multi = nullptr; }
  if (c->pgraph) { (void) hipGraphDestroy (c->pgraph); c->pgraph = nullptr; }
  for (auto &q : c->prof.cap) { if (q.a) (void) hipEventDestroy (q.a); if (q.b) (void) hipEventDestroy (q.b); }
  c->prof.cap.clear ();
  c->graph_valid = false;
}

extern "C" {

const char *pmx_last_error (void) { return g_last_error.c_str (); }

int pmx_step_times (double duration, double step, double curtime, double *times, int cap, int *nsteps)
{
  if (!(step > 0.0)) return fail (PMX_EINVAL, "step must be positive");
  int k = 0;
  for (double time = 0.0; time < duration; time += 0.5*(step+step), curtime += 0.5*(step+step), k++)
    if (times && k < cap) times[k] = curtime;
  if (nsteps) *nsteps = k;
  return PMX_OK;
}

int pmx_create (const pmx_options *opt, pmx_ctx **out)
{
  if (!out) return fail (PMX_EINVAL, "null ctx pointer");
  *out = nullptr;
  pmx_ctx *c = new pmx_ctx ();
  if (opt) c->opt = *opt;
  if (c->opt.precision != PMX_FP64 && c->opt.precision != PMX_MIXED)
  {
    delete c;
    return fail (PMX_EINVAL, "precision must be PMX_FP64 or PMX_MIXED");
  }
  if (c->opt.det_waves != 0 && c->opt.det_waves != 1 && c->opt.det_waves != 2 && c->opt.det_waves != 4 &&
      c->opt.det_waves != 8)
  { delete c; return fail (PMX_EINVAL, "det_waves must be 1, 2, 4 or 8"); }
  if (c->opt.rows_block != 0 && c->opt.rows_block != 64 && c->opt.rows_block != 128 && c->opt.rows_block != 256 &&
      c->opt.rows_block != 512 && c->opt.rows_block != 1024)
  { delete c; return fail (PMX_EINVAL, "rows_block must be 64, 128, 256, 512 or 1024"); }
  if (c->opt.det_waves) c->det_nw = c->opt.det_waves;
  {
    const char *nl = std::getenv ("PMX_NEIGHBOUR_LISTS");
    c->nbr_csr = nl && std::string (nl) == "csr";
    const char *lr = std::getenv ("PMX_LEAF_RECORDS");
    c->rec_packed = lr && std::string (lr) == "packed";
  }
  if (c->opt.rows_block) c->rows_tpb = c->opt.rows_block;
  int ndev = 0;
  if (hipGetDeviceCount (&ndev) != hipSuccess || ndev == 0)
  {
    delete c;
    return fail (PMX_EDEVICE, "no HIP device available");
  }
  c->device = c->opt.device;
  if (hipSetDevice (c->device) != hipSuccess || hipStreamCreateWithFlags (&c->stream, hipStreamNonBlocking) != hipSuccess)
  {
    delete c;
    return fail (PMX_EDEVICE, "cannot open HIP device/stream");
  }
  if (int rc = history_scratch_init (c->device)) { delete c; return rc; }
  if (c->opt.row_capacity > 0) { c->cap = c->opt.row_capacity; c->hcap = c->opt.row_capacity; c->xcap = std::max (8, c->opt.row_capacity / 2); }
  if (c->dflags.alloc (FLAG_COUNT) != hipSuccess || c->ctl.alloc (CTL_COUNT) != hipSuccess)
  { delete c; return fail (PMX_ENOMEM, "out of device memory"); }
  /* on the context's stream, never the legacy one: contexts of other
   * threads may be capturing a step */
  if (hipMemsetAsync (c->ctl.p, 0, CTL_COUNT * sizeof (int), c->stream) != hipSuccess ||
      hipMemsetAsync (c->dflags.p, 0, FLAG_COUNT * sizeof (int), c->stream) != hipSuccess ||
      hipStreamSynchronize (c->stream) != hipSuccess)
  { delete c; return fail (PMX_EDEVICE, "memset failed"); }
  *out = c;
  return PMX_OK;
}

void pmx_destroy (pmx_ctx *c)
{
  if (!c) return;
  (void) hipSetDevice (c->device);
  if (c->stream) (void) hipStreamSynchronize (c->stream);
  dvec<double> *d[] = {&c->angular, &c->linear, &c->rotation, &c->position, &c->inertia, &c->inverse, &c->mass,
                       &c->invm, &c->force, &c->torque, &c->center, &c->radii, &c->tri, &c->iparam, &c->gtab, &c->dtab,
                       &c->lpos, &c->lbox, &c->rowst[0], &c->rowst[1], &c->tree.ext, &c->tree.coord, &c->tree.cell};
  for (auto *p : d) p->release ();
  dvec<int> *ii[] = {&c->flags, &c->part, &c->ellcol, &c->tricol, &c->triobs, &c->pairs, &c->ikind, &c->leafof,
                     &c->slot, &c->lcount, &c->loff, &c->lmem, &c->rowcnt[0], &c->rowcnt[1], &c->hitcnt, &c->xbuf,
                     &c->xcnt, &c->dflags, &c->parell, &c->nosph, &c->tree.order, &c->tree.order2, &c->tree.hist, &c->tree.scanbuf,
                     &c->tree.first, &c->tree.size, &c->tree.split, &c->tree.dim, &c->tree.parent, &c->tree.left,
                     &c->tree.right, &c->tree.nleafkids, &c->tree.leafpar, &c->tree.tricnt, &c->tree.trioff,
                     &c->tree.trilist, &c->tree.nbrcnt, &c->tree.nbroff, &c->tree.nbrlist};
  for (auto *p : ii) p->release ();
  c->rowkey[0].release (); c->rowkey[1].release (); c->hits.release (); c->iso.release ();
  c->tree.code.release (); c->tree.code2.release ();
  c->ctl.release ();
  drop_graph (c);
  if (c->hring) (void) hipHostFree (c->hring);
  if (c->hkey) (void) hipHostFree (c->hkey);
  for (hipEvent_t e : c->rev) if (e) (void) hipEventDestroy (e);
  for (hipEvent_t e : c->wev) if (e) (void) hipEventDestroy (e);
  prof_resolve (*c);
  for (hipEvent_t e : c->prof.pool) (void) hipEventDestroy (e);
  if (c->stream) (void) hipStreamDestroy (c->stream);
  delete c;
}

int pmx_upload (pmx_ctx *c, const pmx_model *m)
{
  if (!c || !m) return fail (PMX_EINVAL, "null argument");
  int rc = validate_model (m);
  if (rc) return rc;
  std::vector<InstOff> off (2);
  off[1] = model_counts (*m);
  return upload_core (c, m, off);
}

int pmx_download (pmx_ctx *c, pmx_model *m)
{
  if (!c || !m) return fail (PMX_EINVAL, "null argument");
  if (m->parnum != c->N || m->ellnum != c->E || m->trinum != c->T)
    return fail (PMX_EINVAL, "model counts differ from the uploaded state");
  (void) hipSetDevice (c->device);
  size_t N = c->N, E = c->E, T = c->T;
  int rc = 0;
  if ((rc = down (*c, m->angular, c->angular, 6*N)) || (rc = down (*c, m->linear, c->linear, 3*N)) ||
      (rc = down (*c, m->rotation, c->rotation, 9*N)) || (rc = down (*c, m->position, c->position, 6*N)) ||
      (rc = down (*c, m->inverse, c->inverse, 9*N)) || (rc = down (*c, m->invm, c->invm, N)) ||
      (rc = down (*c, m->force, c->force, 3*N)) || (rc = down (*c, m->torque, c->torque, 3*N)) ||
      (rc = down (*c, m->center, c->center, 6*E)) || (rc = down (*c, m->tri, c->tri, 9*T)) ||
      (rc = down (*c, m->nodes, c->nodes, 6*(size_t) c->NOD)) || (rc = springs_download (*c, m)) ||
      (rc = jt_download (*c, m)))
    return rc;
  if (m->obsnum == c->O && c->O > 0 &&
      ((rc = down (*c, m->obspnt, c->obspnt, 3*(size_t) c->O)) || (rc = down (*c, m->obsang, c->obsang, 3*(size_t) c->O)) ||
       (rc = down (*c, m->obslin, c->obslin, 3*(size_t) c->O))))
    return rc;
  PMX_HIP (hipStreamSynchronize (c->stream));
  m->curtime = c->curtime; m->curstep = c->curstep; m->stepnum = c->stepnum;
  return PMX_OK;
}

/* build the partitioning tree and its per-leaf lists from the current
 * centres (partitioning_create, partition.ispc:663-746) */
static int rebuild (pmx_ctx *c, bool all)
{
  int rc;
  ProfScope ps (*c, K_TREE, 0.0);
  if ((rc = tree_build (*c, all)) || (rc = tree_trilists (*c)) || (rc = tree_nbrlists (*c)) || (rc = step_prepare (*c)))
    return rc;
  PMX_HIP (hipMemsetAsync (c->ovf.p, 0, sizeof (int) * std::max (c->I, 1), c->stream));
  int zero[2] = {0, 0};
  PMX_HIP (hipMemcpyAsync (c->ctl.p + CTL_HINT, zero, sizeof (int), hipMemcpyHostToDevice, c->stream));
  PMX_HIP (hipMemcpyAsync (c->ctl.p + CTL_ABORT, zero, sizeof (int), hipMemcpyHostToDevice, c->stream));
  PMX_HIP (hipStreamSynchronize (c->stream));
  c->graph_valid = false;
  c->need_bin = true;   /* new leaves: the records are binned again */
  return PMX_OK;
}

/* the error flags of a group of steps (capacity overflows are not errors:
 * they abort their step, which is repeated with larger buffers) */
static int check_flags (pmx_ctx *c, const int *fl)
{
  if (fl[FLAG_TRILIST_OVERFLOW])
    return fail (PMX_ECAPACITY, "partitioning tree deeper than the traversal stack (64 pending nodes)");
  if (fl[FLAG_BAD_KIND]) return fail (PMX_EUNSUPPORTED, "unsupported interaction kind or inconsistent contact rows");
  if (fl[FLAG_JOINT_PIVOT])
    return fail (PMX_EINVAL, "zero pivot in the ball joint system (skyline LU, as amgcl's precondition)");
  c->stats.max_row = std::max (c->stats.max_row, fl[FLAG_MAXROW]);
  return PMX_OK;
}

/* largest per-sphere row, hit, reverse-hit and per-leaf moving triangle
 * capacities: far above any packing of spheres (a sphere can touch at most
 * a few hundred others of a quarter of its radius) */
#define PMX_CAP_LIMIT 4096

/* a step aborted by a contact capacity overflow: the buffers that
 * overflowed grow (doubled, or to the longest row the step formed) and the
 * step is repeated.  The reference's contact lists grow without bound
 * (CONBUF chunks appended by newcon, condet.ispc:52-81); here the rows are
 * dense, so they are re-laid out at the new capacity.  The rows of the
 * previous step (and those of instances frozen at other parities) are in
 * both ping-pong buffers and are kept; hits, reverse hits and per-step
 * lists are recomputed by the repeated step, their counts reset here to the
 * zeros every step starts from. */
static int grow_capacity (pmx_ctx *c, const int *fl)
{
  auto pow2 = [] (int v) { int p = 1; while (p < v) p <<= 1; return p; };
  size_t E = std::max (c->E, 1);
  int cap2 = c->cap, hcap2 = c->hcap, xcap2 = c->xcap, mcap2 = c->mcap;
  if (fl[FLAG_ROW_OVERFLOW]) cap2 = std::max (2 * c->cap, pow2 (fl[FLAG_MAXROW]));
  if (fl[FLAG_HIT_OVERFLOW]) hcap2 = 2 * c->hcap;
  if (fl[FLAG_EXTRA_OVERFLOW]) xcap2 = 2 * c->xcap;
  if (fl[FLAG_MTRI_OVERFLOW]) mcap2 = 2 * c->mcap;
  int tcap2 = fl[FLAG_TREV_OVERFLOW] ? 2 * c->tcap : c->tcap;
  if (cap2 > PMX_CAP_LIMIT || hcap2 > PMX_CAP_LIMIT || xcap2 > PMX_CAP_LIMIT || mcap2 > PMX_CAP_LIMIT ||
      tcap2 > PMX_CAP_LIMIT)
    return fail (PMX_ECAPACITY, "contact capacity beyond " + std::to_string (PMX_CAP_LIMIT) + " per sphere (row " +
                 std::to_string (fl[FLAG_MAXROW]) + ")");
  if (c->opt.verbose)
    fprintf (stderr, "pmx: contact capacity grows: rows %d -> %d, hits %d -> %d, reverse %d -> %d, obstacle lists %d -> %d\n",
             c->cap, cap2, c->hcap, hcap2, c->xcap, xcap2, c->mcap, mcap2);
  PMX_HIP (hipStreamSynchronize (c->stream));
  if (cap2 != c->cap)
  {
    bool mixed = c->opt.precision == PMX_MIXED;
    size_t esz = mixed ? sizeof (float) : sizeof (double);
    for (int b = 0; b < 2; b++)
    {
      dvec<uint32_t> key;
      if (key.alloc ((size_t) cap2 * E)) return fail (PMX_ENOMEM, "out of device memory (contact rows)");
      PMX_HIP (hipMemcpyAsync (key.p, c->rowkey[b].p, sizeof (uint32_t) * c->cap * E, hipMemcpyDeviceToDevice, c->stream));
      void *old = mixed ? (void*) c->rowst32[b].p : (void*) c->rowst[b].p;
      void *nst = nullptr;
      if (hipMalloc (&nst, esz * 3 * cap2 * E) != hipSuccess) return fail (PMX_ENOMEM, "out of device memory (contact rows)");
      for (int k = 0; k < 3; k++)
        PMX_HIP (hipMemcpyAsync ((char*) nst + esz * k * cap2 * E, (const char*) old + esz * k * c->cap * E,
                                 esz * c->cap * E, hipMemcpyDeviceToDevice, c->stream));
      PMX_HIP (hipStreamSynchronize (c->stream));
      std::swap (c->rowkey[b].p, key.p);
      std::swap (c->rowkey[b].n, key.n);
      if (mixed) { c->rowst32[b].release (); c->rowst32[b].p = (float*) nst; c->rowst32[b].n = 3 * cap2 * E; }
      else { c->rowst[b].release (); c->rowst[b].p = (double*) nst; c->rowst[b].n = 3 * cap2 * E; }
    }
    c->cap = cap2;
    if (c->opt.critical && c->crec.alloc (8 * (size_t) cap2 * E))   /* rewritten by the repeated step */
      return fail (PMX_ENOMEM, "out of device memory (contact records)");
  }
  if (tcap2 != c->tcap)   /* a meshed particle's triangle with more contacts than slots: slots of the
                             repeated step */
  {
    size_t T = std::max (c->T, 1);
    if (c->trec_e.alloc ((size_t) tcap2 * T) || c->trec.alloc (6 * (size_t) tcap2 * T))
      return fail (PMX_ENOMEM, "out of device memory (mesh reactions)");
    c->tcap = tcap2;
  }
  if (hcap2 != c->hcap)
  {
    c->hits.release ();
    if (c->hits.alloc ((size_t) hcap2 * E)) return fail (PMX_ENOMEM, "out of device memory (hits)");
    c->hcap = hcap2;
  }
  if (xcap2 != c->xcap)
  {
    c->xbuf.release ();
    if (c->xbuf.alloc ((size_t) xcap2 * E)) return fail (PMX_ENOMEM, "out of device memory (reverse hits)");
    c->xcap = xcap2;
  }
  if (mcap2 != c->mcap)
  {
    c->mcap = mcap2;
    c->mtlist.release ();
    if (c->nmtri > 0 && c->mtlist.alloc ((size_t) std::max (c->tree.nleaves, 1) * mcap2))
      return fail (PMX_ENOMEM, "out of device memory (obstacle lists)");
  }
  /* every step starts from zero hit, reverse-hit and leaf counts */
  PMX_HIP (hipMemsetAsync (c->hitcnt.p, 0, sizeof (int) * E, c->stream));
  PMX_HIP (hipMemsetAsync (c->xcnt.p, 0, sizeof (int) * E, c->stream));
  int rc = clear_leaf_counts (*c);
  if (rc) return rc;
  int zero = 0;
  PMX_HIP (hipMemcpyAsync (c->ctl.p + CTL_ABORT, &zero, sizeof (int), hipMemcpyHostToDevice, c->stream));
  PMX_HIP (hipStreamSynchronize (c->stream));
  c->graph_valid = false;
  c->need_bin = true;   /* the leaf counts were cleared */
  return PMX_OK;
}

/* the overflow flags of an aborted step, handled: back to zero */
static int clear_overflow_flags (pmx_ctx *c)
{
  const int which[] = {FLAG_LEAF_OVERFLOW, FLAG_ROW_OVERFLOW, FLAG_HIT_OVERFLOW, FLAG_EXTRA_OVERFLOW, FLAG_MTRI_OVERFLOW};
  for (int f : which) PMX_HIP (hipMemsetAsync (c->dflags.p + f, 0, sizeof (int), c->stream));
  PMX_HIP (hipStreamSynchronize (c->stream));
  return PMX_OK;
}

/* event record nodes around the profiled scopes of the captured chain:
 * A after the node before the scope and ahead of the scope's first node, B
 * after the scope's last node and ahead of the node that follows it */
static bool place_prof_nodes (pmx_ctx *c)
{
  auto next = [&] (hipGraphNode_t n, hipGraphNode_t *out) -> bool {
    size_t k = 0;
    if (hipGraphNodeGetDependentNodes (n, nullptr, &k) != hipSuccess || k > 1) return false;
    *out = nullptr;
    return k == 0 || hipGraphNodeGetDependentNodes (n, out, &k) == hipSuccess;
  };
  size_t nn = 0;
  if (hipGraphGetRootNodes (c->pgraph, nullptr, &nn) != hipSuccess || nn != 1) return false;
  hipGraphNode_t root = nullptr;
  if (hipGraphGetRootNodes (c->pgraph, &root, &nn) != hipSuccess) return false;
  /* first and following nodes of every scope, from the chain as captured */
  std::vector<hipGraphNode_t> first (c->prof.cap.size ()), follow (c->prof.cap.size ());
  for (size_t k = 0; k < c->prof.cap.size (); k++)
  {
    Prof::Cap &q = c->prof.cap[k];
    if (q.before) { if (!next (q.before, &first[k]) || !first[k]) return false; }
    else first[k] = root;
    if (!next (q.after, &follow[k])) return false;
  }
  for (size_t k = 0; k < c->prof.cap.size (); k++)
  {
    Prof::Cap &q = c->prof.cap[k];
    if (hipEventCreateWithFlags (&q.a, hipEventDisableSystemFence) != hipSuccess ||
        hipEventCreateWithFlags (&q.b, hipEventDisableSystemFence) != hipSuccess)
      return false;
    if (hipGraphAddEventRecordNode (&q.na, c->pgraph, q.before ? &q.before : nullptr, q.before ? 1 : 0, q.a) != hipSuccess ||
        hipGraphAddDependencies (c->pgraph, &q.na, &first[k], 1) != hipSuccess ||
        hipGraphAddEventRecordNode (&q.nb, c->pgraph, &q.after, 1, q.b) != hipSuccess ||
        (follow[k] && hipGraphAddDependencies (c->pgraph, &q.nb, &follow[k], 1) != hipSuccess))
      return false;
  }
  return true;
}

/* capture one step as a graph; with profiling mode 1, event record nodes
 * are put around each profiled kernel so that fresh events can be set on
 * them before every launch (if they cannot be placed, steps are enqueued
 * eagerly instead) */
static int capture (pmx_ctx *c, double h, bool multi)
{
  drop_graph (c);
  if (multi && c->prof.mode == 0)
  {
    /* PMX_GRAPH_STEPS steps in one graph: consecutive launches of a short
     * step each cost the runtime's launch gap, one per graph instead */
    hipGraph_t g = nullptr;
    PMX_HIP (hipStreamBeginCapture (c->stream, hipStreamCaptureModeThreadLocal));
    int rc = PMX_OK;
    for (int k = 0; k < PMX_GRAPH_STEPS && rc == PMX_OK; k++) rc = enqueue_step (*c, h);
    hipError_t e = hipStreamEndCapture (c->stream, &g);
    if (rc) { if (g) (void) hipGraphDestroy (g); return rc; }
    PMX_HIP (e);
    e = hipGraphInstantiate (&c->gexec_multi, g, nullptr, nullptr, 0);
    (void) hipGraphDestroy (g);
    PMX_HIP (e);
  }
  for (int pass = 0; pass < (c->prof.mode == 1 ? 2 : 1); pass++)
  {
    hipGraph_t &g = pass ? c->pgraph : c->graph;
    int mode = c->prof.mode;
    if (pass == 0) c->prof.mode = 0;   /* the plain step: no scopes */
    c->prof.capturing = pass == 1;
    c->prof.err = hipSuccess;
    PMX_HIP (hipStreamBeginCapture (c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step (*c, h);
    hipError_t e = hipStreamEndCapture (c->stream, &g);
    c->prof.capturing = false;
    c->prof.mode = mode;
    if (rc) { drop_graph (c); return rc; }
    PMX_HIP (e);
  }
  if (c->prof.mode == 1)
  {
    c->prof.eager = c->prof.err != hipSuccess || !place_prof_nodes (c);
    (void) hipGetLastError ();
    if (c->prof.eager) { drop_graph (c); return PMX_OK; }
    PMX_HIP (hipGraphInstantiate (&c->pgexec, c->pgraph, nullptr, nullptr, 0));
  }
  PMX_HIP (hipGraphInstantiate (&c->gexec, c->graph, nullptr, nullptr, 0));
  c->graph_valid = true;
  return PMX_OK;
}

/* one replay of the captured step; with profiling mode 1 every
 * PMX_PROF_EVERY-th step replays the copy with event record nodes, which
 * get fresh events first (a launch keeps the events it was given) */
static int launch_step (pmx_ctx *c, int step)
{
  if (!c->pgexec || step % PMX_PROF_EVERY || !c->prof.in_window (step))
  {
    PMX_HIP (hipGraphLaunch (c->gexec, c->stream));
    return PMX_OK;
  }
  for (auto &q : c->prof.cap)
  {
    hipEvent_t a = c->prof.get (), b = c->prof.get ();
    PMX_HIP (hipGraphExecEventRecordNodeSetEvent (c->pgexec, q.na, a));
    PMX_HIP (hipGraphExecEventRecordNodeSetEvent (c->pgexec, q.nb, b));
    c->prof.pend.push_back ({q.kid, a, b, q.bytes, step});
  }
  PMX_HIP (hipGraphLaunch (c->pgexec, c->stream));
  return PMX_OK;
}

} /* extern "C" */

/* the instance has completed its steps of the call: its time bookkeeping
 * (parmec.cpp:2418, 2521-2523) and the row buffer holding its rows */
static void finish_instance (pmx_ctx *c, int i, int nst, double step0, double step1)
{
  for (int k = 0; k < nst; k++) c->icurtime[i] += 0.5*(step0+step1);
  c->icurstep[i] = step1;
  c->istepnum[i]++;
  c->icur[i] = (c->cur0 + nst) & 1;
}

/* dem() over every instance of the context.  durations: one per instance
 * (instances then run different step counts; each is frozen once its own
 * steps are done), or null for `duration` everywhere.  fc holds nfc
 * entries: one forcing for all instances, or one per instance.  done(user,
 * i) is called from inside the call as soon as instance i has finished. */
int dem_core (pmx_ctx *c, const double *durations, double duration, double step, const pmx_forcing *fc, int nfc,
              void (*done_fn) (void*, int), void *user, double *elapsed, DemExtra *x)
{
  if (!c) return fail (PMX_EINVAL, "null ctx");
  if (!c->uploaded) return fail (PMX_EINVAL, "pmx_upload must precede pmx_dem");
  if (!(step > 0.0)) return fail (PMX_EINVAL, "step must be positive");
  int I = c->I;
  const int *steps_in = x ? x->steps : nullptr;
  if (x) x->paused_at = -1;
  /* instances at time 0 and resumed ones (given step counts only): the
   * resumed ones must have taken this step */
  int nfresh = 0;
  for (int i = 0; i < I; i++)
  {
    if (steps_in)
    {
      if (steps_in[i] < 1) return fail (PMX_EINVAL, "step counts must be positive");
      if (c->icurtime[i] != 0.0 && c->icurstep[i] != step)
        return fail (PMX_EINVAL, "a resumed instance steps with the step it has taken so far");
      nfresh += c->icurtime[i] == 0.0;
      continue;
    }
    double d = durations ? durations[i] : duration;
    if (!(d > 0.0)) return fail (PMX_EINVAL, "durations must be positive");
    if (c->icurtime[i] != c->icurtime[0] || c->icurstep[i] != c->icurstep[0])
      return fail (PMX_EINVAL, "instances are at different times after a call with different durations; "
                               "download them and upload a new batch");
  }
  const bool mixed_start = steps_in && nfresh > 0 && nfresh < I;
  if (steps_in && c->sp_offsets) return fail (PMX_EUNSUPPORTED, "step counts with spring stroke offsets");
  if (mixed_start && (c->O > 0 || c->RS > 0))
    return fail (PMX_EUNSUPPORTED, "resumed instances beside new ones: moving obstacles and restraints are not supported");
  (void) hipSetDevice (c->device);
  auto t0 = std::chrono::steady_clock::now ();
  int rc = 0;
  double step0, step1;
  const double curtime = steps_in ? (nfresh > 0 ? 0.0 : c->icurtime[0]) : c->icurtime[0];
  int ctl0[CTL_COUNT] = {0};
  PMX_HIP (hipMemcpyAsync (c->ctl.p, ctl0, sizeof (ctl0), hipMemcpyHostToDevice, c->stream));
  PMX_HIP (hipMemsetAsync (c->dflags.p, 0, FLAG_COUNT * sizeof (int), c->stream));
  c->staggered = false;
  c->first_sel = false;
  if (mixed_start)
  {
    /* euler and shapes of the first call only for the instances at time 0
     * (the resumed ones took theirs in their first context) */
    std::vector<int> fresh (I);
    for (int i = 0; i < I; i++) fresh[i] = c->icurtime[i] == 0.0;
    if ((rc = up (*c, c->isel, fresh.data (), I, "instance selection"))) return rc;
    c->first_sel = true;
  }
  if (curtime == 0.0)
  {
    step0 = step;
    rc = step_euler (*c, 0.5*step0);
    if (!rc) rc = step_shapes (*c);
    c->first_sel = false;
    if (rc || (rc = step_mesh_shapes (*c, false))) return rc;
    if (c->O > 0)
    {
      /* obstaclev (step0/2) and obstacles (step0), parmec.cpp:2403-2405 */
      std::vector<double> t0 (6 * (size_t) c->O, 0.0);
      for (int i = 0; i < I; i++)
      {
        const pmx_forcing *f = fc ? fc + (nfc > 1 ? i : 0) : nullptr;
        int o0 = c->off[i].obs, no = c->off[i+1].obs - o0;
        if (f && f->obstab0) std::memcpy (&t0[6*(size_t) o0], f->obstab0, 6*sizeof (double)*no);
      }
      if ((rc = up (*c, c->obstab, t0.data (), t0.size (), "obstacle table")) ||
          (rc = step_obstacles (*c, c->obstab.p, 0, step0)))
        return rc;
    }
  }
  else step0 = c->icurstep[0];
  step1 = step0;
  /* invert_inertia, then restrain_velocities (parmec.cpp:2409-2411) */
  if ((rc = step_invert (*c)) || (rc = step_restrain_vel (*c))) return rc;

  /* steps of each instance (the reference's accumulation of time by (step0+step1)/2, parmec.cpp:2418) */
  std::vector<int> nst (I, 0);
  int nsteps = 0;
  for (int i = 0; i < I; i++)
  {
    if (steps_in) nst[i] = steps_in[i];
    else pmx_step_times (durations ? durations[i] : duration, step0, curtime, nullptr, 0, &nst[i]);
    nsteps = std::max (nsteps, nst[i]);
  }
  for (int i = 0; i < I; i++) c->staggered = c->staggered || nst[i] != nsteps;
  if (c->staggered && (rc = up (*c, c->istop, nst.data (), I, "instance steps"))) return rc;
  if (c->sp_offsets)
  {
    /* the curtime of each step, where spring stroke offsets read their load
     * curves (one model: offsets are refused in batches) */
    std::vector<double> tt (std::max (nsteps, 1), 0.0);
    int n = 0;
    pmx_step_times (durations ? durations[0] : duration, step0, curtime, tt.data (), nsteps, &n);
    if ((rc = step_times_table (*c, tt.data (), nsteps))) return rc;
  }

  /* per-step tables of the host-evaluated inputs (input.cpp:4424-4503),
   * laid out [step][instance][component]; rows past an instance's own steps
   * are not read */
  {
    size_t ns = std::max (nsteps, 1);
    std::vector<double> g (3 * ns * I, 0.0), d (6 * ns * I, 0.0);
    for (int i = 0; i < I; i++)
    {
      const pmx_forcing *f = fc ? fc + (nfc > 1 ? i : 0) : nullptr;
      for (int k = 0; k < nst[i]; k++)
      {
        for (int q = 0; q < 3; q++) g[3*((size_t) k*I + i) + q] = f && f->gtab ? f->gtab[3*(size_t) k + q] : (f ? f->gravity[q] : 0.0);
        for (int q = 0; q < 6; q++) d[6*((size_t) k*I + i) + q] = f && f->dtab ? f->dtab[6*(size_t) k + q] : (f ? f->damping[q] : 0.0);
      }
    }
    if ((rc = up (*c, c->gtab, g.data (), g.size (), "gravity table")) || (rc = up (*c, c->dtab, d.data (), d.size (), "damping table")))
      return rc;
    c->damped = std::any_of (d.begin (), d.end (), [] (double v) { return v != 0.0; });
    if (c->PRSC > 0)
    {
      /* prescribed values of each history column at each step's curtime,
       * [step][column][6] */
      size_t P = c->PRSC;
      std::vector<double> pt (6 * ns * P, 0.0);
      for (int i = 0; i < I; i++)
      {
        const pmx_forcing *f = fc ? fc + (nfc > 1 ? i : 0) : nullptr;
        int p0 = c->off[i].pcol, np = c->off[i+1].pcol - p0;
        if (!f || !f->prstab || np == 0) continue;
        int stride = f->prscols > 0 ? f->prscols : np;
        if (stride < np) return fail (PMX_EINVAL, "prescribed table has fewer columns than the model's histories");
        for (int k = 0; k < nst[i]; k++)
          std::memcpy (&pt[6*((size_t) k*P + p0)], f->prstab + 6*(size_t) k*stride, 6*sizeof (double)*np);
      }
      if ((rc = up (*c, c->prstab, pt.data (), pt.size (), "prescribed table"))) return rc;
    }
    if (c->O > 0)
    {
      /* obstaclev at curtime+step0 of every step, [step][obstacle][6] */
      size_t O = c->O;
      std::vector<double> ot (6 * ns * O, 0.0);
      for (int i = 0; i < I; i++)
      {
        const pmx_forcing *f = fc ? fc + (nfc > 1 ? i : 0) : nullptr;
        int o0 = c->off[i].obs, no = c->off[i+1].obs - o0;
        if (!f || !f->obstab || no == 0) continue;
        for (int k = 0; k < nst[i]; k++)
          std::memcpy (&ot[6*((size_t) k*O + o0)], f->obstab + 6*(size_t) k*no, 6*sizeof (double)*no);
      }
      if ((rc = up (*c, c->obstab, ot.data (), ot.size (), "obstacle table"))) return rc;
    }
  }
  /* history samples (one model): the row of each step's end state */
  c->his_active = I == 1 && c->nhis > 0 && fc && fc->hisrow && fc->hisrows > 0 && fc->hisout;
  if (c->his_active)
  {
    std::vector<int> hr (std::max (nsteps, 1), -1);
    for (int k = 0; k < nsteps; k++)
    {
      int r = fc->hisrow[k];
      if (r >= fc->hisrows) return fail (PMX_EINVAL, "history row beyond hisrows");
      hr[k] = r < 0 ? -1 : r;
    }
    if ((rc = up (*c, c->hisrow, hr.data (), hr.size (), "history rows"))) return rc;
    if (c->hisbuf.alloc ((size_t) fc->hisrows * c->nhis) != hipSuccess) return fail (PMX_ENOMEM, "out of device memory (histories)");
    PMX_HIP (hipStreamSynchronize (c->stream));
  }
  c->cur0 = c->cur;
  c->graph_valid = false;
  c->stats.max_row = 0;

  long long rebuilds = 0, partial = 0;
  c->fuse_bin = false;
  if (c->E > 0)
  {
    if (c->snap_sel)
    {
      /* resumed instances brought the centres their trees were built from
       * (import_rows): the build takes new snapshots of the others only */
      std::vector<int> fresh (I);
      for (int i = 0; i < I; i++) fresh[i] = c->icurtime[i] == 0.0;
      if ((rc = up (*c, c->isel, fresh.data (), I, "instance selection"))) return rc;
    }
    rc = rebuild (c, true);
    c->snap_sel = false;
    if (rc) return rc;
    rebuilds++;
    /* each step's dynamics bins the centres for the next (slotted records;
     * PMX_FUSED_BIN=0 keeps a separate k_bin at every step's start) */
    const char *fb = getenv ("PMX_FUSED_BIN");
    c->fuse_bin = c->lstride > 0 && !(fb && fb[0] == '0');
    c->bin_until = nsteps;
  }
  /* block shapes for this workload (large models only: small ones are
   * bound by launch latency, not by the shapes): a first pass on the state
   * of the first call, and a second once the context has taken
   * PMX_RETUNE_STEPS steps, when the contact rows have settled (the first
   * steps have no contact of a previous step) */
  auto tune = [&] (int kstep) -> int {
    int r = autotune (*c, step0);
    if (r) return r;
    int z[CTL_COUNT] = {0};   /* hint off: the next bin descends from the roots */
    z[CTL_KSTEP] = kstep;
    PMX_HIP (hipMemcpyAsync (c->ctl.p, z, sizeof (z), hipMemcpyHostToDevice, c->stream));
    PMX_HIP (hipStreamSynchronize (c->stream));
    if ((r = clear_overflow_flags (c))) return r;   /* the timing runs' detections are discarded */
    c->graph_valid = false;
    c->need_bin = true;   /* the tuning runs cleared the leaf counts */
    c->stats.det_waves = c->det_nw;
    c->stats.rows_block = c->rows_tpb;
    return PMX_OK;
  };
  bool tuning = c->opt.tune == 0 && c->E >= 50000;
  if (tuning && c->tuned == 0)
  {
    if ((rc = tune (0))) return rc;
    PMX_HIP (hipMemsetAsync (c->dflags.p, 0, FLAG_COUNT * sizeof (int), c->stream));
    PMX_HIP (hipStreamSynchronize (c->stream));
  }
  c->stats.det_waves = c->det_nw;
  c->stats.rows_block = c->rows_tpb;
  auto tl = std::chrono::steady_clock::now ();
  std::vector<char> reported (I, 0);
  auto report = [&] (int upto) -> int {
    bool any = false;
    for (int i = 0; i < I; i++) any = any || (!reported[i] && nst[i] <= upto);
    if (!any) return PMX_OK;
    /* forces and torques that the step formed where they were used */
    std::vector<int> fin (I, -1);
    for (int i = 0; i < I; i++) if (!reported[i] && nst[i] <= upto && nst[i] > 0) fin[i] = nst[i] - 1;
    int r = step_materialize (*c, fin);
    if (r) return r;
    /* mesh geometry of the current particle state (finished instances are
     * frozen at their final state) */
    r = step_mesh_shapes (*c, false);
    if (r) return r;
    PMX_HIP (hipStreamSynchronize (c->stream));
    for (int i = 0; i < I; i++)
      if (!reported[i] && nst[i] <= upto)
      {
        finish_instance (c, i, nst[i], step0, step1);
        reported[i] = 1;
        if (done_fn) done_fn (user, i);
      }
    return PMX_OK;
  };
  /* Steps go to the device in groups: one captured step replayed as a
   * graph per step, or, with profiling on, each step's kernels enqueued
   * eagerly with an event pair per launch.  After each group the control
   * block and the flags are copied to pinned host memory, and up to
   * PMX_LOOKAHEAD groups are enqueued ahead of the one whose copy the host
   * waits for, so the device does not wait for the host between groups.  A
   * leaf overflow aborts the step it occurs in; the groups enqueued behind
   * it return at once.  Large models take groups of one step (an abort then
   * wastes little), small ones longer groups (per-step copies would cost
   * more than their steps).  The host stops at the steps where instances
   * finish. */
  const int gsize = c->E >= 1000000 ? 1 : c->E >= 100000 ? 8 : 64;
  const int slotn = CTL_COUNT + FLAG_COUNT;
  if (!c->hring)
  {
    if (hipHostMalloc ((void**) &c->hring, sizeof (int) * slotn * PMX_RING) != hipSuccess)
    { c->hring = nullptr; return fail (PMX_ENOMEM, "out of pinned host memory"); }
    for (int k = 0; k < PMX_RING; k++) PMX_HIP (hipEventCreateWithFlags (&c->rev[k], hipEventDisableTiming));
  }
  /* output points: the steps after which the call pauses (the first
   * forcing's list; a batch pauses every instance at the same step) */
  const bool outs = fc && fc->noutstep > 0 && fc->outstep && fc->out_fn;
  int oi = 0;
  int done = 0;
  long long grown = 0;
  std::vector<int> rebuilt_at (I, -1);   /* the last step each instance was re-partitioned for */
  bool pause_req = false;                /* the caller's stop_fn asked for a pause: no more groups */
  while (done < nsteps)
  {
    int next = nsteps;
    for (int i = 0; i < I; i++) if (!reported[i] && nst[i] > done) next = std::min (next, nst[i]);
    if (outs)
    {
      while (oi < fc->noutstep && fc->outstep[oi] + 1 <= done) oi++;
      if (oi < fc->noutstep) next = std::min (next, fc->outstep[oi] + 1);
    }
    if (c->prof.mode != 2 && !c->prof.eager && !c->graph_valid && (rc = capture (c, step0, gsize >= PMX_GRAPH_STEPS)))
      return rc;
    bool eager = c->prof.mode == 2 || (c->prof.mode == 1 && c->prof.eager);
    int ctl[CTL_COUNT] = {0}, fl[FLAG_COUNT] = {0};
    int enq = done, gq = 0, gs = 0;   /* first step not enqueued; groups enqueued, seen */
    int gend[PMX_RING] = {0};         /* first step after each group in flight */
    int err = PMX_OK;
    bool abort = false;
    for (;;)
    {
      while (!pause_req && enq < next && gq - gs < PMX_LOOKAHEAD)
      {
        int n = std::min (gsize, next - enq);
        int q = 0;
        if (c->fuse_bin && c->need_bin)
        {
          /* the records of step `enq`, where no dynamics has binned them
           * (an overflow here aborts that step, as in an unfused step) */
          if ((rc = step_bin (*c))) return rc;
          c->need_bin = false;
        }
        if (!eager && c->gexec_multi && c->win[0] < 0)
          for (; q + PMX_GRAPH_STEPS <= n; q += PMX_GRAPH_STEPS) PMX_HIP (hipGraphLaunch (c->gexec_multi, c->stream));
        for (; q < n; q++)
        {
          /* the device-timed window (pmx_set_window): an event before its
           * first step and after its last, recorded again when a step is
           * repeated */
          if (enq + q == c->win[0]) PMX_HIP (hipEventRecord (c->wev[0], c->stream));
          if (eager)
          {
            c->prof.step = enq + q;
            rc = enqueue_step (*c, step0);
            c->prof.step = -1;
          }
          else
            rc = launch_step (c, enq + q);
          if (rc) return rc;
          if (enq + q == c->win[1]) PMX_HIP (hipEventRecord (c->wev[1], c->stream));
        }
        int *slot = c->hring + (gq % PMX_RING) * slotn;
        PMX_HIP (hipMemcpyAsync (slot, c->ctl.p, sizeof (int) * CTL_COUNT, hipMemcpyDeviceToHost, c->stream));
        PMX_HIP (hipMemcpyAsync (slot + CTL_COUNT, c->dflags.p, sizeof (int) * FLAG_COUNT, hipMemcpyDeviceToHost,
                                 c->stream));
        PMX_HIP (hipEventRecord (c->rev[gq % PMX_RING], c->stream));
        enq += n;
        gend[gq % PMX_RING] = enq;
        gq++;
      }
      if (gs == gq) break;
      PMX_HIP (hipEventSynchronize (c->rev[gs % PMX_RING]));
      const int *slot = c->hring + (gs % PMX_RING) * slotn;
      std::memcpy (ctl, slot, sizeof (ctl));
      std::memcpy (fl, slot + CTL_COUNT, sizeof (fl));
      if (!ctl[CTL_ABORT]) prof_resolve (*c, 0x7FFFFFFF, gend[gs % PMX_RING]);   /* the group's timings */
      gs++;
      if ((err = check_flags (c, fl))) break;
      if ((abort = ctl[CTL_ABORT] != 0)) break;
      if (x && x->stop_fn && !pause_req && x->stop_fn (x->stop_user, ctl[CTL_KSTEP])) pause_req = true;
    }
    PMX_HIP (hipStreamSynchronize (c->stream));   /* the groups behind an abort */
    prof_resolve (*c, ctl[CTL_KSTEP]);
    if (err) return err;
    done = ctl[CTL_KSTEP];
    if (abort)
    {
      /* step `done` was aborted, before it changed any state.  A contact
       * capacity overflow: larger buffers, then the step again.  A leaf
       * overflow: re-partition the overflowing instances from the current
       * centres and repeat the step (parmec.cpp:2420-2427); "Repartitioning
       * failed" when an instance already re-partitioned for this step
       * overflows again, as it would alone (every instance that overflows
       * in a step is flagged by that step's binning) */
      if (fl[FLAG_ROW_OVERFLOW] || fl[FLAG_HIT_OVERFLOW] || fl[FLAG_EXTRA_OVERFLOW] || fl[FLAG_MTRI_OVERFLOW] ||
          fl[FLAG_TREV_OVERFLOW])
      {
        if ((rc = grow_capacity (c, fl))) return rc;
        grown++;
      }
      if (fl[FLAG_LEAF_OVERFLOW])
      {
        std::vector<int> ov (I, 0);
        PMX_HIP (hipMemcpyAsync (ov.data (), c->ovf.p, sizeof (int) * I, hipMemcpyDeviceToHost, c->stream));
        PMX_HIP (hipStreamSynchronize (c->stream));
        for (int i = 0; i < I; i++)
          if (ov[i] && rebuilt_at[i] == done)
            return fail (PMX_EREPART, "Repartitioning failed (instance " + std::to_string (i) + ", step " +
                         std::to_string (done) + ")");
        for (int i = 0; i < I; i++) if (ov[i]) rebuilt_at[i] = done;
        c->prof.step = done;   /* the rebuild is timed with the step it repeats */
        rc = rebuild (c, false);
        c->prof.step = -1;
        if (rc) return rc;
        rebuilds++;
        if (c->tree.last_partial) partial++;
      }
      if ((rc = clear_overflow_flags (c))) return rc;
    }
    if ((rc = report (done))) return rc;
    if (outs && !abort && oi < fc->noutstep && fc->outstep[oi] + 1 == done && done < nsteps)
    {
      /* an output point (the last step's output happens after the loop):
       * forces formed where they were used, mesh geometry, then the caller */
      std::vector<int> fin (I, -1);
      for (int i = 0; i < I; i++) if (!reported[i] && nst[i] >= done) fin[i] = done - 1;
      if ((rc = step_materialize (*c, fin)) || (rc = step_mesh_shapes (*c, false))) return rc;
      PMX_HIP (hipStreamSynchronize (c->stream));
      fc->out_fn (fc->out_user, fc->outstep[oi]);
      (void) hipSetDevice (c->device);
      oi++;
    }
    if (pause_req && done < nsteps) break;
    if (tuning && c->tuned == 1 && !abort && done < nsteps && c->stepped + done >= PMX_RETUNE_STEPS &&
        (rc = tune (done)))
      return rc;
  }
  const bool paused = done < nsteps;
  if (paused)
  {
    /* paused at step `done`: the unfinished instances keep the time of the
     * steps they took (the same additions, parmec.cpp:2418) and their rows,
     * and take no stepnum (their dem() has not returned) */
    for (int i = 0; i < I; i++)
      if (!reported[i])
      {
        for (int k = 0; k < done; k++) c->icurtime[i] += 0.5*(step0+step1);
        c->icurstep[i] = step1;
        c->icur[i] = (c->cur0 + done) & 1;
      }
    nsteps = done;
    if (x) x->paused_at = done;
  }
  else if ((rc = report (nsteps))) return rc;
  if (outs && I == 1 && !paused)
    for (; oi < fc->noutstep; oi++)
      if (fc->outstep[oi] == nsteps - 1)   /* the last step: the finished state */
      {
        c->curtime = c->icurtime[0]; c->curstep = c->icurstep[0]; c->stepnum = c->istepnum[0];
        c->cur = (c->cur0 + nsteps) & 1;
        fc->out_fn (fc->out_user, fc->outstep[oi]);
        (void) hipSetDevice (c->device);
      }
  if (c->his_active)
  {
    PMX_HIP (hipMemcpyAsync (fc->hisout, c->hisbuf.p, sizeof (double) * (size_t) fc->hisrows * c->nhis,
                             hipMemcpyDeviceToHost, c->stream));
    PMX_HIP (hipStreamSynchronize (c->stream));
    c->his_active = false;
  }
  c->stepped += nsteps;
  if (nsteps > 0) { c->crec_valid = c->opt.critical != 0; c->crec_stale = c->opt.critical == 0; }
  c->cur = (c->cur0 + nsteps) & 1;
  c->curtime = c->icurtime[0]; c->curstep = c->icurstep[0]; c->stepnum = c->istepnum[0];
  double dt = std::chrono::duration<double> (std::chrono::steady_clock::now () - t0).count ();
  c->stats.steps = nsteps;
  c->stats.rebuilds = rebuilds;
  c->stats.partial_rebuilds = partial;
  c->stats.grown = grown;
  c->stats.row_capacity = c->cap;
  c->stats.graph_timed = c->prof.mode == 1 && !c->prof.eager;
  c->stats.window_ms = -1.0;
  if (c->win[0] >= 0 && c->win[1] < nsteps)
  {
    float ms = 0.0f;
    if (hipEventElapsedTime (&ms, c->wev[0], c->wev[1]) == hipSuccess) c->stats.window_ms = ms;
  }
  c->win[0] = c->win[1] = -1;
  c->prof.win0 = c->prof.win1 = -1;
  c->stats.leaves = c->tree.nleaves;
  c->stats.leaf_slots = c->lstride;
  c->stats.seconds = std::chrono::duration<double> (std::chrono::steady_clock::now () - tl).count ();
  {
    long long tb[2] = {0, 0};
    if ((rc = row_totals (*c, tb))) return rc;
    c->stats.pairs = tb[0];  /* entries; sphere pairs are held twice */
    c->stats.busy = tb[1];
  }
  if (elapsed) *elapsed = dt;
  return PMX_OK;
}

/* dem() with the adaptive step (parmec.cpp:2315-2530, adaptive in (0, 1]).
 * Each step's length step1 comes from that step's forces, and the step's
 * dynamics needs it, so every step stops once on the host: the first part
 * of the step (binning to the forces, the per-particle estimates and their
 * minimum) is enqueued, the control block, the flags and the minimum are
 * read back, then the rest of the step is enqueued with step0 and step1.
 * The host inputs of step k come from forcing->step_fn at that step's
 * curtime; the tables the kernels read are grown as the steps go. */
static int adaptive_core (pmx_ctx *c, double duration, double step, double adaptive, const pmx_forcing *fc,
                          double *elapsed)
{
  if (!c) return fail (PMX_EINVAL, "null ctx");
  if (!c->uploaded) return fail (PMX_EINVAL, "pmx_upload must precede pmx_dem_adaptive");
  if (!(step > 0.0)) return fail (PMX_EINVAL, "step must be positive");
  if (!(duration > 0.0)) return fail (PMX_EINVAL, "duration must be positive");
  if (!(adaptive > 0.0 && adaptive <= 1.0)) return fail (PMX_EINVAL, "adaptive must lie in (0, 1]");
  if (c->I != 1) return fail (PMX_EUNSUPPORTED, "pmx_dem_adaptive takes single-model contexts");
  (void) hipSetDevice (c->device);
  auto t0 = std::chrono::steady_clock::now ();
  struct Restore
  {
    pmx_ctx *c; int prof;
    ~Restore () { c->adp_on = false; c->prof.mode = prof; }
  } restore {c, c->prof.mode};
  c->prof.mode = 0;   /* launches are eager here; the graph timing does not apply */
  int rc = 0;
  int ctl0[CTL_COUNT] = {0};
  PMX_HIP (hipMemcpyAsync (c->ctl.p, ctl0, sizeof (ctl0), hipMemcpyHostToDevice, c->stream));
  PMX_HIP (hipMemsetAsync (c->dflags.p, 0, FLAG_COUNT * sizeof (int), c->stream));
  c->staggered = false;
  double curtime = c->icurtime[0];
  double step0, step1;
  if (curtime == 0.0)
  {
    step0 = step;
    if ((rc = step_euler (*c, 0.5*step0)) || (rc = step_shapes (*c)) || (rc = step_mesh_shapes (*c, false))) return rc;
    if (c->O > 0)
    {
      /* obstaclev (step0/2) and obstacles (step0), parmec.cpp:2403-2405 */
      std::vector<double> v0 (6 * (size_t) c->O, 0.0);
      if (fc && fc->obstab0) std::memcpy (v0.data (), fc->obstab0, v0.size () * sizeof (double));
      if ((rc = up (*c, c->obstab, v0.data (), v0.size (), "obstacle table")) ||
          (rc = step_obstacles (*c, c->obstab.p, 0, step0)))
        return rc;
    }
  }
  else step0 = c->icurstep[0];
  step1 = step0;
  if ((rc = step_invert (*c)) || (rc = step_restrain_vel (*c))) return rc;
  if (c->adp.alloc (11 * (size_t) std::max (c->N, 1)) || (c->S > 0 && c->sp_adp.alloc (17 * (size_t) c->S)) ||
      c->hmin.alloc (1))
    return fail (PMX_ENOMEM, "out of device memory (adaptive step)");
  const int slotn = CTL_COUNT + FLAG_COUNT;
  if (!c->hring)
  {
    if (hipHostMalloc ((void**) &c->hring, sizeof (int) * slotn * PMX_RING) != hipSuccess)
    { c->hring = nullptr; return fail (PMX_ENOMEM, "out of pinned host memory"); }
    for (int k = 0; k < PMX_RING; k++) PMX_HIP (hipEventCreateWithFlags (&c->rev[k], hipEventDisableTiming));
  }
  if (!c->hkey && hipHostMalloc ((void**) &c->hkey, sizeof (unsigned long long)) != hipSuccess)
  { c->hkey = nullptr; return fail (PMX_ENOMEM, "out of pinned host memory"); }
  c->his_active = c->nhis > 0 && fc && fc->hisout && fc->hisrows > 0;
  if (c->his_active && c->hisbuf.alloc ((size_t) fc->hisrows * c->nhis) != hipSuccess)
    return fail (PMX_ENOMEM, "out of device memory (histories)");

  /* per-step tables, grown as steps are added; the host keeps every row */
  const size_t O = c->O, P = c->PRSC;
  std::vector<double> gh, dh, oh, ph, th;
  std::vector<int> hh;
  int tcap = 0;
  auto ensure = [&] (int k) -> int {
    if (k < tcap) return PMX_OK;
    int ncap = std::max (64, 2 * tcap);
    while (ncap <= k) ncap *= 2;
    gh.resize (3 * (size_t) ncap, 0.0); dh.resize (6 * (size_t) ncap, 0.0);
    oh.resize (6 * O * ncap, 0.0); ph.resize (6 * P * ncap, 0.0); hh.resize (ncap, -1); th.resize (ncap, 0.0);
    int r;
    if ((r = up (*c, c->gtab, gh.data (), gh.size (), "gravity table")) ||
        (r = up (*c, c->dtab, dh.data (), dh.size (), "damping table")) ||
        (O > 0 && (r = up (*c, c->obstab, oh.data (), oh.size (), "obstacle table"))) ||
        (P > 0 && (r = up (*c, c->prstab, ph.data (), ph.size (), "prescribed table"))) ||
        (c->his_active && (r = up (*c, c->hisrow, hh.data (), hh.size (), "history rows"))) ||
        (c->sp_offsets && (r = step_times_table (*c, th.data (), (int) th.size ()))))
      return r;
    tcap = ncap;
    return PMX_OK;
  };
  c->cur0 = c->cur;
  c->graph_valid = false;
  c->stats.max_row = 0;
  c->adp_on = true;
  c->fuse_bin = false;
  c->damped = true;      /* damping rows come step by step */   /* each step is enqueued in two parts: k_bin at its start */
  long long rebuilds = 0, partial = 0, grown = 0;
  if (c->E > 0)
  {
    if ((rc = rebuild (c, true))) return rc;
    rebuilds++;
  }
  std::vector<double> obsv (6 * std::max (O, (size_t) 1)), prsv (6 * std::max (P, (size_t) 1));
  double time = 0.0;
  int k = 0, rebuilt_at = -1;
  for (; time < duration; time += 0.5*(step0+step1), curtime += 0.5*(step0+step1), step0 = step1, k++)
  {
    pmx_step_inputs in;
    for (int q = 0; q < 3; q++) in.gravity[q] = fc ? fc->gravity[q] : 0.0;
    for (int q = 0; q < 6; q++) in.damping[q] = fc ? fc->damping[q] : 0.0;
    std::fill (obsv.begin (), obsv.end (), 0.0);
    std::fill (prsv.begin (), prsv.end (), 0.0);
    in.obsvel = obsv.data ();
    in.prs = prsv.data ();
    in.hisrow = -1;
    in.output = 0;
    if (fc && fc->step_fn)
    {
      int r = fc->step_fn (fc->step_user, k, curtime, step0, &in);
      (void) hipSetDevice (c->device);
      if (r) return fail (PMX_EINVAL, "step_fn failed at step " + std::to_string (k));
    }
    if (in.hisrow >= 0 && (!c->his_active || in.hisrow >= fc->hisrows))
      return fail (PMX_EINVAL, "history row beyond hisrows");
    if ((rc = ensure (k))) return rc;
    std::memcpy (&gh[3*(size_t) k], in.gravity, 3 * sizeof (double));
    std::memcpy (&dh[6*(size_t) k], in.damping, 6 * sizeof (double));
    PMX_HIP (hipMemcpyAsync (c->gtab.p + 3*(size_t) k, &gh[3*(size_t) k], 3 * sizeof (double), hipMemcpyHostToDevice, c->stream));
    PMX_HIP (hipMemcpyAsync (c->dtab.p + 6*(size_t) k, &dh[6*(size_t) k], 6 * sizeof (double), hipMemcpyHostToDevice, c->stream));
    if (O > 0)
    {
      std::memcpy (&oh[6*O*k], obsv.data (), 6 * O * sizeof (double));
      PMX_HIP (hipMemcpyAsync (c->obstab.p + 6*O*k, &oh[6*O*k], 6 * O * sizeof (double), hipMemcpyHostToDevice, c->stream));
    }
    if (P > 0)
    {
      std::memcpy (&ph[6*P*k], prsv.data (), 6 * P * sizeof (double));
      PMX_HIP (hipMemcpyAsync (c->prstab.p + 6*P*k, &ph[6*P*k], 6 * P * sizeof (double), hipMemcpyHostToDevice, c->stream));
    }
    if (c->his_active)
    {
      hh[k] = in.hisrow;
      PMX_HIP (hipMemcpyAsync (c->hisrow.p + k, &hh[k], sizeof (int), hipMemcpyHostToDevice, c->stream));
    }
    if (c->sp_offsets)
    {
      th[k] = curtime;
      PMX_HIP (hipMemcpyAsync (c->ttab.p + k, &th[k], sizeof (double), hipMemcpyHostToDevice, c->stream));
    }
    int ctl[CTL_COUNT], fl[FLAG_COUNT];
    for (;;)
    {
      if ((rc = enqueue_step_forces (*c, step0))) return rc;
      int *slot = c->hring;
      PMX_HIP (hipMemcpyAsync (slot, c->ctl.p, sizeof (int) * CTL_COUNT, hipMemcpyDeviceToHost, c->stream));
      PMX_HIP (hipMemcpyAsync (slot + CTL_COUNT, c->dflags.p, sizeof (int) * FLAG_COUNT, hipMemcpyDeviceToHost, c->stream));
      PMX_HIP (hipMemcpyAsync (c->hkey, c->hmin.p, sizeof (unsigned long long), hipMemcpyDeviceToHost, c->stream));
      PMX_HIP (hipStreamSynchronize (c->stream));
      std::memcpy (ctl, slot, sizeof (ctl));
      std::memcpy (fl, slot + CTL_COUNT, sizeof (fl));
      if ((rc = check_flags (c, fl))) return rc;
      if (!ctl[CTL_ABORT]) break;
      /* the step was aborted before it changed any state (as in dem_core):
       * its sequence number moves on, then larger buffers or a new tree,
       * and the step again */
      if ((rc = step_end (*c))) return rc;
      if (fl[FLAG_ROW_OVERFLOW] || fl[FLAG_HIT_OVERFLOW] || fl[FLAG_EXTRA_OVERFLOW] || fl[FLAG_MTRI_OVERFLOW] ||
          fl[FLAG_TREV_OVERFLOW])
      {
        if ((rc = grow_capacity (c, fl))) return rc;
        grown++;
      }
      if (fl[FLAG_LEAF_OVERFLOW])
      {
        if (rebuilt_at == k) return fail (PMX_EREPART, "Repartitioning failed (step " + std::to_string (k) + ")");
        rebuilt_at = k;
        if ((rc = rebuild (c, false))) return rc;
        rebuilds++;
        if (c->tree.last_partial) partial++;
      }
      if ((rc = clear_overflow_flags (c))) return rc;
    }
    /* adaptive_timestep (parmec.ispc:391-420): adaptive times the smallest
     * estimate, or the current step when no particle is loaded */
    double hmin;
    std::memcpy (&hmin, c->hkey, sizeof (double));
    step1 = hmin < 1.7976931348623157e308 ? adaptive*hmin : step0;
    if (!(step1 > 0.0) || step1 != step1)
      return fail (PMX_EINVAL, "adaptive step " + std::to_string (step1) + " at step " + std::to_string (k));
    if ((rc = enqueue_step_motion (*c, step0, step1))) return rc;
    if (in.output && fc && fc->out_fn)
    {
      /* an output point: forces formed where they were used, mesh geometry,
       * then the caller */
      if ((rc = step_materialize (*c, std::vector<int> (1, k))) || (rc = step_mesh_shapes (*c, false))) return rc;
      PMX_HIP (hipStreamSynchronize (c->stream));
      fc->out_fn (fc->out_user, k);
      (void) hipSetDevice (c->device);
    }
  }
  const int nsteps = k;
  {
    /* a zero joint pivot in the last step's motion */
    int fl[FLAG_COUNT];
    PMX_HIP (hipMemcpyAsync (c->hring, c->dflags.p, sizeof (int) * FLAG_COUNT, hipMemcpyDeviceToHost, c->stream));
    PMX_HIP (hipStreamSynchronize (c->stream));
    std::memcpy (fl, c->hring, sizeof (fl));
    if ((rc = check_flags (c, fl))) return rc;
  }
  if (nsteps > 0 && (rc = step_materialize (*c, std::vector<int> (1, nsteps - 1)))) return rc;
  if ((rc = step_mesh_shapes (*c, false))) return rc;
  if (c->his_active)
  {
    PMX_HIP (hipMemcpyAsync (fc->hisout, c->hisbuf.p, sizeof (double) * (size_t) fc->hisrows * c->nhis,
                             hipMemcpyDeviceToHost, c->stream));
    c->his_active = false;
  }
  PMX_HIP (hipStreamSynchronize (c->stream));
  /* curstep = step1, stepnum++ (parmec.cpp:2521-2523) */
  c->icurtime[0] = curtime;
  c->icurstep[0] = step1;
  c->istepnum[0]++;
  c->stepped += nsteps;
  if (nsteps > 0) { c->crec_valid = c->opt.critical != 0; c->crec_stale = c->opt.critical == 0; }
  c->cur = (c->cur0 + nsteps) & 1;
  c->icur[0] = c->cur;
  c->curtime = c->icurtime[0]; c->curstep = c->icurstep[0]; c->stepnum = c->istepnum[0];
  double dt = std::chrono::duration<double> (std::chrono::steady_clock::now () - t0).count ();
  c->stats.steps = nsteps;
  c->stats.rebuilds = rebuilds;
  c->stats.partial_rebuilds = partial;
  c->stats.grown = grown;
  c->stats.row_capacity = c->cap;
  c->stats.graph_timed = 0;
  c->stats.window_ms = -1.0;
  c->win[0] = c->win[1] = -1;
  c->prof.win0 = c->prof.win1 = -1;
  c->stats.leaves = c->tree.nleaves;
  c->stats.leaf_slots = c->lstride;
  c->stats.seconds = dt;
  {
    long long tb[2] = {0, 0};
    if ((rc = row_totals (*c, tb))) return rc;
    c->stats.pairs = tb[0];
    c->stats.busy = tb[1];
  }
  if (elapsed) *elapsed = dt;
  return PMX_OK;
}

extern "C" {

int pmx_dem_adaptive (pmx_ctx *c, double duration, double step, double adaptive, const pmx_forcing *fc,
                      double *elapsed)
{
  return adaptive_core (c, duration, step, adaptive, fc, elapsed);
}

int pmx_dem (pmx_ctx *c, double duration, double step, const pmx_forcing *fc, double *elapsed)
{
  return dem_core (c, nullptr, duration, step, fc, 1, nullptr, nullptr, elapsed);
}

int pmx_dem_batch (pmx_ctx *c, double duration, double step, const pmx_forcing *fc, double *elapsed)
{
  if (!c) return fail (PMX_EINVAL, "null ctx");
  return dem_core (c, nullptr, duration, step, fc, c->I, nullptr, nullptr, elapsed);
}

int pmx_contacts (pmx_ctx *c, pmx_pair *out, size_t cap, size_t *count)
{
  if (!c) return fail (PMX_EINVAL, "null ctx");
  (void) hipSetDevice (c->device);
  std::vector<pmx_pair> all;
  for (int i = 0; i < c->I; i++)
  {
    int rc = contacts_range (*c, i, false, all);
    if (rc) return rc;
  }
  if (count) *count = all.size ();
  if (out) std::memcpy (out, all.data (), std::min (cap, all.size ()) * sizeof (pmx_pair));
  return PMX_OK;
}

int pmx_set_history (pmx_ctx *c, const pmx_history *h)
{
  if (!c) return fail (PMX_EINVAL, "null ctx");
  (void) hipSetDevice (c->device);
  if (!c->uploaded) return fail (PMX_EINVAL, "no model uploaded");
  return history_set (*c, h);
}

int pmx_history_sample (pmx_ctx *c, double *out)
{
  if (!c || !out) return fail (PMX_EINVAL, "null argument");
  (void) hipSetDevice (c->device);
  return history_now (*c, out);
}

int pmx_set_profiling (pmx_ctx *c, int on)
{
  if (!c) return fail (PMX_EINVAL, "null ctx");
  (void) hipSetDevice (c->device);
  if (on < 0 || on > 2) return fail (PMX_EINVAL, "profiling mode must be 0, 1 or 2");
  PMX_HIP (hipStreamSynchronize (c->stream));
  prof_resolve (*c);
  if (on != c->prof.mode) drop_graph (c);   /* the graph holds event record nodes in mode 1 only */
  c->prof.mode = on;
  c->prof.eager = false;
  for (int k = 0; k < K_COUNT; k++) { c->prof.ms[k] = 0.0; c->prof.bytes[k] = 0.0; c->prof.n[k] = 0; }
  return PMX_OK;
}

int pmx_get_profile (pmx_ctx *c, pmx_kernel_time *out, int cap, int *count)
{
  static const char *names[K_COUNT] = {"bin", "scatter", "detect", "force_init", "rows", "dynamics", "shapes",
                                       "tree_build", "tree_trilists", "springs", "spring_acc"};
  if (!c) return fail (PMX_EINVAL, "null ctx");
  (void) hipSetDevice (c->device);
  PMX_HIP (hipStreamSynchronize (c->stream));
  prof_resolve (*c);
  if (count) *count = K_COUNT;
  for (int k = 0; k < K_COUNT && out && k < cap; k++)
  {
    std::memset (out[k].name, 0, sizeof (out[k].name));
    std::strncpy (out[k].name, names[k], sizeof (out[k].name) - 1);
    out[k].launches = c->prof.n[k];
    out[k].ms = c->prof.ms[k];
    out[k].bytes = c->prof.bytes[k];
  }
  return PMX_OK;
}

int pmx_set_window (pmx_ctx *c, int first, int last)
{
  if (!c) return fail (PMX_EINVAL, "null ctx");
  if (first < 0) { c->win[0] = c->win[1] = -1; c->prof.win0 = c->prof.win1 = -1; return PMX_OK; }
  if (last < first) return fail (PMX_EINVAL, "window: last < first");
  (void) hipSetDevice (c->device);
  for (hipEvent_t &e : c->wev)
    if (!e) PMX_HIP (hipEventCreate (&e));
  c->win[0] = c->prof.win0 = first;
  c->win[1] = c->prof.win1 = last;
  return PMX_OK;
}

int pmx_get_stats (pmx_ctx *c, pmx_stats *s)
{
  if (!c || !s) return fail (PMX_EINVAL, "null argument");
  *s = c->stats;
  return PMX_OK;
}

} /* extern "C" */
