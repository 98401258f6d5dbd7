/*
 * pmx_serve.hip -- resident request loop over the batched step path.
 *
 * Requests are independent simulations (a model, a duration, a step and
 * the per-step forcing tables), the unit of a parameter sweep.  A scheduler
 * thread admits them in arrival order while they fit the HBM budget:
 * requests queued together (within the gather window, up to max_batch) are
 * concatenated into one batch context, and each context is stepped by its
 * own worker thread on its own HIP stream, so batches admitted at different
 * times overlap on the device.  A batch holds one kind of request (spheres,
 * or bodies joined by springs), so that each of its launches serves one
 * kind of work and batches of the two kinds overlap on the device.  Inside
 * a batch every request runs its own
 * step count (pmx_dem_batch_durations): it is frozen when done and its state
 * is copied back at once, so a short request does not wait for a long one.
 * A batch's device memory is released when its last request has finished.
 *
 * Each request's result is bitwise the result of running it alone with
 * pmx_upload + pmx_dem (one dem call of the request's duration): instances
 * of a batch never interact and keep their own partitioning trees.
 *
 * No reference counterpart: Parmec runs one model per process; this is the
 * resident form of the loop of separate DEM runs a sweep script performs
 * (python/acc_sweep.py).
 */
#include "pmx_internal.h"
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace {

int fail (int code, const std::string &msg) { pmx_set_error (msg); return code; }

double now_s (std::chrono::steady_clock::time_point t0)
{
  return std::chrono::duration<double> (std::chrono::steady_clock::now () - t0).count ();
}

struct Req
{
  long long id = 0;
  Merged g;                       /* owned copy of the model; the result is written here */
  double duration = 0.0, step = 0.0;
  int nsteps = 0;
  std::vector<double> gtab, dtab, obstab, obstab0, prstab;
  pmx_forcing f {};
  size_t bytes = 0;
  double t_submit = 0.0, t_admit = -1.0, t_done = -1.0;
  int status = PMX_REQ_QUEUED;
  int batch = 0;                  /* requests in its batch */
  long long pairs = 0;
  std::string err;
  /* continuous batching: steps taken in earlier contexts, and the contact
   * rows and tree snapshot carried from the last one */
  int done_steps = 0;
  std::unique_ptr<ResumeRows> rr;
};

struct Group
{
  std::vector<Req*> reqs;         /* a request is the waiter's once reported: never touched again */
  std::vector<char> reported;
  size_t bytes = 0;
  std::thread th;
  bool finished = false;
  int kind = 0;
  bool rolling = false;           /* takes in queued requests of its kind while it runs */
  int nlive = 0;                  /* requests it runs now (rolling; under the server's lock) */
};

/* device bytes of one request in a context (rows, broad phase, state) */
size_t estimate_bytes (const pmx_model &m, int precision, int cap)
{
  size_t E = m.ellnum, N = m.parnum, T = m.trinum, S = m.sprnum;
  size_t st = precision == PMX_MIXED ? 4 : 8;
  size_t per_sphere = 2 * (size_t) cap * (4 + 3 * st) + 4 * (size_t) cap + 2 * (size_t) cap   /* rows, hits, reverse hits */
                      + 48 + 24 + 8 + 32 + 24 + 24 + 200                                       /* state, records, tree */
                      + 120;                                                                   /* slotted records */
  return E * per_sphere + N * 420 + T * 120 + S * 460 + (size_t) m.nodnum * 56 + (size_t) m.facnum * 16 +
         ((size_t) 1 << 20);
}

/* the kind of work a request brings: spheres (contact detection and rows),
 * bodies with springs, torsion springs or mesh faces */
int request_kind (const pmx_model &m)
{
  return (m.ellnum > 0 ? 1 : 0) | (m.sprnum > 0 || m.trqsprnum > 0 || m.facnum > 0 ? 2 : 0);
}

/* a request a running batch can take in and carry over (continuous
 * batching): it starts at time 0, and nothing in it depends on the call's
 * first step beyond euler and shapes (moving obstacles, restraints) */
bool rollable (const pmx_model &m)
{
  return m.curtime == 0.0 && m.obsnum == 0 && m.rstnum == 0;
}

void copy_outputs (const pmx_model &s, pmx_model *d)
{
  auto cp = [] (double *dst, const double *src, size_t n) { if (dst && src && n) std::memcpy (dst, src, n * sizeof (double)); };
  size_t N = s.parnum, E = s.ellnum, T = s.trinum, S = s.sprnum, NOD = s.nodnum, O = s.obsnum;
  cp (d->angular, s.angular, 6*N); cp (d->linear, s.linear, 3*N); cp (d->rotation, s.rotation, 9*N);
  cp (d->position, s.position, 6*N); cp (d->inverse, s.inverse, 9*N); cp (d->invm, s.invm, N);
  cp (d->force, s.force, 3*N); cp (d->torque, s.torque, 3*N); cp (d->center, s.center, 6*E);
  cp (d->tri, s.tri, 9*T); cp (d->nodes, s.nodes, 6*NOD);
  cp (d->obspnt, s.obspnt, 3*O); cp (d->obsang, s.obsang, 3*O); cp (d->obslin, s.obslin, 3*O);
  cp (d->sprpnt, s.sprpnt, 12*S); cp (d->sprdir, s.sprdir, 6*S); cp (d->yield, s.yield, 2*S);
  cp (d->sprsdsp, s.sprsdsp, 3*S); cp (d->stroke, s.stroke, 3*S); cp (d->sprfrc, s.sprfrc, 3*S);
  if (d->sprflg && s.sprflg && S) std::memcpy (d->sprflg, s.sprflg, S * sizeof (int));
  d->curtime = s.curtime; d->curstep = s.curstep; d->stepnum = s.stepnum;
}

} /* namespace */

struct pmx_server
{
  pmx_server_options opt {};
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Req*> queue;
  std::map<long long, Req*> reqs;
  std::vector<Group*> groups;
  size_t used = 0;
  long long next_id = 1;
  bool stop = false;
  std::thread sched;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now ();
  long long done = 0, failed = 0, batches = 0;
  double particle_steps = 0.0;
  bool longest_first = false;   /* batches take the queued requests with the most steps first */
};

namespace {

struct DoneCtx { pmx_server *s; Group *g; pmx_ctx *c; };

void on_done (void *user, int i)
{
  DoneCtx *d = (DoneCtx*) user;
  Req *r = d->g->reqs[i];
  int rc = pmx_download_instance (d->c, i, &r->g.m);
  size_t npairs = 0;
  if (!rc) rc = pmx_contacts_instance (d->c, i, nullptr, 0, &npairs);
  std::string err = rc ? pmx_last_error () : "";
  {
    std::lock_guard<std::mutex> lk (d->s->mu);
    r->t_done = now_s (d->s->t0);
    r->pairs = (long long) npairs;
    r->status = rc ? PMX_REQ_FAILED : PMX_REQ_DONE;
    r->err = err;
    if (rc) d->s->failed++;
    else { d->s->done++; d->s->particle_steps += (double) r->g.m.parnum * r->nsteps; }
    d->g->reported[i] = 1;
  }
  d->s->cv.notify_all ();
}

void worker (pmx_server *s, Group *g)
{
  pmx_options o {};
  o.precision = s->opt.precision;
  o.device = s->opt.device;
  o.row_capacity = s->opt.row_capacity;
  pmx_ctx *c = nullptr;
  int n = (int) g->reqs.size ();
  int rc = pmx_create (&o, &c);
  if (!rc)
  {
    std::vector<pmx_model> ms (n);
    std::vector<double> dur (n);
    std::vector<pmx_forcing> fs (n);
    for (int i = 0; i < n; i++) { ms[i] = g->reqs[i]->g.m; dur[i] = g->reqs[i]->duration; fs[i] = g->reqs[i]->f; }
    rc = pmx_upload_batch (c, ms.data (), n);
    if (!rc)
    {
      DoneCtx d {s, g, c};
      double el = 0.0;
      rc = pmx_dem_batch_durations (c, dur.data (), g->reqs[0]->step, fs.data (), on_done, &d, &el);
    }
  }
  std::string err = rc ? pmx_last_error () : "";
  if (c) pmx_destroy (c);
  {
    std::lock_guard<std::mutex> lk (s->mu);
    for (size_t i = 0; i < g->reqs.size (); i++)
      if (!g->reported[i])
      {
        Req *r = g->reqs[i];
        r->status = PMX_REQ_FAILED;
        r->err = err.empty () ? "batch ended before the request finished" : err;
        r->t_done = now_s (s->t0);
        s->failed++;
        g->reported[i] = 1;
      }
    s->used -= g->bytes;
    g->finished = true;
  }
  s->cv.notify_all ();
}

/* continuous batching: the batch runs its requests in one context; at most
 * every repack_s seconds, when requests of its kind are queued, the call
 * pauses after a group of steps, the unfinished requests are carried over
 * (state, contact rows, tree snapshot, time), the queued ones join, and the
 * call resumes over all of them.  Each request's result stays bitwise its
 * solo result: its tree is rebuilt from the centres it was built from, its
 * rows and time go on from the step where it paused. */
struct RollCtx
{
  pmx_server *s;
  Group *g;
  pmx_ctx *c;
  std::vector<Req*> *live;
  std::vector<char> *reported;    /* a reported request is the waiter's: never touched again */
  double step;
  std::chrono::steady_clock::time_point t_round;
};

void on_done_roll (void *user, int i)
{
  RollCtx *d = (RollCtx*) user;
  Req *r = (*d->live)[i];
  int rc = pmx_download_instance (d->c, i, &r->g.m);
  size_t npairs = 0;
  if (!rc) rc = pmx_contacts_instance (d->c, i, nullptr, 0, &npairs);
  std::string err = rc ? pmx_last_error () : "";
  {
    std::lock_guard<std::mutex> lk (d->s->mu);
    r->t_done = now_s (d->s->t0);
    r->pairs = (long long) npairs;
    r->status = rc ? PMX_REQ_FAILED : PMX_REQ_DONE;
    r->err = err;
    r->rr.reset ();
    if (rc) d->s->failed++;
    else { d->s->done++; d->s->particle_steps += (double) r->g.m.parnum * r->nsteps; }
    d->s->used -= r->bytes;   /* its memory is back in the budget */
    d->g->bytes -= r->bytes;
    d->g->nlive--;
    (*d->reported)[i] = 1;
  }
  d->s->cv.notify_all ();
}

/* take the queued requests of the batch's kind that fit the budget (the
 * lock held) */
void take_queued (pmx_server *s, Group *g, std::vector<Req*> &into, size_t room)
{
  double t = now_s (s->t0);
  std::vector<Req*> order (s->queue.begin (), s->queue.end ());
  if (s->longest_first)
    std::stable_sort (order.begin (), order.end (), [] (const Req *a, const Req *b) { return a->nsteps > b->nsteps; });
  size_t n0 = into.size ();
  for (Req *r : order)
  {
    if (into.size () >= room) break;
    if (!rollable (r->g.m) || request_kind (r->g.m) != g->kind || r->step != into[0]->step) continue;
    if (s->used + r->bytes > s->opt.hbm_budget) break;
    s->used += r->bytes;
    g->bytes += r->bytes;
    r->status = PMX_REQ_RUNNING;
    r->t_admit = t;
    into.push_back (r);
  }
  for (size_t k = n0; k < into.size (); k++) s->queue.erase (std::find (s->queue.begin (), s->queue.end (), into[k]));
}

int roll_pause (void *user, int)
{
  RollCtx *d = (RollCtx*) user;
  double age = std::chrono::duration<double> (std::chrono::steady_clock::now () - d->t_round).count ();
  if (age < d->s->opt.repack_s) return 0;
  std::lock_guard<std::mutex> lk (d->s->mu);
  if ((int) d->live->size () >= d->s->opt.max_batch) return 0;
  for (Req *r : d->s->queue)
    if (rollable (r->g.m) && request_kind (r->g.m) == d->g->kind && r->step == d->step &&
        d->s->used + r->bytes <= d->s->opt.hbm_budget)
      return 1;
  return 0;
}

void rolling_worker (pmx_server *s, Group *g)
{
  pmx_options o {};
  o.precision = s->opt.precision;
  o.device = s->opt.device;
  o.row_capacity = s->opt.row_capacity;
  pmx_ctx *c = nullptr;
  int rc = pmx_create (&o, &c);
  std::vector<Req*> live = g->reqs;
  std::vector<char> reported (live.size (), 0);
  std::string err;
  while (!rc && !live.empty ())
  {
    int n = (int) live.size ();
    double step = live[0]->step;
    reported.assign (n, 0);
    /* models at time 0 for the upload (the resumed ones get their time from
     * their carried state); forcing tables from each request's own step */
    std::vector<pmx_model> ms (n);
    std::vector<pmx_forcing> fs (n);
    std::vector<int> steps (n);
    std::vector<const ResumeRows*> rs (n, nullptr);
    for (int i = 0; i < n; i++)
    {
      Req *r = live[i];
      ms[i] = r->g.m;
      ms[i].curtime = 0.0; ms[i].curstep = 0.0; ms[i].stepnum = 0;
      size_t k = r->done_steps;
      fs[i] = r->f;
      if (fs[i].gtab) fs[i].gtab += 3*k;
      if (fs[i].dtab) fs[i].dtab += 6*k;
      if (fs[i].prstab) fs[i].prstab += 6*k*(size_t) (fs[i].prscols > 0 ? fs[i].prscols : 0);
      steps[i] = r->nsteps - r->done_steps;
      rs[i] = r->rr.get ();
      r->batch = n;
    }
    if ((rc = pmx_upload_batch (c, ms.data (), n)) || (rc = import_rows (*c, rs))) break;
    RollCtx d {s, g, c, &live, &reported, step, std::chrono::steady_clock::now ()};
    DemExtra x;
    x.steps = steps.data ();
    x.stop_fn = roll_pause;
    x.stop_user = &d;
    double el = 0.0;
    rc = dem_core (c, nullptr, 0.0, step, fs.data (), n, on_done_roll, &d, &el, &x);
    if (rc) break;
    if (x.paused_at < 0) break;   /* every request finished */
    /* carry the unfinished requests over; the queued ones of this kind join */
    std::vector<Req*> next;
    for (int i = 0; i < n && !rc; i++)
    {
      bool was;
      {
        std::lock_guard<std::mutex> lk (s->mu);
        was = reported[i] != 0;
      }
      if (was) continue;   /* finished in this round (and maybe freed by its waiter) */
      Req *r = live[i];
      std::unique_ptr<ResumeRows> rr (new ResumeRows);
      if ((rc = pmx_download_instance (c, i, &r->g.m)) || (rc = export_rows (*c, i, *rr))) break;
      rr->curtime = r->g.m.curtime; rr->curstep = r->g.m.curstep; rr->stepnum = r->g.m.stepnum;
      r->rr = std::move (rr);
      r->done_steps += x.paused_at;
      next.push_back (r);
    }
    if (rc) break;
    {
      std::lock_guard<std::mutex> lk (s->mu);
      if (!next.empty ()) take_queued (s, g, next, (size_t) s->opt.max_batch);
      g->nlive = (int) next.size ();
    }
    live.swap (next);
    reported.assign (live.size (), 0);
  }
  if (rc) err = pmx_last_error ();
  if (c) pmx_destroy (c);
  {
    std::lock_guard<std::mutex> lk (s->mu);
    for (size_t i = 0; i < live.size (); i++)
      if (!reported[i])
      {
        Req *r = live[i];
        r->status = PMX_REQ_FAILED;
        r->err = err.empty () ? "batch ended before the request finished" : err;
        r->t_done = now_s (s->t0);
        r->rr.reset ();
        s->failed++;
      }
    s->used -= g->bytes;
    g->finished = true;
  }
  s->cv.notify_all ();
}

void scheduler (pmx_server *s)
{
  std::unique_lock<std::mutex> lk (s->mu);
  for (;;)
  {
    /* the lock is released while a finished worker is joined; every state
     * change made meanwhile (stop, freed memory, new requests) is read again
     * below with the lock held up to the next wait, so no wake-up is lost */
    for (size_t k = 0; k < s->groups.size (); )
    {
      Group *g = s->groups[k];
      if (g->finished)
      {
        lk.unlock ();
        g->th.join ();
        lk.lock ();
        delete g;
        s->groups.erase (s->groups.begin () + k);
      }
      else k++;
    }
    if (s->stop) break;
    int active = (int) s->groups.size ();
    if (!s->queue.empty () && active < s->opt.max_active)
    {
      /* gather: let requests arriving within the window join one batch */
      double wait = s->opt.gather_s - (now_s (s->t0) - s->queue.front ()->t_submit);
      if (wait > 0.0 && (int) s->queue.size () < s->opt.max_batch)
      {
        s->cv.wait_for (lk, std::chrono::duration<double> (wait));
        continue;
      }
      /* the batch: the oldest request and, in arrival order, the queued
       * requests of its kind (same step and time state, same mix of
       * spheres and bodies with springs or meshes: one batch's launches then
       * serve one kind of work, and batches of the other kind run beside it
       * on their own streams) while they fit the budget */
      std::vector<Req*> take;
      size_t bytes = 0;
      const bool roll = s->opt.repack_s > 0.0;
      auto joins_running = [&] (const Req *r) {
        /* a running continuous batch of its kind will take it in */
        if (!roll || !rollable (r->g.m)) return false;
        for (Group *g : s->groups)
          if (g->rolling && !g->finished && g->kind == request_kind (r->g.m) && g->nlive < s->opt.max_batch)
            return true;
        return false;
      };
      /* arrival order, or (longest_first) the requests with the most
       * steps first: they set how long the stream takes */
      std::vector<Req*> order (s->queue.begin (), s->queue.end ());
      if (s->longest_first)
        std::stable_sort (order.begin (), order.end (), [] (const Req *a, const Req *b) { return a->nsteps > b->nsteps; });
      for (Req *r : order)
      {
        if ((int) take.size () >= s->opt.max_batch) break;
        if (joins_running (r) ||
            (!take.empty () && (r->step != take[0]->step || r->g.m.curtime != take[0]->g.m.curtime ||
                                r->g.m.curstep != take[0]->g.m.curstep || r->g.m.stepnum != take[0]->g.m.stepnum ||
                                request_kind (r->g.m) != request_kind (take[0]->g.m) ||
                                (roll && rollable (r->g.m) != rollable (take[0]->g.m)))))
          continue;
        if (s->used + bytes + r->bytes > s->opt.hbm_budget) break;   /* in order: wait for memory */
        take.push_back (r);
        bytes += r->bytes;
      }
      for (Req *r : take) s->queue.erase (std::find (s->queue.begin (), s->queue.end (), r));
      if (take.empty () && !s->queue.empty () && joins_running (s->queue.front ()))
      {
        /* everything queued waits for a running batch of its kind */
        s->cv.wait (lk);
        continue;
      }
      if (take.empty ())
      {
        if (active == 0)
        {
          /* nothing runs and the head still does not fit: it never will */
          Req *r = s->queue.front ();
          s->queue.pop_front ();
          r->status = PMX_REQ_FAILED;
          r->err = "request exceeds the HBM budget";
          r->t_done = now_s (s->t0);
          s->failed++;
          s->cv.notify_all ();
          continue;
        }
        s->cv.wait (lk);
        continue;
      }
      Group *g = new Group;
      g->reqs = take;
      g->reported.assign (take.size (), 0);
      g->bytes = bytes;
      g->kind = request_kind (take[0]->g.m);
      g->rolling = roll && rollable (take[0]->g.m);
      g->nlive = (int) take.size ();
      s->used += bytes;
      double t = now_s (s->t0);
      for (Req *r : take) { r->status = PMX_REQ_RUNNING; r->t_admit = t; r->batch = (int) take.size (); }
      s->batches++;
      s->groups.push_back (g);
      g->th = std::thread (g->rolling ? rolling_worker : worker, s, g);
      continue;
    }
    s->cv.wait (lk);
  }
}

} /* namespace */

extern "C" {

int pmx_server_create (const pmx_server_options *opt, pmx_server **out)
{
  if (!out) return fail (PMX_EINVAL, "null server pointer");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount (&ndev) != hipSuccess || ndev == 0) return fail (PMX_EDEVICE, "no HIP device available");
  pmx_server *s = new pmx_server ();
  if (opt) s->opt = *opt;
  if (s->opt.precision != PMX_FP64 && s->opt.precision != PMX_MIXED)
  { delete s; return fail (PMX_EINVAL, "precision must be PMX_FP64 or PMX_MIXED"); }
  if (s->opt.hbm_budget == 0) s->opt.hbm_budget = (size_t) 200 << 30;
  if (s->opt.max_batch <= 0) s->opt.max_batch = 256;
  if (s->opt.max_active <= 0) s->opt.max_active = 16;
  if (s->opt.gather_s < 0.0) s->opt.gather_s = 0.0;
  s->longest_first = s->opt.longest_first != 0;
  s->sched = std::thread (scheduler, s);
  *out = s;
  return PMX_OK;
}

int pmx_server_submit (pmx_server *s, const pmx_model *m, double duration, double step, const pmx_forcing *f,
                       long long *id)
{
  if (!s || !m) return fail (PMX_EINVAL, "null argument");
  if (!(duration > 0.0) || !(step > 0.0)) return fail (PMX_EINVAL, "duration and step must be positive");
  Req *r = new Req ();
  std::vector<InstOff> off;
  int rc = merge_models (m, 1, r->g, off);
  if (rc) { delete r; return rc; }
  pmx_model &c = r->g.m;
  if (!c.force) c.force = r->g.D (3*(size_t) c.parnum);
  if (!c.torque) c.torque = r->g.D (3*(size_t) c.parnum);
  r->duration = duration;
  r->step = step;
  double h = m->curtime == 0.0 ? step : m->curstep;
  pmx_step_times (duration, h, m->curtime, nullptr, 0, &r->nsteps);
  size_t ns = r->nsteps, O = m->obsnum;
  for (int d = 0; d < 3; d++) r->f.gravity[d] = f ? f->gravity[d] : 0.0;
  for (int d = 0; d < 6; d++) r->f.damping[d] = f ? f->damping[d] : 0.0;
  if (f && f->gtab) { r->gtab.assign (f->gtab, f->gtab + 3*ns); r->f.gtab = r->gtab.data (); }
  if (f && f->dtab) { r->dtab.assign (f->dtab, f->dtab + 6*ns); r->f.dtab = r->dtab.data (); }
  if (f && f->obstab && O) { r->obstab.assign (f->obstab, f->obstab + 6*ns*O); r->f.obstab = r->obstab.data (); }
  if (f && f->obstab0 && O) { r->obstab0.assign (f->obstab0, f->obstab0 + 6*O); r->f.obstab0 = r->obstab0.data (); }
  size_t PR = off[1].pcol;
  if (f && f->prstab && PR)
  {
    size_t stride = f->prscols > 0 ? (size_t) f->prscols : PR;
    r->prstab.assign (f->prstab, f->prstab + 6*ns*stride);
    r->f.prstab = r->prstab.data ();
    r->f.prscols = (int) stride;
  }
  r->bytes = estimate_bytes (c, s->opt.precision, s->opt.row_capacity > 0 ? s->opt.row_capacity : 32);
  {
    std::lock_guard<std::mutex> lk (s->mu);
    r->id = s->next_id++;
    r->t_submit = now_s (s->t0);
    s->reqs[r->id] = r;
    s->queue.push_back (r);
    if (id) *id = r->id;
  }
  s->cv.notify_all ();
  return PMX_OK;
}

int pmx_server_wait (pmx_server *s, long long id, double timeout, pmx_model *out, pmx_request_info *info)
{
  if (!s) return fail (PMX_EINVAL, "null server");
  std::unique_lock<std::mutex> lk (s->mu);
  auto it = s->reqs.find (id);
  if (it == s->reqs.end ()) return fail (PMX_EINVAL, "unknown request id");
  Req *r = it->second;
  auto ready = [r] { return r->status == PMX_REQ_DONE || r->status == PMX_REQ_FAILED; };
  if (timeout < 0.0) s->cv.wait (lk, ready);
  else if (!s->cv.wait_for (lk, std::chrono::duration<double> (timeout), ready))
    return fail (PMX_EBUSY, "request not finished");
  if (info)
  {
    info->id = r->id; info->status = r->status; info->batch = r->batch; info->steps = r->nsteps;
    info->particles = r->g.m.parnum; info->pairs = r->pairs; info->bytes = r->bytes;
    info->submit_s = r->t_submit; info->admit_s = r->t_admit; info->done_s = r->t_done;
  }
  int rc = PMX_OK;
  if (r->status == PMX_REQ_FAILED) rc = fail (PMX_EDEVICE, "request failed: " + r->err);
  else if (out)
  {
    InstOff a = model_counts (*out), b = model_counts (r->g.m);
    if (a.par != b.par || a.ell != b.ell || a.tri != b.tri || a.nod != b.nod || a.spr != b.spr || a.obs != b.obs)
      rc = fail (PMX_EINVAL, "result view counts differ from the request's model");
    else copy_outputs (r->g.m, out);
  }
  s->reqs.erase (it);
  delete r;
  return rc;
}

int pmx_server_get_stats (pmx_server *s, pmx_server_stats *st)
{
  if (!s || !st) return fail (PMX_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk (s->mu);
  st->queued = (long long) s->queue.size ();
  st->running = 0;
  for (auto &kv : s->reqs) st->running += kv.second->status == PMX_REQ_RUNNING;
  st->done = s->done;
  st->failed = s->failed;
  st->batches = s->batches;
  st->active_batches = (int) s->groups.size ();
  st->bytes_in_use = s->used;
  st->particle_steps = s->particle_steps;
  st->seconds = now_s (s->t0);
  return PMX_OK;
}

void pmx_server_destroy (pmx_server *s)
{
  if (!s) return;
  {
    std::unique_lock<std::mutex> lk (s->mu);
    s->cv.wait (lk, [s] {
      bool busy = !s->queue.empty ();
      for (Group *g : s->groups) busy = busy || !g->finished;
      return !busy; });
    s->stop = true;
  }
  s->cv.notify_all ();
  s->sched.join ();
  for (Group *g : s->groups) { if (g->th.joinable ()) g->th.join (); delete g; }
  for (auto &kv : s->reqs) delete kv.second;
  delete s;
}

} /* extern "C" */
