/*
 * pmx.h -- C ABI of the MI355X-native Parmec DEM step engine (libpmx.so).
 *
 * The engine replaces the body of Parmec's dem() step loop for spheres,
 * static/moving obstacle triangles, granular contacts and SPRING elements.
 * All arrays are owned by the caller and laid out like Parmec's global SoA
 * buffers (parmec.h:66-283): a k-component per-item array is one contiguous
 * block of k*count doubles, component-major (component c of item i at
 * c*count + i).  3x3 matrices are column-major, element (r,c) at 3c+r
 * (macros.h:906-950).
 *
 * Errors: every entry point returns PMX_OK (0) or a negative PMX_E* code and
 * leaves a message for pmx_last_error() (thread-local).  Contexts are not
 * thread safe; one HIP stream per context.
 */
#ifndef PMX_H
#define PMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PMX_CUTOFF 64   /* radix-tree terminal node size, partition.h:29 */
#define PMX_LSIZE 96    /* leaf capacity before re-partitioning, partition.h:31 */

/* surface interaction parameters, constants.h:37 */
enum { PMX_SPRING = 0, PMX_DAMPER, PMX_FRISTAT, PMX_FRIDYN, PMX_FRIROL, PMX_FRIDRIL, PMX_KSKN, PMX_NIPARAM };

/* interaction kinds, constants.h:35 */
enum { PMX_GRANULAR_FORCE = 0, PMX_USER_FORCE = 1 };

/* spring types and flags, constants.h:61-63 */
enum { PMX_SPRING_NONLINEAR_ELASTIC = 0, PMX_SPRING_GENERAL_NONLINEAR = 1 };
enum { PMX_SPRING_YIELDED = 1, PMX_SPRING_UNLOADING = 2, PMX_SPRDIR_CONSTANT = 4,
       PMX_SPRDIR_PLANAR = 8, PMX_SPRDIR_FOLLOWER = 16, PMX_SPRDIR_PROJECT = 32 };

/* error codes */
enum { PMX_OK = 0, PMX_EINVAL = -1, PMX_ENOMEM = -2, PMX_EDEVICE = -3,
       PMX_ECAPACITY = -4, PMX_EUNSUPPORTED = -5, PMX_EREPART = -6, PMX_EBUSY = -7 };

/* arithmetic of the contact/force stage */
enum { PMX_FP64 = 0,     /* plain FP64 drop-in */
       PMX_MIXED = 1 };  /* FP32 contact geometry and forces, FP64 integration */

/*
 * State view: the SoA buffers of parmec.h (namespace parmec) that the step
 * path reads and writes.  Counts follow parmec.h:36-283.
 */
typedef struct pmx_model
{
  /* particles, parmec.h:66-84 */
  int parnum;
  double *angular;   /* [6][parnum]: 0-2 body-frame Omega, 3-5 spatial omega */
  double *linear;    /* [3][parnum] */
  double *rotation;  /* [9][parnum] column-major */
  double *position;  /* [6][parnum]: 0-2 current, 3-5 reference */
  double *inertia;   /* [9][parnum] referential inertia */
  double *inverse;   /* [9][parnum] inverse inertia (rewritten by every dem call) */
  double *mass;      /* [parnum] */
  double *invm;      /* [parnum] (rewritten by every dem call) */
  double *force;     /* [3][parnum] (output) */
  double *torque;    /* [3][parnum] (output) */
  int *flags;        /* [parnum] */

  /* ellipsoids (spheres: radii[1] < 0), parmec.h:86-96 */
  int ellnum;
  double *center;    /* [6][ellnum]: 0-2 current, 3-5 reference */
  double *radii;     /* [3][ellnum] */
  int *part;         /* [ellnum] ellipsoid -> particle */
  int *ellcol;       /* [ellnum] colour */

  /* triangles, parmec.h:98-102 */
  int trinum;
  double *tri;       /* [9][trinum]: vertex v, component c at (3v+c)*trinum + t */
  int *tricol;       /* [trinum] */
  int *triobs;       /* [trinum]: -1 static, <= -2 moving obstacle, >= 0 mesh particle */

  /* moving obstacles, parmec.h:104-112 */
  int obsnum;
  int *trirng;       /* [2*obsnum] triangle ranges */
  double *obspnt;    /* [3*obsnum] */
  double *obslin;    /* [3*obsnum] (per step input) */
  double *obsang;    /* [3*obsnum] (per step input) */

  /* mesh nodes and faces (meshed particles), parmec.h:114-130 */
  int nodnum;
  double *nodes;     /* [6][nodnum]: 0-2 current, 3-5 reference */
  int *nodpart;      /* [nodnum] */
  int facnum;
  int *facnod;       /* [3][facnum] */
  int *factri;       /* [facnum] face -> triangle */

  /* surface pairing, parmec.h:140-150 */
  int pairnum;
  int *pairs;        /* [2*pairnum] colour pairs (index 0 = default) */
  int *ikind;        /* [pairnum] */
  double *iparam;    /* [PMX_NIPARAM][pairnum] */

  /* translational springs, parmec.h:152-200 (sorted by (part1, part2)) */
  int sprnum;
  int *sprtype;      /* [sprnum] */
  int *sprpart;      /* [2][sprnum] */
  double *sprpnt;    /* [2][6][sprnum]: point p (0-2 current, 3-5 reference) at (6p+c)*sprnum + s */
  int *spridx;       /* [sprnum+1] into spring curve table */
  double *spring;    /* [2][sprcrv]: stroke, force */
  int sprcrv;        /* number of spring curve points */
  int *dashidx;      /* [sprnum+1] into dashpot table */
  double *dashpot;   /* [2][dashcrv] */
  int dashcrv;
  int *unidx;        /* [sprnum+1] into unload table */
  double *unload;    /* [2][uncrv] */
  int uncrv;
  double *yield;     /* [2][sprnum] */
  double *sprdir;    /* [6][sprnum] */
  int *sprflg;       /* [sprnum] */
  double *sprfric;   /* [sprnum] */
  double *sprkskn;   /* [sprnum] */
  double *sprsdsp;   /* [3][sprnum] */
  double *stroke0;   /* [sprnum] */
  double *stroke;    /* [3][sprnum] */
  double *sprfrc;    /* [3][sprnum] */
  int *unspring;     /* [sprnum] (-2: no unspring action) */
  int *sproffset;    /* [sprnum] (-1: no load curve offset) */

  /* prescribed motion, parmec.h:244-251 (PRESCRIBE, input.cpp:2647-2795):
   * per entry the particle and the kind of its linear and angular parts
   * (-1 none, 0 velocity, 1 acceleration); values per step in
   * pmx_forcing.prstab */
  int prsnum;
  int *prspart;      /* [prsnum] */
  int *prskind;      /* [2][prsnum]: linear kind, angular kind */
  int *prscol;       /* [prsnum] column of the entry's history in pmx_forcing.prstab (entries driven by one
                        history share it), or NULL: column = entry index */

  /* time bookkeeping, parmec.h:44-49 */
  double curtime;
  double curstep;
  int stepnum;       /* dem() calls so far (parmec.cpp:2523); UNSPRING tests run when stepnum % nsteps == 0 */

  /* load curves: time series as look-up tables (lcurve_from_time_series,
   * parmec.cpp:1239-1269), read by spring stroke offsets (sproffset) and by
   * UNSPRING's unloading curves (unaction) */
  int lcnum;
  int *lcidx;        /* [lcnum+1] into lcurve */
  double *lcurve;    /* [2][lccrv]: time (or stroke), value */
  int lccrv;

  /* torsion springs, parmec.cpp:180-198 (TORSION_SPRING, input.cpp:1791-2080),
   * sorted by (part1, part2); the roll, pitch and yaw curves of all springs
   * share one table per kind */
  int trqsprnum;
  int *trqsprpart;   /* [2][trqsprnum] (part2 = -1: a spring to the ground) */
  double *trqzdir0;  /* [3][trqsprnum] reference z (roll) direction in part1's body frame */
  double *trqxdir0;  /* [3][trqsprnum] reference x direction */
  int *krpyidx;      /* [3][trqsprnum+1] roll, pitch, yaw stiffness curves in krpy */
  double *krpy;      /* [2][krpycrv]: angle, torque */
  int krpycrv;
  int *drpyidx;      /* [3][trqsprnum+1] damper curves in drpy; one point: a critical damping ratio */
  double *drpy;      /* [2][drpycrv]: angular velocity, torque (or ratio, ratio) */
  int drpycrv;
  int *trqcone;      /* [trqsprnum] cone flags: 1 roll, 2 pitch, 4 yaw (constants.h:66-68) */
  double *trqzdir1;  /* [3][trqsprnum] (output) current z direction */
  double *trqxdir1;  /* [3][trqsprnum] (output) current x direction */
  double *trqrpy;    /* [3][trqsprnum] (state) roll, pitch, yaw angles */
  double *trqrpytot; /* [3][trqsprnum] (output) total torques conjugate with the angles */
  double *trqrpyspr; /* [3][trqsprnum] (output) spring torques without the damper parts */

  /* UNSPRING conditions (input.cpp:2083-2295; unspring_task, forces.ispc:2058-2127);
   * spring numbers here are indices of the sorted springs */
  int unsprnum;
  int *tspridx;      /* [unsprnum+1] into tsprings: the springs each condition tests */
  int *tsprings;
  int *mspridx;      /* [unsprnum+1] into msprings: the springs it switches off */
  int *msprings;
  int *actidx;       /* [unsprnum+1] into activate: the inactive springs it switches on */
  int *activate;
  double *unlim;     /* [2][unsprnum] lower, upper limit */
  int *unent;        /* [unsprnum] PMX_HIS_STROKE, _F, _SF or _FF */
  int *unop;         /* [unsprnum] PMX_OP_* */
  int *unabs;        /* [unsprnum] test |value| */
  int *nsteps;       /* [unsprnum] test when stepnum % nsteps == 0 */
  int *nfreq;        /* [unsprnum] (state) failures left before the condition triggers; 0: triggered */
  int *unaction;     /* [unsprnum] -1: zero force, >= 0: unloading load curve */

  /* RESTRAIN (input.cpp:2458-2644; restrain.ispc:30-154) */
  int rstnum;
  int *rstpart;      /* [rstnum] */
  double *rstlin;    /* [9][rstnum] up to three orthonormal restrained linear directions (zeros: none) */
  double *rstang;    /* [9][rstnum] restrained angular directions */

  /* BALL_JOINT (input.cpp:2298-2333; joints.cpp:83-567) */
  int jnum;
  int *jpart;        /* [2][jnum] (part2 = -1: a joint to the ground) */
  double *jpoint;    /* [3][jnum] joint point in the reference configuration */
  double *jreac;     /* [3][jnum] (output) joint reactions */
} pmx_model;

/* UNSPRING operators, constants.h:70 */
enum { PMX_OP_SUM = 0, PMX_OP_MAX = 1, PMX_OP_MIN = 2 };

/*
 * Per-step external inputs.  Parmec evaluates gravity and global damping
 * callbacks on the host at every step (input.cpp:4424-4503); the engine takes
 * them as tables evaluated at each step's curtime (pmx_step_times gives the
 * times), or as constants when the table pointer is NULL.
 */
struct pmx_step_inputs;

typedef struct pmx_forcing
{
  double gravity[3];         /* constant gravity when gtab == NULL */
  const double *gtab;        /* [nsteps][3] gravity per step, or NULL */
  double damping[6];         /* constant global damping when dtab == NULL */
  const double *dtab;        /* [nsteps][6] damping per step, or NULL */
  const double *obstab;      /* [nsteps][obsnum][6] obstacle (angular, linear) velocity at curtime+h, or NULL */
  const double *obstab0;     /* [obsnum][6] obstacle velocity at t = h/2 (first call only), or NULL */
  const double *prstab;      /* [nsteps][prscols][6] prescribed linear 0-2 and angular 3-5 values of each
                                history column at each step's curtime (input.cpp:4247-4421), or NULL (zeros) */
  int prscols;               /* columns per step of prstab (0: the model's, max prscol + 1 or prsnum) */
  const int *hisrow;         /* [nsteps] row of the history output that samples the state at the end of
                                each step, or -1 (parmec.cpp:2509-2514); NULL: no sampling */
  int hisrows;               /* rows of hisout */
  double *hisout;            /* [hisrows][histories] values written by pmx_dem (output_history,
                                output.cpp:5143-5520); TIME entries are left to the caller */
  const int *outstep;        /* [noutstep] ascending steps of this call (0-based) after which pmx_dem pauses
                                and calls out_fn(out_user, step) with the state complete on the device
                                (forces, mesh nodes and triangles), so that pmx_download gives the state
                                dem() writes at its output points (output_files, parmec.cpp:2502-2507);
                                single-model contexts; NULL: no pause */
  int noutstep;
  void (*out_fn) (void *user, int step);
  void *out_user;
  /* pmx_dem_adaptive only (the steps and their times are known one at a
   * time): called before step k with that step's curtime and length step0,
   * it fills the step's host inputs; NULL: gravity and damping above, still
   * obstacles, no prescribed values, no history sample, no output pause */
  int (*step_fn) (void *user, int k, double curtime, double step0, struct pmx_step_inputs *in);
  void *step_user;
} pmx_forcing;

/* the host inputs of one step of an adaptive dem call (read_gravity_and_damping,
 * obstaclev, prescribe_*, the history and output rules at that step's curtime;
 * parmec.cpp:2432, 2449-2509) */
typedef struct pmx_step_inputs
{
  double gravity[3];
  double damping[6];
  double *obsvel;            /* [obsnum][6] obstacle (angular, linear) velocity at curtime + step0 */
  double *prs;               /* [prscols][6] prescribed values of each history column at curtime */
  int hisrow;                /* row of hisout that samples this step's end state, or -1 */
  int output;                /* 1: pause after this step and call out_fn (as at an output point) */
} pmx_step_inputs;

/* time-history entities (HISTORY, input.cpp:3157-3555); particle entities
 * average over their list (or take one particle's body point), spring
 * entities over a list of spring numbers */
enum
{
  PMX_HIS_TIME = 0, PMX_HIS_PX, PMX_HIS_PY, PMX_HIS_PZ, PMX_HIS_PL, PMX_HIS_DX, PMX_HIS_DY, PMX_HIS_DZ, PMX_HIS_DL,
  PMX_HIS_VX, PMX_HIS_VY, PMX_HIS_VZ, PMX_HIS_VL, PMX_HIS_OX, PMX_HIS_OY, PMX_HIS_OZ, PMX_HIS_OL,
  PMX_HIS_FX, PMX_HIS_FY, PMX_HIS_FZ, PMX_HIS_FL, PMX_HIS_TX, PMX_HIS_TY, PMX_HIS_TZ, PMX_HIS_TL,
  PMX_HIS_LENGTH, PMX_HIS_STROKE, PMX_HIS_F, PMX_HIS_SF, PMX_HIS_FF, PMX_HIS_SS,
  PMX_HIS_ZDIR_X, PMX_HIS_ZDIR_Y, PMX_HIS_ZDIR_Z, PMX_HIS_XDIR_X, PMX_HIS_XDIR_Y, PMX_HIS_XDIR_Z,
  PMX_HIS_ROLL, PMX_HIS_PITCH, PMX_HIS_YAW, PMX_HIS_TRQTOT_R, PMX_HIS_TRQTOT_P, PMX_HIS_TRQTOT_Y,
  PMX_HIS_TRQSPR_R, PMX_HIS_TRQSPR_P, PMX_HIS_TRQSPR_Y, PMX_HIS_JREAC_X, PMX_HIS_JREAC_Y, PMX_HIS_JREAC_Z,
  PMX_HIS_JREAC_L, PMX_HIS_COUNT
};

typedef struct pmx_history
{
  int count;                 /* histories */
  const int *entity;         /* [count] PMX_HIS_* */
  const int *off;            /* [count+1] offsets into list */
  const int *list;           /* particle numbers; spring numbers for LENGTH .. SS; torsion spring
                                numbers for ZDIR_X .. TRQSPR_Y; joint numbers for JREAC_* */
  const int *haspoint;       /* [count] 1: one particle's body point `point` (HIS_LIST|HIS_POINT), or NULL */
  const double *point;       /* [count][3] referential point, or NULL */
} pmx_history;

typedef struct pmx_options
{
  int precision;       /* PMX_FP64 or PMX_MIXED */
  int device;          /* HIP device ordinal */
  int row_capacity;    /* contacts per sphere row (0: default 32) */
  int verbose;
  int tune;            /* 0: pick the detection block shape by timing it (models of 50,000
                          spheres or more: at the first dem call and again after 256
                          steps); 1: the fixed det_waves below */
  int det_waves;       /* tune = 1: waves per detection block, 1, 2, 4 or 8 (0: 4) */
  int rows_block;      /* spheres per contact-row block (per list of spheres in contact), 64, 128,
                          256, 512 or 1024 (0: 1024) */
  int critical;        /* 1: every step records its contacts (point, direction, stiffness and
                          damping coefficient per row entry, 64 B each), for pmx_critical after
                          a dem call; 0: pmx_critical's per-particle estimate is refused once
                          steps were taken (it needs the last step's contacts) */
} pmx_options;

/* one contact of the canonical pair set */
typedef struct pmx_pair
{
  int master;        /* master ellipsoid */
  int slave;         /* slave ellipsoid, or -(triangle+1) */
  double state[3];   /* tangential spring displacement */
} pmx_pair;

typedef struct pmx_stats
{
  long long steps;         /* steps taken by the last pmx_dem call */
  long long rebuilds;      /* partitioning tree (re)builds in the last call */
  long long pairs;         /* live contact pairs after the last step */
  int leaves;              /* leaves of the current partitioning tree */
  int max_row;             /* largest contact row seen */
  double seconds;          /* wall time of the last pmx_dem step loop */
  int det_waves;           /* block shapes in use: waves per detection block, */
  int rows_block;          /* threads per contact-row block */
  int leaf_slots;          /* leaf-record layout: PMX_LSIZE slots per leaf, or 0 (packed) */
  long long partial_rebuilds;  /* rebuilds of the last call that re-sorted only the overflowing
                                  instances of a batch (counted in rebuilds too) */
  int graph_timed;         /* profiling mode 1: the last call's kernel timings came from event
                              record nodes of the replayed step graph (1) or from eager steps (0) */
  long long grown;         /* steps of the last call repeated after a contact capacity grew */
  int row_capacity;        /* contacts per sphere row now (grows as contacts need) */
  long long busy;          /* spheres with a contact after the last step */
  double window_ms;        /* device time of the steps pmx_set_window named in the last call, or -1 */
} pmx_stats;

/* per-kernel timing collected with HIP events on the context's stream */
typedef struct pmx_kernel_time
{
  char name[32];
  long long launches;
  double ms;               /* summed event time */
  double bytes;            /* summed algorithmic HBM bytes of the launches */
} pmx_kernel_time;

typedef struct pmx_ctx pmx_ctx;

/* Number of steps dem(duration, step) takes from the given curtime
 * (parmec.cpp:2418 accumulation `time < duration`), and the curtime value at
 * the start of each step if times != NULL (capacity cap). */
int pmx_step_times (double duration, double step, double curtime, double *times, int cap, int *nsteps);

/* Create a context on a device (replaces parmec::init, parmec.cpp:2227). */
int pmx_create (const pmx_options *opt, pmx_ctx **ctx);

/* Copy a state view to the device; resets contact history when the
 * particle/ellipsoid/triangle counts changed (parmec.cpp:288-1398 buffers). */
int pmx_upload (pmx_ctx *ctx, const pmx_model *model);

/* Run dem(duration, step) with output disabled (parmec.cpp:2315-2530).
 * elapsed receives the wall seconds of the step loop (timer.h:36-49). */
int pmx_dem (pmx_ctx *ctx, double duration, double step, const pmx_forcing *forcing, double *elapsed);

/* Copy the device state back into the caller's view. */
int pmx_download (pmx_ctx *ctx, pmx_model *model);

/*
 * Batches: count independent models stepped together (many small instances
 * fill the GPU where one would not).  The models are concatenated on the
 * device; each instance keeps its own partitioning tree, pairing table and
 * forcing, and its result is bitwise the result of stepping it alone.  All
 * instances must share curtime, curstep and stepnum.  No reference
 * counterpart: Parmec runs one model per process (parmec.h:36-283 globals),
 * and a sweep is a loop of separate runs (python/acc_sweep.py:23-131).
 */
int pmx_upload_batch (pmx_ctx *ctx, const pmx_model *models, int count);

/* dem(duration, step) for every instance; forcing holds one entry per
 * instance (or NULL: no gravity, no damping). */
int pmx_dem_batch (pmx_ctx *ctx, double duration, double step, const pmx_forcing *forcing, double *elapsed);

/* Copy every instance's state back into its own view. */
int pmx_download_batch (pmx_ctx *ctx, pmx_model *models, int count);

/* Contact pairs of one instance, numbered as in that instance's model. */
int pmx_contacts_instance (pmx_ctx *ctx, int instance, pmx_pair *out, size_t cap, size_t *count);

/* One instance's state into its own view. */
int pmx_download_instance (pmx_ctx *ctx, int instance, pmx_model *model);

/* dem() with one duration per instance: every instance takes its own step
 * count (the reference's accumulation for its duration) and is frozen once
 * done; done(user, i) is called from inside the call as soon as instance i
 * has finished, when pmx_download_instance and pmx_contacts_instance already
 * give its final state.  Afterwards the instances are at different times:
 * further dem calls on the context are refused.  No reference counterpart
 * (used by the request loop, pmx_server_*). */
typedef void (*pmx_done_fn) (void *user, int instance);
int pmx_dem_batch_durations (pmx_ctx *ctx, const double *durations, double step, const pmx_forcing *forcing,
                             pmx_done_fn done, void *user, double *elapsed);

/* Canonical contact pair set, sorted by (master, slave). */
int pmx_contacts (pmx_ctx *ctx, pmx_pair *out, size_t cap, size_t *count);

int pmx_get_stats (pmx_ctx *ctx, pmx_stats *stats);

/* Per-kernel event timing: 0 off; 1 every 8th step replays a copy of the
 * step graph with a HIP event pair around each kernel as event record nodes
 * (fresh events per launch; the other steps replay the plain graph, since
 * timing events slow the kernels around them); 2 every step enqueued
 * eagerly with an event pair per kernel.  Any call clears the totals;
 * launches of a step aborted by a leaf overflow are not counted.  Timing
 * stands in for the reference's dem() wall clock (timer.h:36-49). */
int pmx_set_profiling (pmx_ctx *ctx, int on);

/* Per-kernel totals since profiling was enabled; count receives the number
 * of entries (at most cap are written). */
int pmx_get_profile (pmx_ctx *ctx, pmx_kernel_time *out, int cap, int *count);

/* Measurement hook (no counterpart in the reference): time steps first..last
 * (0-based) of the next dem call on the device timeline, from an event
 * recorded on the context's stream before step `first` to one after step
 * `last` (a step repeated after an abort records its event again, so the
 * window holds the repeat and the rebuild before it), and, with profiling
 * on, time only those steps' kernels.  The result is pmx_stats.window_ms;
 * the window applies to that one call.  first < 0 clears it. */
int pmx_set_window (pmx_ctx *ctx, int first, int last);

/* Time histories of the resident model (single model contexts): the specs
 * are copied; pmx_upload clears them.  Particles of force and torque
 * histories keep their force and torque stored every step.  NULL or count
 * 0 removes them.  Replaces the registration half of HISTORY
 * (input.cpp:3157-3555); the sampling is pmx_forcing.hisrow. */
int pmx_set_history (pmx_ctx *ctx, const pmx_history *h);

/* The histories' values of the current state into out[count] (the sample
 * dem() takes before its first step when curtime is 0, parmec.cpp:2386-2394). */
int pmx_history_sample (pmx_ctx *ctx, double *out);

/* Energy and momentum of one instance's particles (instance -1: all of a
 * single model or of a batch at one time), out[9]: translational and
 * rotational kinetic energy, gravitational potential energy -sum m g.x for
 * the gravity g (NULL: zero), linear momentum (3) and angular momentum about
 * the origin (3), sum of x cross m v + L J Omega.  A diagnostic for energy
 * and momentum histories; no reference counterpart (the reference's
 * HISTORY has no energy entity, output.cpp:5143-5520). */
int pmx_energy (pmx_ctx *ctx, int instance, const double *g, double *out);

/* pmx_history_sample without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  The sample passes through one
 * static device buffer per GPU (PMX_HIS_NOALLOC_MAX doubles, held until the
 * process ends) shared by every context on that GPU, so calls that overlap
 * on one GPU race; more than PMX_HIS_NOALLOC_MAX histories is PMX_ECAPACITY. */
#define PMX_HIS_NOALLOC_MAX 4096
int pmx_history_sample_noalloc (pmx_ctx *ctx, double *out);

void pmx_destroy (pmx_ctx *ctx);

/* critical time step estimates, CRITICAL (input.cpp:3034-3154): per particle
 * (critical_perparticle, parmec.ispc:547-888; the contacts of the last step
 * taken and the springs) or per spring (critical_perspring,
 * parmec.ispc:485-535); h, o, r receive N or sprnum values: the step
 * estimate, the circular frequency and the damping ratio (o, r may be
 * NULL).  A particle that nothing loads has h = DBL_MAX.  The caller sorts
 * them and takes the smallest, as CRITICAL does.  Single-model contexts. */
enum { PMX_CRIT_PARTICLE = 0, PMX_CRIT_SPRING = 1 };
int pmx_critical (pmx_ctx *ctx, int which, double *h, double *o, double *r);

/* dem() with adaptive in (0, 1] (parmec.cpp:2315-2530, 2444-2447): every
 * step estimates each particle's critical step from what loads it in that
 * step (contacts_acc_task and the springs, forces.ispc:453-596, 1203-1411;
 * time_step_task, parmec.ispc:275-420) and takes step1 = adaptive times the
 * smallest, or step0 when no particle is loaded; the step's dynamics and
 * obstacles use step0 and step1 and time advances by (step0+step1)/2.  The
 * step count is known only as the call runs, so the host inputs come from
 * forcing->step_fn step by step.  Single-model contexts. */
int pmx_dem_adaptive (pmx_ctx *ctx, double duration, double step, double adaptive, const pmx_forcing *forcing,
                      double *elapsed);

const char *pmx_last_error (void);

/*
 * Resident request loop (serving).  Requests are independent simulations:
 * a model, one dem(duration, step) call and its forcing tables.  A scheduler
 * thread admits queued requests in arrival order while they fit the HBM
 * budget; requests queued within gather_s of each other (up to max_batch,
 * same step and time state) form one batch context stepped by its own worker
 * thread and HIP stream, so batches admitted at different times overlap on
 * the device (at most max_active batches at once).  Each request is frozen
 * and copied back as soon as its own steps are done; its result is bitwise
 * that of pmx_upload + pmx_dem on it alone.  No reference counterpart: the
 * resident form of a sweep's loop of separate runs (python/acc_sweep.py).
 */
enum { PMX_REQ_QUEUED = 0, PMX_REQ_RUNNING = 1, PMX_REQ_DONE = 2, PMX_REQ_FAILED = 3 };

typedef struct pmx_server_options
{
  int precision;        /* PMX_FP64 or PMX_MIXED */
  int device;
  int row_capacity;     /* 0: default */
  int max_batch;        /* requests per batch context (0: 256) */
  int max_active;       /* batch contexts stepping at once (0: 16) */
  double gather_s;      /* how long the oldest queued request waits for others to batch with */
  size_t hbm_budget;    /* bytes of device memory the admitted requests may use (0: 200 GiB) */
  double repack_s;      /* continuous batching: > 0, a running batch takes in queued requests of its kind
                         * (pausing at a step, carrying its unfinished requests over with their contact rows
                         * and tree snapshots) at most this often; 0: a batch runs until its last request ends */
  int longest_first;    /* nonzero: a batch takes the queued requests with the most steps first (they set how long
                         * a stream takes); 0: arrival order */
} pmx_server_options;

typedef struct pmx_request_info
{
  long long id;
  int status;           /* PMX_REQ_* */
  int batch;            /* requests in the batch it ran in */
  long long steps;
  long long particles;
  long long pairs;      /* contact pairs at the end */
  size_t bytes;         /* device memory charged to the budget */
  double submit_s, admit_s, done_s;   /* seconds since the server started */
} pmx_request_info;

typedef struct pmx_server_stats
{
  long long queued, running, done, failed, batches;
  int active_batches;
  size_t bytes_in_use;
  double particle_steps;   /* of the finished requests */
  double seconds;          /* since the server started */
} pmx_server_stats;

typedef struct pmx_server pmx_server;

int pmx_server_create (const pmx_server_options *opt, pmx_server **server);

/* Queue a request (the model and the forcing tables are copied). */
int pmx_server_submit (pmx_server *server, const pmx_model *model, double duration, double step,
                       const pmx_forcing *forcing, long long *id);

/* Wait for a request (timeout < 0: no limit; PMX_EBUSY when it runs out),
 * copy its final state into out (same counts as the submitted model; NULL
 * arrays are skipped) and release it. */
int pmx_server_wait (pmx_server *server, long long id, double timeout, pmx_model *out, pmx_request_info *info);

int pmx_server_get_stats (pmx_server *server, pmx_server_stats *stats);

/* Finish the admitted and queued requests, then stop. */
void pmx_server_destroy (pmx_server *server);

#ifdef __cplusplus
}
#endif

#endif
