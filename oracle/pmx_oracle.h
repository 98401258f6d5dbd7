/*
 * pmx_oracle.h -- CPU oracle of Parmec's DEM step path (TEST INFRASTRUCTURE).
 *
 * This is a scalar C++ restatement of the reference ISPC kernels, used only by
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg, as the
 * checker.  It is never linked into or called by the product library.
 *
 * Canonical choices where the reference is nondeterministic (documented per
 * function in pmx_oracle.cpp):
 *   - Morton codes are sorted stably (partition.ispc:97-243 is unstable);
 *   - tangential contact state starts at zero and moves with its pair
 *     (condet.ispc:52-81 leaves it uninitialised, forces.ispc:329-364 does not
 *     move it on compaction);
 *   - per-particle force/torque sums run sequentially over the particle's
 *     contacts ordered by the partner's identity (forces.ispc:492-577 sums in
 *     lane/task-dependent order); a meshed particle sums the reactions on its
 *     triangles triangle by triangle, spheres in index order within one.
 */
#ifndef PMX_ORACLE_H
#define PMX_ORACLE_H

#include "../include/pmx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

int orc_create (int nthreads, orc_ctx **ctx);
int orc_upload (orc_ctx *ctx, const pmx_model *model);
int orc_dem (orc_ctx *ctx, double duration, double step, const pmx_forcing *forcing, double *elapsed);
int orc_dem_adaptive (orc_ctx *ctx, double duration, double step, double adaptive, const pmx_forcing *forcing,
                      double *elapsed);
int orc_critical (orc_ctx *ctx, int which, double *h, double *o, double *r);
void orc_trig (int fn, const double *x, double *y, int n);   /* 0 sin, 1 cos, 2 acos (pmx_trig.h) */
double orc_time_step (const double kact[3], double kmax, double emax, const double krot[6], double mass,
                      const double inertia[9], double *o, double *r);
double orc_extremum_eigenvalue (const double A[9], int minimum);
int orc_download (orc_ctx *ctx, pmx_model *model);
int orc_contacts (orc_ctx *ctx, pmx_pair *out, size_t cap, size_t *count);
int orc_get_stats (orc_ctx *ctx, pmx_stats *stats);
void orc_destroy (orc_ctx *ctx);
const char *orc_last_error (void);

/* pure-function hooks for known-answer tests */
void orc_expmap (const double omega[3], double lambda[9]);
double orc_triangle_sphere (const double a[3], const double b[3], const double c[3],
                            const double p[3], double r, double point[3], double normal[3]);
/* leaf-reachability pair set of a freshly built tree over the given centres
 * (sphere-sphere pairs with positive depth, as (master, slave) ellipsoid
 * indices assuming ellipsoid i belongs to particle i); returns count, writes
 * at most cap pairs */
long long orc_tree_pairs (int n, const double *center, const double *radius, int *out, long long cap);
/* leaf of each centre in a fresh tree (depth-first leaf numbering);
 * returns the leaf count; *overflow receives the leaf overflow count */
int orc_tree_leaves (int n, const double *center, int *leafof, int *overflow);
/* leaves of a fresh tree reached by each box (lo xyz, hi xyz), in visiting
 * order, indexed by off[nbox+1]; returns the total */
long long orc_tree_reach (int n, const double *center, int nbox, const double *box, long long *off, int *leaves,
                          long long cap);
/* Morton code of one point given the extents (partition.ispc:69-94) */
unsigned int orc_morton (const double p[3], const double extents[6]);
/* amgcl's skyline_lu (Cuthill-McKee order, Crout on the non-zero profile)
 * of a dense n x n row-major matrix: x = A^-1 b; returns 1 on a zero pivot */
int orc_skyline_lu (int n, const double *A, const double *b, double *x);
/* amgcl's Cuthill-McKee order of a CSR pattern */
void orc_cuthill_mckee (int n, const int *ptr, const int *col, int *perm);

#ifdef __cplusplus
}
#endif

#endif
