/*
 * orc.h -- CPU FP64 restatement of the reference BSAM adaptive FAS path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the oracle that the HIP engine is
 * checked against. Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it; the product library never links or calls it.
 *
 * Every routine restates the reference algorithm in plain C with the same
 * floating-point operation order as the Fortran source it cites, so that on
 * the same inputs results agree to rounding (bitwise when neither side
 * contracts multiply-adds). Arrays use the reference storage layout:
 * column-major, variable index slowest, lower bound 1-mbc for q/qold/qc and
 * 1 for f/rf/ftmp (lib/bsamstorage.f90:46-123).
 */
#ifndef BSAM_ORC_H
#define BSAM_ORC_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXLEV 24

enum { ORC_Q = 0, ORC_QOLD = 1, ORC_QC = 2, ORC_F = 3, ORC_RF = 4, ORC_FTMP = 5, ORC_QRTE = 6 };

typedef struct orc_patch {
  int id;              /* caller's id (ngrid in the harness) */
  int level;
  int parent;          /* index into the hierarchy's patch array, -1 if none */
  int mx[3];
  int mb[3][2];        /* mbounds: footprint in parent-local coarse indices */
  int mg[3][2];        /* mglobal: bounds in the level's global index space */
  int mthbc[6];
  double dx;
  double xlower[3];
  int cmx[3];
  /* element strides and lower bounds of q/qold, qc and f/rf/ftmp */
  long qs[4], cs[4], fs[4];
  int ql[3], cl[3];
  long nq, nc, nf;     /* element counts */
  double *q, *qold, *qc, *f, *rf, *ftmp;
  /* relative truncation error qrte(1:cmx, 1:nrvars) and the error flags
     errorflags(1:mx) (AllocFields, lib/bsamstorage.f90:75-97) */
  long nr, nfl;
  double *qrte;
  int *errorflags;
} orc_patch;

typedef struct orc_hier {
  int ndims, nrvars, mbc;
  int minlevel, finestlevel;
  int npatches, cap;
  orc_patch *p;
  /* per-level patch lists in level-list order (ApplyOnLevel order) */
  int lev_off[ORC_MAXLEV + 1];
  int *lev_list;
  /* problem and solver parameters */
  double dt, eps2;
  double fc_alpha, fc_shift;   /* Dfc = fc_alpha*(c-fc_shift)^3 */
  double fe_num, fe_den;       /* Dfe = ((c-fc_shift)*fe_num)/fe_den */
  int nsmoothingpasses, maxvcycles;
  double qerrortol;
  int periodic;                /* root has periodic faces (set at commit) */
  double qpo[2];               /* quasi-periodic offsets qpo(1:nrvars) */
  int noff, offidx[13];        /* PeriodicSetup's offset list */
  int rootmx[3];
  char err[256];
} orc_hier;

orc_hier *orc_create(int ndims, int nrvars, int mbc, int minlevel);
void orc_destroy(orc_hier *h);
/* Patches must be added level by level (ascending), each level in
 * level-list order. Returns the patch index or -1. */
int orc_add_patch(orc_hier *h, int id, int level, int parent, const int *mx,
                  const int *mb, const int *mg, const int *mthbc, double dx,
                  const double *xlower);
int orc_commit(orc_hier *h);
void orc_set_problem(orc_hier *h, double dt, double eps2, double fc_alpha,
                     double fc_shift, double fe_num, double fe_den);
void orc_set_qpo(orc_hier *h, double qpo1, double qpo2);
void orc_set_solver(orc_hier *h, int nsmoothingpasses, int maxvcycles,
                    double qerrortol);
double *orc_field(orc_hier *h, int patch, int field);
long orc_field_len(orc_hier *h, int patch, int field);
const char *orc_error(orc_hier *h);

/* Reference operations (see orc.c for the file:line each one follows). */
int orc_set_ghost(orc_hier *h, int level, int ipass);
int orc_relax_level(orc_hier *h, int level, int color);
int orc_restrict_level(orc_hier *h, int level);
int orc_coarse_ghost_level(orc_hier *h, int level);
int orc_coarse_load_level(orc_hier *h, int level);
int orc_correct_level(orc_hier *h, int level);
int orc_fill_down(orc_hier *h, int field);
int orc_source(orc_hier *h);
int orc_update_source(orc_hier *h);
int orc_vcycle(orc_hier *h, int level);
double orc_error_afas(orc_hier *h, int level);
int orc_copy_qold(orc_hier *h);
/* EstimateLevelErrors' truncation error (lib/bsamroutines.f90:971-981) on
 * `level`: RestrictSolution, SetGhost(level-1, 0), GetCoarseGhostPoints,
 * qold = q on level and level-1, RelativeTruncationError into qrte. */
int orc_truncation_error(orc_hier *h, int level);
/* SetErrFlags2D/3D (lib/bsamroutines.f90:1093-1157) on every patch of
 * `level` from its qrte, with qtolerance(level). */
int orc_error_flags(orc_hier *h, int level, double qtolerance);
/* SetErrFlagsUser2D/3D (problem.f90:135-163, :425-452) on every patch of
 * `level` from q (ghosts as SetGhost(level, 0) leaves them). */
int orc_error_flags_user(orc_hier *h, int level, double qtolerance);
/* Mass and (2D) energy after a step (AfterStepStatistics, AfterStep2D/3D). */
int orc_after_step_stats(orc_hier *h, double *out);
int *orc_error_flag_field(orc_hier *h, int patch);
long orc_error_flag_len(orc_hier *h, int patch);
/* MultigridIterations; writes up to maxhist residuals, returns #cycles. */
int orc_multigrid_iterations(orc_hier *h, double *hist, int maxhist);
/* One time step on a fixed hierarchy: ghosts, FillDown, qold copy, solve. */
int orc_time_step(orc_hier *h, double *hist, int maxhist);

#ifdef __cplusplus
}
#endif
#endif
