/*
 * orc.c -- CPU FP64 restatement of the reference BSAM adaptive FAS path.
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h). Each function names the reference
 * file:line it restates. Floating-point expressions keep the reference's
 * left-to-right evaluation order; build with -ffp-contract=off.
 */
#include "orc.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Coarse-fine interpolation weights, lib/boundary.f90:45-54. */
static const double A1 = -1.0 / 5.0, A2 = 2.0 / 3.0, A3 = 8.0 / 15.0;
static const double B1 = 5.0 / 32.0, B2 = 15.0 / 16.0, B3 = -3.0 / 32.0;
static const double C1 = 9.0 / 16.0, C2 = 3.0 / 16.0, C3 = 3.0 / 16.0,
                    C4 = 1.0 / 16.0;

static int fail(orc_hier *h, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(h->err, sizeof h->err, fmt, ap);
  va_end(ap);
  return -1;
}

/* ---- array addressing ---------------------------------------------------- */
#define QX(p, i, j, k, v)                                                   \
  ((long)((i) - (p)->ql[0]) + (long)((j) - (p)->ql[1]) * (p)->qs[1] +       \
   (long)((k) - (p)->ql[2]) * (p)->qs[2] + (long)((v) - 1) * (p)->qs[3])
#define CX(p, i, j, k, v)                                                   \
  ((long)((i) - (p)->cl[0]) + (long)((j) - (p)->cl[1]) * (p)->cs[1] +       \
   (long)((k) - (p)->cl[2]) * (p)->cs[2] + (long)((v) - 1) * (p)->cs[3])
#define FX(p, i, j, k, v)                                                   \
  ((long)((i) - 1) + (long)((j) - 1) * (p)->fs[1] +                         \
   (long)((k) - 1) * (p)->fs[2] + (long)((v) - 1) * (p)->fs[3])
#define Q(p, i, j, k, v) ((p)->q[QX(p, i, j, k, v)])
#define QC(p, i, j, k, v) ((p)->qc[CX(p, i, j, k, v)])
#define F(p, i, j, k, v) ((p)->f[FX(p, i, j, k, v)])
#define RF(p, i, j, k, v) ((p)->rf[FX(p, i, j, k, v)])
#define FT(p, i, j, k, v) ((p)->ftmp[FX(p, i, j, k, v)])

/* ---- hierarchy ----------------------------------------------------------- */
orc_hier *orc_create(int ndims, int nrvars, int mbc, int minlevel) {
  orc_hier *h = calloc(1, sizeof *h);
  if (!h) return NULL;
  h->ndims = ndims;
  h->nrvars = nrvars;
  h->mbc = mbc;
  h->minlevel = minlevel;
  h->finestlevel = minlevel;
  h->cap = 64;
  h->p = calloc((size_t)h->cap, sizeof *h->p);
  h->nsmoothingpasses = 2;
  h->maxvcycles = 20;
  h->qerrortol = 1e-6;
  return h;
}

void orc_destroy(orc_hier *h) {
  if (!h) return;
  for (int n = 0; n < h->npatches; n++) {
    orc_patch *p = &h->p[n];
    free(p->q); free(p->qold); free(p->qc); free(p->f); free(p->rf);
    free(p->ftmp);
    free(p->qrte);
    free(p->errorflags);
  }
  free(h->p);
  free(h->lev_list);
  free(h);
}

/* Shapes follow AllocFields (lib/bsamstorage.f90:46-123); q and qold start
 * at zero (:116-117). f, rf, ftmp and qc are zeroed here too although the
 * reference leaves them undefined: every read of them follows a write. */
int orc_add_patch(orc_hier *h, int id, int level, int parent, const int *mx,
                  const int *mb, const int *mg, const int *mthbc, double dx,
                  const double *xlower) {
  if (level - h->minlevel >= ORC_MAXLEV) return fail(h, "level out of range");
  if (h->npatches == h->cap) {
    h->cap *= 2;
    orc_patch *np = realloc(h->p, (size_t)h->cap * sizeof *np);
    if (!np) return fail(h, "out of memory");
    h->p = np;
  }
  orc_patch *p = &h->p[h->npatches];
  memset(p, 0, sizeof *p);
  p->id = id;
  p->level = level;
  p->parent = parent;
  int nd = h->ndims, mbc = h->mbc;
  for (int d = 0; d < 3; d++) {
    p->mx[d] = d < nd ? mx[d] : 1;
    p->mb[d][0] = d < nd ? mb[2 * d] : 1;
    p->mb[d][1] = d < nd ? mb[2 * d + 1] : 1;
    p->mg[d][0] = d < nd ? mg[2 * d] : 1;
    p->mg[d][1] = d < nd ? mg[2 * d + 1] : 1;
    p->xlower[d] = d < nd ? xlower[d] : 0.0;
    p->cmx[d] = (p->mx[d] % 2 == 1) ? 1 : p->mx[d] / 2;
    if (d >= nd) p->cmx[d] = 1;
  }
  for (int f = 0; f < 6; f++) p->mthbc[f] = f < 2 * nd ? mthbc[f] : 1;
  p->dx = dx;
  long n[3], c[3];
  for (int d = 0; d < 3; d++) {
    int ext = d < nd;
    n[d] = ext ? p->mx[d] + 2 * mbc : 1;
    c[d] = ext ? p->cmx[d] + 2 * mbc : 1;
    p->ql[d] = ext ? 1 - mbc : 1;
    p->cl[d] = ext ? 1 - mbc : 1;
  }
  p->qs[0] = 1; p->qs[1] = n[0]; p->qs[2] = n[0] * n[1];
  p->qs[3] = n[0] * n[1] * n[2];
  p->cs[0] = 1; p->cs[1] = c[0]; p->cs[2] = c[0] * c[1];
  p->cs[3] = c[0] * c[1] * c[2];
  p->fs[0] = 1; p->fs[1] = p->mx[0]; p->fs[2] = (long)p->mx[0] * p->mx[1];
  p->fs[3] = (long)p->mx[0] * p->mx[1] * p->mx[2];
  p->nq = p->qs[3] * h->nrvars;
  p->nc = p->cs[3] * h->nrvars;
  p->nf = p->fs[3] * h->nrvars;
  p->q = calloc((size_t)p->nq, sizeof(double));
  p->qold = calloc((size_t)p->nq, sizeof(double));
  p->qc = calloc((size_t)p->nc, sizeof(double));
  p->f = calloc((size_t)p->nf, sizeof(double));
  p->rf = calloc((size_t)p->nf, sizeof(double));
  p->ftmp = calloc((size_t)p->nf, sizeof(double));
  p->nr = (long)p->cmx[0] * p->cmx[1] * p->cmx[2] * h->nrvars;
  p->nfl = p->fs[3];
  p->qrte = calloc((size_t)p->nr, sizeof(double));
  p->errorflags = calloc((size_t)p->nfl, sizeof(int));
  if (!p->q || !p->qold || !p->qc || !p->f || !p->rf || !p->ftmp || !p->qrte ||
      !p->errorflags)
    return fail(h, "out of memory");
  if (level > h->finestlevel) h->finestlevel = level;
  return h->npatches++;
}

int orc_commit(orc_hier *h) {
  free(h->lev_list);
  h->lev_list = malloc(sizeof(int) * (size_t)(h->npatches + 1));
  int nl = h->finestlevel - h->minlevel + 1, k = 0;
  for (int l = 0; l < nl; l++) {
    h->lev_off[l] = k;
    for (int n = 0; n < h->npatches; n++)
      if (h->p[n].level == l + h->minlevel) h->lev_list[k++] = n;
  }
  h->lev_off[nl] = k;
  for (int n = 0; n < h->npatches; n++) {
    orc_patch *p = &h->p[n];
    if (p->level > h->minlevel && (p->parent < 0 || p->parent >= h->npatches ||
                                   h->p[p->parent].level != p->level - 1))
      return fail(h, "patch %d: bad parent", n);
  }
  /* PeriodicSetup (lib/boundary.f90:893-963): offset vectors by the root's
   * periodic faces; mxmax(level) = root mx * 2^level (bsamroutines.f90:373-377) */
  h->periodic = 0;
  h->noff = 0;
  for (int n = 0; n < h->npatches; n++) {
    orc_patch *r = &h->p[n];
    if (r->level != 0) continue;
    int px = r->mthbc[0] == 2, py = r->mthbc[2] == 2, pz = h->ndims == 3 && r->mthbc[4] == 2;
    static const int t13[13] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13};
    const int *list = NULL;
    static const int l124[4] = {1, 2, 4, 5}, l1[1] = {1}, l1389[4] = {1, 3, 8, 9},
                     l2367[4] = {2, 3, 6, 7}, l2[1] = {2}, l3[1] = {3};
    if (px && py && pz) { list = t13; h->noff = 13; }
    else if (px && py) { list = l124; h->noff = 4; }
    else if (px && pz) { list = l1389; h->noff = 4; }
    else if (px) { list = l1; h->noff = 1; }
    else if (py && pz) { list = l2367; h->noff = 4; }
    else if (py) { list = l2; h->noff = 1; }
    else if (pz) { list = l3; h->noff = 1; }
    for (int k = 0; k < h->noff; k++) h->offidx[k] = list[k];
    h->periodic = h->noff > 0;
    for (int d = 0; d < 3; d++) h->rootmx[d] = r->mx[d];
  }
  return 0;
}

void orc_set_problem(orc_hier *h, double dt, double eps2, double fc_alpha,
                     double fc_shift, double fe_num, double fe_den) {
  h->dt = dt; h->eps2 = eps2; h->fc_alpha = fc_alpha; h->fc_shift = fc_shift;
  h->fe_num = fe_num; h->fe_den = fe_den;
}

/* qpo(1:nrvars), lib/nodeinfodef.f90:101 (0 unless set) */
void orc_set_qpo(orc_hier *h, double qpo1, double qpo2) {
  h->qpo[0] = qpo1;
  h->qpo[1] = qpo2;
}

void orc_set_solver(orc_hier *h, int nsmoothingpasses, int maxvcycles,
                    double qerrortol) {
  h->nsmoothingpasses = nsmoothingpasses; h->maxvcycles = maxvcycles;
  h->qerrortol = qerrortol;
}

double *orc_field(orc_hier *h, int n, int field) {
  if (n < 0 || n >= h->npatches) return NULL;
  orc_patch *p = &h->p[n];
  switch (field) {
    case ORC_Q: return p->q;
    case ORC_QOLD: return p->qold;
    case ORC_QC: return p->qc;
    case ORC_F: return p->f;
    case ORC_RF: return p->rf;
    case ORC_FTMP: return p->ftmp;
    case ORC_QRTE: return p->qrte;
  }
  return NULL;
}

long orc_field_len(orc_hier *h, int n, int field) {
  if (n < 0 || n >= h->npatches) return -1;
  orc_patch *p = &h->p[n];
  switch (field) {
    case ORC_Q: case ORC_QOLD: return p->nq;
    case ORC_QC: return p->nc;
    case ORC_F: case ORC_RF: case ORC_FTMP: return p->nf;
    case ORC_QRTE: return p->nr;
  }
  return -1;
}

int *orc_error_flag_field(orc_hier *h, int n) {
  return n < 0 || n >= h->npatches ? NULL : h->p[n].errorflags;
}
long orc_error_flag_len(orc_hier *h, int n) {
  return n < 0 || n >= h->npatches ? -1 : h->p[n].nfl;
}

const char *orc_error(orc_hier *h) { return h->err; }

static int nlev(orc_hier *h, int level) {
  int l = level - h->minlevel;
  return h->lev_off[l + 1] - h->lev_off[l];
}
static orc_patch *lev_patch(orc_hier *h, int level, int k) {
  return &h->p[h->lev_list[h->lev_off[level - h->minlevel] + k]];
}
static int level_ok(orc_hier *h, int level) {
  return level >= h->minlevel && level <= h->finestlevel;
}

/* ---- CH elemental terms, problems/simple_cahn-hilliard/problem.f90:760-813.
 * The shift/scale parameters make the same code serve the reference module
 * (shift 0.5, alpha, Dfe = (c-0.5)*alpha/4) and a +-1 double well
 * (shift 0, alpha 1, Dfe = c). Mob == 1 (:760-768). */
static inline double dfc(const orc_hier *h, double c) {
  double cmh = c - h->fc_shift;
  return cmh * cmh * cmh * h->fc_alpha;
}
static inline double d2fc(const orc_hier *h, double c) {
  double cmh = c - h->fc_shift;
  return 3.0 * cmh * cmh * h->fc_alpha;
}
static inline double dfe(const orc_hier *h, double c) {
  return (c - h->fc_shift) * h->fe_num / h->fe_den;
}

/* ---- RelaxGrid2D / RelaxGrid3D, problem.f90:164-215 and :453-518 --------- */
static void relax_patch(orc_hier *h, orc_patch *p, int ipass) {
  double hh = p->dx, h2 = hh * hh;
  double tmp1 = h->dt / h2, tmp2 = h->eps2 / h2;
  if (h->ndims == 2) {
    double tmp3 = 4.0 * h->eps2 / h2;
    double a12 = tmp1 * (((1.0 + 1.0) + 1.0) + 1.0);
    for (int j = 1; j <= p->mx[1]; j++)
      for (int i = 1 + (j + ipass) % 2; i <= p->mx[0]; i += 2) {
        double c = Q(p, i, j, 1, 1);
        double a21 = -d2fc(h, c) - tmp3;
        double b1 = F(p, i, j, 1, 1) +
                    tmp1 * (((Q(p, i - 1, j, 1, 2) + Q(p, i + 1, j, 1, 2)) +
                             Q(p, i, j - 1, 1, 2)) + Q(p, i, j + 1, 1, 2));
        double b2 = ((F(p, i, j, 1, 2) + dfc(h, c)) - d2fc(h, c) * c) -
                    tmp2 * (((Q(p, i + 1, j, 1, 1) + Q(p, i - 1, j, 1, 1)) +
                             Q(p, i, j + 1, 1, 1)) + Q(p, i, j - 1, 1, 1));
        double det = 1.0 - a12 * a21;
        Q(p, i, j, 1, 1) = (b1 - a12 * b2) / det;
        Q(p, i, j, 1, 2) = (b2 - a21 * b1) / det;
      }
  } else {
    double tmp3 = 6.0 * h->eps2 / h2;
    double a12 = tmp1 * (((((1.0 + 1.0) + 1.0) + 1.0) + 1.0) + 1.0);
    for (int k = 1; k <= p->mx[2]; k++)
      for (int j = 1; j <= p->mx[1]; j++)
        for (int i = 1 + (k + j + ipass) % 2; i <= p->mx[0]; i += 2) {
          double c = Q(p, i, j, k, 1);
          double a21 = -d2fc(h, c) - tmp3;
          double b1 =
              F(p, i, j, k, 1) +
              tmp1 * (((((Q(p, i - 1, j, k, 2) + Q(p, i + 1, j, k, 2)) +
                         Q(p, i, j - 1, k, 2)) + Q(p, i, j + 1, k, 2)) +
                       Q(p, i, j, k - 1, 2)) + Q(p, i, j, k + 1, 2));
          double b2 =
              ((F(p, i, j, k, 2) + dfc(h, c)) - d2fc(h, c) * c) -
              tmp2 * (((((Q(p, i + 1, j, k, 1) + Q(p, i - 1, j, k, 1)) +
                         Q(p, i, j + 1, k, 1)) + Q(p, i, j - 1, k, 1)) +
                       Q(p, i, j, k + 1, 1)) + Q(p, i, j, k - 1, 1));
          double det = 1.0 - a12 * a21;
          Q(p, i, j, k, 1) = (b1 - a12 * b2) / det;
          Q(p, i, j, k, 2) = (b2 - a21 * b1) / det;
        }
  }
}

/* ---- Operator2D/3D (problem.f90:231-257, :534-561) with ULapMob#D
 * (:304-329, :607-642), ULap#D and UDiv#D (lib/gridutilities.f90:42-87,
 * :345-404). a(...) evaluates the operator on an array with lower bound 0
 * given by an accessor; used for q (fine) and for qc (coarse, FAS). -------- */
typedef double (*acc_fn)(const orc_patch *p, int i, int j, int k, int v);
static double acc_q(const orc_patch *p, int i, int j, int k, int v) {
  return Q(p, i, j, k, v);
}
static double acc_qc(const orc_patch *p, int i, int j, int k, int v) {
  return QC(p, i, j, k, v);
}

static void operator_at(const orc_hier *h, const orc_patch *p, acc_fn a,
                        double hh, int i, int j, int k, double *r1,
                        double *r2) {
  double h2 = hh * hh, tmp1 = h->dt / h2, tmp2 = h->eps2 / h2;
  double c = a(p, i, j, k, 1), m = a(p, i, j, k, 2);
  if (h->ndims == 2) {
    double lapmob = (((a(p, i + 1, j, 1, 2) - m) - (m - a(p, i - 1, j, 1, 2))) +
                     (a(p, i, j + 1, 1, 2) - m)) - (m - a(p, i, j - 1, 1, 2));
    double lap = (((a(p, i + 1, j, 1, 1) + a(p, i - 1, j, 1, 1)) +
                   a(p, i, j + 1, 1, 1)) + a(p, i, j - 1, 1, 1)) - 4.0 * c;
    *r1 = c - tmp1 * lapmob;
    *r2 = (m - dfc(h, c)) + tmp2 * lap;
  } else {
    double lapmob =
        (((((a(p, i + 1, j, k, 2) - m) - (m - a(p, i - 1, j, k, 2))) +
           (a(p, i, j + 1, k, 2) - m)) - (m - a(p, i, j - 1, k, 2))) +
         (a(p, i, j, k + 1, 2) - m)) - (m - a(p, i, j, k - 1, 2));
    double lap = (((((a(p, i + 1, j, k, 1) + a(p, i - 1, j, k, 1)) +
                     a(p, i, j + 1, k, 1)) + a(p, i, j - 1, k, 1)) +
                   a(p, i, j, k + 1, 1)) + a(p, i, j, k - 1, 1)) - 6.0 * c;
    *r1 = c - tmp1 * lapmob;
    *r2 = (m - dfc(h, c)) + tmp2 * lap;
  }
}

/* Residual, lib/afasroutines.f90:799-840: rf = f - A_h(q). */
static void residual_patch(orc_hier *h, orc_patch *p) {
  int kmax = h->ndims == 3 ? p->mx[2] : 1;
  for (int k = 1; k <= kmax; k++)
    for (int j = 1; j <= p->mx[1]; j++)
      for (int i = 1; i <= p->mx[0]; i++) {
        double r1, r2;
        operator_at(h, p, acc_q, p->dx, i, j, k, &r1, &r2);
        RF(p, i, j, k, 1) = F(p, i, j, k, 1) - r1;
        RF(p, i, j, k, 2) = F(p, i, j, k, 2) - r2;
      }
}

/* Restriction2D/3D, lib/gridutilities.f90:89-110 and :406-433. */
static double restrict_at(const orc_hier *h, const double *a, const long *s,
                          const int *lo, int ci, int cj, int ck, int v) {
#define AX(i, j, k)                                                        \
  a[(long)((i) - lo[0]) + (long)((j) - lo[1]) * s[1] +                     \
    (long)((k) - lo[2]) * s[2] + (long)(v - 1) * s[3]]
  int i2 = 2 * ci, j2 = 2 * cj;
  if (h->ndims == 2)
    return 0.25 * (((AX(i2, j2, 1) + AX(i2 - 1, j2, 1)) + AX(i2, j2 - 1, 1)) +
                   AX(i2 - 1, j2 - 1, 1));
  int k2 = 2 * ck;
  return 0.125 *
         (((((((AX(i2, j2, k2) + AX(i2 - 1, j2, k2)) + AX(i2, j2 - 1, k2)) +
              AX(i2 - 1, j2 - 1, k2)) + AX(i2, j2, k2 - 1)) +
            AX(i2 - 1, j2, k2 - 1)) + AX(i2, j2 - 1, k2 - 1)) +
          AX(i2 - 1, j2 - 1, k2 - 1));
#undef AX
}

static const int ONE3[3] = {1, 1, 1};

/* RestrictSolution, lib/afasroutines.f90:600-644. */
static void restrict_solution_patch(orc_hier *h, orc_patch *p) {
  orc_patch *pa = &h->p[p->parent];
  int ck = h->ndims == 3 ? p->cmx[2] : 1;
  for (int v = 1; v <= h->nrvars; v++)
    for (int k = 1; k <= ck; k++)
      for (int j = 1; j <= p->cmx[1]; j++)
        for (int i = 1; i <= p->cmx[0]; i++)
          QC(p, i, j, k, v) = restrict_at(h, p->q, p->qs, p->ql, i, j, k, v);
  for (int v = 1; v <= h->nrvars; v++)
    for (int k = 1; k <= ck; k++)
      for (int j = 1; j <= p->cmx[1]; j++)
        for (int i = 1; i <= p->cmx[0]; i++)
          Q(pa, p->mb[0][0] + i - 1, p->mb[1][0] + j - 1,
            h->ndims == 3 ? p->mb[2][0] + k - 1 : 1, v) = QC(p, i, j, k, v);
}

/* FillDownLevel, lib/afasroutines.f90:241-310. field 1 = f, 2 = q. */
static void fill_down_patch(orc_hier *h, orc_patch *p, int field) {
  orc_patch *pa = &h->p[p->parent];
  int ck = h->ndims == 3 ? p->cmx[2] : 1;
  for (int v = 1; v <= h->nrvars; v++)
    for (int k = 1; k <= ck; k++)
      for (int j = 1; j <= p->cmx[1]; j++)
        for (int i = 1; i <= p->cmx[0]; i++) {
          int pi = p->mb[0][0] + i - 1, pj = p->mb[1][0] + j - 1;
          int pk = h->ndims == 3 ? p->mb[2][0] + k - 1 : 1;
          if (field == 1)
            F(pa, pi, pj, pk, v) = restrict_at(h, p->f, p->fs, ONE3, i, j, k, v);
          else
            Q(pa, pi, pj, pk, v) = restrict_at(h, p->q, p->qs, p->ql, i, j, k, v);
        }
}

/* GetCoarseGhostPoints, lib/boundary.f90:1203-1269. */
static void coarse_ghost_patch(orc_hier *h, orc_patch *p) {
  orc_patch *pa = &h->p[p->parent];
  int nd = h->ndims;
  int cm[3] = {p->cmx[0], p->cmx[1], nd == 3 ? p->cmx[2] : 1};
  int il[3], ih[3], jl[3], jh[3], mb0[3];
  for (int d = 0; d < 3; d++) {
    int ext = d < nd;
    il[d] = ext ? 0 : 1;
    ih[d] = ext ? cm[d] + 1 : 1;
    jl[d] = ext ? p->mb[d][0] - 1 : 1;
    jh[d] = ext ? p->mb[d][1] + 1 : 1;
    mb0[d] = ext ? p->mb[d][0] : 1;
  }
  for (int v = 1; v <= h->nrvars; v++) {
    /* faces 1 and 2 span il:ih in dims 2 and 3 */
    for (int k = il[2]; k <= ih[2]; k++)
      for (int j = il[1]; j <= ih[1]; j++) {
        int pj = jl[1] + (j - il[1]), pk = jl[2] + (k - il[2]);
        QC(p, il[0], j, k, v) = Q(pa, jl[0], pj, pk, v);
        QC(p, ih[0], j, k, v) = Q(pa, jh[0], pj, pk, v);
      }
    /* faces 3 and 4: 1:cmx(1) x il(3):ih(3) */
    for (int k = il[2]; k <= ih[2]; k++)
      for (int i = 1; i <= cm[0]; i++) {
        int pi = mb0[0] + i - 1, pk = jl[2] + (k - il[2]);
        QC(p, i, il[1], k, v) = Q(pa, pi, jl[1], pk, v);
        QC(p, i, ih[1], k, v) = Q(pa, pi, jh[1], pk, v);
      }
    if (nd == 3)
      for (int j = 1; j <= cm[1]; j++)
        for (int i = 1; i <= cm[0]; i++) {
          int pi = mb0[0] + i - 1, pj = mb0[1] + j - 1;
          QC(p, i, j, il[2], v) = Q(pa, pi, pj, jl[2], v);
          QC(p, i, j, ih[2], v) = Q(pa, pi, pj, jh[2], v);
        }
  }
}

/* CoarseLoadingFunction, lib/afasroutines.f90:646-714:
 * parent%f(mb) = Restriction(f - A_h(q)) + A_2h(qc) with ch = parent dx. */
static void coarse_load_patch(orc_hier *h, orc_patch *p) {
  orc_patch *pa = &h->p[p->parent];
  residual_patch(h, p);
  int ck = h->ndims == 3 ? p->cmx[2] : 1;
  for (int k = 1; k <= ck; k++)
    for (int j = 1; j <= p->cmx[1]; j++)
      for (int i = 1; i <= p->cmx[0]; i++) {
        double r1, r2;
        operator_at(h, p, acc_qc, pa->dx, i, j, k, &r1, &r2);
        int pi = p->mb[0][0] + i - 1, pj = p->mb[1][0] + j - 1;
        int pk = h->ndims == 3 ? p->mb[2][0] + k - 1 : 1;
        F(pa, pi, pj, pk, 1) = restrict_at(h, p->rf, p->fs, ONE3, i, j, k, 1) + r1;
        F(pa, pi, pj, pk, 2) = restrict_at(h, p->rf, p->fs, ONE3, i, j, k, 2) + r2;
      }
}

/* RelativeTruncationError, lib/afasroutines.f90:716-797:
 * qrte(1:cmx) = Restriction(A_h(q)) - A_2h(qc), the coarse operator with
 * ch = parent dx; A_h(q) goes to a temporary as in the reference (rf is not
 * touched). */
#define QR(p, i, j, k, v)                                                  \
  ((p)->qrte[(long)((i) - 1) + (long)((j) - 1) * (p)->cmx[0] +             \
             (long)((k) - 1) * (p)->cmx[0] * (p)->cmx[1] +                 \
             (long)((v) - 1) * (p)->cmx[0] * (p)->cmx[1] * (p)->cmx[2]])
static void truncation_error_patch(orc_hier *h, orc_patch *p) {
  orc_patch *pa = &h->p[p->parent];
  double *ah = malloc((size_t)p->nf * sizeof(double));
  if (!ah) return;
  int kmax = h->ndims == 3 ? p->mx[2] : 1;
  for (int k = 1; k <= kmax; k++)
    for (int j = 1; j <= p->mx[1]; j++)
      for (int i = 1; i <= p->mx[0]; i++) {
        double r1, r2;
        operator_at(h, p, acc_q, p->dx, i, j, k, &r1, &r2);
        ah[FX(p, i, j, k, 1)] = r1;
        ah[FX(p, i, j, k, 2)] = r2;
      }
  int ck = h->ndims == 3 ? p->mx[2] / 2 : 1;
  for (int k = 1; k <= ck; k++)
    for (int j = 1; j <= p->mx[1] / 2; j++)
      for (int i = 1; i <= p->mx[0] / 2; i++) {
        double c1, c2;
        operator_at(h, p, acc_qc, pa->dx, i, j, k, &c1, &c2);
        QR(p, i, j, k, 1) = restrict_at(h, ah, p->fs, ONE3, i, j, k, 1) - c1;
        QR(p, i, j, k, 2) = restrict_at(h, ah, p->fs, ONE3, i, j, k, 2) - c2;
      }
  free(ah);
}

/* SetErrFlags2D/3D, lib/bsamroutines.f90:1093-1157: flag the 2^d children
 * of every coarse cell whose largest |qrte| over the variables exceeds
 * qtolerance / h^d. */
static void error_flags_patch(orc_hier *h, orc_patch *p, double qtol) {
  const int nd = h->ndims;
  const double hh = p->dx;
  double tol = qtol / hh / hh;
  if (nd == 3) tol = tol / hh;
  for (long e = 0; e < p->nfl; e++) p->errorflags[e] = 0;
  int ck = nd == 3 ? p->cmx[2] : 1;
  for (int i = 1; i <= p->cmx[0]; i++)
    for (int j = 1; j <= p->cmx[1]; j++)
      for (int k = 1; k <= ck; k++) {
        double m = 0.0;
        for (int v = 1; v <= h->nrvars; v++) {
          double a = fabs(QR(p, i, j, k, v));
          if (v == 1 || a > m) m = a;
        }
        if (m > tol)
          for (int c = 0; c < (nd == 3 ? 2 : 1); c++)
            for (int b = 0; b < 2; b++)
              for (int a = 0; a < 2; a++) {
                int fi = 2 * i - a, fj = 2 * j - b, fk = nd == 3 ? 2 * k - c : 1;
                p->errorflags[FX(p, fi, fj, fk, 1)] = 1;
              }
      }
}

/* SetErrFlagsUser2D/3D (problems/simple_cahn-hilliard/problem.f90:135-163,
 * :425-452), ErrFlag's errflaguser case (lib/bsamroutines.f90:1061-1075):
 * every interior cell whose undivided central gradient of variable 1,
 * sqrt((q(i+1)-q(i-1))^2 + (q(j+1)-q(j-1))^2 [+ (q(k+1)-q(k-1))^2]), is
 * strictly above qtolerance(level) is flagged; all others are cleared. */
static void error_flags_user_patch(orc_hier *h, orc_patch *p, double qtol) {
  const int nd = h->ndims;
  int kz = nd == 3 ? p->mx[2] : 1;
  for (int k = 1; k <= kz; k++)
    for (int j = 1; j <= p->mx[1]; j++)
      for (int i = 1; i <= p->mx[0]; i++) {
        double gx = Q(p, i + 1, j, k, 1) - Q(p, i - 1, j, k, 1);
        double gy = Q(p, i, j + 1, k, 1) - Q(p, i, j - 1, k, 1);
        double s2 = gx * gx + gy * gy;
        if (nd == 3) {
          double gz = Q(p, i, j, k + 1, 1) - Q(p, i, j, k - 1, 1);
          s2 = s2 + gz * gz;
        }
        p->errorflags[FX(p, i, j, k, 1)] = fabs(sqrt(s2)) > qtol ? 1 : 0;
      }
}

/* The double well of the problem module, Ff(c) = alpha c^2 (c-1)^2 / 4
 * (problem.f90:769-780), written for minima lo, hi: the hierarchy's
 * Dfc = fc_alpha (c-s)^3 and Dfe = (fe_num/fe_den)(c-s) are its convex and
 * concave parts when lo, hi = s -+ sqrt((fe_num/fe_den)/fc_alpha) (0 and 1
 * for the reference's constants, so (c-lo) is c exactly). */
static void well_minima(const orc_hier *h, double *lo, double *hi) {
  double w = sqrt(h->fe_num / h->fe_den / h->fc_alpha);
  *lo = h->fc_shift - w;
  *hi = h->fc_shift + w;
}

static inline double ff(const orc_hier *h, double lo, double hi, double c) {
  double c0 = c - lo, cm1 = c - hi;
  return c0 * c0 * cm1 * cm1 * h->fc_alpha / 4.0;
}

/* Energy2D (problem.f90:330-361) of c(0:m1+1, 0:m2+1) = variable 1 of `src`
 * from cell (i0, j0) on, grid spacing hh. Sums run in array element order. */
static double energy2d(const orc_hier *h, const orc_patch *src, int i0, int j0, int m1,
                       int m2, double hh) {
  double lo, hi;
  well_minima(h, &lo, &hi);
#define C2(a, b) Q(src, i0 + (a), j0 + (b), 1, 1)
  const double h2 = hh * hh;
  double sff = 0.0;
  for (int j = 1; j <= m2; j++)
    for (int i = 1; i <= m1; i++) sff += ff(h, lo, hi, C2(i, j));
  double sg = 0.0;
  for (int j = 1; j <= m2; j++)
    for (int i = 1; i <= m1; i++) {
      double a = C2(i + 1, j) - C2(i, j);       /* f1(i, j) */
      double b = C2(i, j) - C2(i - 1, j);       /* f1(i-1, j) */
      double c = C2(i, j + 1) - C2(i, j);       /* f2(i, j) */
      double d = C2(i, j) - C2(i, j - 1);       /* f2(i, j-1) */
      sg += a * a + b * b + c * c + d * d;
    }
#undef C2
  const double tmp = h->eps2 / h2 / 4.0;
  return h2 * (sff + tmp * sg);
}

/* AfterStep2D/3D (problem.f90:102-134, :398-424) through AfterStepStatistics
 * (lib/bsamroutines.f90:896-939): mass (variable 1) and, in 2D, energy of the
 * composite grid, each refined patch adding its own sum and removing its
 * parent's cells under it. */
static void after_step_patch(orc_hier *h, orc_patch *p, double out[2]) {
  const int nd = h->ndims;
  const orc_patch *pa = &h->p[p->parent];
  const double hh = p->dx, ch = hh * 2.0;
  int kz = nd == 3 ? p->mx[2] : 1;
  double sq = 0.0;
  for (int k = 1; k <= kz; k++)
    for (int j = 1; j <= p->mx[1]; j++)
      for (int i = 1; i <= p->mx[0]; i++) sq += Q(p, i, j, k, 1);
  const int cm[3] = {p->mx[0] / 2, p->mx[1] / 2, nd == 3 ? p->mx[2] / 2 : 1};
  double sc = 0.0;
  if (p->level != 0) {
    int ck = nd == 3 ? cm[2] : 1;
    for (int k = 1; k <= ck; k++)
      for (int j = 1; j <= cm[1]; j++)
        for (int i = 1; i <= cm[0]; i++)
          sc += Q(pa, p->mb[0][0] - 1 + i, p->mb[1][0] - 1 + j,
                  nd == 3 ? p->mb[2][0] - 1 + k : 1, 1);
  }
  if (nd == 2) {
    const double h2 = hh * hh, ch2 = ch * ch;
    if (p->level == 0) {
      out[0] = out[0] + h2 * sq;
      out[1] = out[1] + energy2d(h, p, 0, 0, p->mx[0], p->mx[1], hh);
    } else {
      out[0] = out[0] + h2 * sq - ch2 * sc;
      out[1] = out[1] + energy2d(h, p, 0, 0, p->mx[0], p->mx[1], hh) -
               energy2d(h, pa, p->mb[0][0] - 1, p->mb[1][0] - 1, cm[0], cm[1], ch);
    }
  } else {
    const double h3 = hh * hh * hh, ch3 = ch * ch * ch;
    if (p->level == 0) out[0] = out[0] + h3 * sq;
    else out[0] = out[0] + h3 * sq - ch3 * sc;
  }
}

/* CorrectFine, lib/afasroutines.f90:842-963, mbc = 1: the coarse correction
 * (parent q minus the restricted qc, ghost ring included) is added to q over
 * 0:mx+1 with BiLinProlongationP1 / TriLinProlongationP1
 * (lib/gridutilities.f90:137-175, :474-529). */
static void correct_patch(orc_hier *h, orc_patch *p) {
  orc_patch *pa = &h->p[p->parent];
  int nd = h->ndims;
  int cm[3] = {p->cmx[0], p->cmx[1], nd == 3 ? p->cmx[2] : 1};
  int k0 = nd == 3 ? 0 : 1, k1 = nd == 3 ? cm[2] + 1 : 1;
  for (int v = 1; v <= h->nrvars; v++)
    for (int k = k0; k <= k1; k++)
      for (int j = 0; j <= cm[1] + 1; j++)
        for (int i = 0; i <= cm[0] + 1; i++) {
          int pk = nd == 3 ? p->mb[2][0] - 1 + k : 1;
          QC(p, i, j, k, v) =
              Q(pa, p->mb[0][0] - 1 + i, p->mb[1][0] - 1 + j, pk, v) -
              QC(p, i, j, k, v);
        }
  int mxx = p->mx[0], myy = p->mx[1];
  if (nd == 2) {
    /* b(0:mx+1, 0:cmy+1): x pass first, then y pass divided by 16 */
    long bn = (long)(mxx + 2) * (cm[1] + 2);
    double *b = malloc(sizeof(double) * (size_t)bn);
    for (int v = 1; v <= h->nrvars; v++) {
      for (int j = 0; j <= cm[1] + 1; j++)
        for (int kk = 0; kk <= cm[0]; kk++) {
          double a0 = QC(p, kk, j, 1, v), a1 = QC(p, kk + 1, j, 1, v);
          b[2 * kk + (long)j * (mxx + 2)] = 3.0 * a0 + a1;
          b[2 * kk + 1 + (long)j * (mxx + 2)] = a0 + 3.0 * a1;
        }
      for (int kk = 0; kk <= cm[1]; kk++)
        for (int i = 0; i <= mxx + 1; i++) {
          double b0 = b[i + (long)kk * (mxx + 2)],
                 b1 = b[i + (long)(kk + 1) * (mxx + 2)];
          Q(p, i, 2 * kk, 1, v) = Q(p, i, 2 * kk, 1, v) + (3.0 * b0 + b1) / 16.0;
          Q(p, i, 2 * kk + 1, 1, v) =
              Q(p, i, 2 * kk + 1, 1, v) + (b0 + 3.0 * b1) / 16.0;
        }
    }
    free(b);
  } else {
    int mzz = p->mx[2];
    long nbx = mxx + 2, nby = cm[1] + 2, ncy = myy + 2, ncz = cm[2] + 2;
    double *b = malloc(sizeof(double) * (size_t)(nbx * nby * ncz));
    double *c = malloc(sizeof(double) * (size_t)(nbx * ncy * ncz));
    for (int v = 1; v <= h->nrvars; v++) {
      for (int k = 0; k <= cm[2] + 1; k++)
        for (int j = 0; j <= cm[1] + 1; j++)
          for (int kk = 0; kk <= cm[0]; kk++) {
            double a0 = QC(p, kk, j, k, v), a1 = QC(p, kk + 1, j, k, v);
            b[2 * kk + j * nbx + k * nbx * nby] = 3.0 * a0 + a1;
            b[2 * kk + 1 + j * nbx + k * nbx * nby] = a0 + 3.0 * a1;
          }
      for (int k = 0; k <= cm[2] + 1; k++)
        for (int kk = 0; kk <= cm[1]; kk++)
          for (int i = 0; i <= mxx + 1; i++) {
            double b0 = b[i + kk * nbx + k * nbx * nby],
                   b1 = b[i + (kk + 1) * nbx + k * nbx * nby];
            c[i + (2 * kk) * nbx + k * nbx * ncy] = 3.0 * b0 + b1;
            c[i + (2 * kk + 1) * nbx + k * nbx * ncy] = b0 + 3.0 * b1;
          }
      for (int kk = 0; kk <= cm[2]; kk++)
        for (int j = 0; j <= myy + 1; j++)
          for (int i = 0; i <= mxx + 1; i++) {
            double c0 = c[i + j * nbx + kk * nbx * ncy],
                   c1 = c[i + j * nbx + (kk + 1) * nbx * ncy];
            Q(p, i, j, 2 * kk, v) = Q(p, i, j, 2 * kk, v) + (3.0 * c0 + c1) / 64.0;
            Q(p, i, j, 2 * kk + 1, v) =
                Q(p, i, j, 2 * kk + 1, v) + (c0 + 3.0 * c1) / 64.0;
          }
    }
    (void)mzz;
    free(b);
    free(c);
  }
}

/* ---- Interpolate2D, lib/boundary.f90:135-323 ------------------------------ */
static void interpolate2d(orc_hier *h, orc_patch *p, int ipass) {
  orc_patch *pa = &h->p[p->parent];
  int mx = p->mx[0], my = p->mx[1];
  int mb11 = p->mb[0][0], mb12 = p->mb[0][1], mb21 = p->mb[1][0],
      mb22 = p->mb[1][1];
  int red = ipass == 1 || ipass == 0, black = ipass == 2 || ipass == 0;
#define QF(i, j) Q(p, i, j, 1, v)
#define CQ(i, j) Q(pa, i, j, 1, v)
  for (int v = 1; v <= h->nrvars; v++) {
    if (red) {
      QF(mx + 1, my + 1) = A1 * QF(mx - 1, my - 1) + A2 * QF(mx, my) +
                           A3 * CQ(mb12 + 1, mb22 + 1);
      QF(0, 0) = A1 * QF(2, 2) + A2 * QF(1, 1) + A3 * CQ(mb11 - 1, mb21 - 1);
    }
    if (black) {
      QF(mx + 1, 0) = A1 * QF(mx - 1, 2) + A2 * QF(mx, 1) +
                      A3 * CQ(mb12 + 1, mb21 - 1);
      QF(0, my + 1) = A1 * QF(2, my - 1) + A2 * QF(1, my) +
                      A3 * CQ(mb11 - 1, mb22 + 1);
    }
    int t = p->mthbc[0];
    if (t == 2 || t == 999) {
      if (red) {
        for (int j = 2; j <= my - 2; j += 2)
          QF(0, j) = A1 * QF(2, j + 2) + A2 * QF(1, j + 1) +
                     A3 * CQ(mb11 - 1, mb21 + (j - 2) / 2);
        QF(0, my) = A1 * QF(2, my) + A2 * QF(1, my) +
                    A3 * (B1 * CQ(mb11 - 1, mb22 + 1) + B2 * CQ(mb11 - 1, mb22) +
                          B3 * CQ(mb11 - 1, mb22 - 1));
      }
      if (black) {
        for (int j = 3; j <= my - 1; j += 2)
          QF(0, j) = A1 * QF(2, j - 2) + A2 * QF(1, j - 1) +
                     A3 * CQ(mb11 - 1, mb21 + (j - 1) / 2);
        QF(0, 1) = A1 * QF(2, 1) + A2 * QF(1, 1) +
                   A3 * (B1 * CQ(mb11 - 1, mb21 - 1) + B2 * CQ(mb11 - 1, mb21) +
                         B3 * CQ(mb11 - 1, mb21 + 1));
      }
    }
    t = p->mthbc[1];
    if (t == 2 || t == 999) {
      if (red) {
        for (int j = 3; j <= my - 1; j += 2)
          QF(mx + 1, j) = A1 * QF(mx - 1, j - 2) + A2 * QF(mx, j - 1) +
                          A3 * CQ(mb12 + 1, mb21 + (j - 1) / 2);
        QF(mx + 1, 1) = A1 * QF(mx - 1, 1) + A2 * QF(mx, 1) +
                        A3 * (B1 * CQ(mb12 + 1, mb21 - 1) +
                              B2 * CQ(mb12 + 1, mb21) + B3 * CQ(mb12 + 1, mb21 + 1));
      }
      if (black) {
        for (int j = 2; j <= my - 2; j += 2)
          QF(mx + 1, j) = A1 * QF(mx - 1, j + 2) + A2 * QF(mx, j + 1) +
                          A3 * CQ(mb12 + 1, mb21 + (j - 2) / 2);
        QF(mx + 1, my) = A1 * QF(mx - 1, my) + A2 * QF(mx, my) +
                         A3 * (B1 * CQ(mb12 + 1, mb22 + 1) +
                               B2 * CQ(mb12 + 1, mb22) + B3 * CQ(mb12 + 1, mb22 - 1));
      }
    }
    t = p->mthbc[2];
    if (t == 2 || t == 999) {
      if (red) {
        for (int i = 2; i <= mx - 2; i += 2)
          QF(i, 0) = A1 * QF(i + 2, 2) + A2 * QF(i + 1, 1) +
                     A3 * CQ(mb11 + (i - 2) / 2, mb21 - 1);
        QF(mx, 0) = A1 * QF(mx, 2) + A2 * QF(mx, 1) +
                    A3 * (B1 * CQ(mb12 + 1, mb21 - 1) + B2 * CQ(mb12, mb21 - 1) +
                          B3 * CQ(mb12 - 1, mb21 - 1));
      }
      if (black) {
        for (int i = 3; i <= mx - 1; i += 2)
          QF(i, 0) = A1 * QF(i - 2, 2) + A2 * QF(i - 1, 1) +
                     A3 * CQ(mb11 + (i - 1) / 2, mb21 - 1);
        QF(1, 0) = A1 * QF(1, 2) + A2 * QF(1, 1) +
                   A3 * (B1 * CQ(mb11 - 1, mb21 - 1) + B2 * CQ(mb11, mb21 - 1) +
                         B3 * CQ(mb11 + 1, mb21 - 1));
      }
    }
    t = p->mthbc[3];
    if (t == 2 || t == 999) {
      if (red) {
        for (int i = 3; i <= mx - 1; i += 2)
          QF(i, my + 1) = A1 * QF(i - 2, my - 1) + A2 * QF(i - 1, my) +
                          A3 * CQ(mb11 + (i - 1) / 2, mb22 + 1);
        QF(1, my + 1) = A1 * QF(1, my - 1) + A2 * QF(1, my) +
                        A3 * (B1 * CQ(mb11 - 1, mb22 + 1) +
                              B2 * CQ(mb11, mb22 + 1) + B3 * CQ(mb11 + 1, mb22 + 1));
      }
      if (black) {
        for (int i = 2; i <= mx - 2; i += 2)
          QF(i, my + 1) = A1 * QF(i + 2, my - 1) + A2 * QF(i + 1, my) +
                          A3 * CQ(mb11 + (i - 2) / 2, mb22 + 1);
        QF(mx, my + 1) = A1 * QF(mx, my - 1) + A2 * QF(mx, my) +
                         A3 * (B1 * CQ(mb12 + 1, mb22 + 1) +
                               B2 * CQ(mb12, mb22 + 1) + B3 * CQ(mb12 - 1, mb22 + 1));
      }
    }
  }
#undef QF
#undef CQ
}

int orc_interpolate3d(orc_hier *h, orc_patch *p, int ipass);

/* GhostOverlap, lib/boundary.f90:346-415: both directions of a same-level
 * pair; each copies interior cells of one patch into the ghost-extended
 * overlap of the other. */
static void ghost_overlap(orc_hier *h, orc_patch *g1, orc_patch *g2) {
  int nd = h->ndims, mbc = h->mbc;
  int ov[3][2], o1[3][2], o2[3][2];
  for (int n = 0; n < 3; n++) {
    int a1 = n < nd ? g1->mg[n][0] : 1, a2 = n < nd ? g1->mg[n][1] : 1;
    int b1 = n < nd ? g2->mg[n][0] : 1, b2 = n < nd ? g2->mg[n][1] : 1;
    ov[n][0] = (a1 - mbc > b1 - mbc) ? a1 - mbc : b1 - mbc;
    ov[n][1] = (a2 + mbc < b2 + mbc) ? a2 + mbc : b2 + mbc;
    if (n < nd && ov[n][1] - ov[n][0] < 0) return;
    o1[n][0] = ov[n][0] > b1 ? ov[n][0] : b1;
    o1[n][1] = ov[n][1] < b2 ? ov[n][1] : b2;
    o2[n][0] = ov[n][0] > a1 ? ov[n][0] : a1;
    o2[n][1] = ov[n][1] < a2 ? ov[n][1] : a2;
  }
  int g1lo[3], g2lo[3];
  for (int n = 0; n < 3; n++) {
    g1lo[n] = n < nd ? g1->mg[n][0] : 1;
    g2lo[n] = n < nd ? g2->mg[n][0] : 1;
  }
  for (int v = 1; v <= h->nrvars; v++)
    for (int k = o1[2][0]; k <= o1[2][1]; k++)
      for (int j = o1[1][0]; j <= o1[1][1]; j++)
        for (int i = o1[0][0]; i <= o1[0][1]; i++)
          Q(g1, i - g1lo[0] + 1, j - g1lo[1] + 1, k - g1lo[2] + 1, v) =
              Q(g2, i - g2lo[0] + 1, j - g2lo[1] + 1, k - g2lo[2] + 1, v);
  for (int v = 1; v <= h->nrvars; v++)
    for (int k = o2[2][0]; k <= o2[2][1]; k++)
      for (int j = o2[1][0]; j <= o2[1][1]; j++)
        for (int i = o2[0][0]; i <= o2[0][1]; i++)
          Q(g2, i - g2lo[0] + 1, j - g2lo[1] + 1, k - g2lo[2] + 1, v) =
              Q(g1, i - g1lo[0] + 1, j - g1lo[1] + 1, k - g1lo[2] + 1, v);
}

/* SetBC2D/SetBC3D (lib/boundary.f90:643-891) with the CH module's UserBC2D/3D
 * (problem.f90:646-756) for codes 10..99. Faces are applied in order
 * 1..2*ndims over the full ghost-extended tangential range, so later faces
 * overwrite shared corner ghosts. */
static int set_bc_patch(orc_hier *h, orc_patch *p) {
  int nd = h->ndims, mbc = h->mbc;
  int lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    lo[d] = d < nd ? 1 - mbc : 1;
    hi[d] = d < nd ? p->mx[d] + mbc : 1;
  }
  for (int face = 0; face < 2 * nd; face++) {
    int t = p->mthbc[face];
    if (t == 2 || t == 999) continue;
    if (t == 3) return fail(h, "SetBC: mthbc = 3 not in use");
    int mode;
    if (t == 1) mode = 0;                 /* reflect */
    else if (t >= 10 && t <= 99) {
      /* UserBC2D: faces 1,2 give 2-q, faces 3,4 reflect; UserBC3D: 2-q */
      mode = (nd == 3 || face < 2) ? 1 : 0;
    } else return fail(h, "SetBC: boundary code %d not coded", t);
    int d = face / 2, right = face % 2;
    int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    for (int v = 1; v <= h->nrvars; v++)
      for (int ibc = 1; ibc <= mbc; ibc++)
        for (int b = lo[tb]; b <= hi[tb]; b++)
          for (int a = lo[ta]; a <= hi[ta]; a++) {
            int dst = right ? p->mx[d] + ibc : 1 - ibc;
            int src = right ? p->mx[d] - ibc + 1 : ibc;
            int di[3], si[3];
            di[d] = dst; si[d] = src;
            di[ta] = si[ta] = a;
            di[tb] = si[tb] = b;
            double val = Q(p, si[0], si[1], si[2], v);
            Q(p, di[0], di[1], di[2], v) = mode ? 2.0 - val : val;
          }
  }
  return 0;
}

/* PeriodicBC, lib/boundary.f90:417-486: a patch periodic on both sides of a
 * dimension wraps its own interior into its ghosts, dimensions in order, over
 * the full ghost-extended range; in dimension 1 the high ghosts take
 * +qpo(n), the low ones -qpo(n) (:441-442, :460-461). */
static void periodic_bc(orc_hier *h, orc_patch *p) {
  int nd = h->ndims;
  for (int d = 0; d < nd; d++) {
    if (!(p->mthbc[2 * d] == 2 && p->mthbc[2 * d + 1] == 2)) continue;
    int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    int alo = p->ql[ta], ahi = ta < nd ? p->mx[ta] + h->mbc : 1;
    int blo = p->ql[tb], bhi = tb < nd ? p->mx[tb] + h->mbc : 1;
    for (int v = 1; v <= h->nrvars; v++)
      for (int ibc = 1; ibc <= h->mbc; ibc++)
        for (int b = blo; b <= bhi; b++)
          for (int a = alo; a <= ahi; a++) {
            int hi[3], lo[3], shi[3], slo[3];
            hi[d] = p->mx[d] + ibc; shi[d] = ibc;
            lo[d] = 1 - ibc; slo[d] = p->mx[d] - ibc + 1;
            hi[ta] = lo[ta] = shi[ta] = slo[ta] = a;
            hi[tb] = lo[tb] = shi[tb] = slo[tb] = b;
            if (d == 0) {
              Q(p, hi[0], hi[1], hi[2], v) = Q(p, shi[0], shi[1], shi[2], v) + h->qpo[v - 1];
              Q(p, lo[0], lo[1], lo[2], v) = Q(p, slo[0], slo[1], slo[2], v) - h->qpo[v - 1];
            } else {
              Q(p, hi[0], hi[1], hi[2], v) = Q(p, shi[0], shi[1], shi[2], v);
              Q(p, lo[0], lo[1], lo[2], v) = Q(p, slo[0], slo[1], slo[2], v);
            }
          }
  }
}

/* GetPeriodicOffsets, lib/boundary.f90:965-1044: the offset vector of index
 * idx; its qoffset is qpo times the sign of the offset's x component. */
static void periodic_offset(const orc_hier *h, int level, int idx, int off[3]) {
  int W = h->rootmx[0] << level, H = h->rootmx[1] << level, D = h->rootmx[2] << level;
  static const int tab[13][3] = {{1, 0, 0},  {0, 1, 0},  {0, 0, 1},  {1, 1, 0},  {-1, 1, 0},
                                 {0, 1, 1},  {0, 1, -1}, {1, 0, 1},  {-1, 0, 1}, {1, 1, 1},
                                 {1, 1, -1}, {-1, 1, 1}, {1, -1, 1}};
  off[0] = tab[idx - 1][0] * W;
  off[1] = tab[idx - 1][1] * H;
  off[2] = tab[idx - 1][2] * D;
}

/* PeriodicGhostOverlap, lib/boundary.f90:510-596: for every periodic image
 * of g2 (g2 shifted by an offset), the ghost-extended overlap with g1 is
 * filled both ways as in GhostOverlap. */
static void periodic_overlap(orc_hier *h, int level, orc_patch *g1, orc_patch *g2) {
  int nd = h->ndims, mbc = h->mbc;
  for (int o = 0; o < h->noff; o++) {
    int off[3];
    periodic_offset(h, level, h->offidx[o], off);
    int a1[3], a2[3], b1[3], b2[3], ov[3][2], o1[3][2], o2[3][2];
    int empty = 0;
    for (int n = 0; n < 3; n++) {
      a1[n] = n < nd ? g1->mg[n][0] : 1; a2[n] = n < nd ? g1->mg[n][1] : 1;
      b1[n] = n < nd ? g2->mg[n][0] + off[n] : 1; b2[n] = n < nd ? g2->mg[n][1] + off[n] : 1;
      ov[n][0] = (a1[n] - mbc > b1[n] - mbc) ? a1[n] - mbc : b1[n] - mbc;
      ov[n][1] = (a2[n] + mbc < b2[n] + mbc) ? a2[n] + mbc : b2[n] + mbc;
      if (n < nd && ov[n][1] - ov[n][0] < 0) empty = 1;
      o1[n][0] = ov[n][0] > b1[n] ? ov[n][0] : b1[n];
      o1[n][1] = ov[n][1] < b2[n] ? ov[n][1] : b2[n];
      o2[n][0] = ov[n][0] > a1[n] ? ov[n][0] : a1[n];
      o2[n][1] = ov[n][1] < a2[n] ? ov[n][1] : a2[n];
    }
    if (empty) continue;
    /* qoffset(offset, n): +qpo(n) for an x offset > 0, -qpo(n) for < 0, else 0
       (as a value: 0.0 when the x offset is 0, whatever qpo) */
    for (int v = 1; v <= h->nrvars; v++) {
      const double qo = off[0] > 0 ? h->qpo[v - 1] : off[0] < 0 ? -h->qpo[v - 1] : 0.0;
      for (int k = o1[2][0]; k <= o1[2][1]; k++)
        for (int j = o1[1][0]; j <= o1[1][1]; j++)
          for (int i = o1[0][0]; i <= o1[0][1]; i++)
            Q(g1, i - a1[0] + 1, j - a1[1] + 1, k - a1[2] + 1, v) =
                Q(g2, i - b1[0] + 1, j - b1[1] + 1, k - b1[2] + 1, v) + qo;
    }
    for (int v = 1; v <= h->nrvars; v++) {
      const double qo = off[0] > 0 ? h->qpo[v - 1] : off[0] < 0 ? -h->qpo[v - 1] : 0.0;
      for (int k = o2[2][0]; k <= o2[2][1]; k++)
        for (int j = o2[1][0]; j <= o2[1][1]; j++)
          for (int i = o2[0][0]; i <= o2[0][1]; i++)
            Q(g2, i - b1[0] + 1, j - b1[1] + 1, k - b1[2] + 1, v) =
                Q(g1, i - a1[0] + 1, j - a1[1] + 1, k - a1[2] + 1, v) - qo;
    }
  }
}

/* SetGhost, lib/boundary.f90:58-89. */
int orc_set_ghost(orc_hier *h, int level, int ipass) {
  if (!level_ok(h, level)) return fail(h, "set_ghost: bad level %d", level);
  int n = nlev(h, level);
  if (level > 0) {
    for (int a = 0; a < n; a++) {
      orc_patch *p = lev_patch(h, level, a);
      if (h->ndims == 2) interpolate2d(h, p, ipass);
      else if (orc_interpolate3d(h, p, ipass)) return -1;
    }
    /* ApplyOnLevelPairs order, lib/treeops.f90:574-610 */
    for (int a = 0; a < n; a++)
      for (int b = a + 1; b < n; b++)
        ghost_overlap(h, lev_patch(h, level, a), lev_patch(h, level, b));
  }
  if (h->periodic) {
    for (int a = 0; a < n; a++) periodic_bc(h, lev_patch(h, level, a));
    if (level > 0)
      for (int a = 0; a < n; a++)
        for (int b = a + 1; b < n; b++) {
          periodic_overlap(h, level, lev_patch(h, level, a), lev_patch(h, level, b));
          periodic_overlap(h, level, lev_patch(h, level, b), lev_patch(h, level, a));
        }
  }
  for (int a = 0; a < n; a++)
    if (set_bc_patch(h, lev_patch(h, level, a))) return -1;
  return 0;
}

int orc_relax_level(orc_hier *h, int level, int color) {
  if (!level_ok(h, level)) return fail(h, "relax: bad level");
  for (int a = 0; a < nlev(h, level); a++)
    relax_patch(h, lev_patch(h, level, a), color);
  return 0;
}

#define FOR_LEVEL_CHILDREN(fn)                                               \
  if (!level_ok(h, level) || level <= h->minlevel)                           \
    return fail(h, #fn ": bad level %d", level);                             \
  for (int a = 0; a < nlev(h, level); a++) fn(h, lev_patch(h, level, a));    \
  return 0;

int orc_restrict_level(orc_hier *h, int level) {
  FOR_LEVEL_CHILDREN(restrict_solution_patch)
}
int orc_coarse_ghost_level(orc_hier *h, int level) {
  FOR_LEVEL_CHILDREN(coarse_ghost_patch)
}
int orc_coarse_load_level(orc_hier *h, int level) {
  FOR_LEVEL_CHILDREN(coarse_load_patch)
}
int orc_correct_level(orc_hier *h, int level) {
  FOR_LEVEL_CHILDREN(correct_patch)
}

int orc_truncation_error(orc_hier *h, int level) {
  if (!level_ok(h, level) || level <= h->minlevel)
    return fail(h, "truncation_error: bad level %d", level);
  if (orc_restrict_level(h, level) || orc_set_ghost(h, level - 1, 0) ||
      orc_coarse_ghost_level(h, level))
    return -1;
  for (int l = level; l >= level - 1; l--)
    for (int a = 0; a < nlev(h, l); a++) {
      orc_patch *p = lev_patch(h, l, a);
      for (long e = 0; e < p->nq; e++) p->qold[e] = p->q[e];
    }
  for (int a = 0; a < nlev(h, level); a++) truncation_error_patch(h, lev_patch(h, level, a));
  return 0;
}

int orc_error_flags(orc_hier *h, int level, double qtolerance) {
  if (!level_ok(h, level)) return fail(h, "error_flags: bad level %d", level);
  for (int a = 0; a < nlev(h, level); a++)
    error_flags_patch(h, lev_patch(h, level, a), qtolerance);
  return 0;
}

int orc_error_flags_user(orc_hier *h, int level, double qtolerance) {
  if (!level_ok(h, level)) return fail(h, "bad level %d", level);
  for (int a = 0; a < nlev(h, level); a++)
    error_flags_user_patch(h, lev_patch(h, level, a), qtolerance);
  return 0;
}

/* TakeTimeSteps' statistics loop (lib/bsamroutines.f90:690-697): levels
 * finest down to the root, patches in level-list order. */
int orc_after_step_stats(orc_hier *h, double *out) {
  out[0] = out[1] = 0.0;
  for (int level = h->finestlevel; level >= 0; level--)
    for (int a = 0; a < nlev(h, level); a++) after_step_patch(h, lev_patch(h, level, a), out);
  return 0;
}

/* FillDown, lib/afasroutines.f90:223-239. */
int orc_fill_down(orc_hier *h, int field) {
  for (int level = h->finestlevel; level >= h->minlevel + 1; level--) {
    for (int a = 0; a < nlev(h, level); a++)
      fill_down_patch(h, lev_patch(h, level, a), field);
    if (field == 2 && orc_set_ghost(h, level - 1, 0)) return -1;
  }
  return 0;
}

/* GetSourceFunction -> Source2D/3D (lib/afasroutines.f90:312-367,
 * problem.f90:258-280): ftmp = (qold_1, -Dfe(qold_1)) on levels 0..finest. */
int orc_source(orc_hier *h) {
  for (int level = h->finestlevel; level >= 0; level--)
    for (int a = 0; a < nlev(h, level); a++) {
      orc_patch *p = lev_patch(h, level, a);
      int kmax = h->ndims == 3 ? p->mx[2] : 1;
      for (int k = 1; k <= kmax; k++)
        for (int j = 1; j <= p->mx[1]; j++)
          for (int i = 1; i <= p->mx[0]; i++) {
            double c = p->qold[QX(p, i, j, k, 1)];
            FT(p, i, j, k, 1) = c;
            FT(p, i, j, k, 2) = -dfe(h, c);
          }
    }
  return 0;
}

/* UpdateSourceFunction (lib/afasroutines.f90:427-486) with SourceUpdate = 0
 * (problem.f90:281-303): f = ftmp + 0. */
int orc_update_source(orc_hier *h) {
  for (int level = h->finestlevel; level >= 0; level--)
    for (int a = 0; a < nlev(h, level); a++) {
      orc_patch *p = lev_patch(h, level, a);
      for (long e = 0; e < p->nf; e++) p->f[e] = p->ftmp[e] + 0.0;
    }
  return 0;
}

/* LevelRelax, lib/afasroutines.f90:523-547. */
static int level_relax(orc_hier *h, int level) {
  for (int pass = 1; pass <= h->nsmoothingpasses; pass++)
    for (int c = 1; c <= 2; c++) {
      orc_relax_level(h, level, c);
      if (orc_set_ghost(h, level, c)) return -1;
    }
  return 0;
}

/* AFASVcycle, lib/afasroutines.f90:488-521. */
int orc_vcycle(orc_hier *h, int level) {
  if (level_relax(h, level)) return -1;
  if (level > h->minlevel) {
    orc_restrict_level(h, level);
    if (orc_set_ghost(h, level - 1, 0)) return -1;
    orc_coarse_ghost_level(h, level);
    orc_coarse_load_level(h, level);
    if (orc_vcycle(h, level - 1)) return -1;
    orc_correct_level(h, level);
    if (orc_set_ghost(h, level, 0)) return -1;
    if (level_relax(h, level)) return -1;
  }
  return 0;
}

/* ErrorAFAS with errortype 1 (L2Error, SquareSum), lib/afasroutines.f90:
 * 108-221: sqrt(sum rf^2 / (nrvars * cells)) over the given level. */
double orc_error_afas(orc_hier *h, int level) {
  double s = 0.0, cnt = 0.0;
  for (int a = 0; a < nlev(h, level); a++) {
    orc_patch *p = lev_patch(h, level, a);
    residual_patch(h, p);
    double ps = 0.0;
    for (long e = 0; e < p->nf; e++) ps = ps + p->rf[e] * p->rf[e];
    s = ps + s;
    long cells = (long)p->mx[0] * p->mx[1] * (h->ndims == 3 ? p->mx[2] : 1);
    cnt = (double)(h->nrvars * cells) + cnt;
  }
  return sqrt(s / cnt);
}

int orc_copy_qold(orc_hier *h) {
  for (int n = 0; n < h->npatches; n++)
    memcpy(h->p[n].qold, h->p[n].q, sizeof(double) * (size_t)h->p[n].nq);
  return 0;
}

/* MultigridIterations, lib/afasroutines.f90:38-106 (maux = 0: the
 * UpdateAuxInVcycle calls are no-ops, :406). */
int orc_multigrid_iterations(orc_hier *h, double *hist, int maxhist) {
  if (orc_fill_down(h, 2)) return -1;
  orc_source(h);
  orc_update_source(h);
  orc_fill_down(h, 1);
  int it;
  for (it = 1; it <= h->maxvcycles; it++) {
    if (orc_vcycle(h, h->finestlevel)) return -1;
    if (orc_fill_down(h, 2)) return -1;
    orc_update_source(h);
    orc_fill_down(h, 1);
    double r = orc_error_afas(h, h->finestlevel);
    if (hist && it <= maxhist) hist[it - 1] = r;
    if (r < h->qerrortol) return it;
  }
  return h->maxvcycles;
}

/* One time step on a fixed hierarchy, following TakeTimeSteps
 * (lib/bsamroutines.f90:648-723) without the regrid: ghost refresh per level
 * (as AMR's SetGhost calls do), FillDown, qold = q, then the solve. */
int orc_time_step(orc_hier *h, double *hist, int maxhist) {
  for (int level = 0; level <= h->finestlevel; level++)
    if (orc_set_ghost(h, level, 0)) return -1;
  if (orc_fill_down(h, 2)) return -1;
  orc_copy_qold(h);
  return orc_multigrid_iterations(h, hist, maxhist);
}

/* Interpolate3D is restated in orc3d.c. */
#include "orc3d.inc"
