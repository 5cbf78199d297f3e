// kernels.hip -- gfx950 kernels of the FAS multigrid engine (FP64 path).
//
// Every kernel works on one level of the hierarchy at a time: a launch covers
// all patches of the level through a tile list (one 256-thread workgroup per
// tile, tiles of uneven patches mixed in one launch). Field storage is the
// reference layout, so a wavefront reads whole rows of q, f and qc.
//
// Floating-point expressions keep the reference's evaluation order and the
// library is built without multiply-add contraction, so field results equal
// the FP64 CPU restatement bit for bit.
#include "kernels.hpp"

namespace bsam {

// Tile extents of the level kernels (256 threads).
template <int ND> struct TileShape;
template <> struct TileShape<2> { static constexpr int X = 32, Y = 8, Z = 1; };
template <> struct TileShape<3> { static constexpr int X = 8, Y = 8, Z = 4; };

// ---- CH elemental terms (problem.f90:760-813) ----------------------------
__device__ __forceinline__ double dfc(const Problem &p, double c) {
  double cmh = c - p.fc_shift;
  return cmh * cmh * cmh * p.fc_alpha;
}
__device__ __forceinline__ double d2fc(const Problem &p, double c) {
  double cmh = c - p.fc_shift;
  return 3.0 * cmh * cmh * p.fc_alpha;
}
__device__ __forceinline__ double dfe(const Problem &p, double c) {
  return (c - p.fc_shift) * p.fe_num / p.fe_den;
}

// ---- array views -----------------------------------------------------------
// A ghost-extended array with lower bound 0: element (i,j,k,v) at
// base + i + px*(j + py*k) + v*vs (k = 0 in 2D).
struct View {
  double *p;
  long long px, pl, vs;
  __device__ __forceinline__ long long at(int i, int j, int k) const {
    return i + px * j + pl * k;
  }
};

__device__ __forceinline__ View qview(const Dev &d, const DPatch &P) {
  return View{d.q + P.q, P.px, (long long)P.px * P.py, P.qvs};
}
__device__ __forceinline__ View qcview(const Dev &d, const DPatch &P) {
  return View{d.qc + P.qc, P.cmx + 2, (long long)(P.cmx + 2) * (P.cmy + 2), P.cvs};
}
// Interior f-type arrays (lower bound 1).
__device__ __forceinline__ long long fat(const DPatch &P, int i, int j, int k) {
  return (i - 1) + (long long)P.mx * ((j - 1) + (long long)P.my * (k - 1));
}

// Operator2D/3D (problem.f90:231-257, :534-561) at cell c of view a with
// spacing hh: r1 = phi - dt/h^2 ULapMob(phi,mu), r2 = mu - Dfc(phi) +
// eps^2/h^2 ULap(phi), with the undivided stencils of gridutilities.f90.
template <int ND>
__device__ __forceinline__ void op_at(const Problem &prm, const View &a, double hh,
                                      long long c, double &r1, double &r2) {
  double h2 = hh * hh, tmp1 = prm.dt / h2, tmp2 = prm.eps2 / h2;
  const double *q1 = a.p, *q2 = a.p + a.vs;
  double phi = q1[c], mu = q2[c];
  long long sx = 1, sy = a.px, sz = a.pl;
  double lapmob, lap;
  if (ND == 2) {
    lapmob = (((q2[c + sx] - mu) - (mu - q2[c - sx])) + (q2[c + sy] - mu)) -
             (mu - q2[c - sy]);
    lap = (((q1[c + sx] + q1[c - sx]) + q1[c + sy]) + q1[c - sy]) - 4.0 * phi;
  } else {
    lapmob = (((((q2[c + sx] - mu) - (mu - q2[c - sx])) + (q2[c + sy] - mu)) -
               (mu - q2[c - sy])) + (q2[c + sz] - mu)) - (mu - q2[c - sz]);
    lap = (((((q1[c + sx] + q1[c - sx]) + q1[c + sy]) + q1[c - sy]) + q1[c + sz]) +
           q1[c - sz]) - 6.0 * phi;
  }
  r1 = phi - tmp1 * lapmob;
  r2 = (mu - dfc(prm, phi)) + tmp2 * lap;
}

// Pointwise nonlinear Gauss-Seidel update of one cell: one Newton step on the
// 2x2 (phi, mu) system, Cramer's rule (RelaxGrid2D/3D, problem.f90:164-215,
// :453-518; Mob = 1).
template <int ND>
__device__ __forceinline__ void relax_cell(const Dev &d, const DPatch &P, int i, int j,
                                           int k) {
  const Problem &prm = d.prm;
  View a = qview(d, P);
  long long c = a.at(i, j, k);
  double *q1 = a.p, *q2 = a.p + a.vs;
  long long fo = P.f + fat(P, i, j, ND == 3 ? k : 1);
  double h2 = P.dx * P.dx;
  double tmp1 = prm.dt / h2, tmp2 = prm.eps2 / h2;
  double phi = q1[c];
  double a12, a21, b1, b2;
  long long sx = 1, sy = a.px, sz = a.pl;
  if (ND == 2) {
    double tmp3 = 4.0 * prm.eps2 / h2;
    a12 = tmp1 * 4.0;
    a21 = -d2fc(prm, phi) - tmp3;
    b1 = d.f[fo] + tmp1 * (((q2[c - sx] + q2[c + sx]) + q2[c - sy]) + q2[c + sy]);
    b2 = ((d.f[fo + P.fvs] + dfc(prm, phi)) - d2fc(prm, phi) * phi) -
         tmp2 * (((q1[c + sx] + q1[c - sx]) + q1[c + sy]) + q1[c - sy]);
  } else {
    double tmp3 = 6.0 * prm.eps2 / h2;
    a12 = tmp1 * 6.0;
    a21 = -d2fc(prm, phi) - tmp3;
    b1 = d.f[fo] + tmp1 * (((((q2[c - sx] + q2[c + sx]) + q2[c - sy]) + q2[c + sy]) +
                            q2[c - sz]) + q2[c + sz]);
    b2 = ((d.f[fo + P.fvs] + dfc(prm, phi)) - d2fc(prm, phi) * phi) -
         tmp2 * (((((q1[c + sx] + q1[c - sx]) + q1[c + sy]) + q1[c - sy]) + q1[c + sz]) +
                 q1[c - sz]);
  }
  double det = 1.0 - a12 * a21;
  q1[c] = (b1 - a12 * b2) / det;
  q2[c] = (b2 - a21 * b1) / det;
}

// Restriction2D/3D (gridutilities.f90:89-110, :406-433) of the child cells
// under coarse cell (ci,cj,ck) of a ghost-extended (lower bound 0) view.
template <int ND>
__device__ __forceinline__ double restrict_view(const View &a, int v, int ci, int cj,
                                                int ck) {
  const double *x = a.p + v * a.vs;
  int i2 = 2 * ci, j2 = 2 * cj;
  if (ND == 2)
    return 0.25 * (((x[a.at(i2, j2, 0)] + x[a.at(i2 - 1, j2, 0)]) +
                    x[a.at(i2, j2 - 1, 0)]) + x[a.at(i2 - 1, j2 - 1, 0)]);
  int k2 = 2 * ck;
  return 0.125 *
         (((((((x[a.at(i2, j2, k2)] + x[a.at(i2 - 1, j2, k2)]) + x[a.at(i2, j2 - 1, k2)]) +
              x[a.at(i2 - 1, j2 - 1, k2)]) + x[a.at(i2, j2, k2 - 1)]) +
            x[a.at(i2 - 1, j2, k2 - 1)]) + x[a.at(i2, j2 - 1, k2 - 1)]) +
          x[a.at(i2 - 1, j2 - 1, k2 - 1)]);
}
// Same for an interior (lower bound 1) f-type array.
template <int ND>
__device__ __forceinline__ double restrict_f(const double *x, const DPatch &P, int ci,
                                             int cj, int ck) {
  int i2 = 2 * ci, j2 = 2 * cj;
  if (ND == 2)
    return 0.25 * (((x[fat(P, i2, j2, 1)] + x[fat(P, i2 - 1, j2, 1)]) +
                    x[fat(P, i2, j2 - 1, 1)]) + x[fat(P, i2 - 1, j2 - 1, 1)]);
  int k2 = 2 * ck;
  return 0.125 *
         (((((((x[fat(P, i2, j2, k2)] + x[fat(P, i2 - 1, j2, k2)]) +
               x[fat(P, i2, j2 - 1, k2)]) + x[fat(P, i2 - 1, j2 - 1, k2)]) +
             x[fat(P, i2, j2, k2 - 1)]) + x[fat(P, i2 - 1, j2, k2 - 1)]) +
           x[fat(P, i2, j2 - 1, k2 - 1)]) + x[fat(P, i2 - 1, j2 - 1, k2 - 1)]);
}

// ---- per-cell operations ----------------------------------------------------
// Each function performs the reference operation for one cell; the level
// kernels below map one tile of cells to a workgroup, and the tail kernels
// loop the same functions over whole single-patch levels in one workgroup.

// One ghost-cell update of a SetGhost program.
__device__ __forceinline__ void ghost_op(const Dev &d, const GhostOp &g) {
  double *q = d.q;
  for (int v = 0; v < d.nrvars; v++) {
    const double *f = q + (long long)v * g.fvs;
    const double *c = q + (long long)v * g.cvs;
    double val;
    switch (g.kind) {
      case 0:
        val = f[g.f0];
        break;
      case 1:
        val = BSAM_A1 * f[g.f0] + BSAM_A2 * f[g.f1] + BSAM_A3 * c[g.c[0]];
        break;
      case 2:
        val = BSAM_A1 * f[g.f0] + BSAM_A2 * f[g.f1] +
              BSAM_A3 * (BSAM_B1 * c[g.c[0]] + BSAM_B2 * c[g.c[1]] + BSAM_B3 * c[g.c[2]]);
        break;
      default:
        val = BSAM_A1 * f[g.f0] + BSAM_A2 * f[g.f1] +
              BSAM_A3 * (BSAM_C1 * c[g.c[0]] + BSAM_C2 * c[g.c[1]] + BSAM_C3 * c[g.c[2]] +
                         BSAM_C4 * c[g.c[3]]);
        break;
    }
    for (int r = 0; r < g.nflip; r++) val = 2.0 - val;
    q[g.dst + (long long)v * g.dvs] = val;
  }
}

// RestrictSolution (afasroutines.f90:600-644) for coarse cell (i,j,k): qc =
// R(q) and the parent's q under the child takes the same value.
template <int ND>
__device__ __forceinline__ void restrict_cell(const Dev &d, const DPatch &P, int i, int j,
                                              int k) {
  const DPatch &PA = d.P[P.parent];
  View a = qview(d, P), cq = qcview(d, P), pq = qview(d, PA);
  long long pc = pq.at(P.mb[0][0] + i - 1, P.mb[1][0] + j - 1,
                       ND == 3 ? P.mb[2][0] + k - 1 : 0);
  long long cc = cq.at(i, j, k);
  for (int v = 0; v < d.nrvars; v++) {
    double r = restrict_view<ND>(a, v, i, j, k);
    cq.p[cc + v * cq.vs] = r;
    pq.p[pc + v * pq.vs] = r;
  }
}

// FillDownLevel (afasroutines.f90:241-310): parent X(mbounds) = R(child X)
// for X = f (field 1) or q (field 2).
template <int ND>
__device__ __forceinline__ void filldown_cell(const Dev &d, const DPatch &P, int i, int j,
                                              int k, int field) {
  const DPatch &PA = d.P[P.parent];
  int pi = P.mb[0][0] + i - 1, pj = P.mb[1][0] + j - 1,
      pk = ND == 3 ? P.mb[2][0] + k - 1 : 1;
  if (field == 1) {
    for (int v = 0; v < d.nrvars; v++)
      d.f[PA.f + fat(PA, pi, pj, pk) + v * PA.fvs] =
          restrict_f<ND>(d.f + P.f + v * P.fvs, P, i, j, k);
  } else {
    View a = qview(d, P), pq = qview(d, PA);
    long long pc = pq.at(pi, pj, ND == 3 ? pk : 0);
    for (int v = 0; v < d.nrvars; v++) pq.p[pc + v * pq.vs] = restrict_view<ND>(a, v, i, j, k);
  }
}

// GetCoarseGhostPoints (boundary.f90:1203-1269): ring cell (ci,cj,ck) of qc
// <- parent (mb1-1+ci, mb2-1+cj, mb3-1+ck); (i,j,k) ranges over 0..cmx+1.
template <int ND>
__device__ __forceinline__ void coarse_ghost_cell(const Dev &d, const DPatch &P, int i, int j,
                                                  int k) {
  bool ring = i == 0 || i == P.cmx + 1 || j == 0 || j == P.cmy + 1 ||
              (ND == 3 && (k == 0 || k == P.cmz + 1));
  if (!ring) return;
  const DPatch &PA = d.P[P.parent];
  View cq = qcview(d, P), pq = qview(d, PA);
  long long pc = pq.at(P.mb[0][0] - 1 + i, P.mb[1][0] - 1 + j,
                       ND == 3 ? P.mb[2][0] - 1 + k : 0);
  long long cc = cq.at(i, j, k);
  for (int v = 0; v < d.nrvars; v++) cq.p[cc + v * cq.vs] = pq.p[pc + v * pq.vs];
}

// CoarseLoadingFunction (afasroutines.f90:646-714): parent f(mbounds) =
// R(f - A_h q) + A_2h(qc), the FAS right-hand side, with the fine residual
// recomputed in registers for the 2^d children of coarse cell (i,j,k).
template <int ND>
__device__ __forceinline__ void coarse_load_cell(const Dev &d, const DPatch &P, int i, int j,
                                                 int k) {
  const DPatch &PA = d.P[P.parent];
  View a = qview(d, P), cq = qcview(d, P);
  const double *f1 = d.f + P.f, *f2 = d.f + P.f + P.fvs;
  constexpr int NC = ND == 2 ? 4 : 8;
  double r1[NC], r2[NC];
#pragma unroll
  for (int s = 0; s < NC; s++) {
    int fi = 2 * i - (s & 1), fj = 2 * j - ((s >> 1) & 1);
    int fk = ND == 3 ? 2 * k - ((s >> 2) & 1) : 0;
    double o1, o2;
    op_at<ND>(d.prm, a, P.dx, a.at(fi, fj, fk), o1, o2);
    long long fo = fat(P, fi, fj, ND == 3 ? fk : 1);
    r1[s] = f1[fo] - o1;
    r2[s] = f2[fo] - o2;
  }
  double R1, R2;
  if (ND == 2) {
    R1 = 0.25 * (((r1[0] + r1[1]) + r1[2]) + r1[3]);
    R2 = 0.25 * (((r2[0] + r2[1]) + r2[2]) + r2[3]);
  } else {
    R1 = 0.125 * (((((((r1[0] + r1[1]) + r1[2]) + r1[3]) + r1[4]) + r1[5]) + r1[6]) + r1[7]);
    R2 = 0.125 * (((((((r2[0] + r2[1]) + r2[2]) + r2[3]) + r2[4]) + r2[5]) + r2[6]) + r2[7]);
  }
  double c1, c2;
  op_at<ND>(d.prm, cq, PA.dx, cq.at(i, j, k), c1, c2);
  int pi = P.mb[0][0] + i - 1, pj = P.mb[1][0] + j - 1,
      pk = ND == 3 ? P.mb[2][0] + k - 1 : 1;
  long long po = PA.f + fat(PA, pi, pj, pk);
  d.f[po] = R1 + c1;
  d.f[po + PA.fvs] = R2 + c2;
}

// CorrectFine (afasroutines.f90:842-963, mbc = 1): q += P1(parent q - qc) at
// fine cell (i,j,k) of 0..mx+1, the separable (3,1)/4 prolongation of
// BiLinProlongationP1 / TriLinProlongationP1 (gridutilities.f90:137-175,
// :474-529) evaluated in the reference's x, y(, z) pass order.
template <int ND>
__device__ __forceinline__ void correct_cell(const Dev &d, const DPatch &P, int i, int j,
                                             int k) {
  const DPatch &PA = d.P[P.parent];
  View a = qview(d, P), cq = qcview(d, P), pq = qview(d, PA);
  int kx = i >> 1, ky = j >> 1, kz = ND == 3 ? k >> 1 : 0;
  bool ex = (i & 1) == 0, ey = (j & 1) == 0, ez = (k & 1) == 0;
  int o0 = P.mb[0][0] - 1, o1 = P.mb[1][0] - 1, o2 = ND == 3 ? P.mb[2][0] - 1 : 0;
  long long qc0 = a.at(i, j, k);
  for (int v = 0; v < d.nrvars; v++) {
    const double *cp = cq.p + v * cq.vs, *pp = pq.p + v * pq.vs;
    auto corr = [&](int ci, int cj, int ck) {
      return pp[pq.at(o0 + ci, o1 + cj, ND == 3 ? o2 + ck : 0)] - cp[cq.at(ci, cj, ck)];
    };
    auto xpass = [&](int cj, int ck) {
      double a0 = corr(kx, cj, ck), a1 = corr(kx + 1, cj, ck);
      return ex ? 3.0 * a0 + a1 : a0 + 3.0 * a1;
    };
    double p;
    if (ND == 2) {
      double b0 = xpass(ky, 0), b1 = xpass(ky + 1, 0);
      p = ey ? (3.0 * b0 + b1) / 16.0 : (b0 + 3.0 * b1) / 16.0;
    } else {
      auto ypass = [&](int ck) {
        double b0 = xpass(ky, ck), b1 = xpass(ky + 1, ck);
        return ey ? 3.0 * b0 + b1 : b0 + 3.0 * b1;
      };
      double c0 = ypass(kz), c1 = ypass(kz + 1);
      p = ez ? (3.0 * c0 + c1) / 64.0 : (c0 + 3.0 * c1) / 64.0;
    }
    double *qv = a.p + v * a.vs;
    qv[qc0] = qv[qc0] + p;
  }
}

// GetSourceFunction + UpdateSourceFunction for the CH module (Source2D/3D,
// SourceUpdate = 0; problem.f90:258-303): ftmp = (qold_1, -Dfe(qold_1)) and
// f = ftmp + 0 when full != 0; only f = ftmp + 0 otherwise.
template <int ND>
__device__ __forceinline__ void source_cell(const Dev &d, const DPatch &P, int i, int j,
                                            int k, int full) {
  long long fo = P.f + fat(P, i, j, ND == 3 ? k : 1);
  if (full) {
    double c = d.qold[P.q + i + (long long)P.px * (j + (long long)P.py * k)];
    d.ftmp[fo] = c;
    d.ftmp[fo + P.fvs] = -dfe(d.prm, c);
  }
  for (int v = 0; v < d.nrvars; v++) d.f[fo + v * P.fvs] = d.ftmp[fo + v * P.fvs] + 0.0;
}

// ---- tile decode -----------------------------------------------------------
template <int ND>
__device__ __forceinline__ void tile_cell(const Tile &t, int &i, int &j, int &k) {
  using S = TileShape<ND>;
  int tid = threadIdx.x;
  i = t.i0 + tid % S::X;
  j = t.j0 + (tid / S::X) % S::Y;
  k = ND == 3 ? t.k0 + tid / (S::X * S::Y) : 0;
}

// ---- level kernels ---------------------------------------------------------
template <int ND>
__global__ __launch_bounds__(256) void k_relax(Dev d, const Tile *tiles, int ipass) {
  Tile t = tiles[blockIdx.x];
  const DPatch &P = d.P[t.patch];
  int i, j, k;
  tile_cell<ND>(t, i, j, k);
  if (i > P.mx || j > P.my || (ND == 3 && k > P.mz)) return;
  // RelaxGrid's loop start 1+mod(j+ipass,2) (and +k in 3D) selects the
  // cells with i+j(+k)+ipass odd.
  if (((i + j + (ND == 3 ? k : 0) + ipass) & 1) == 0) return;
  relax_cell<ND>(d, P, i, j, k);
}

// One SetGhost(level, ipass) program: every ghost-cell update is independent.
__global__ __launch_bounds__(256) void k_ghost(Dev d, const GhostOp *ops, int n) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  ghost_op(d, ops[e]);
}

template <int ND>
__global__ __launch_bounds__(256) void k_restrict(Dev d, const Tile *tiles) {
  Tile t = tiles[blockIdx.x];
  const DPatch &P = d.P[t.patch];
  int i, j, k;
  tile_cell<ND>(t, i, j, k);
  if (i > P.cmx || j > P.cmy || (ND == 3 && k > P.cmz)) return;
  restrict_cell<ND>(d, P, i, j, k);
}

template <int ND>
__global__ __launch_bounds__(256) void k_filldown(Dev d, const Tile *tiles, int field) {
  Tile t = tiles[blockIdx.x];
  const DPatch &P = d.P[t.patch];
  int i, j, k;
  tile_cell<ND>(t, i, j, k);
  if (i > P.cmx || j > P.cmy || (ND == 3 && k > P.cmz)) return;
  filldown_cell<ND>(d, P, i, j, k, field);
}

// Tiles cover the ghost-extended coarse box 0..cmx+1.
template <int ND>
__global__ __launch_bounds__(256) void k_coarse_ghost(Dev d, const Tile *tiles) {
  Tile t = tiles[blockIdx.x];
  const DPatch &P = d.P[t.patch];
  int i, j, k;
  tile_cell<ND>(t, i, j, k);
  if (i > P.cmx + 1 || j > P.cmy + 1 || (ND == 3 && k > P.cmz + 1)) return;
  coarse_ghost_cell<ND>(d, P, i, j, k);
}

template <int ND>
__global__ __launch_bounds__(256) void k_coarse_load(Dev d, const Tile *tiles) {
  Tile t = tiles[blockIdx.x];
  const DPatch &P = d.P[t.patch];
  int i, j, k;
  tile_cell<ND>(t, i, j, k);
  if (i > P.cmx || j > P.cmy || (ND == 3 && k > P.cmz)) return;
  coarse_load_cell<ND>(d, P, i, j, k);
}

// Tiles cover the ghost-extended fine box 0..mx+1.
template <int ND>
__global__ __launch_bounds__(256) void k_correct(Dev d, const Tile *tiles) {
  Tile t = tiles[blockIdx.x];
  const DPatch &P = d.P[t.patch];
  int i, j, k;
  tile_cell<ND>(t, i, j, k);
  if (i > P.mx + 1 || j > P.my + 1 || (ND == 3 && k > P.mz + 1)) return;
  correct_cell<ND>(d, P, i, j, k);
}

template <int ND>
__global__ __launch_bounds__(256) void k_source(Dev d, const Tile *tiles, int full) {
  Tile t = tiles[blockIdx.x];
  const DPatch &P = d.P[t.patch];
  int i, j, k;
  tile_cell<ND>(t, i, j, k);
  if (i > P.mx || j > P.my || (ND == 3 && k > P.mz)) return;
  source_cell<ND>(d, P, i, j, k, full);
}

// ---- single-workgroup tail: the single-patch coarse levels ------------------
// Below the root every level is one patch with physical boundaries only. The
// tail kernels run the part of the V-cycle (or of FillDown) on levels
// lo..hi of that chain in one workgroup, synchronising with barriers instead
// of launches.
template <int ND, class F>
__device__ __forceinline__ void for_cells(int ex, int ey, int ez, int base, F f) {
  int n = ex * ey * ez;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    int i = e % ex, j = (e / ex) % ey, k = e / (ex * ey);
    f(i + base, j + base, ND == 3 ? k + base : 0);
  }
}

template <int ND>
__device__ void tail_ghost(const Dev &d, const TailLevel &L, int ipass) {
  const GhostOp *ops = L.gp[ipass];
  for (int e = threadIdx.x; e < L.ngp[ipass]; e += blockDim.x) ghost_op(d, ops[e]);
  __syncthreads();
}

template <int ND>
__device__ void tail_relax(const Dev &d, const TailLevel &L, int nsmooth) {
  const DPatch &P = d.P[L.patch];
  for (int pass = 0; pass < nsmooth; pass++)
    for (int color = 1; color <= 2; color++) {
      for_cells<ND>(P.mx, P.my, ND == 3 ? P.mz : 1, 1, [&](int i, int j, int k) {
        if (((i + j + (ND == 3 ? k : 0) + color) & 1) == 1) relax_cell<ND>(d, P, i, j, k);
      });
      __syncthreads();
      tail_ghost<ND>(d, L, color);
    }
}

// AFASVcycle(hi) (afasroutines.f90:488-521) restricted to levels lo..hi of
// the single-patch chain; levels are indexed from lo.
template <int ND>
__global__ __launch_bounds__(1024) void k_tail_vcycle(Dev d, const TailLevel *lv, int nlev,
                                                       int nsmooth) {
  for (int l = nlev - 1; l >= 1; l--) {
    const DPatch &P = d.P[lv[l].patch];
    int cz = ND == 3 ? P.cmz : 1;
    tail_relax<ND>(d, lv[l], nsmooth);
    for_cells<ND>(P.cmx, P.cmy, cz, 1,
                  [&](int i, int j, int k) { restrict_cell<ND>(d, P, i, j, k); });
    __syncthreads();
    tail_ghost<ND>(d, lv[l - 1], 0);
    for_cells<ND>(P.cmx + 2, P.cmy + 2, ND == 3 ? P.cmz + 2 : 1, 0,
                  [&](int i, int j, int k) { coarse_ghost_cell<ND>(d, P, i, j, k); });
    __syncthreads();
    for_cells<ND>(P.cmx, P.cmy, cz, 1,
                  [&](int i, int j, int k) { coarse_load_cell<ND>(d, P, i, j, k); });
    __syncthreads();
  }
  tail_relax<ND>(d, lv[0], nsmooth);
  for (int l = 1; l < nlev; l++) {
    const DPatch &P = d.P[lv[l].patch];
    for_cells<ND>(P.mx + 2, P.my + 2, ND == 3 ? P.mz + 2 : 1, 0,
                  [&](int i, int j, int k) { correct_cell<ND>(d, P, i, j, k); });
    __syncthreads();
    tail_ghost<ND>(d, lv[l], 0);
    tail_relax<ND>(d, lv[l], nsmooth);
  }
}

// FillDown (afasroutines.f90:223-239) from level hi down to lo of the chain:
// restriction into the parent, and for q a SetGhost(parent, 0).
template <int ND>
__global__ __launch_bounds__(1024) void k_tail_filldown(Dev d, const TailLevel *lv, int nlev,
                                                         int field) {
  for (int l = nlev - 1; l >= 1; l--) {
    const DPatch &P = d.P[lv[l].patch];
    for_cells<ND>(P.cmx, P.cmy, ND == 3 ? P.cmz : 1, 1,
                  [&](int i, int j, int k) { filldown_cell<ND>(d, P, i, j, k, field); });
    __syncthreads();
    if (field == 2) tail_ghost<ND>(d, lv[l - 1], 0);
  }
}

// Residual of the finest level (Residual, afasroutines.f90:799-840) reduced
// to one sum of squares per tile; deterministic (fixed order) reduction.
__device__ __forceinline__ double block_sum(double s) {
  __shared__ double red[16];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = s;
  __syncthreads();
  double tot = 0.0;
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    for (int q = 0; q < nw; q++) tot += red[q];
  }
  return tot;
}

template <int ND>
__global__ __launch_bounds__(256) void k_resnorm(Dev d, const Tile *tiles, double *partial,
                                                  int write_rf) {
  Tile t = tiles[blockIdx.x];
  const DPatch &P = d.P[t.patch];
  int i, j, k;
  tile_cell<ND>(t, i, j, k);
  double s = 0.0;
  if (i <= P.mx && j <= P.my && (ND == 2 || k <= P.mz)) {
    View a = qview(d, P);
    double o1, o2;
    op_at<ND>(d.prm, a, P.dx, a.at(i, j, k), o1, o2);
    long long fo = P.f + fat(P, i, j, ND == 3 ? k : 1);
    double r1 = d.f[fo] - o1, r2 = d.f[fo + P.fvs] - o2;
    if (write_rf) {
      d.rf[fo] = r1;
      d.rf[fo + P.fvs] = r2;
    }
    s = r1 * r1 + r2 * r2;
  }
  double tot = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Sum the per-tile partials in index order (strided, then a fixed tree) and
// finish ErrorAFAS: sqrt(sum / count) into hist[state[0]]; state[0] counts
// the calls and state[1] is set once a residual falls below tol.
__global__ __launch_bounds__(256) void k_finalize(const double *partial, int n, double count,
                                                   double tol, double *hist, int *state) {
  double s = 0.0;
  for (int e = threadIdx.x; e < n; e += blockDim.x) s += partial[e];
  double tot = block_sum(s);
  if (threadIdx.x == 0) {
    double r = sqrt(tot / count);
    int it = state[0];
    hist[it] = r;
    state[0] = it + 1;
    if (r < tol) state[1] = 1;
  }
}

__global__ __launch_bounds__(256) void k_copy(double *dst, const double *src, long long n) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; e < n; e += stride) dst[e] = src[e];
}

// ---- launchers -------------------------------------------------------------
static inline dim3 grid1(int n) { return dim3((unsigned)n); }

hipError_t launch_relax(int nd, const Dev &d, const Tile *tiles, int nt, int ipass,
                        hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (nd == 2) k_relax<2><<<grid1(nt), 256, 0, s>>>(d, tiles, ipass);
  else k_relax<3><<<grid1(nt), 256, 0, s>>>(d, tiles, ipass);
  return hipGetLastError();
}

hipError_t launch_ghost(const Dev &d, const GhostOp *ops, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_ghost<<<dim3((n + 255) / 256), 256, 0, s>>>(d, ops, n);
  return hipGetLastError();
}

hipError_t launch_restrict(int nd, const Dev &d, const Tile *tiles, int nt, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (nd == 2) k_restrict<2><<<grid1(nt), 256, 0, s>>>(d, tiles);
  else k_restrict<3><<<grid1(nt), 256, 0, s>>>(d, tiles);
  return hipGetLastError();
}

hipError_t launch_filldown(int nd, const Dev &d, const Tile *tiles, int nt, int field,
                           hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (nd == 2) k_filldown<2><<<grid1(nt), 256, 0, s>>>(d, tiles, field);
  else k_filldown<3><<<grid1(nt), 256, 0, s>>>(d, tiles, field);
  return hipGetLastError();
}

hipError_t launch_coarse_ghost(int nd, const Dev &d, const Tile *tiles, int nt,
                               hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (nd == 2) k_coarse_ghost<2><<<grid1(nt), 256, 0, s>>>(d, tiles);
  else k_coarse_ghost<3><<<grid1(nt), 256, 0, s>>>(d, tiles);
  return hipGetLastError();
}

hipError_t launch_coarse_load(int nd, const Dev &d, const Tile *tiles, int nt,
                              hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (nd == 2) k_coarse_load<2><<<grid1(nt), 256, 0, s>>>(d, tiles);
  else k_coarse_load<3><<<grid1(nt), 256, 0, s>>>(d, tiles);
  return hipGetLastError();
}

hipError_t launch_correct(int nd, const Dev &d, const Tile *tiles, int nt, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (nd == 2) k_correct<2><<<grid1(nt), 256, 0, s>>>(d, tiles);
  else k_correct<3><<<grid1(nt), 256, 0, s>>>(d, tiles);
  return hipGetLastError();
}

hipError_t launch_source(int nd, const Dev &d, const Tile *tiles, int nt, int full,
                         hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (nd == 2) k_source<2><<<grid1(nt), 256, 0, s>>>(d, tiles, full);
  else k_source<3><<<grid1(nt), 256, 0, s>>>(d, tiles, full);
  return hipGetLastError();
}

hipError_t launch_resnorm(int nd, const Dev &d, const Tile *tiles, int nt, double *partial,
                          int write_rf, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (nd == 2) k_resnorm<2><<<grid1(nt), 256, 0, s>>>(d, tiles, partial, write_rf);
  else k_resnorm<3><<<grid1(nt), 256, 0, s>>>(d, tiles, partial, write_rf);
  return hipGetLastError();
}

hipError_t launch_finalize(const double *partial, int n, double count, double tol,
                           double *hist, int *state, hipStream_t s) {
  k_finalize<<<1, 256, 0, s>>>(partial, n, count, tol, hist, state);
  return hipGetLastError();
}

hipError_t launch_copy(double *dst, const double *src, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_copy<<<dim3((unsigned)blocks), 256, 0, s>>>(dst, src, n);
  return hipGetLastError();
}

hipError_t launch_tail_vcycle(int nd, const Dev &d, const TailLevel *lv, int nlev, int nsmooth,
                              int threads, hipStream_t s) {
  if (nlev <= 0) return hipSuccess;
  if (nd == 2) k_tail_vcycle<2><<<1, threads, 0, s>>>(d, lv, nlev, nsmooth);
  else k_tail_vcycle<3><<<1, threads, 0, s>>>(d, lv, nlev, nsmooth);
  return hipGetLastError();
}

hipError_t launch_tail_filldown(int nd, const Dev &d, const TailLevel *lv, int nlev, int field,
                                int threads, hipStream_t s) {
  if (nlev <= 1) return hipSuccess;
  if (nd == 2) k_tail_filldown<2><<<1, threads, 0, s>>>(d, lv, nlev, field);
  else k_tail_filldown<3><<<1, threads, 0, s>>>(d, lv, nlev, field);
  return hipGetLastError();
}

int tile_x(int nd) { return nd == 2 ? TileShape<2>::X : TileShape<3>::X; }
int tile_y(int nd) { return nd == 2 ? TileShape<2>::Y : TileShape<3>::Y; }
int tile_z(int nd) { return nd == 2 ? TileShape<2>::Z : TileShape<3>::Z; }

}  // namespace bsam
