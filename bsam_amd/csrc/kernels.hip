// kernels.hip -- gfx950 kernels of the FAS multigrid engine.
//
// Every level kernel covers all patches of one level through a tile list (one
// 256-thread workgroup per tile; tiles of uneven patches mixed in one launch).
// Field storage keeps the reference layout, so a wavefront reads whole rows.
//
// Each operation is instantiated twice:
//  - FP64 on the solution (T = double, DEF = false): floating-point
//    expressions keep the reference's evaluation order and the library is
//    built without multiply-add contraction, so fields equal the FP64 CPU
//    restatement bit for bit;
//  - FP32 on the correction of a defect-correction step (T = float,
//    DEF = true): the operator becomes A(qb + e) - A(qb), whose cubic term is
//    evaluated as alpha*e*(3a^2 + 3ae + e^2) (a = qb - shift) so that FP32
//    keeps its relative accuracy in e however small e gets.
#include "kernels.hpp"

#include <type_traits>

namespace bsam {

// Level kernels run 256-thread workgroups over tiles of 2^tile_lx x 2^tile_ly
// (x 256 / 2^(tile_lx + tile_ly) in 3D) cells; the shape is chosen per
// context (Dev::tile_lx, tile_ly).

// ---- CH elemental terms (problem.f90:760-813; Mob = 1) -------------------
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
// Dfc(b + e) - Dfc(b) and D2fc(b + e) in the precision of e.
template <class T>
__device__ __forceinline__ T dfc_delta(const Problem &p, T b, T e) {
  T a = b - (T)p.fc_shift;
  return (T)p.fc_alpha * e * ((T)3 * a * a + (T)3 * a * e + e * e);
}
template <class T>
__device__ __forceinline__ T d2fc_at(const Problem &p, T b, T e) {
  T x = b - (T)p.fc_shift + e;
  return (T)3 * x * x * (T)p.fc_alpha;
}

// ---- field sets and level references ------------------------------------------
template <class T> struct Fld {
  T *q, *qc, *f;     // solution-layout, qc-layout and f-layout pools
  const T *qb;       // FP32 base point (defect form), q layout
};
template <class T> __device__ __forceinline__ Fld<T> fields(const Dev &d);
template <> __device__ __forceinline__ Fld<double> fields<double>(const Dev &d) {
  return Fld<double>{d.q, d.qc, d.f, nullptr};
}
template <> __device__ __forceinline__ Fld<float> fields<float>(const Dev &d) {
  return Fld<float>{d.e, d.qce, d.fe, d.qb};
}

// A group (independent sub-problem) that has stopped iterating: the kernels
// of a solve leave its patches and ghost cells untouched.
__device__ __forceinline__ bool stopped(const Dev &d, int group) {
  return d.gstate && d.gstate[4 * group + 3];
}
__device__ __forceinline__ bool stopped_patch(const Dev &d, int patch) {
  // with one group the flag's address needs no patch record
  return d.gstate && d.gstate[4 * (d.one_group ? 0 : d.P[patch].group) + 3];
}

// One patch's arrays, each addressed from its own origin: q(0,0,0,1),
// qc(0,0,0,1), f(1,1,1,1) and the defect base qb(0,0,0). The arrays may live in
// the global pools or, in the tail kernels, in LDS; the layout (pitches and
// variable strides of the patch record P) is the same in both. P stays a
// reference into the patch table (scalar loads) rather than a register copy.
template <class T> struct Lvl {
  T *q, *qc, *f;
  const T *qb;
  const DPatch &P;
};
template <class T> __device__ __forceinline__ Lvl<T> level_ref(const Dev &d, int patch) {
  Fld<T> F = fields<T>(d);
  const DPatch &P = d.P[patch];
  return Lvl<T>{F.q + P.q, F.qc + P.qc, F.f + P.f, F.qb ? F.qb + P.q : nullptr, P};
}

// A ghost-extended array with lower bound 0 in the device layout (rows
// split by the parity of i): element (i,j,k,v) at
// p + (i & 1)*xo + (i >> 1) + rs*j + pl*k + v*vs (k = 0 in 2D). The y
// neighbours of a cell are rs apart, the z neighbours pl; the x neighbours
// lie in the other half of the row.
template <class T> struct View {
  T *p;
  long long xo, rs, pl, vs;
  __device__ __forceinline__ long long at(int i, int j, int k) const {
    return (i & 1) * xo + (i >> 1) + rs * j + pl * k;
  }
};
template <class T> __device__ __forceinline__ View<T> view_of(T *p, long long n, long long ny,
                                                              long long vs) {
  const Split s = split_of(n, ny);
  return View<T>{p, s.xo, s.rs, s.pl, vs};
}
template <class T> __device__ __forceinline__ View<T> qview(const Lvl<T> &L) {
  return view_of<T>(L.q, L.P.px, L.P.py, L.P.qvs);
}
template <class T> __device__ __forceinline__ View<T> qcview(const Lvl<T> &L) {
  return view_of<T>(L.qc, L.P.cmx + 2, L.P.cmy + 2, L.P.cvs);
}
// Interior f-type arrays (lower bound 1; planes split by the parity of i - 1).
__device__ __forceinline__ long long fat(const DPatch &P, int i, int j, int k) {
  return scell(split_of(P.mx, P.my), i - 1, j - 1, k - 1);
}
// Ghost-extended cell (i, j, k) of a patch's q-layout arrays, relative to q(0,0,0,1).
__device__ __forceinline__ long long qat(const DPatch &P, int i, int j, int k) {
  return qcell(P.px, P.py, i, j, k);
}

// Operator2D/3D (problem.f90:231-257, :534-561) at cell c of view a on grid
// G: r1 = phi - dt/h^2 ULapMob(phi,mu), r2 = mu - Dfc(phi) +
// eps^2/h^2 ULap(phi), with the undivided stencils of gridutilities.f90.
// Defect form: Dfc(phi) becomes Dfc(base + phi) - Dfc(base).
// The same from gathered values: a1/a2 hold phi/mu at the cell, +x, -x, +y,
// -y (+z, -z).
template <int ND, class T, bool DEF>
__device__ __forceinline__ void op_vals(const Problem &prm, T base, const DPatch &G, const T *a1,
                                        const T *a2, T &r1, T &r2) {
  T tmp1 = (T)G.dt_h2, tmp2 = (T)G.eps2_h2;
  T phi = a1[0], mu = a2[0];
  T lapmob, lap;
  if (ND == 2) {
    lapmob = (((a2[1] - mu) - (mu - a2[2])) + (a2[3] - mu)) - (mu - a2[4]);
    lap = (((a1[1] + a1[2]) + a1[3]) + a1[4]) - (T)4 * phi;
  } else {
    lapmob = (((((a2[1] - mu) - (mu - a2[2])) + (a2[3] - mu)) - (mu - a2[4])) +
              (a2[5] - mu)) - (mu - a2[6]);
    lap = (((((a1[1] + a1[2]) + a1[3]) + a1[4]) + a1[5]) + a1[6]) - (T)6 * phi;
  }
  r1 = phi - tmp1 * lapmob;
  if (DEF) r2 = (mu - dfc_delta<T>(prm, base, phi)) + tmp2 * lap;
  else r2 = (mu - (T)dfc(prm, (double)phi)) + tmp2 * lap;
}
template <int ND, class T, bool DEF>
__device__ __forceinline__ void op_at(const Problem &prm, const View<T> &a, T base,
                                      const DPatch &G, int i, int j, int k, T &r1, T &r2) {
  const T *q1 = a.p, *q2 = a.p + a.vs;
  const long long c = a.at(i, j, k);
  long long at[7] = {c, a.at(i + 1, j, k), a.at(i - 1, j, k), c + a.rs, c - a.rs, c + a.pl,
                     c - a.pl};
  T a1[2 * ND + 1], a2[2 * ND + 1];
#pragma unroll
  for (int s = 0; s < 2 * ND + 1; s++) {
    a1[s] = q1[at[s]];
    a2[s] = q2[at[s]];
  }
  op_vals<ND, T, DEF>(prm, base, G, a1, a2, r1, r2);
}

// Pointwise nonlinear Gauss-Seidel update of one cell: one Newton step on the
// 2x2 (phi, mu) system, Cramer's rule (RelaxGrid2D/3D, problem.f90:164-215,
// :453-518; Mob = 1). Defect form: the Newton step linearises
// Dfc(base + e) about the current e.
// The update of cell (i,j,k) reading q through view `a` (the patch's arrays,
// or a staged copy of its neighbourhood) and writing the new values to *o1,
// *o2; `base` is the FP32 defect form's base point at the cell.
template <int ND, class T, bool DEF>
__device__ __forceinline__ void relax_core(const Problem &prm, const DPatch &P, const View<T> &a,
                                           int i, int j, int k, T *o1, T *o2, T base, T fv1,
                                           T fv2) {
  const long long c = a.at(i, j, k);
  const long long cxp = a.at(i + 1, j, k), cxm = a.at(i - 1, j, k);   // the other half-row
  const T *q1 = a.p, *q2 = a.p + a.vs;
  T tmp1 = (T)P.dt_h2, tmp2 = (T)P.eps2_h2, tmp3 = (T)P.lap_h2;
  T phi = q1[c];
  T dc, d2;
  if (DEF) {
    dc = dfc_delta<T>(prm, base, phi);
    d2 = d2fc_at<T>(prm, base, phi);
  } else {
    dc = (T)dfc(prm, (double)phi);
    d2 = (T)d2fc(prm, (double)phi);
  }
  T a12, a21, b1, b2;
  long long sy = a.rs, sz = a.pl;
  if (ND == 2) {
    a12 = tmp1 * (T)4;
    a21 = -d2 - tmp3;
    b1 = fv1 + tmp1 * (((q2[cxm] + q2[cxp]) + q2[c - sy]) + q2[c + sy]);
    b2 = ((fv2 + dc) - d2 * phi) -
         tmp2 * (((q1[cxp] + q1[cxm]) + q1[c + sy]) + q1[c - sy]);
  } else {
    a12 = tmp1 * (T)6;
    a21 = -d2 - tmp3;
    b1 = fv1 + tmp1 * (((((q2[cxm] + q2[cxp]) + q2[c - sy]) + q2[c + sy]) +
                         q2[c - sz]) + q2[c + sz]);
    b2 = ((fv2 + dc) - d2 * phi) -
         tmp2 * (((((q1[cxp] + q1[cxm]) + q1[c + sy]) + q1[c - sy]) + q1[c + sz]) +
                 q1[c - sz]);
  }
  T det = (T)1 - a12 * a21;
  *o1 = (b1 - a12 * b2) / det;
  *o2 = (b2 - a21 * b1) / det;
}
template <int ND, class T, bool DEF>
__device__ __forceinline__ void relax_cell_f(const Problem &prm, const Lvl<T> &L, int i, int j,
                                             int k, T fv1, T fv2) {
  View<T> a = qview(L);
  const long long c = a.at(i, j, k);
  relax_core<ND, T, DEF>(prm, L.P, a, i, j, k, a.p + c, a.p + a.vs + c, DEF ? L.qb[c] : (T)0,
                         fv1, fv2);
}
template <int ND, class T, bool DEF>
__device__ __forceinline__ void relax_cell(const Problem &prm, const Lvl<T> &L, int i, int j,
                                           int k) {
  long long fo = fat(L.P, i, j, ND == 3 ? k : 1);
  relax_cell_f<ND, T, DEF>(prm, L, i, j, k, L.f[fo], L.f[fo + L.P.fvs]);
}

// Restriction2D/3D (gridutilities.f90:89-110, :406-433) of the child cells
// under coarse cell (ci,cj,ck) of a ghost-extended (lower bound 0) view.
template <int ND, class T>
__device__ __forceinline__ T restrict_view(const View<T> &a, int v, int ci, int cj, int ck) {
  const T *x = a.p + v * a.vs;
  int i2 = 2 * ci, j2 = 2 * cj;
  if (ND == 2)
    return (T)0.25 * (((x[a.at(i2, j2, 0)] + x[a.at(i2 - 1, j2, 0)]) +
                       x[a.at(i2, j2 - 1, 0)]) + x[a.at(i2 - 1, j2 - 1, 0)]);
  int k2 = 2 * ck;
  return (T)0.125 *
         (((((((x[a.at(i2, j2, k2)] + x[a.at(i2 - 1, j2, k2)]) + x[a.at(i2, j2 - 1, k2)]) +
              x[a.at(i2 - 1, j2 - 1, k2)]) + x[a.at(i2, j2, k2 - 1)]) +
            x[a.at(i2 - 1, j2, k2 - 1)]) + x[a.at(i2, j2 - 1, k2 - 1)]) +
          x[a.at(i2 - 1, j2 - 1, k2 - 1)]);
}
// Same for an interior (lower bound 1) f-type array.
template <int ND, class T>
__device__ __forceinline__ T restrict_f(const T *x, const DPatch &P, int ci, int cj, int ck) {
  int i2 = 2 * ci, j2 = 2 * cj;
  if (ND == 2)
    return (T)0.25 * (((x[fat(P, i2, j2, 1)] + x[fat(P, i2 - 1, j2, 1)]) +
                       x[fat(P, i2, j2 - 1, 1)]) + x[fat(P, i2 - 1, j2 - 1, 1)]);
  int k2 = 2 * ck;
  return (T)0.125 *
         (((((((x[fat(P, i2, j2, k2)] + x[fat(P, i2 - 1, j2, k2)]) +
               x[fat(P, i2, j2 - 1, k2)]) + x[fat(P, i2 - 1, j2 - 1, k2)]) +
             x[fat(P, i2, j2, k2 - 1)]) + x[fat(P, i2 - 1, j2, k2 - 1)]) +
           x[fat(P, i2, j2 - 1, k2 - 1)]) + x[fat(P, i2 - 1, j2 - 1, k2 - 1)]);
}

// ---- per-cell operations ----------------------------------------------------
// Each function performs the reference operation for one cell of a child
// patch C (and its parent PA where the operation couples the two); the level
// kernels below map one tile of cells to a workgroup, and the tail kernels
// loop the same functions over whole single-patch levels in one workgroup.

__device__ __forceinline__ Qpo qpo_of(const Problem &prm) { return Qpo{prm.qpo[0], prm.qpo[1]}; }

// The value of one ghost-cell update of a SetGhost program: kind 0 copies
// source 0; kinds 1-3 are the quadratic coarse-fine interpolations of
// Interpolate2D/3D over two fine sources (0, 1) and one, three or four
// coarse sources (2..5), with the weights of lib/boundary.f90:45-54. A
// boundary "2 - x" map is affine; in the defect form only its linear part,
// -x, applies (flip_count(fb) times), after the periodic offset of code
// qpo_code(fb) for variable v (add_qpo). src(n) loads source n; every caller
// (global pools, short records, an isolated patch's LDS) evaluates the same
// expression tree.
template <class T, bool DEF, class Src, bool QPO = true>
__device__ __forceinline__ T ghost_value(int kind, int fb, Src src, Qpo qpo, int v) {
  T val;
  switch (kind) {
    case 0:
      val = src(0);
      break;
    case 1:
      val = (T)BSAM_A1 * src(0) + (T)BSAM_A2 * src(1) + (T)BSAM_A3 * src(2);
      break;
    case 2:
      val = (T)BSAM_A1 * src(0) + (T)BSAM_A2 * src(1) +
            (T)BSAM_A3 * ((T)BSAM_B1 * src(2) + (T)BSAM_B2 * src(3) + (T)BSAM_B3 * src(4));
      break;
    default:
      val = (T)BSAM_A1 * src(0) + (T)BSAM_A2 * src(1) +
            (T)BSAM_A3 * ((T)BSAM_C1 * src(2) + (T)BSAM_C2 * src(3) + (T)BSAM_C3 * src(4) +
                          (T)BSAM_C4 * src(5));
      break;
  }
  // the periodic offset, before the maps (a constant: none on a correction)
  if (!DEF && QPO) val = add_qpo<T>(val, qpo_code(fb), qpo, v);
  for (int r = 0; r < flip_count(fb); r++) val = DEF ? -val : (T)2 - val;
  return val;
}
// One ghost-cell update of a SetGhost program, offsets relative to q and
// variables vs apart.
template <class T, bool DEF, bool QPO = true>
__device__ __forceinline__ void ghost_op(int nrvars, T *q, long long vs, const GhostOp &g,
                                         Qpo qpo) {
  for (int v = 0; v < nrvars; v++) {
    const T *f = q + (long long)v * vs;
    auto src = [&](int n) { return f[n == 0 ? g.f0 : n == 1 ? g.f1 : g.c[n - 2]]; };
    const T val = ghost_value<T, DEF, decltype(src), QPO>(g.kind, g.nflip, src, qpo, v);
    q[g.dst + (long long)v * vs] = val;
  }
}
// The same update from its short record (kinds 0 and 1).
template <class T, bool DEF>
__device__ __forceinline__ void ghost_op_s(int nrvars, T *q, long long vs, const GhostS &g,
                                           Qpo qpo) {
  for (int v = 0; v < nrvars; v++) {
    const T *f = q + (long long)v * vs;
    const T val = ghost_value<T, DEF>(g.kind, g.nflip, [&](int n) {
      return f[g.kind == 0 || n == 2 ? g.src : g.dst + (n == 0 ? g.d0 : g.d1)];
    }, qpo, v);
    q[g.dst + (long long)v * vs] = val;
  }
}

// RestrictSolution (afasroutines.f90:600-644) for coarse cell (i,j,k): qc =
// R(q) and the parent's q under the child takes the same value.
template <int ND, class T>
__device__ __forceinline__ void restrict_cell(int nrvars, const Lvl<T> &C, const Lvl<T> &PA,
                                              int i, int j, int k, bool both_first = false) {
  const DPatch &P = C.P;
  View<T> a = qview(C), cq = qcview(C), pq = qview(PA);
  long long pc = pq.at(P.mb[0][0] + i - 1, P.mb[1][0] + j - 1,
                       ND == 3 ? P.mb[2][0] + k - 1 : 0);
  long long cc = cq.at(i, j, k);
  if (both_first && nrvars == 2) {
    // both variables' children loaded before either store
    const T r0 = restrict_view<ND, T>(a, 0, i, j, k), r1 = restrict_view<ND, T>(a, 1, i, j, k);
    cq.p[cc] = r0;
    pq.p[pc] = r0;
    cq.p[cc + cq.vs] = r1;
    pq.p[pc + pq.vs] = r1;
    return;
  }
  for (int v = 0; v < nrvars; v++) {
    T r = restrict_view<ND, T>(a, v, i, j, k);
    cq.p[cc + v * cq.vs] = r;
    pq.p[pc + v * pq.vs] = r;
  }
}

// FillDownLevel (afasroutines.f90:241-310): parent X(mbounds) = R(child X)
// for X = f (field 1) or q (field 2).
template <int ND, class T>
__device__ __forceinline__ void filldown_cell(int nrvars, const Lvl<T> &C, const Lvl<T> &PA,
                                              int i, int j, int k, int field,
                                              bool both_first = false) {
  const DPatch &P = C.P;
  int pi = P.mb[0][0] + i - 1, pj = P.mb[1][0] + j - 1,
      pk = ND == 3 ? P.mb[2][0] + k - 1 : 1;
  if (both_first && nrvars == 2) {
    // both variables' children loaded before either store
    if (field == 1) {
      const long long po = fat(PA.P, pi, pj, pk);
      const T r0 = restrict_f<ND, T>(C.f, P, i, j, k), r1 = restrict_f<ND, T>(C.f + P.fvs, P, i, j, k);
      PA.f[po] = r0;
      PA.f[po + PA.P.fvs] = r1;
    } else {
      View<T> a = qview(C), pq = qview(PA);
      const long long pc = pq.at(pi, pj, ND == 3 ? pk : 0);
      const T r0 = restrict_view<ND, T>(a, 0, i, j, k), r1 = restrict_view<ND, T>(a, 1, i, j, k);
      pq.p[pc] = r0;
      pq.p[pc + pq.vs] = r1;
    }
    return;
  }
  if (field == 1) {
    for (int v = 0; v < nrvars; v++)
      PA.f[fat(PA.P, pi, pj, pk) + v * PA.P.fvs] =
          restrict_f<ND, T>(C.f + v * P.fvs, P, i, j, k);
  } else {
    View<T> a = qview(C), pq = qview(PA);
    long long pc = pq.at(pi, pj, ND == 3 ? pk : 0);
    for (int v = 0; v < nrvars; v++) pq.p[pc + v * pq.vs] = restrict_view<ND, T>(a, v, i, j, k);
  }
}

// GetCoarseGhostPoints (boundary.f90:1203-1269): ring cell (ci,cj,ck) of qc
// <- parent (mb1-1+ci, mb2-1+cj, mb3-1+ck); (i,j,k) ranges over 0..cmx+1.
template <int ND, class T>
__device__ __forceinline__ void coarse_ghost_cell(int nrvars, const Lvl<T> &C,
                                                  const Lvl<T> &PA, int i, int j, int k) {
  const DPatch &P = C.P;
  bool ring = i == 0 || i == P.cmx + 1 || j == 0 || j == P.cmy + 1 ||
              (ND == 3 && (k == 0 || k == P.cmz + 1));
  if (!ring) return;
  View<T> cq = qcview(C), pq = qview(PA);
  long long pc = pq.at(P.mb[0][0] - 1 + i, P.mb[1][0] - 1 + j,
                       ND == 3 ? P.mb[2][0] - 1 + k : 0);
  long long cc = cq.at(i, j, k);
  for (int v = 0; v < nrvars; v++) cq.p[cc + v * cq.vs] = pq.p[pc + v * pq.vs];
}

// R(f - A_h q) at coarse cell (i,j,k): the fine residual of its 2^d children,
// child s at (2i - (s&1), 2j - (s>>1 & 1)(, 2k - (s>>2 & 1))), restricted as
// Restriction2D/3D does (gridutilities.f90:89-110, :406-433), the sum taken in
// child order and scaled by 1/2^d. The children are visited one at a time and
// summed as they come, which is the same sum in the same order, so only one
// fine stencil is live in registers.
template <int ND, class T, bool DEF>
__device__ __forceinline__ void restricted_residual(const Problem &prm, const Lvl<T> &C,
                                                    const View<T> &a, int i, int j, int k,
                                                    T &R1, T &R2) {
  const DPatch &P = C.P;
  const T *f1 = C.f, *f2 = C.f + P.fvs;
  constexpr int NC = ND == 2 ? 4 : 8;
  T s1 = (T)0, s2 = (T)0;
#pragma unroll 1
  for (int s = 0; s < NC; s++) {
    int fi = 2 * i - (s & 1), fj = 2 * j - ((s >> 1) & 1);
    int fk = ND == 3 ? 2 * k - ((s >> 2) & 1) : 0;
    T base = DEF ? C.qb[a.at(fi, fj, fk)] : (T)0;
    T o1, o2;
    op_at<ND, T, DEF>(prm, a, base, P, fi, fj, fk, o1, o2);
    long long fo = fat(P, fi, fj, ND == 3 ? fk : 1);
    const T r1 = f1[fo] - o1, r2 = f2[fo] - o2;
    s1 = s == 0 ? r1 : s1 + r1;
    s2 = s == 0 ? r2 : s2 + r2;
  }
  R1 = (T)(ND == 2 ? 0.25 : 0.125) * s1;
  R2 = (T)(ND == 2 ? 0.25 : 0.125) * s2;
}

// CoarseLoadingFunction (afasroutines.f90:646-714): parent f(mbounds) =
// R(f - A_h q) + A_2h(qc), the FAS right-hand side, with the fine residual
// recomputed in registers for the 2^d children of coarse cell (i,j,k).
template <int ND, class T, bool DEF>
__device__ __forceinline__ void coarse_load_cell(const Problem &prm, const Lvl<T> &C,
                                                 const Lvl<T> &PA, int i, int j, int k) {
  const DPatch &P = C.P;
  View<T> a = qview(C), cq = qcview(C);
  T R1, R2;
  restricted_residual<ND, T, DEF>(prm, C, a, i, j, k, R1, R2);
  int pi = P.mb[0][0] + i - 1, pj = P.mb[1][0] + j - 1,
      pk = ND == 3 ? P.mb[2][0] + k - 1 : 1;
  T base = (T)0;
  if (DEF) base = PA.qb[qat(PA.P, pi, pj, ND == 3 ? pk : 0)];
  T c1, c2;
  op_at<ND, T, DEF>(prm, cq, base, PA.P, i, j, ND == 3 ? k : 0, c1, c2);
  long long po = fat(PA.P, pi, pj, pk);
  PA.f[po] = R1 + c1;
  PA.f[po + PA.P.fvs] = R2 + c2;
}

// RelativeTruncationError (afasroutines.f90:716-797) at coarse cell (i,j,k)
// of patch C: qrte = R(A_h(q)) - A_2h(qc), the fine operator at the 2^d
// children restricted in Restriction2D/3D's order, the coarse one on qc with
// the parent's spacing.
template <int ND>
__device__ __forceinline__ void trunc_error_cell(const Dev &d, const Lvl<double> &C,
                                                 const DPatch &PA, int i, int j, int k) {
  const DPatch &P = C.P;
  View<double> a = qview(C), cq = qcview(C);
  constexpr int NC = ND == 2 ? 4 : 8;
  double r1[NC], r2[NC];
#pragma unroll
  for (int s = 0; s < NC; s++) {
    int fi = 2 * i - (s & 1), fj = 2 * j - ((s >> 1) & 1);
    int fk = ND == 3 ? 2 * k - ((s >> 2) & 1) : 0;
    op_at<ND, double, false>(d.prm, a, 0.0, P, fi, fj, fk, r1[s], r2[s]);
  }
  double R1, R2;
  if constexpr (ND == 2) {
    R1 = 0.25 * (((r1[0] + r1[1]) + r1[2]) + r1[3]);
    R2 = 0.25 * (((r2[0] + r2[1]) + r2[2]) + r2[3]);
  } else {
    R1 = 0.125 * (((((((r1[0] + r1[1]) + r1[2]) + r1[3]) + r1[4]) + r1[5]) + r1[6]) + r1[7]);
    R2 = 0.125 * (((((((r2[0] + r2[1]) + r2[2]) + r2[3]) + r2[4]) + r2[5]) + r2[6]) + r2[7]);
  }
  double c1, c2;
  op_at<ND, double, false>(d.prm, cq, 0.0, PA, i, j, ND == 3 ? k : 0, c1, c2);
  const int cx = P.mx & 1 ? 1 : P.mx >> 1, cy = P.my & 1 ? 1 : P.my >> 1;
  long long o = P.qr + (i - 1) + (long long)cx * ((j - 1) + (long long)cy * (ND == 3 ? k - 1 : 0));
  d.qrte[o] = R1 - c1;
  d.qrte[o + P.rvs] = R2 - c2;
}

// The coarse part of CoarseLoadingFunction at interior coarse cell (i,j,k)
// once the restricted fine residual (R1, R2) is known: parent f(mbounds) =
// R + A_2h(qc), the coarse stencil reading qc inside the box and, on the
// ring, the parent cells GetCoarseGhostPoints copies there (the same values).
template <int ND, class T, bool DEF>
__device__ __forceinline__ void coarse_load_from(const Problem &prm, const Lvl<T> &C,
                                                 const Lvl<T> &PA, int i, int j, int k, T R1,
                                                 T R2) {
  const DPatch &P = C.P;
  auto on_ring = [&](int a, int b, int cc) {
    return a == 0 || a == P.cmx + 1 || b == 0 || b == P.cmy + 1 ||
           (ND == 3 && (cc == 0 || cc == P.cmz + 1));
  };
  View<T> cq = qcview(C), pq = qview(PA);
  int pi = P.mb[0][0] + i - 1, pj = P.mb[1][0] + j - 1,
      pk = ND == 3 ? P.mb[2][0] + k - 1 : 1;
  T base = (T)0;
  if (DEF) base = PA.qb[qat(PA.P, pi, pj, ND == 3 ? pk : 0)];
  const int di[7] = {0, 1, -1, 0, 0, 0, 0}, dj[7] = {0, 0, 0, 1, -1, 0, 0},
            dk[7] = {0, 0, 0, 0, 0, 1, -1};
  T a1[2 * ND + 1], a2[2 * ND + 1];
#pragma unroll
  for (int s = 0; s < 2 * ND + 1; s++) {
    int ci = i + di[s], cj = j + dj[s], ck = ND == 3 ? k + dk[s] : 0;
    if (on_ring(ci, cj, ck)) {
      long long pc = pq.at(P.mb[0][0] - 1 + ci, P.mb[1][0] - 1 + cj,
                           ND == 3 ? P.mb[2][0] - 1 + ck : 0);
      a1[s] = pq.p[pc];
      a2[s] = pq.p[pc + pq.vs];
    } else {
      long long cc = cq.at(ci, cj, ck);
      a1[s] = cq.p[cc];
      a2[s] = cq.p[cc + cq.vs];
    }
  }
  T c1, c2;
  op_vals<ND, T, DEF>(prm, base, PA.P, a1, a2, c1, c2);
  long long po = fat(PA.P, pi, pj, pk);
  PA.f[po] = R1 + c1;
  PA.f[po + PA.P.fvs] = R2 + c2;
}

// GetCoarseGhostPoints and CoarseLoadingFunction for one cell of the
// ghost-extended coarse box (i,j,k in 0..cmx+1): ring cells copy the parent
// into qc; interior cells form their children's restricted residual and the
// coarse part (coarse_load_from).
template <int ND, class T, bool DEF>
__device__ __forceinline__ void coarse_ghost_load_cell(const Problem &prm, int nrvars,
                                                       const Lvl<T> &C, const Lvl<T> &PA, int i,
                                                       int j, int k) {
  const DPatch &P = C.P;
  if (i == 0 || i == P.cmx + 1 || j == 0 || j == P.cmy + 1 ||
      (ND == 3 && (k == 0 || k == P.cmz + 1))) {
    coarse_ghost_cell<ND, T>(nrvars, C, PA, i, j, k);
    return;
  }
  T R1, R2;
  restricted_residual<ND, T, DEF>(prm, C, qview(C), i, j, k, R1, R2);
  coarse_load_from<ND, T, DEF>(prm, C, PA, i, j, k, R1, R2);
}

// CorrectFine (afasroutines.f90:842-963, mbc = 1): q += P1(parent q - qc) at
// fine cell (i,j,k) of 0..mx+1, the separable (3,1)/4 prolongation of
// BiLinProlongationP1 / TriLinProlongationP1 (gridutilities.f90:137-175,
// :474-529) evaluated in the reference's x, y(, z) pass order.
template <int ND, class T>
__device__ __forceinline__ void correct_cell(int nrvars, const Lvl<T> &C, const Lvl<T> &PA,
                                             int i, int j, int k) {
  const DPatch &P = C.P;
  View<T> a = qview(C), cq = qcview(C), pq = qview(PA);
  int kx = i >> 1, ky = j >> 1, kz = ND == 3 ? k >> 1 : 0;
  bool ex = (i & 1) == 0, ey = (j & 1) == 0, ez = (k & 1) == 0;
  int o0 = P.mb[0][0] - 1, o1 = P.mb[1][0] - 1, o2 = ND == 3 ? P.mb[2][0] - 1 : 0;
  long long qc0 = a.at(i, j, k);
  for (int v = 0; v < nrvars; v++) {
    const T *cp = cq.p + v * cq.vs, *pp = pq.p + v * pq.vs;
    auto corr = [&](int ci, int cj, int ck) {
      return pp[pq.at(o0 + ci, o1 + cj, ND == 3 ? o2 + ck : 0)] - cp[cq.at(ci, cj, ck)];
    };
    auto xpass = [&](int cj, int ck) {
      T a0 = corr(kx, cj, ck), a1 = corr(kx + 1, cj, ck);
      return ex ? (T)3 * a0 + a1 : a0 + (T)3 * a1;
    };
    T p;
    if (ND == 2) {
      T b0 = xpass(ky, 0), b1 = xpass(ky + 1, 0);
      p = ey ? ((T)3 * b0 + b1) / (T)16 : (b0 + (T)3 * b1) / (T)16;
    } else {
      auto ypass = [&](int ck) {
        T b0 = xpass(ky, ck), b1 = xpass(ky + 1, ck);
        return ey ? (T)3 * b0 + b1 : b0 + (T)3 * b1;
      };
      T c0 = ypass(kz), c1 = ypass(kz + 1);
      p = ez ? ((T)3 * c0 + c1) / (T)64 : (c0 + (T)3 * c1) / (T)64;
    }
    T *qv = a.p + v * a.vs;
    qv[qc0] = qv[qc0] + p;
  }
}

// The same for the 2^d fine cells (2m..2m+1, 2n..2n+1(, 2o..2o+1)) of a
// ghost-extended fine box, which all read the coarse corrections of cells
// m..m+1, n..n+1(, o..o+1): each correction is formed once and every fine
// cell is then updated with correct_cell's arithmetic.
template <int ND, class T>
__device__ __forceinline__ void correct_block(int nrvars, const Lvl<T> &C, const Lvl<T> &PA,
                                              int m, int n, int o) {
  const DPatch &P = C.P;
  View<T> a = qview(C), cq = qcview(C), pq = qview(PA);
  int o0 = P.mb[0][0] - 1, o1 = P.mb[1][0] - 1, o2 = ND == 3 ? P.mb[2][0] - 1 : 0;
  constexpr int NZ = ND == 3 ? 2 : 1;
  for (int v = 0; v < nrvars; v++) {
    const T *cp = cq.p + v * cq.vs, *pp = pq.p + v * pq.vs;
    T cr[2][2][NZ];   // cr[dx][dy][dz] = parent - qc at coarse (m+dx, n+dy, o+dz)
#pragma unroll
    for (int dz = 0; dz < NZ; dz++)
#pragma unroll
      for (int dy = 0; dy < 2; dy++)
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
          int ci = m + dx, cj = n + dy, ck = ND == 3 ? o + dz : 0;
          cr[dx][dy][dz] = pp[pq.at(o0 + ci, o1 + cj, ND == 3 ? o2 + ck : 0)] - cp[cq.at(ci, cj, ck)];
        }
    T *qv = a.p + v * a.vs;
#pragma unroll
    for (int fz = 0; fz < NZ; fz++)
#pragma unroll
      for (int fy = 0; fy < 2; fy++)
#pragma unroll
        for (int fx = 0; fx < 2; fx++) {
          int i = 2 * m + fx, j = 2 * n + fy, k = ND == 3 ? 2 * o + fz : 0;
          if (i > P.mx + 1 || j > P.my + 1 || (ND == 3 && k > P.mz + 1)) continue;
          bool ex = fx == 0, ey = fy == 0, ez = fz == 0;
          auto xpass = [&](int dy, int dz) {
            T a0 = cr[0][dy][dz], a1 = cr[1][dy][dz];
            return ex ? (T)3 * a0 + a1 : a0 + (T)3 * a1;
          };
          T p;
          if constexpr (ND == 2) {
            T b0 = xpass(0, 0), b1 = xpass(1, 0);
            p = ey ? ((T)3 * b0 + b1) / (T)16 : (b0 + (T)3 * b1) / (T)16;
          } else {
            auto ypass = [&](int dz) {
              T b0 = xpass(0, dz), b1 = xpass(1, dz);
              return ey ? (T)3 * b0 + b1 : b0 + (T)3 * b1;
            };
            T c0 = ypass(0), c1 = ypass(1);
            p = ez ? ((T)3 * c0 + c1) / (T)64 : (c0 + (T)3 * c1) / (T)64;
          }
          long long q0 = a.at(i, j, k);
          qv[q0] = qv[q0] + p;
        }
  }
}

// correct_block with every variable's loads issued before any store (one
// round of loads per thread instead of one per variable), for the level
// kernel; the tail keeps correct_block, which needs fewer registers.
template <int ND, class T>
__device__ __forceinline__ void correct_block_all(int nrvars, const Lvl<T> &C, const Lvl<T> &PA,
                                              int m, int n, int o) {
  const DPatch &P = C.P;
  View<T> a = qview(C), cq = qcview(C), pq = qview(PA);
  int o0 = P.mb[0][0] - 1, o1 = P.mb[1][0] - 1, o2 = ND == 3 ? P.mb[2][0] - 1 : 0;
  constexpr int NZ = ND == 3 ? 2 : 1, NF = 4 * NZ;
  // the fine cells' offsets (-1 past the box) and the coarse corrections of
  // every variable, all loaded before any store
  long long fo[NF];
#pragma unroll
  for (int fz = 0; fz < NZ; fz++)
#pragma unroll
    for (int fy = 0; fy < 2; fy++)
#pragma unroll
      for (int fx = 0; fx < 2; fx++) {
        int i = 2 * m + fx, j = 2 * n + fy, k = ND == 3 ? 2 * o + fz : 0;
        const bool in = !(i > P.mx + 1 || j > P.my + 1 || (ND == 3 && k > P.mz + 1));
        fo[fx + 2 * (fy + 2 * fz)] = in ? a.at(i, j, k) : -1;
      }
  auto one = [&](int v, T (&cr)[2][2][NZ], T (&qf)[NF]) {
    const T *cp = cq.p + v * cq.vs, *pp = pq.p + v * pq.vs;
    const T *qv = a.p + v * a.vs;
#pragma unroll
    for (int dz = 0; dz < NZ; dz++)
#pragma unroll
      for (int dy = 0; dy < 2; dy++)
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
          int ci = m + dx, cj = n + dy, ck = ND == 3 ? o + dz : 0;
          cr[dx][dy][dz] = pp[pq.at(o0 + ci, o1 + cj, ND == 3 ? o2 + ck : 0)] - cp[cq.at(ci, cj, ck)];
        }
#pragma unroll
    for (int e = 0; e < NF; e++) qf[e] = fo[e] >= 0 ? qv[fo[e]] : (T)0;
  };
  auto put = [&](int v, const T (&cr)[2][2][NZ], const T (&qf)[NF]) {
    T *qv = a.p + v * a.vs;
#pragma unroll
    for (int fz = 0; fz < NZ; fz++)
#pragma unroll
      for (int fy = 0; fy < 2; fy++)
#pragma unroll
        for (int fx = 0; fx < 2; fx++) {
          const int e = fx + 2 * (fy + 2 * fz);
          if (fo[e] < 0) continue;
          bool ex = fx == 0, ey = fy == 0, ez = fz == 0;
          auto xpass = [&](int dy, int dz) {
            T a0 = cr[0][dy][dz], a1 = cr[1][dy][dz];
            return ex ? (T)3 * a0 + a1 : a0 + (T)3 * a1;
          };
          T p;
          if constexpr (ND == 2) {
            T b0 = xpass(0, 0), b1 = xpass(1, 0);
            p = ey ? ((T)3 * b0 + b1) / (T)16 : (b0 + (T)3 * b1) / (T)16;
          } else {
            auto ypass = [&](int dz) {
              T b0 = xpass(0, dz), b1 = xpass(1, dz);
              return ey ? (T)3 * b0 + b1 : b0 + (T)3 * b1;
            };
            T c0 = ypass(0), c1 = ypass(1);
            p = ez ? ((T)3 * c0 + c1) / (T)64 : (c0 + (T)3 * c1) / (T)64;
          }
          qv[fo[e]] = qf[e] + p;
        }
  };
  if (nrvars == 2) {
    T cr0[2][2][NZ], cr1[2][2][NZ], q0[NF], q1[NF];
    one(0, cr0, q0);
    one(1, cr1, q1);
    put(0, cr0, q0);
    put(1, cr1, q1);
    return;
  }
  for (int v = 0; v < nrvars; v++) {
    T cr[2][2][NZ], qf[NF];
    one(v, cr, qf);
    put(v, cr, qf);
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
    double c = d.qold[P.q + qat(P, i, j, k)];
    d.ftmp[fo] = c;
    d.ftmp[fo + P.fvs] = -dfe(d.prm, c);
  }
  for (int v = 0; v < d.nrvars; v++) d.f[fo + v * P.fvs] = d.ftmp[fo + v * P.fvs] + 0.0;
}

// Start of a defect-correction step on cell (i,j,k) of 0..mx+1: the FP32
// right-hand side fe = f - A(q) (FP64 residual, interior cells), the base
// point qb = q_1 and a zero correction e.
template <int ND>
__device__ __forceinline__ void defect_setup_cell(const Dev &d, const DPatch &P, int i, int j,
                                                  int k) {
  Lvl<double> L{d.q + P.q, nullptr, nullptr, nullptr, P};
  View<double> a = qview(L);
  long long c = a.at(i, j, k);
  d.qb[P.q + c] = (float)d.q[P.q + c];
  d.e[P.q + c] = 0.f;
  d.e[P.q + c + P.qvs] = 0.f;
  bool inner = i >= 1 && i <= P.mx && j >= 1 && j <= P.my &&
               (ND == 2 || (k >= 1 && k <= P.mz));
  if (!inner) return;
  double o1, o2;
  op_at<ND, double, false>(d.prm, a, 0.0, P, i, j, k, o1, o2);
  long long fo = P.f + fat(P, i, j, ND == 3 ? k : 1);
  d.fe[fo] = (float)(d.f[fo] - o1);
  d.fe[fo + P.fvs] = (float)(d.f[fo + P.fvs] - o2);
}

// End of a defect-correction step: q += e on interior cell (i,j,k).
template <int ND>
__device__ __forceinline__ void apply_correction_cell(const Dev &d, const DPatch &P, int i,
                                                      int j, int k) {
  long long c = P.q + qat(P, i, j, k);
  d.q[c] = d.q[c] + (double)d.e[c];
  d.q[c + P.qvs] = d.q[c + P.qvs] + (double)d.e[c + P.qvs];
}

// ---- tile decode -----------------------------------------------------------
template <int ND>
__device__ __forceinline__ void tile_cell(const Dev &d, const Tile &t, int &i, int &j, int &k) {
  int tid = threadIdx.x, lx = d.tile_lx, ly = d.tile_ly;
  i = t.i0 + (tid & ((1 << lx) - 1));
  j = t.j0 + ((tid >> lx) & ((1 << ly) - 1));
  k = ND == 3 ? t.k0 + (tid >> (lx + ly)) : 0;
}
// The same cells with the lanes of each tile row ordered by the parity of i:
// the first half of the row's lanes take the cells of i0's parity, the
// second half the others, so one load or store of the row's lanes touches a
// contiguous run of each half-row in the split layout. For kernels whose
// thread works on one fine cell.
template <int ND>
__device__ __forceinline__ void tile_cell_px(const Dev &d, const Tile &t, int &i, int &j,
                                             int &k) {
  int tid = threadIdx.x, lx = d.tile_lx, ly = d.tile_ly;
  const int l = tid & ((1 << lx) - 1), hx = lx - 1;
  i = t.i0 + 2 * (l & ((1 << hx) - 1)) + (l >> hx);
  j = t.j0 + ((tid >> lx) & ((1 << ly) - 1));
  k = ND == 3 ? t.k0 + (tid >> (lx + ly)) : 0;
}
// Workgroups are dispatched round-robin over the chip's 8 XCDs, each with its
// own L2. A kernel launched over xcd_grid(nt) blocks runs tile
// (b mod 8) * ceil(nt / 8) + b / 8 in block b, so each XCD walks a contiguous
// run of the tile list (neighbouring tiles of a patch, whose halo rows and
// cache lines overlap); blocks past the list return.
constexpr int NXCD = 8;
inline unsigned xcd_grid(int nt) { return (unsigned)(NXCD * ((nt + NXCD - 1) / NXCD)); }
__device__ __forceinline__ int xcd_tile(int nt) {
  const int per = (nt + NXCD - 1) / NXCD;
  return (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD);
}

// The relaxation kernels run half-size workgroups (RELAX_THREADS = half a
// tile): thread tid takes the (tid mod TX/2)-th cell of colour `color` (i + j
// (+ k) + color odd, the cells RelaxGrid updates in that pass) in its row of
// the tile, so no lane idles on the other colour. Tiles are at least 2 wide.
constexpr int RELAX_THREADS = 128;
template <int ND>
__device__ __forceinline__ void tile_color_cell(const Dev &d, const Tile &t, int color, int &i,
                                                int &j, int &k) {
  int tid = threadIdx.x, hx = d.tile_lx - 1, ly = d.tile_ly;
  int ti = tid & ((1 << hx) - 1);
  j = t.j0 + ((tid >> hx) & ((1 << ly) - 1));
  k = ND == 3 ? t.k0 + (tid >> (hx + ly)) : 0;
  int p = (j + k + color + 1) & 1;   // parity of the colour's i in this row
  i = t.i0 + ((p - t.i0) & 1) + 2 * ti;
}
// Linear tiles: tile t covers LIN_TILE consecutive cells (x fastest, then y,
// z) of an ex x ey (x ez) box of one patch, starting at cell t.i0 of the
// box, so a launch over a level's boxes keeps every lane busy but those of
// each box's last tile. Returns false for lanes past the box.
template <int ND>
__device__ __forceinline__ bool lin_cell(const Tile &t, int ex, int ey, int ez, int &i, int &j,
                                         int &k) {
  const int e = t.i0 + (int)threadIdx.x;
  if (e >= ex * ey * (ND == 3 ? ez : 1)) return false;
  i = e % ex;
  const int r = e / ex;
  j = ND == 3 ? r % ey : r;
  k = ND == 3 ? r / ey : 0;
  return true;
}
// Linear tiles in device order: tile t covers LIN_TILE consecutive device
// positions of a patch's ghost-extended q array (rows of px cells split by
// the parity of x), from position t.i0, so consecutive lanes touch
// consecutive addresses. Returns the cell (i, j, k), false past the array or
// on a row's padding slot.
template <int ND>
__device__ __forceinline__ bool lin_qcell(const Tile &t, const DPatch &P, int &i, int &j,
                                          int &k) {
  const Split s = split_of(P.px, P.py);
  const long long e = t.i0 + (long long)threadIdx.x;
  if (e >= s.pl * (ND == 3 ? P.mz + 2 : 1)) return false;
  k = ND == 3 ? (int)(e / s.pl) : 0;
  const long long m = e - (long long)k * s.pl;
  j = (int)(m / s.rs);
  const long long x = m - (long long)j * s.rs;
  i = (int)(x < s.xo ? 2 * x : 2 * (x - s.xo) + 1);
  return i < P.px;
}
template <int ND>
__device__ __forceinline__ bool in_box(int i, int j, int k, int ex, int ey, int ez) {
  return i <= ex && j <= ey && (ND == 2 || k <= ez);
}

// ---- level kernels ---------------------------------------------------------
template <int ND, class T, bool DEF>
__global__ __launch_bounds__(RELAX_THREADS) void k_relax(Dev d, const Tile *tiles, int nt,
                                                          int ipass) {
  const int b = xcd_tile(nt);
  if (b >= nt) return;
  Tile t = tiles[b];
  if (stopped_patch(d, t.patch)) return;
  Lvl<T> L = level_ref<T>(d, t.patch);
  int i, j, k;
  // RelaxGrid's loop start 1+mod(j+ipass,2) (and +k in 3D) selects the
  // cells with i+j(+k)+ipass odd.
  tile_color_cell<ND>(d, t, ipass, i, j, k);
  if (!in_box<ND>(i, j, k, L.P.mx, L.P.my, L.P.mz)) return;
  relax_cell<ND, T, DEF>(d.prm, L, i, j, k);
}

// Ghost updates a workgroup applies for its own tile: list g's updates of tile
// b (the same update may be applied by several tiles, each writing the same
// value), then a workgroup barrier if there were any.
template <class T>
__device__ __forceinline__ void ghost_tile(const Dev &d, const GhostList &g, int b) {
  if (!g.ss) return;
  const int s0 = g.ss[b], s1 = g.ss[b + 1], l0 = g.sl[b], l1 = g.sl[b + 1];
  if (s1 > s0 || l1 > l0) {
    T *q = fields<T>(d).q;
    constexpr bool DEF = std::is_same<T, float>::value;
    for (int e = s0 + (int)threadIdx.x; e < s1; e += blockDim.x)
      ghost_op_s<T, DEF>(d.nrvars, q, d.qvs, g.s[e], qpo_of(d.prm));
    for (int e = l0 + (int)threadIdx.x; e < l1; e += blockDim.x)
      ghost_op<T, DEF>(d.nrvars, q, d.qvs, g.l[e], qpo_of(d.prm));
    __syncthreads();
  }
}
// A list's updates spread over the whole grid (no ordering among them),
// skipping those of stopped groups.
template <class T>
__device__ __forceinline__ void ghost_alongside(const Dev &d, const GhostList &g) {
  T *q = fields<T>(d).q;
  constexpr bool DEF = std::is_same<T, float>::value;
  const int stride = gridDim.x * blockDim.x, e0 = blockIdx.x * blockDim.x + threadIdx.x;
  for (int e = e0; e < g.ns; e += stride)
    if (!stopped(d, g.s[e].group)) ghost_op_s<T, DEF>(d.nrvars, q, d.qvs, g.s[e], qpo_of(d.prm));
  for (int e = e0; e < g.nl; e += stride)
    if (!stopped(d, g.l[e].group)) ghost_op<T, DEF>(d.nrvars, q, d.qvs, g.l[e], qpo_of(d.prm));
}

// SetGhost(level, ipass) followed by the relaxation of colour `color`, in one
// launch: each workgroup first applies the ghost updates its tile's cells read
// (its share of list `pre`, chosen on the host so that no update reads a cell
// another tile relaxes), then relaxes its tile. With a `post` list, the
// workgroup then synchronises and applies its share of it:
// updates of the next SetGhost(level, color) that read cells this tile has
// just relaxed (chosen on the host so that no other tile touches their
// cells in this launch). A list with null starts: no updates. The tile, both
// lists' per-tile starts and each thread's first record of both lists are
// loaded at once at the start, so the chain of dependent loads ahead of the
// post updates is short.
//
// With STAGE, the tile's cells and their one-cell halo (both variables) are
// first copied into LDS in the same parity-split row layout (a box of
// (tx+2) x (ty+2) (x (tz+2)) cells whose first cell i0 - 1 is even, so a
// cell's parity is the same in the box), with coalesced loads of whole
// half-rows; the relaxation then reads its stencil from LDS and writes the
// new values to the global pool. The arithmetic is relax_core's either way.
extern __shared__ __align__(16) unsigned char relax_lds[];
template <int ND, class T, bool DEF, bool STAGE>
__global__ __launch_bounds__(RELAX_THREADS) void k_relax_fused(Dev d, const Tile *tiles, int nt,
                                                                int color, GhostList pre,
                                                                GhostList post) {
  const int b = xcd_tile(nt);
  if (b >= nt) return;
  // the tile and both lists' shares of it, loads independent of one another
  const Tile t = tiles[b];
  int s0 = 0, s1 = 0, l0 = 0, l1 = 0, ps0 = 0, ps1 = 0, pl0 = 0, pl1 = 0;
  if (pre.ss) {
    s0 = pre.ss[b]; s1 = pre.ss[b + 1]; l0 = pre.sl[b]; l1 = pre.sl[b + 1];
  }
  if (post.ss) {
    ps0 = post.ss[b]; ps1 = post.ss[b + 1]; pl0 = post.sl[b]; pl1 = post.sl[b + 1];
  }
  // this thread's first update of each list, loaded now (records only)
  const int se = s0 + (int)threadIdx.x, pe = ps0 + (int)threadIdx.x;
  GhostS sr{}, pr{};
  if (se < s1) sr = pre.s[se];
  if (pe < ps1) pr = post.s[pe];
  if (stopped_patch(d, t.patch)) return;   // the tile's updates are its patch's group's
  T *q = fields<T>(d).q;
  Lvl<T> L = level_ref<T>(d, t.patch);
  int i, j, k;
  tile_color_cell<ND>(d, t, color, i, j, k);
  // every thread reaches the barriers below (they depend on the block only)
  const bool mine = in_box<ND>(i, j, k, L.P.mx, L.P.my, L.P.mz);
  if (s1 > s0 || l1 > l0) {
    if (se < s1) ghost_op_s<T, DEF>(d.nrvars, q, d.qvs, sr, qpo_of(d.prm));
    for (int e = se + blockDim.x; e < s1; e += blockDim.x)
      ghost_op_s<T, DEF>(d.nrvars, q, d.qvs, pre.s[e], qpo_of(d.prm));
    for (int e = l0 + (int)threadIdx.x; e < l1; e += blockDim.x)
      ghost_op<T, DEF>(d.nrvars, q, d.qvs, pre.l[e], qpo_of(d.prm));
    __syncthreads();
  }
  if constexpr (!STAGE) {
    if (mine) relax_cell<ND, T, DEF>(d.prm, L, i, j, k);
  } else {
    const int tx = 1 << d.tile_lx, ty = 1 << d.tile_ly;
    const int tz = ND == 3 ? 256 >> (d.tile_lx + d.tile_ly) : 1;
    const int bx = tx + 2, by = ty + 2, bz = ND == 3 ? tz + 2 : 1;
    const int hb = bx >> 1, ncell = bx * by * bz;
    const View<T> g = qview(L);
    T *box = (T *)relax_lds;
    // box cell (li, lj, lk) is patch cell (i0 - 1 + li, j0 - 1 + lj, k0 - 1 + lk);
    // cells past the patch's ghost ring are never read and stay unset
    const int o0 = t.i0 - 1, o1 = t.j0 - 1, o2 = ND == 3 ? t.k0 - 1 : 0;
    for (int e = threadIdx.x; e < 2 * ncell; e += blockDim.x) {
      const int v = e >= ncell, r = e - v * ncell;
      const int lk = r / (bx * by), m = r - lk * (bx * by), lj = m / bx, x = m - lj * bx;
      const int li = x < hb ? 2 * x : 2 * (x - hb) + 1;
      const int gi = o0 + li, gj = o1 + lj, gk = o2 + lk;
      if (gi <= L.P.mx + 1 && gj <= L.P.my + 1 && (ND == 2 || gk <= L.P.mz + 1))
        box[e] = g.p[g.at(gi, gj, gk) + v * g.vs];
    }
    __syncthreads();
    if (mine) {
      const View<T> a{box, hb, bx, (long long)bx * by, (long long)ncell};
      const long long c = g.at(i, j, k);
      const long long fo = fat(L.P, i, j, ND == 3 ? k : 1);
      relax_core<ND, T, DEF>(d.prm, L.P, a, i - o0, j - o1, ND == 3 ? k - o2 : 0, g.p + c,
                             g.p + g.vs + c, DEF ? L.qb[c] : (T)0, L.f[fo], L.f[fo + L.P.fvs]);
    }
  }
  if (ps1 > ps0 || pl1 > pl0) {
    __syncthreads();
    if (pe < ps1) ghost_op_s<T, DEF>(d.nrvars, q, d.qvs, pr, qpo_of(d.prm));
    for (int e = pe + blockDim.x; e < ps1; e += blockDim.x)
      ghost_op_s<T, DEF>(d.nrvars, q, d.qvs, post.s[e], qpo_of(d.prm));
    for (int e = pl0 + (int)threadIdx.x; e < pl1; e += blockDim.x)
      ghost_op<T, DEF>(d.nrvars, q, d.qvs, post.l[e], qpo_of(d.prm));
  }
}

// The whole smoothing of an isolated patch (IsoPatch) in one workgroup: the
// patch's ghost-extended q and the parent cells its SetGhost programs read
// are copied into LDS; then, as level_relax orders them, per pass colour 1
// and colour 2 are relaxed, each after the pending program (SetGhost(0)
// first when ghost0_first, then the colour just relaxed), and SetGhost(2)
// after the last sweep when final2; the array goes back to the pool. No
// other patch of the level reads or writes a cell of this patch, and the
// parent level does not change while the level is smoothed, so the result
// equals the level-wide sequence bit for bit. One workgroup per patch.
extern __shared__ __align__(16) unsigned char iso_lds[];
constexpr int ISO_THREADS = 1024;
constexpr int ISO_LOAD = 8;   // array elements each thread has in flight when staging
constexpr int ISO_FREG = 2;   // cells per colour whose f each thread keeps in registers
constexpr int ISO_RREG = 2;   // updates of the next program each thread prefetches
template <int ND, class T, bool DEF>
__global__ __launch_bounds__(ISO_THREADS) void k_smooth_iso(Dev d, const IsoPatch *iso,
                                                             const int *stage, const IsoRec *rec,
                                                             const unsigned long long *srec,
                                                             int npasses, int ghost0_first,
                                                             int final2) {
  const IsoPatch &I = iso[blockIdx.x];   // read in place (scalar loads)
  if (stopped_patch(d, I.patch)) return;
  const DPatch &P = d.P[I.patch];
  const Fld<T> F = fields<T>(d);
  T *box = (T *)iso_lds;
  const int qn = (int)P.qn, W = qn + I.ns, tid = threadIdx.x, nth = blockDim.x;
  T *gq = F.q + P.q;
  const T *f = F.f + P.f;
  // the colour's cells: slot sl of row r = sl / hx is the t-th cell of the
  // colour's parity in that row (i + j + k + color odd, tile_color_cell)
  const int hx = (P.mx + 1) >> 1, rows = P.my * (ND == 3 ? P.mz : 1), slots = hx * rows;
  auto cell = [&](int sl, int color, int &i, int &j, int &k) __attribute__((always_inline)) {
    const int r = sl / hx, t = sl - r * hx;
    j = 1 + (ND == 3 ? r % P.my : r);
    k = ND == 3 ? 1 + r / P.my : 0;
    i = 2 * t + 2 - ((j + k + color + 1) & 1);
    return i <= P.mx;
  };
  // every load of the staging and of the first cells' f in flight at once
  T fr[2][ISO_FREG][2];
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int u = 0; u < ISO_FREG; u++) {
      int i, j, k;
      const int sl = tid + u * nth;
      fr[c][u][0] = fr[c][u][1] = (T)0;
      if (sl < slots && cell(sl, c + 1, i, j, k)) {
        const long long fo = fat(P, i, j, ND == 3 ? k : 1);
        fr[c][u][0] = f[fo];
        fr[c][u][1] = f[fo + P.fvs];
      }
    }
  for (int e0 = 0; e0 < W; e0 += ISO_LOAD * nth) {
    T v[ISO_LOAD][2];
#pragma unroll
    for (int u = 0; u < ISO_LOAD; u++) {
      const int e = e0 + u * nth + tid;
      if (e < qn) {
        v[u][0] = gq[e];
        v[u][1] = gq[e + d.qvs];
      } else if (e < W) {
        const long long o = stage[I.stage0 + e - qn];
        v[u][0] = F.q[o];
        v[u][1] = F.q[o + d.qvs];
      }
    }
#pragma unroll
    for (int u = 0; u < ISO_LOAD; u++) {
      const int e = e0 + u * nth + tid;
      if (e < W) {
        box[e] = v[u][0];
        box[W + e] = v[u][1];
      }
    }
  }
  __syncthreads();
  // a program's first ISO_RREG short records per thread are fetched before
  // the relaxation that precedes it; the 16-byte records (dwords: dst |
  // kind << 16 | nflip << 24, then the sources two to a dword, low half
  // first) are read as they are applied
  const uint4 *rw = reinterpret_cast<const uint4 *>(rec);
  unsigned long long pre[ISO_RREG];
  auto fetch = [&](int p) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < ISO_RREG; u++) {
      const int r = I.srec[p] + u * nth + tid;
      if (r < I.srec[p + 1]) pre[u] = srec[r];
    }
  };
  auto put = [&](unsigned dst, int kind, int nflip, auto src) __attribute__((always_inline)) {
#pragma unroll
    for (int v = 0; v < 2; v++) {
      const T *b = box + v * W;
      box[v * W + dst] =
          ghost_value<T, DEF>(kind, nflip, [&](int n) { return b[src(n)]; }, qpo_of(d.prm), v);
    }
  };
  auto apply_s = [&](const unsigned long long w) __attribute__((always_inline)) {
    constexpr unsigned m = (1u << ISO_SHORT_BITS) - 1;
    put((unsigned)w & m, (int)(w >> 52) & 1, (int)(w >> 53) & 63,
        [&](int n) { return (unsigned)(w >> (ISO_SHORT_BITS * (n + 1))) & m; });
  };
  auto apply_l = [&](const uint4 g) __attribute__((always_inline)) {
    const unsigned wv[3] = {g.y, g.z, g.w};
    put(g.x & 0xffffu, (g.x >> 16) & 0xff, g.x >> 24,
        [&](int n) { return (wv[n >> 1] >> (16 * (n & 1))) & 0xffffu; });
  };
  auto program = [&](int p) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < ISO_RREG; u++)
      if (I.srec[p] + u * nth + tid < I.srec[p + 1]) apply_s(pre[u]);
    for (int r = I.srec[p] + ISO_RREG * nth + tid; r < I.srec[p + 1]; r += nth) apply_s(srec[r]);
    for (int r = I.rec[p] + tid; r < I.rec[p + 1]; r += nth) apply_l(rw[r]);
    __syncthreads();
  };
  const View<T> a = view_of<T>(box, P.px, P.py, W);
  auto relax_color = [&](const T (&fc)[ISO_FREG][2], int color) {
    auto one = [&](int sl, T f1, T f2) __attribute__((always_inline)) {
      int i, j, k;
      if (!cell(sl, color, i, j, k)) return;
      const long long c = a.at(i, j, k);
      relax_core<ND, T, DEF>(d.prm, P, a, i, j, k, box + c, box + W + c,
                             DEF ? F.qb[P.q + c] : (T)0, f1, f2);
    };
#pragma unroll
    for (int u = 0; u < ISO_FREG; u++)
      if (tid + u * nth < slots) one(tid + u * nth, fc[u][0], fc[u][1]);
    for (int sl = tid + ISO_FREG * nth; sl < slots; sl += nth) {
      int i, j, k;
      if (!cell(sl, color, i, j, k)) continue;
      const long long fo = fat(P, i, j, ND == 3 ? k : 1);
      one(sl, f[fo], f[fo + P.fvs]);
    }
    __syncthreads();
  };
  // level_relax's order: per pass colour 1 then 2, each after the pending
  // program; SetGhost(2) after the last sweep
  int pending = ghost0_first ? 0 : -1;
  if (pending >= 0) fetch(pending);
  for (int pass = 0; pass < npasses; pass++) {
    if (pending >= 0) program(pending);
    fetch(1);
    relax_color(fr[0], 1);
    program(1);
    if (pass + 1 < npasses || final2) fetch(2);
    relax_color(fr[1], 2);
    pending = 2;
    if (pass + 1 < npasses) continue;
    if (final2) program(2);
  }
  for (int e0 = 0; e0 < qn; e0 += ISO_LOAD * nth)
#pragma unroll
    for (int u = 0; u < ISO_LOAD; u++) {
      const int e = e0 + u * nth + tid;
      if (e < qn) {
        gq[e] = box[e];
        gq[e + d.qvs] = box[W + e];
      }
    }
}

// One SetGhost(level, ipass) program: every ghost-cell update is independent.
// (one thread per update, short records first)
template <class T, bool DEF>
__global__ __launch_bounds__(256) void k_ghost(Dev d, GhostList g) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < g.ns) {
    if (!stopped(d, g.s[e].group)) ghost_op_s<T, DEF>(d.nrvars, fields<T>(d).q, d.qvs, g.s[e], qpo_of(d.prm));
  } else if (e - g.ns < g.nl) {
    const GhostOp &u = g.l[e - g.ns];
    if (!stopped(d, u.group)) ghost_op<T, DEF>(d.nrvars, fields<T>(d).q, d.qvs, u, qpo_of(d.prm));
  }
}

template <int ND, class T>
__global__ __launch_bounds__(256) void k_restrict(Dev d, const Tile *tiles, int nt,
                                                  GhostList gops) {
  const int b = xcd_tile(nt);
  if (b < nt && !stopped_patch(d, tiles[b].patch)) {
    const Tile t = tiles[b];
    Lvl<T> C = level_ref<T>(d, t.patch);
    int i, j, k;
    tile_cell<ND>(d, t, i, j, k);
    if (in_box<ND>(i, j, k, C.P.cmx, C.P.cmy, C.P.cmz))
      restrict_cell<ND, T>(d.nrvars, C, level_ref<T>(d, C.P.parent), i, j, k, true);
  }
  // the riding program reads no parent cell this launch writes and writes
  // fine ghost cells it does not read, so it runs after the work
  ghost_alongside<T>(d, gops);
}

template <int ND, class T>
__global__ __launch_bounds__(256) void k_filldown(Dev d, const Tile *tiles, int nt, int field,
                                                  GhostList gops) {
  const int b = xcd_tile(nt);
  if (b < nt && !stopped_patch(d, tiles[b].patch)) {
    const Tile t = tiles[b];
    Lvl<T> C = level_ref<T>(d, t.patch);
    int i, j, k;
    tile_cell<ND>(d, t, i, j, k);
    if (in_box<ND>(i, j, k, C.P.cmx, C.P.cmy, C.P.cmz))
      filldown_cell<ND, T>(d.nrvars, C, level_ref<T>(d, C.P.parent), i, j, k, field, true);
  }
  ghost_alongside<T>(d, gops);   // as in k_restrict
}

// Tiles cover the ghost-extended coarse box 0..cmx+1.
template <int ND, class T>
__global__ __launch_bounds__(256) void k_coarse_ghost(Dev d, const Tile *tiles, int nt) {
  const int b = xcd_tile(nt);
  if (b >= nt) return;
  Tile t = tiles[b];
  if (stopped_patch(d, t.patch)) return;
  Lvl<T> C = level_ref<T>(d, t.patch);
  int i, j, k;
  tile_cell<ND>(d, t, i, j, k);
  if (!in_box<ND>(i, j, k, C.P.cmx + 1, C.P.cmy + 1, C.P.cmz + 1)) return;
  coarse_ghost_cell<ND, T>(d.nrvars, C, level_ref<T>(d, C.P.parent), i, j, k);
}

template <int ND, class T, bool DEF>
__global__ __launch_bounds__(256) void k_coarse_load(Dev d, const Tile *tiles, int nt) {
  const int b = xcd_tile(nt);
  if (b >= nt) return;
  Tile t = tiles[b];
  if (stopped_patch(d, t.patch)) return;
  Lvl<T> C = level_ref<T>(d, t.patch);
  int i, j, k;
  tile_cell<ND>(d, t, i, j, k);
  if (!in_box<ND>(i, j, k, C.P.cmx, C.P.cmy, C.P.cmz)) return;
  coarse_load_cell<ND, T, DEF>(d.prm, C, level_ref<T>(d, C.P.parent), i, j, k);
}

// The same launch over linear tiles of the ghost-extended coarse box
// 0..cmx+1 (LIN_TILE cells each), one thread per coarse cell, which forms the
// residuals of its children one after another (restricted_residual): fewer,
// fuller workgroups on levels of small patches. A SetGhost program of the
// parent may ride along: `rest` grid-stride, and per tile the updates it
// reads (`tops`).
template <int ND, class T, bool DEF>
__global__ __launch_bounds__(256) void k_coarse_ghost_load_lin(Dev d, const Tile *tiles, int nt,
                                                           GhostList rest, GhostList tops) {
  const int b = xcd_tile(nt);
  // the tile's updates are of its patch's parent, in the same group
  if (b < nt && !stopped_patch(d, tiles[b].patch)) {
    const Tile t = tiles[b];
    ghost_tile<T>(d, tops, b);
    Lvl<T> C = level_ref<T>(d, t.patch);
    int i, j, k;
    if (lin_cell<ND>(t, C.P.cmx + 2, C.P.cmy + 2, C.P.cmz + 2, i, j, k))
      coarse_ghost_load_cell<ND, T, DEF>(d.prm, d.nrvars, C, level_ref<T>(d, C.P.parent), i, j,
                                         k);
  }
  // the updates no tile reads, after the work (they touch nothing it uses)
  ghost_alongside<T>(d, rest);
}

// GetCoarseGhostPoints + CoarseLoadingFunction (afasroutines.f90:646-714)
// of a level in one launch. Tiles 0..nres-1 are fine tiles (cl_lx, cl_ly,
// cl_lz): each thread forms the fine residual f - A_h(q) of one cell into
// LDS; after a barrier one thread per coarse cell of the tile sums its 2^d
// children in the reference's child order (restricted_residual's sum),
// evaluates the coarse operator on qc (the parent on the ring, the values the
// ring copy stores) and writes the parent's f. Tiles nres.. are linear tiles
// over the ring cells of the ghost-extended coarse box (ring_cell), which
// copy the parent into qc. A SetGhost program of the parent may ride along:
// `rest` grid-stride, and per tile the updates it reads (`tops`).
template <int ND, class T, bool DEF>
__global__ __launch_bounds__(256) void k_coarse_ghost_load(Dev d, const Tile *tiles, int nt,
                                                           int nres, GhostList rest,
                                                           GhostList tops) {
  constexpr int LX = ND == 3 ? 4 : 5, LY = 3, LZ = ND == 3 ? 1 : 0;
  constexpr int TX = 1 << LX, TY = 1 << LY, HX = TX / 2, HY = TY / 2;
  constexpr int HZ = ND == 3 ? (1 << LZ) / 2 : 1;
  static_assert(TX * TY * (1 << LZ) == 256, "256 fine cells per tile");
  __shared__ T res[2][256];
  const int b = xcd_tile(nt);
  // the tile's updates are of its patch's parent, in the same group
  if (b < nt && !stopped_patch(d, tiles[b].patch)) {
    const Tile t = tiles[b];
    ghost_tile<T>(d, tops, b);
    const Lvl<T> C = level_ref<T>(d, t.patch);
    const Lvl<T> PA = level_ref<T>(d, C.P.parent);
    const DPatch &P = C.P;
    if (b < nres) {
      // this thread's fine cell, the lanes of a row ordered by the parity of i
      const int tid = threadIdx.x, l = tid & (TX - 1);
      const int li = 2 * (l & (HX - 1)) + (l >> (LX - 1)), lj = (tid >> LX) & (TY - 1);
      const int lk = ND == 3 ? tid >> (LX + LY) : 0;
      const int i = t.i0 + li, j = t.j0 + lj, k = ND == 3 ? t.k0 + lk : 0;
      const int at = li + TX * (lj + TY * lk);
      if (i <= P.mx && j <= P.my && (ND == 2 || k <= P.mz)) {
        const View<T> a = qview(C);
        const T base = DEF ? C.qb[a.at(i, j, k)] : (T)0;
        T o1, o2;
        op_at<ND, T, DEF>(d.prm, a, base, P, i, j, k, o1, o2);
        const long long fo = fat(P, i, j, ND == 3 ? k : 1);
        res[0][at] = C.f[fo] - o1;
        res[1][at] = C.f[fo + P.fvs] - o2;
      }
      __syncthreads();
      if (tid < HX * HY * HZ) {
        const int ca = tid % HX, cb = (tid / HX) % HY, cc = tid / (HX * HY);
        const int ci = (t.i0 + 1) / 2 + ca, cj = (t.j0 + 1) / 2 + cb;
        const int ck = ND == 3 ? (t.k0 + 1) / 2 + cc : 0;
        if (ci <= P.cmx && cj <= P.cmy && (ND == 2 || ck <= P.cmz)) {
          constexpr int NC = ND == 2 ? 4 : 8;
          T s1 = (T)0, s2 = (T)0;
#pragma unroll
          for (int s = 0; s < NC; s++) {
            const int x = 2 * ca + 1 - (s & 1), y = 2 * cb + 1 - ((s >> 1) & 1);
            const int z = ND == 3 ? 2 * cc + 1 - ((s >> 2) & 1) : 0;
            const int e = x + TX * (y + TY * z);
            s1 = s == 0 ? res[0][e] : s1 + res[0][e];
            s2 = s == 0 ? res[1][e] : s2 + res[1][e];
          }
          const T R1 = (T)(ND == 2 ? 0.25 : 0.125) * s1, R2 = (T)(ND == 2 ? 0.25 : 0.125) * s2;
          coarse_load_from<ND, T, DEF>(d.prm, C, PA, ci, cj, ck, R1, R2);
        }
      }
    } else {
      const long long e = t.i0 + (long long)threadIdx.x;
      const int cz = ND == 3 ? P.cmz : 1;
      if (e < ring_count(ND, P.cmx, P.cmy, cz)) {
        int i, j, k;
        ring_cell(ND, P.cmx, P.cmy, cz, e, i, j, k);
        coarse_ghost_cell<ND, T>(d.nrvars, C, PA, i, j, k);
      }
    }
  }
  // the updates no tile reads, after the work (they touch nothing it uses)
  ghost_alongside<T>(d, rest);
}

// Linear tiles over the 2x2(x2) fine blocks (m, n(, o)) = 0..cmx of the
// ghost-extended fine box 0..mx+1: each thread corrects the 2^d fine cells of
// one block, which share their coarse corrections (correct_block). A SetGhost
// program of the parent may ride along as in k_coarse_ghost_load.
template <int ND, class T>
__global__ __launch_bounds__(256) void k_correct(Dev d, const Tile *tiles, int nt,
                                                 GhostList rest, GhostList tops) {
  const int b = xcd_tile(nt);
  if (b < nt && !stopped_patch(d, tiles[b].patch)) {
    const Tile t = tiles[b];
    ghost_tile<T>(d, tops, b);
    Lvl<T> C = level_ref<T>(d, t.patch);
    int m, n, o;
    if (lin_cell<ND>(t, C.P.cmx + 1, C.P.cmy + 1, C.P.cmz + 1, m, n, o))
      correct_block_all<ND, T>(d.nrvars, C, level_ref<T>(d, C.P.parent), m, n, o);
  }
  // the updates no tile reads, after the work (they touch nothing it uses)
  ghost_alongside<T>(d, rest);
}

// RelativeTruncationError over the coarse interior tiles of a level.
template <int ND>
__global__ __launch_bounds__(256) void k_trunc_error(Dev d, const Tile *tiles, int nt) {
  const int b = xcd_tile(nt);
  if (b >= nt) return;
  Tile t = tiles[b];
  Lvl<double> C = level_ref<double>(d, t.patch);
  int i, j, k;
  tile_cell<ND>(d, t, i, j, k);
  const DPatch &P = C.P;
  if (!in_box<ND>(i, j, k, P.mx >> 1, P.my >> 1, P.mz >> 1)) return;
  trunc_error_cell<ND>(d, C, d.P[P.parent], i, j, k);
}

// SetErrFlags2D/3D (bsamroutines.f90:1093-1157) over the interior tiles of a
// level: a fine cell is flagged when its coarse cell's largest |qrte| over
// the variables exceeds tol (= qtolerance / h^d); cells outside the coarse
// cells' children are cleared.
template <int ND>
__global__ __launch_bounds__(256) void k_error_flags(Dev d, const Tile *tiles, int nt,
                                                     double tol) {
  const int b = xcd_tile(nt);
  if (b >= nt) return;
  Tile t = tiles[b];
  const DPatch &P = d.P[t.patch];
  int i, j, k;
  tile_cell_px<ND>(d, t, i, j, k);
  if (!in_box<ND>(i, j, k, P.mx, P.my, P.mz)) return;
  const int ci = (i + 1) >> 1, cj = (j + 1) >> 1, ck = ND == 3 ? (k + 1) >> 1 : 1;
  int flag = 0;
  if (ci <= P.mx >> 1 && cj <= P.my >> 1 && (ND == 2 || ck <= P.mz >> 1)) {
    const int cx = P.mx & 1 ? 1 : P.mx >> 1, cy = P.my & 1 ? 1 : P.my >> 1;
    long long o =
        P.qr + (ci - 1) + (long long)cx * ((cj - 1) + (long long)cy * (ND == 3 ? ck - 1 : 0));
    double m = fabs(d.qrte[o]);
    for (int v = 1; v < d.nrvars; v++) {
      double a = fabs(d.qrte[o + v * P.rvs]);
      if (a > m) m = a;
    }
    flag = m > tol ? 1 : 0;
  }
  d.eflags[P.fl + (i - 1) + (long long)P.mx * ((j - 1) + (long long)P.my * (ND == 3 ? k - 1 : 0))] =
      flag;
}

// SetErrFlagsUser2D/3D (problems/simple_cahn-hilliard/problem.f90:135-163,
// :425-452) over the interior tiles of a level: a cell is flagged where the
// undivided central gradient of variable 1, sqrt(gx^2 + gy^2 [+ gz^2]), is
// strictly above qtolerance(level), and cleared elsewhere.
template <int ND>
__global__ __launch_bounds__(256) void k_error_flags_user(Dev d, const Tile *tiles, int nt,
                                                          double qtol) {
  const int b = xcd_tile(nt);
  if (b >= nt) return;
  Tile t = tiles[b];
  const DPatch &P = d.P[t.patch];
  int i, j, k;
  tile_cell_px<ND>(d, t, i, j, k);
  if (!in_box<ND>(i, j, k, P.mx, P.my, P.mz)) return;
  const View<double> a = view_of<double>(d.q + P.q, P.px, P.py, P.qvs);
  const long long c = a.at(i, j, ND == 3 ? k : 0);
  const double *q = a.p + c;
  double gx = a.p[a.at(i + 1, j, ND == 3 ? k : 0)] - a.p[a.at(i - 1, j, ND == 3 ? k : 0)];
  double gy = q[a.rs] - q[-a.rs];
  double s2 = gx * gx + gy * gy;
  if constexpr (ND == 3) {
    double gz = q[a.pl] - q[-a.pl];
    s2 = s2 + gz * gz;
  }
  d.eflags[P.fl + (i - 1) + (long long)P.mx * ((j - 1) + (long long)P.my * (ND == 3 ? k - 1 : 0))] =
      fabs(sqrt(s2)) > qtol ? 1 : 0;
}

// The sums of AfterStep2D/3D (problem.f90:102-134, :398-424), one task per
// workgroup: a run of rows (2D) or planes (3D) of a patch's interior (part
// 0) or of its parent's cells under it (part 1). Each task writes three
// sums: of q1; and in 2D the two sums of Energy2D (problem.f90:330-361) --
// of Ff(c) and of the four squared differences per cell (ghost ring read).
// Each sum is reduced in a fixed order (strided per thread, then the wave
// and workgroup trees), so repeated calls agree bit for bit.
__device__ __forceinline__ double well_ff(double c, double lo, double hi, double alpha) {
  double c0 = c - lo, cm1 = c - hi;
  return c0 * c0 * cm1 * cm1 * alpha / 4.0;
}

__device__ __forceinline__ double block_sum(double v, double *lds) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[wv] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < nw; w++) t += lds[w];
  return t;
}

template <int ND>
__global__ __launch_bounds__(256) void k_after_step(Dev d, const StatTask *tasks, double lo,
                                                    double hi, double *out) {
  __shared__ double lds[4];
  const StatTask T = tasks[blockIdx.x];
  const DPatch &P = d.P[T.patch];
  const double alpha = d.prm.fc_alpha;
  // the box c(1..m1, 1..m2[, 1..m3]) of the task's array: cell (i, j, k) of
  // the box is the array's cell (o0 + i, o1 + j, o2 + k), rows of rx cells
  const double *base;
  long long rx, ry;
  int m1, m2, o0, o1, o2;
  if (T.part == 0) {
    base = d.q + P.q;
    rx = P.px;
    ry = P.py;
    m1 = P.mx;
    m2 = P.my;
    o0 = o1 = o2 = 0;
  } else {
    const DPatch &A = d.P[P.parent];
    rx = A.px;
    ry = A.py;
    base = d.q + A.q;
    m1 = P.mx / 2;
    m2 = P.my / 2;
    o0 = P.mb[0][0] - 1;
    o1 = P.mb[1][0] - 1;
    o2 = ND == 3 ? P.mb[2][0] - 1 : 0;
  }
  // rows r0..r1-1 (2D: j; 3D: k) of the box
  const int per = m1 * (ND == 3 ? m2 : 1);
  const int n = per * (T.r1 - T.r0);
  double sq = 0.0, sff = 0.0, sg = 0.0;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const int r = T.r0 + e / per, w = e % per;
    const int i = 1 + w % m1;
    const int j = ND == 3 ? 1 + w / m1 : r;
    const int k = ND == 3 ? r : 0;
    const Split sp = split_of(rx, ry);
    const double *c = base + scell(sp, o0 + i, o1 + j, o2 + k);
    sq += c[0];
    if constexpr (ND == 2) {
      sff += well_ff(c[0], lo, hi, alpha);
      const double cxp = base[scell(sp, o0 + i + 1, o1 + j, 0)];
      const double cxm = base[scell(sp, o0 + i - 1, o1 + j, 0)];
      double a = cxp - c[0], b = c[0] - cxm, u = c[sp.rs] - c[0], v = c[0] - c[-sp.rs];
      sg += a * a + b * b + u * u + v * v;
    }
  }
  double t0 = block_sum(sq, lds), t1 = block_sum(sff, lds), t2 = block_sum(sg, lds);
  if (threadIdx.x == 0) {
    out[3 * blockIdx.x] = t0;
    out[3 * blockIdx.x + 1] = t1;
    out[3 * blockIdx.x + 2] = t2;
  }
}

// InitFields for a grid created by the regrid (lib/bsamroutines.f90:1685-1772,
// mbc = 1), one workgroup per patch of the list: qc(0:cmx+1) = the parent's
// q under the patch with one ring, and q(0:mx+1) = BiLinProlongationP1MC /
// TriLinProlongationP1MC of it (lib/gridutilities.f90:219-279, :588-686):
// linear 3:1 interpolation dimension by dimension (/16 in 2D, /64 in 3D),
// then every interior cell adds its coarse cell's mass correction
// a - (sum of its 2^d children's predictions)/2^d. qold = q.
__device__ __forceinline__ double lin31(double a0, double a1, bool even) {
  return even ? 3.0 * a0 + a1 : a0 + 3.0 * a1;
}

// The parent's cells under a patch: a(i, j, k) is the parent's q at
// (b0 + i, b1 + j, b2 + k) (b the cell one below the patch's footprint).
struct ParentBox {
  const double *q;
  long long px, py;
  int b0, b1, b2;
  __device__ __forceinline__ double operator()(int i, int j, int k) const {
    return q[qcell(px, py, b0 + i, b1 + j, b2 + k)];
  }
};

template <int ND>
__device__ __forceinline__ double p1_pred(const ParentBox &a, int i, int j, int k) {
  const int m = i >> 1, n = j >> 1;
  const bool ei = (i & 1) == 0, ej = (j & 1) == 0;
  if constexpr (ND == 2) {
    const double b0 = lin31(a(m, n, 0), a(m + 1, n, 0), ei);
    const double b1 = lin31(a(m, n + 1, 0), a(m + 1, n + 1, 0), ei);
    return lin31(b0, b1, ej) / 16.0;
  } else {
    const int o = k >> 1;
    const bool ek = (k & 1) == 0;
    double cz[2];
    for (int t = 0; t < 2; t++) {
      const double b0 = lin31(a(m, n, o + t), a(m + 1, n, o + t), ei);
      const double b1 = lin31(a(m, n + 1, o + t), a(m + 1, n + 1, o + t), ei);
      cz[t] = lin31(b0, b1, ej);
    }
    return lin31(cz[0], cz[1], ek) / 64.0;
  }
}

template <int ND>
__global__ __launch_bounds__(256) void k_init_fields(Dev d, const int *patches) {
  const DPatch &P = d.P[patches[blockIdx.x]];
  const DPatch &A = d.P[P.parent];
  const int cx = P.mx / 2, cy = P.my / 2, cz = ND == 3 ? P.mz / 2 : 0;
  for (int v = 0; v < d.nrvars; v++) {
    // a(0,0,0): the parent cell one below the patch's footprint
    const ParentBox a{d.q + A.q + v * A.qvs, A.px, A.py, P.mb[0][0] - 1, P.mb[1][0] - 1,
                      ND == 3 ? P.mb[2][0] - 1 : 0};
    double *qc = d.qc + P.qc + v * P.cvs;
    const int nc = (cx + 2) * (cy + 2) * (ND == 3 ? cz + 2 : 1);
    for (int e = threadIdx.x; e < nc; e += blockDim.x) {
      const int i = e % (cx + 2), j = (e / (cx + 2)) % (cy + 2);
      const int k = ND == 3 ? e / ((cx + 2) * (cy + 2)) : 0;
      qc[qcell(cx + 2, cy + 2, i, j, k)] = a(i, j, k);
    }
    double *q = d.q + P.q + v * P.qvs;
    double *qo = d.qold + P.q + v * P.qvs;
    const int ex = P.mx + 2, ey = P.my + 2, ez = ND == 3 ? P.mz + 2 : 1;
    const int nf = ex * ey * ez;
    for (int e = threadIdx.x; e < nf; e += blockDim.x) {
      const int i = e % ex, j = (e / ex) % ey, k = ND == 3 ? e / (ex * ey) : 0;
      double p = p1_pred<ND>(a, i, j, k);
      const bool inner = i >= 1 && i <= P.mx && j >= 1 && j <= P.my &&
                         (ND == 2 || (k >= 1 && k <= P.mz));
      if (inner) {
        const int ic = (i + 1) >> 1, jc = (j + 1) >> 1, kc = ND == 3 ? (k + 1) >> 1 : 0;
        double sum;
        if constexpr (ND == 2) {
          sum = ((p1_pred<2>(a, 2 * ic, 2 * jc, 0) + p1_pred<2>(a, 2 * ic - 1, 2 * jc, 0)) +
                 p1_pred<2>(a, 2 * ic, 2 * jc - 1, 0)) +
                p1_pred<2>(a, 2 * ic - 1, 2 * jc - 1, 0);
          p = p + (a(ic, jc, 0) - 0.25 * sum);
        } else {
          sum = 0.0;
          const int ci[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
          for (int t = 0; t < 8; t++) {
            double c = p1_pred<3>(a, 2 * ic - ci[t][0], 2 * jc - ci[t][1], 2 * kc - ci[t][2]);
            sum = t == 0 ? c : sum + c;
          }
          p = p + (a(ic, jc, kc) - 0.125 * sum);
        }
      }
      q[qat(P, i, j, k)] = p;
      qo[qat(P, i, j, k)] = p;
    }
  }
}

// Box copies between two q pools (TransferValues' overlaps), one workgroup
// per box.
__global__ __launch_bounds__(256) void k_box_copy(double *dst, const double *src,
                                                  const BoxCopy *boxes) {
  const BoxCopy B = boxes[blockIdx.x];
  const long long n = (long long)B.nx * B.ny * B.nz * B.nv;
  for (long long e = threadIdx.x; e < n; e += blockDim.x) {
    const int i = (int)(e % B.nx);
    const int j = (int)((e / B.nx) % B.ny);
    const int k = (int)((e / ((long long)B.nx * B.ny)) % B.nz);
    const int v = (int)(e / ((long long)B.nx * B.ny * B.nz));
    dst[B.d + qcell(B.dpx, B.dpy, B.di + i, B.dj + j, B.dk + k) + B.dsv * v] =
        src[B.s + qcell(B.spx, B.spy, B.si + i, B.sj + j, B.sk + k) + B.ssv * v];
  }
}

template <int ND>
__global__ __launch_bounds__(256) void k_source(Dev d, const Tile *tiles, int nt, int full) {
  const int b = xcd_tile(nt);
  if (b >= nt) return;
  Tile t = tiles[b];
  const DPatch &P = d.P[t.patch];
  int i, j, k;
  tile_cell_px<ND>(d, t, i, j, k);
  if (!in_box<ND>(i, j, k, P.mx, P.my, P.mz)) return;
  source_cell<ND>(d, P, i, j, k, full);
}

// Linear tiles in device order over the ghost-extended fine box 0..mx+1.
template <int ND>
__global__ __launch_bounds__(256) void k_defect_setup(Dev d, const Tile *tiles, int nt) {
  const int b = xcd_tile(nt);
  if (b >= nt) return;
  Tile t = tiles[b];
  if (stopped_patch(d, t.patch)) return;
  const DPatch &P = d.P[t.patch];
  int i, j, k;
  if (!lin_qcell<ND>(t, P, i, j, k)) return;
  defect_setup_cell<ND>(d, P, i, j, k);
}

template <int ND>
__global__ __launch_bounds__(256) void k_apply_correction(Dev d, const Tile *tiles, int nt) {
  const int b = xcd_tile(nt);
  if (b >= nt) return;
  Tile t = tiles[b];
  if (stopped_patch(d, t.patch)) return;
  const DPatch &P = d.P[t.patch];
  int i, j, k;
  tile_cell_px<ND>(d, t, i, j, k);
  if (!in_box<ND>(i, j, k, P.mx, P.my, P.mz)) return;
  apply_correction_cell<ND>(d, P, i, j, k);
}

// ---- single-workgroup tail: the single-patch coarse levels ------------------
// Below the root every level is one patch with physical boundaries only. The
// tail kernels run the part of the V-cycle (or of FillDown) on levels
// 0..nlev-1 of that chain (coarsest first) in one workgroup, synchronising
// with barriers instead of launches. The V-cycle kernel keeps the arrays a
// level places in LDS (TailLevel::lds, per field set) there for the whole
// cycle: copied in at the start and out at the end; every other array is used
// in place in the global pools. The ghost programs of these levels copy
// within their own patch (patch-local offsets). Per visit of a level, each
// thread loads its ghost updates and its cells' right-hand sides into
// registers once, so the barrier-separated phases touch only LDS, registers
// and the level arrays.
extern __shared__ __align__(16) unsigned char tail_lds[];

constexpr int TAIL_GR = 1;   // ghost updates per thread and pass held in registers
constexpr int TAIL_FR = 2;   // cells per thread and colour whose f is held in registers

// ceil(log2(x)) for x >= 1 (uniform values: scalar work).
__device__ __forceinline__ int clog2(int x) { return x <= 1 ? 0 : 32 - __builtin_clz(x - 1); }

// Calls f(i, j, k) for the cells of an ex x ey (x ez) box (lower corner
// `base`), spread over the workgroup. The box is walked as if padded to
// powers of two in x and y, so a flat index maps to (i, j, k) with shifts
// and masks instead of integer divisions; padding lanes do nothing.
template <int ND, class F>
__device__ __forceinline__ void for_cells(int ex, int ey, int ez, int base, F f) {
  const int lx = clog2(ex), ly = clog2(ey);
  const int n = (ez << (lx + ly));
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    int i = e & ((1 << lx) - 1), j = (e >> lx) & ((1 << ly) - 1), k = e >> (lx + ly);
    if (i < ex && j < ey) f(i + base, j + base, ND == 3 ? k + base : 0);
  }
}

// A tail level's arrays. With IN_LDS the level's q and qc (and the FP32 base
// point) sit in LDS and are addressed from the LDS base itself, so the
// compiler emits LDS instructions for them; its f may sit in LDS or in the
// global pool. Without, every array is used in place in the global pools.
template <class T, bool IN_LDS>
__device__ __forceinline__ Lvl<T> tail_ref(const Dev &d, const TailLevel &t, int set) {
  Fld<T> F = fields<T>(d);
  const DPatch &P = t.P[set];
  Lvl<T> L{F.q + P.q, F.qc + P.qc, F.f + P.f, F.qb ? F.qb + P.q : nullptr, P};
  if constexpr (IN_LDS) {
    const int *o = t.lds[set];
    L.q = (T *)(tail_lds + o[0]);
    L.qc = (T *)(tail_lds + o[1]);
    if (o[2] >= 0) L.f = (T *)(tail_lds + o[2]);
    if (std::is_same<T, float>::value) L.qb = (const T *)(tail_lds + o[3]);
  }
  return L;
}

// The level table stays in global memory; the tail kernels take it as a
// restrict pointer, so the compiler reads it with scalar loads, which do not
// wait behind the vector loads in flight.
constexpr int TAIL_MAXLEV = 24;

// Moves arrays between the global pools and LDS as listed by a host-built
// table of 64-element chunks (the tail levels' LDS-resident arrays: q, qc, f,
// qb in; q, qc, f out). Each wave takes every nw-th chunk, one element per
// lane, U chunks per round: lane u < U reads the round's u-th table entry
// with one vector load, and readlane hands each entry to the whole wave as a
// uniform base. The data loads are branch-free (lane index clamped into the
// chunk) so that they issue back to back; only the stores are predicated.
// Staging in ends with a barrier.
template <class T, bool in>
__device__ void tail_stage(const TailCopy *tab, int nchunks) {
  constexpr int U = 16;
  static_assert(sizeof(TailCopy) == 16, "one int4 per table entry");
  typedef __attribute__((address_space(1))) T GT;
  typedef __attribute__((address_space(3))) T LT;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  LT *lds = (LT *)tail_lds;
  for (int c0 = wave; c0 < nchunks; c0 += U * nw) {
    const __attribute__((address_space(1))) int *ent =
        (const __attribute__((address_space(1))) int *)tab +
        4 * min(c0 + (lane & (U - 1)) * nw, nchunks - 1);
    const int4 e = make_int4(ent[0], ent[1], ent[2], ent[3]);
    T v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const unsigned long long g = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane(e.y, u)
                                    << 32) |
                                   (unsigned)__builtin_amdgcn_readlane(e.x, u);
      const int o = __builtin_amdgcn_readlane(e.z, u) / (int)sizeof(T);
      const int k = max(0, min(lane, __builtin_amdgcn_readlane(e.w, u) - 1));
      if constexpr (in) v[u] = ((const GT *)g)[k];
      else v[u] = lds[o + k];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const unsigned long long g = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane(e.y, u)
                                    << 32) |
                                   (unsigned)__builtin_amdgcn_readlane(e.x, u);
      const int o = __builtin_amdgcn_readlane(e.z, u) / (int)sizeof(T);
      const int n = c0 + u * nw < nchunks ? __builtin_amdgcn_readlane(e.w, u) : 0;
      if (lane < n) {
        if constexpr (in) lds[o + lane] = v[u];
        else ((GT *)g)[lane] = v[u];
      }
    }
  }
  if constexpr (in) __syncthreads();
}

// One SetGhost program of a tail level, this thread's share of its updates
// held in registers as loaded (destination, and source | nflip byte << 26, which
// the host stores in f1), so that the loads stay in flight until the program
// is applied.
struct TailPass {
  int n;
  const GhostOp *ops;
  int dst[TAIL_GR], srcf[TAIL_GR];
};
struct TailGhost {
  TailPass p0, p1, p2;   // ipass 0, 1, 2
};
// The ghost programs live in global memory; loading through a global-space
// pointer keeps the loads off the LDS wait counter.
typedef const __attribute__((address_space(1))) GhostOp *GlobalOps;
__device__ __forceinline__ void tail_pass_load(int n, const GhostOp *ops, TailPass &g) {
  g.n = n;
  g.ops = ops;
  GlobalOps gops = (GlobalOps)ops;
#pragma unroll
  for (int r = 0; r < TAIL_GR; r++) {
    int e = threadIdx.x + r * blockDim.x;
    g.dst[r] = g.srcf[r] = 0;
    if (e < n) {
      g.dst[r] = gops[e].dst;
      g.srcf[r] = gops[e].f1;
    }
  }
}
__device__ __forceinline__ void tail_ghost_load(const TailLevel &t, TailGhost &g) {
  tail_pass_load(t.ngp[0], t.gp[0], g.p0);
  tail_pass_load(t.ngp[1], t.gp[1], g.p1);
  tail_pass_load(t.ngp[2], t.gp[2], g.p2);
}
// SetGhost on a tail level: patch-local copies with "2 - x" maps and
// quasi-periodic offsets (the ghost_op kind 0 update).
template <class T, bool DEF, bool QPO>
__device__ __forceinline__ void tail_ghost(int nrvars, const Lvl<T> &L, const TailPass &g,
                                           Qpo qpo) {
  const long long vs = L.P.qvs;
#pragma unroll
  for (int r = 0; r < TAIL_GR; r++) {
    if ((int)(threadIdx.x + r * blockDim.x) < g.n) {
      const int src = g.srcf[r] & 0x03ffffff, fb = (int)((unsigned)g.srcf[r] >> 26);
      for (int v = 0; v < nrvars; v++) {
        auto sv = [&](int) { return L.q[src + v * vs]; };
        const T val = ghost_value<T, DEF, decltype(sv), QPO>(0, fb, sv, qpo, v);
        L.q[g.dst[r] + v * vs] = val;
      }
    }
  }
  for (int e = threadIdx.x + TAIL_GR * blockDim.x; e < g.n; e += blockDim.x)
    ghost_op<T, DEF, QPO>(nrvars, L.q, vs, g.ops[e], qpo);
  __syncthreads();
}

// Cells of one colour on a level, colour-major: each row holds hx =
// ceil(mx/2) cells of the colour, i = i0 + 2t with i0 chosen so that
// i + j (+ k) + color is odd (the cells RelaxGrid updates in that pass). Rows
// and t are padded to powers of two (lh = log2 of the padded hx, ly of the
// padded my) so that a flat index e < ncells maps with shifts and masks;
// color_cell returns false for padding and past the row's end.
struct ColorMap {
  int lh, ly, n;
};
template <int ND>
__device__ __forceinline__ ColorMap color_map(const DPatch &P) {
  ColorMap m;
  m.lh = clog2((P.mx + 1) >> 1);
  m.ly = clog2(P.my);
  m.n = (ND == 3 ? P.mz : 1) << (m.lh + m.ly);
  return m;
}
template <int ND>
__device__ __forceinline__ bool color_cell(const DPatch &P, const ColorMap &m, int color, int e,
                                           int &i, int &j, int &k) {
  int t = e & ((1 << m.lh) - 1);
  j = ((e >> m.lh) & ((1 << m.ly) - 1)) + 1;
  k = ND == 3 ? (e >> (m.lh + m.ly)) + 1 : 0;
  i = 1 + ((j + k + color) & 1) + 2 * t;
  return i <= P.mx && j <= P.my;
}

// Element i of an array in LDS (lds) or in global memory, read with an LDS or
// a global load: a flat load would also wait for the global loads in flight.
template <class T>
__device__ __forceinline__ T tail_load(const T *p, long long i, bool lds) {
  if (lds) return ((const __attribute__((address_space(3))) T *)p)[i];
  return ((const __attribute__((address_space(1))) T *)p)[i];
}

// nsmooth red-black passes on a tail level. Threads walk the cells of each
// colour; the first TAIL_FR cells per thread and colour keep their f (in LDS
// when f_lds) in registers for the whole visit.
template <int ND, class T, bool DEF, bool QPO>
__device__ void tail_relax(const Problem &prm, int nrvars, const Lvl<T> &L, bool f_lds,
                           const TailGhost &g, int nsmooth) {
  const DPatch &P = L.P;
  const ColorMap m = color_map<ND>(P);
  const int n = m.n;
  T f1[2][TAIL_FR], f2[2][TAIL_FR];
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int r = 0; r < TAIL_FR; r++) {
      int e = threadIdx.x + r * blockDim.x, i, j, k;
      f1[c][r] = f2[c][r] = (T)0;
      if (e < n && color_cell<ND>(P, m, c + 1, e, i, j, k)) {
        long long fo = fat(P, i, j, ND == 3 ? k : 1);
        f1[c][r] = tail_load(L.f, fo, f_lds);
        f2[c][r] = tail_load(L.f, fo + P.fvs, f_lds);
      }
    }
  auto half = [&](int c) {
#pragma unroll
    for (int r = 0; r < TAIL_FR; r++) {
      int e = threadIdx.x + r * blockDim.x, i, j, k;
      if (e < n && color_cell<ND>(P, m, c + 1, e, i, j, k))
        relax_cell_f<ND, T, DEF>(prm, L, i, j, k, f1[c][r], f2[c][r]);
      __builtin_amdgcn_sched_barrier(0);
    }
    for (int e = threadIdx.x + TAIL_FR * blockDim.x; e < n; e += blockDim.x) {
      int i, j, k;
      if (color_cell<ND>(P, m, c + 1, e, i, j, k)) relax_cell<ND, T, DEF>(prm, L, i, j, k);
    }
    __syncthreads();
  };
  for (int pass = 0; pass < nsmooth; pass++) {
    half(0);
    tail_ghost<T, DEF, QPO>(nrvars, L, g.p1, qpo_of(prm));
    half(1);
    tail_ghost<T, DEF, QPO>(nrvars, L, g.p2, qpo_of(prm));
  }
}

// The tail's visit of level C = lv[l] on the way down: smoothing, then
// RestrictSolution, SetGhost(parent, 0), GetCoarseGhostPoints and
// CoarseLoadingFunction into the parent PA = lv[l - 1].
template <int ND, class T, bool DEF, bool QPO, bool LC, bool LP>
__device__ void tail_down(const Dev &d, const TailLevel *__restrict__ lv, int l, int set,
                          int nsmooth, const TailGhost &gc, const TailPass &gp) {
  const int nv = d.nrvars;
  const Problem prm = d.prm;
  Lvl<T> C = tail_ref<T, LC>(d, lv[l], set), PA = tail_ref<T, LP>(d, lv[l - 1], set);
  const DPatch &P = C.P;
  int cz = ND == 3 ? P.cmz : 1;
  tail_relax<ND, T, DEF, QPO>(prm, nv, C, LC && lv[l].lds[set][2] >= 0, gc, nsmooth);
  for_cells<ND>(P.cmx, P.cmy, cz, 1,
                [&](int i, int j, int k) { restrict_cell<ND, T>(nv, C, PA, i, j, k); });
  __syncthreads();
  tail_ghost<T, DEF, QPO>(nv, PA, gp, qpo_of(d.prm));
  for_cells<ND>(P.cmx + 2, P.cmy + 2, ND == 3 ? P.cmz + 2 : 1, 0,
                [&](int i, int j, int k) { coarse_ghost_cell<ND, T>(nv, C, PA, i, j, k); });
  __syncthreads();
  for_cells<ND>(P.cmx, P.cmy, cz, 1,
                [&](int i, int j, int k) { coarse_load_cell<ND, T, DEF>(prm, C, PA, i, j, k); });
  __syncthreads();
}

// On the way up: CorrectFine from the parent, SetGhost(C, 0), smoothing.
template <int ND, class T, bool DEF, bool QPO, bool LC, bool LP>
__device__ void tail_up(const Dev &d, const TailLevel *__restrict__ lv, int l, int set,
                        int nsmooth, const TailGhost &gc) {
  const int nv = d.nrvars;
  Lvl<T> C = tail_ref<T, LC>(d, lv[l], set), PA = tail_ref<T, LP>(d, lv[l - 1], set);
  const DPatch &P = C.P;
  for_cells<ND>(P.cmx + 1, P.cmy + 1, ND == 3 ? P.cmz + 1 : 1, 0,
                [&](int m, int n, int o) { correct_block<ND, T>(nv, C, PA, m, n, o); });
  __syncthreads();
  tail_ghost<T, DEF, QPO>(nv, C, gc.p0, qpo_of(d.prm));
  tail_relax<ND, T, DEF, QPO>(d.prm, nv, C, LC && lv[l].lds[set][2] >= 0, gc, nsmooth);
}

template <int ND, class T, bool DEF, bool QPO, bool LB>
__device__ void tail_bottom(const Dev &d, const TailLevel *__restrict__ lv, int set, int nsmooth,
                            const TailGhost &gb) {
  Lvl<T> B = tail_ref<T, LB>(d, lv[0], set);
  tail_relax<ND, T, DEF, QPO>(d.prm, d.nrvars, B, LB && lv[0].lds[set][2] >= 0, gb, nsmooth);
}

// AFASVcycle(top) (afasroutines.f90:488-521) on levels 0..nlev-1 of the
// single-patch chain. Levels holding their q in LDS form a prefix of the
// chain (coarsest first), so a visit pairs LDS/LDS, global/LDS or
// global/global child and parent.
// QPO: the chain's SetGhost programs hold periodic offsets (without them
// the offset code is left out, which keeps the kernel within its registers).
template <int ND, class T, bool DEF, bool QPO>
__global__ __launch_bounds__(1024) void k_tail_vcycle(Dev d, const TailLevel *__restrict__ lv,
                                                       int nlev, int nsmooth, int lds,
                                                       const TailCopy *cin, int nin,
                                                       const TailCopy *cout, int nout) {
  const int set = DEF ? 1 : 0;
  if (stopped_patch(d, lv[0].patch)) return;   // one chain, one group
  auto in_lds = [&](int l) { return lds && lv[l].lds[set][0] >= 0; };
  // each level's ghost updates are loaded one visit ahead, so their latency
  // overlaps the previous level's work
  TailGhost gcur, gnext;
  tail_ghost_load(lv[nlev - 1], gcur);
  if (lds) tail_stage<T, true>(cin, nin);
  for (int l = nlev - 1; l >= 1; l--) {
    tail_ghost_load(lv[l - 1], gnext);
    if (in_lds(l)) tail_down<ND, T, DEF, QPO, true, true>(d, lv, l, set, nsmooth, gcur, gnext.p0);
    else if (in_lds(l - 1))
      tail_down<ND, T, DEF, QPO, false, true>(d, lv, l, set, nsmooth, gcur, gnext.p0);
    else tail_down<ND, T, DEF, QPO, false, false>(d, lv, l, set, nsmooth, gcur, gnext.p0);
    gcur = gnext;
  }
  if (nlev > 1) tail_ghost_load(lv[1], gnext);
  if (in_lds(0)) tail_bottom<ND, T, DEF, QPO, true>(d, lv, set, nsmooth, gcur);
  else tail_bottom<ND, T, DEF, QPO, false>(d, lv, set, nsmooth, gcur);
  for (int l = 1; l < nlev; l++) {
    gcur = gnext;
    if (l + 1 < nlev) tail_ghost_load(lv[l + 1], gnext);
    if (in_lds(l)) tail_up<ND, T, DEF, QPO, true, true>(d, lv, l, set, nsmooth, gcur);
    else if (in_lds(l - 1)) tail_up<ND, T, DEF, QPO, false, true>(d, lv, l, set, nsmooth, gcur);
    else tail_up<ND, T, DEF, QPO, false, false>(d, lv, l, set, nsmooth, gcur);
  }
  if (lds) tail_stage<T, false>(cout, nout);
}

// FillDown (afasroutines.f90:223-239) from level nlev-1 down to 0 of the
// chain, in the global pools: restriction into the parent, and for q a
// SetGhost(parent, 0).
template <int ND>
__global__ __launch_bounds__(1024) void k_tail_filldown(Dev d, const TailLevel *lv, int nlev,
                                                         int field) {
  const int nv = d.nrvars;
  if (stopped_patch(d, lv[0].patch)) return;
  for (int l = nlev - 1; l >= 1; l--) {
    Lvl<double> C = level_ref<double>(d, lv[l].patch), PA = level_ref<double>(d, lv[l - 1].patch);
    TailPass gp;
    tail_pass_load(lv[l - 1].ngp[0], lv[l - 1].gp[0], gp);
    const DPatch &P = C.P;
    for_cells<ND>(P.cmx, P.cmy, ND == 3 ? P.cmz : 1, 1, [&](int i, int j, int k) {
      filldown_cell<ND, double>(nv, C, PA, i, j, k, field);
    });
    __syncthreads();
    if (field == 2) tail_ghost<double, false, true>(nv, PA, gp, qpo_of(d.prm));
  }
}

// ---- norms -------------------------------------------------------------------
// Fixed-order (deterministic) workgroup sum; the total is valid in thread 0.
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

// Per-tile sum of squares of the residual f - A(q) (Residual,
// afasroutines.f90:799-840; what = 0) or of f itself (what = 1).
template <int ND>
__global__ __launch_bounds__(256) void k_resnorm(Dev d, const Tile *tiles, int nt, double *partial,
                                                  int what, int comps) {
  const int b = xcd_tile(nt);
  if (b >= nt) return;
  Tile t = tiles[b];
  const DPatch &P = d.P[t.patch];
  int i, j, k;
  tile_cell_px<ND>(d, t, i, j, k);
  double s = 0.0, s2 = 0.0;
  if (in_box<ND>(i, j, k, P.mx, P.my, P.mz)) {
    long long fo = P.f + fat(P, i, j, ND == 3 ? k : 1);
    double r1 = d.f[fo], r2 = d.f[fo + P.fvs];
    if (what == 0) {
      Lvl<double> L{d.q + P.q, nullptr, nullptr, nullptr, P};
      View<double> a = qview(L);
      double o1, o2;
      op_at<ND, double, false>(d.prm, a, 0.0, P, i, j, k, o1, o2);
      r1 = r1 - o1;
      r2 = r2 - o2;
    }
    s = comps == 2 ? r1 * r1 : r1 * r1 + r2 * r2;
    s2 = r2 * r2;
  }
  double tot = block_sum(s);
  if (threadIdx.x == 0) partial[b] = tot;
  if (comps == 2) {   // errortype 2: the second component's sum after the first's
    __syncthreads();
    tot = block_sum(s2);
    if (threadIdx.x == 0) partial[nt + b] = tot;
  }
}

// Sum the per-tile partials in index order (strided, then a fixed tree) and
// finish ErrorAFAS: sqrt(sum / count) into hist[state[0]] (errortype 1), or
// with comps 2 the larger of the two components' sqrt(sum / count), the
// second component's partials following the first's (errortype 2,
// lib/afasroutines.f90:128-142; a NaN in either wins); state[0] counts the
// calls, state[1] is set once a residual falls below tol and state[2] once a
// residual is not finite, state[3] (stopped) once either is. One workgroup
// per group: group g (blockIdx.x) sums its tiles granges[2g] ..
// granges[2g] + granges[2g+1] - 1 of the n partials with count gcount[g], in
// the order a solve of that group alone sums its own, and writes
// hist[g * hstride + state[4g]] and state[4g .. 4g+3]; without granges the
// single group takes all n tiles and `count`. A group already stopped keeps
// its history and state.
__global__ __launch_bounds__(1024) void k_finalize(const double *partial, int n, double count,
                                                   const int *granges, const double *gcount,
                                                   double tol, double *hist, int hstride,
                                                   int *state, int comps) {
  const int g = blockIdx.x;
  int *st = state + 4 * g;
  if (st[3]) return;
  const int t0 = granges ? granges[2 * g] : 0, m = granges ? granges[2 * g + 1] : n;
  const double cnt = granges ? gcount[g] : count;
  double s = 0.0;
  for (int e = threadIdx.x; e < m; e += blockDim.x) s += partial[t0 + e];
  double tot = block_sum(s);
  double tot2 = 0.0;
  if (comps == 2) {
    __syncthreads();
    s = 0.0;
    for (int e = threadIdx.x; e < m; e += blockDim.x) s += partial[n + t0 + e];
    tot2 = block_sum(s);
  }
  if (threadIdx.x == 0) {
    double r = sqrt(tot / cnt);
    if (comps == 2) {
      const double r2 = sqrt(tot2 / cnt);
      r = (isnan(r) || r >= r2) ? r : r2;
    }
    int it = st[0];
    hist[(long long)g * hstride + it] = r;
    st[0] = it + 1;
    if (r < tol) st[1] = 1;
    if (!isfinite(r)) st[2] = 1;
    if (st[1] || st[2]) st[3] = 1;
  }
}

__global__ __launch_bounds__(256) void k_copy(double *dst, const double *src, long long n) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; e < n; e += stride) dst[e] = src[e];
}

// ---- launchers -------------------------------------------------------------

#define BSAM_LAUNCH_ND_PREC(KERNEL, nd, prec, grid, block, ...)                        \
  do {                                                                                 \
    if (prec == PREC_SOLUTION_FP64) {                                                  \
      if (nd == 2) KERNEL<2, double, false><<<grid, block, 0, s>>>(__VA_ARGS__);       \
      else KERNEL<3, double, false><<<grid, block, 0, s>>>(__VA_ARGS__);               \
    } else {                                                                           \
      if (nd == 2) KERNEL<2, float, true><<<grid, block, 0, s>>>(__VA_ARGS__);         \
      else KERNEL<3, float, true><<<grid, block, 0, s>>>(__VA_ARGS__);                 \
    }                                                                                  \
  } while (0)
#define BSAM_LAUNCH_ND_T(KERNEL, nd, prec, grid, block, ...)                           \
  do {                                                                                 \
    if (prec == PREC_SOLUTION_FP64) {                                                  \
      if (nd == 2) KERNEL<2, double><<<grid, block, 0, s>>>(__VA_ARGS__);              \
      else KERNEL<3, double><<<grid, block, 0, s>>>(__VA_ARGS__);                      \
    } else {                                                                           \
      if (nd == 2) KERNEL<2, float><<<grid, block, 0, s>>>(__VA_ARGS__);               \
      else KERNEL<3, float><<<grid, block, 0, s>>>(__VA_ARGS__);                       \
    }                                                                                  \
  } while (0)
#define BSAM_LAUNCH_ND(KERNEL, nd, grid, block, ...)                                   \
  do {                                                                                 \
    if (nd == 2) KERNEL<2><<<grid, block, 0, s>>>(__VA_ARGS__);                        \
    else KERNEL<3><<<grid, block, 0, s>>>(__VA_ARGS__);                                \
  } while (0)

hipError_t launch_relax(int nd, int prec, const Dev &d, const Tile *tiles, int nt, int ipass,
                        hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (d.tile_lx < 1) return hipErrorInvalidValue;
  BSAM_LAUNCH_ND_PREC(k_relax, nd, prec, xcd_grid(nt), RELAX_THREADS, d, tiles, nt, ipass);
  return hipGetLastError();
}

hipError_t launch_relax_fused(int nd, int prec, const Dev &d, const Tile *tiles, int nt,
                              int color, const GhostList &pre, const GhostList &post,
                              bool stage, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (d.tile_lx < 1) return hipErrorInvalidValue;
  const dim3 grid(xcd_grid(nt));
  const int tx = 1 << d.tile_lx, ty = 1 << d.tile_ly;
  const int tz = nd == 3 ? 256 >> (d.tile_lx + d.tile_ly) : 1;
  const size_t cells = (size_t)(tx + 2) * (ty + 2) * (nd == 3 ? tz + 2 : 1);
  const size_t lds = 2 * cells * (prec == PREC_SOLUTION_FP64 ? 8 : 4);
  if (prec == PREC_SOLUTION_FP64) {
    if (stage) {
      if (nd == 2)
        k_relax_fused<2, double, false, true><<<grid, RELAX_THREADS, lds, s>>>(
            d, tiles, nt, color, pre, post);
      else
        k_relax_fused<3, double, false, true><<<grid, RELAX_THREADS, lds, s>>>(
            d, tiles, nt, color, pre, post);
    } else {
      if (nd == 2)
        k_relax_fused<2, double, false, false><<<grid, RELAX_THREADS, 0, s>>>(
            d, tiles, nt, color, pre, post);
      else
        k_relax_fused<3, double, false, false><<<grid, RELAX_THREADS, 0, s>>>(
            d, tiles, nt, color, pre, post);
    }
  } else {
    if (stage) {
      if (nd == 2)
        k_relax_fused<2, float, true, true><<<grid, RELAX_THREADS, lds, s>>>(
            d, tiles, nt, color, pre, post);
      else
        k_relax_fused<3, float, true, true><<<grid, RELAX_THREADS, lds, s>>>(
            d, tiles, nt, color, pre, post);
    } else {
      if (nd == 2)
        k_relax_fused<2, float, true, false><<<grid, RELAX_THREADS, 0, s>>>(
            d, tiles, nt, color, pre, post);
      else
        k_relax_fused<3, float, true, false><<<grid, RELAX_THREADS, 0, s>>>(
            d, tiles, nt, color, pre, post);
    }
  }
  return hipGetLastError();
}

hipError_t launch_ghost(int prec, const Dev &d, const GhostList &ops, hipStream_t s) {
  const int n = ops.ns + ops.nl;
  if (n <= 0) return hipSuccess;
  dim3 g((n + 255) / 256);
  if (prec == PREC_SOLUTION_FP64) k_ghost<double, false><<<g, 256, 0, s>>>(d, ops);
  else k_ghost<float, true><<<g, 256, 0, s>>>(d, ops);
  return hipGetLastError();
}

hipError_t launch_restrict(int nd, int prec, const Dev &d, const Tile *tiles, int nt,
                           const GhostList &gops, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  BSAM_LAUNCH_ND_T(k_restrict, nd, prec, xcd_grid(nt), 256, d, tiles, nt, gops);
  return hipGetLastError();
}

hipError_t launch_filldown(int nd, const Dev &d, const Tile *tiles, int nt, int field,
                           const GhostList &gops, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (nd == 2) k_filldown<2, double><<<xcd_grid(nt), 256, 0, s>>>(d, tiles, nt, field, gops);
  else k_filldown<3, double><<<xcd_grid(nt), 256, 0, s>>>(d, tiles, nt, field, gops);
  return hipGetLastError();
}

hipError_t launch_coarse_ghost(int nd, int prec, const Dev &d, const Tile *tiles, int nt,
                               hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  BSAM_LAUNCH_ND_T(k_coarse_ghost, nd, prec, xcd_grid(nt), 256, d, tiles, nt);
  return hipGetLastError();
}

hipError_t launch_coarse_load(int nd, int prec, const Dev &d, const Tile *tiles, int nt,
                              hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  BSAM_LAUNCH_ND_PREC(k_coarse_load, nd, prec, xcd_grid(nt), 256, d, tiles, nt);
  return hipGetLastError();
}

hipError_t launch_coarse_ghost_load(int nd, int prec, const Dev &d, const Tile *tiles, int nt,
                                   int nres, const GhostList &rest, const GhostList &tops,
                                   hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (nres < 0)
    BSAM_LAUNCH_ND_PREC(k_coarse_ghost_load_lin, nd, prec, xcd_grid(nt), 256, d, tiles, nt, rest,
                        tops);
  else
    BSAM_LAUNCH_ND_PREC(k_coarse_ghost_load, nd, prec, xcd_grid(nt), 256, d, tiles, nt, nres,
                        rest, tops);
  return hipGetLastError();
}

hipError_t launch_correct(int nd, int prec, const Dev &d, const Tile *tiles, int nt,
                          const GhostList &rest, const GhostList &tops, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  BSAM_LAUNCH_ND_T(k_correct, nd, prec, xcd_grid(nt), 256, d, tiles, nt, rest, tops);
  return hipGetLastError();
}

hipError_t launch_source(int nd, const Dev &d, const Tile *tiles, int nt, int full,
                         hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  BSAM_LAUNCH_ND(k_source, nd, xcd_grid(nt), 256, d, tiles, nt, full);
  return hipGetLastError();
}

hipError_t launch_trunc_error(int nd, const Dev &d, const Tile *tiles, int nt, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  BSAM_LAUNCH_ND(k_trunc_error, nd, xcd_grid(nt), 256, d, tiles, nt);
  return hipGetLastError();
}

hipError_t launch_error_flags(int nd, const Dev &d, const Tile *tiles, int nt, double tol,
                              hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  BSAM_LAUNCH_ND(k_error_flags, nd, xcd_grid(nt), 256, d, tiles, nt, tol);
  return hipGetLastError();
}

hipError_t launch_error_flags_user(int nd, const Dev &d, const Tile *tiles, int nt,
                                   double qtol, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  BSAM_LAUNCH_ND(k_error_flags_user, nd, xcd_grid(nt), 256, d, tiles, nt, qtol);
  return hipGetLastError();
}

hipError_t launch_after_step(int nd, const Dev &d, const StatTask *tasks, int n, double lo,
                             double hi, double *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  BSAM_LAUNCH_ND(k_after_step, nd, dim3((unsigned)n), 256, d, tasks, lo, hi, out);
  return hipGetLastError();
}

hipError_t launch_init_fields(int nd, const Dev &d, const int *patches, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  BSAM_LAUNCH_ND(k_init_fields, nd, dim3((unsigned)n), 256, d, patches);
  return hipGetLastError();
}

hipError_t launch_box_copy(double *dst, const double *src, const BoxCopy *boxes, int n,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_box_copy<<<dim3((unsigned)n), 256, 0, s>>>(dst, src, boxes);
  return hipGetLastError();
}

hipError_t launch_defect_setup(int nd, const Dev &d, const Tile *tiles, int nt,
                               hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  BSAM_LAUNCH_ND(k_defect_setup, nd, xcd_grid(nt), 256, d, tiles, nt);
  return hipGetLastError();
}

hipError_t launch_apply_correction(int nd, const Dev &d, const Tile *tiles, int nt,
                                   hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  BSAM_LAUNCH_ND(k_apply_correction, nd, xcd_grid(nt), 256, d, tiles, nt);
  return hipGetLastError();
}

hipError_t launch_resnorm(int nd, const Dev &d, const Tile *tiles, int nt, double *partial,
                          int what, int comps, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  BSAM_LAUNCH_ND(k_resnorm, nd, xcd_grid(nt), 256, d, tiles, nt, partial, what, comps);
  return hipGetLastError();
}

hipError_t launch_finalize(const double *partial, int n, double count, double tol,
                           double *hist, int *state, int comps, hipStream_t s,
                           int ngroups, const int *granges, const double *gcount, int hstride) {
  // one workgroup reads all partials of its group: a full-size group keeps
  // 16 waves of loads in flight
  k_finalize<<<ngroups, 1024, 0, s>>>(partial, n, count, granges, gcount, tol, hist, hstride,
                                      state, comps);
  return hipGetLastError();
}

hipError_t launch_copy(double *dst, const double *src, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_copy<<<dim3((unsigned)blocks), 256, 0, s>>>(dst, src, n);
  return hipGetLastError();
}

hipError_t launch_tail_vcycle(int nd, int prec, const Dev &d, const TailLevel *lv, int nlev,
                              int nsmooth, int threads, size_t lds, const TailCopy *cin,
                              int nin, const TailCopy *cout, int nout, bool qpo, hipStream_t s) {
  if (nlev <= 0) return hipSuccess;
  if (nlev > TAIL_MAXLEV) return hipErrorInvalidValue;
  int use = lds > 0;
  if (prec == PREC_SOLUTION_FP64) {
    if (nd == 2) {
      if (qpo)
        k_tail_vcycle<2, double, false, true><<<1, threads, lds, s>>>(d, lv, nlev, nsmooth, use,
                                                                     cin, nin, cout, nout);
      else
        k_tail_vcycle<2, double, false, false><<<1, threads, lds, s>>>(d, lv, nlev, nsmooth, use,
                                                                      cin, nin, cout, nout);
    } else {
      if (qpo)
        k_tail_vcycle<3, double, false, true><<<1, threads, lds, s>>>(d, lv, nlev, nsmooth, use,
                                                                     cin, nin, cout, nout);
      else
        k_tail_vcycle<3, double, false, false><<<1, threads, lds, s>>>(d, lv, nlev, nsmooth, use,
                                                                      cin, nin, cout, nout);
    }
  } else {
    if (nd == 2)
      k_tail_vcycle<2, float, true, false><<<1, threads, lds, s>>>(d, lv, nlev, nsmooth, use, cin,
                                                                  nin, cout, nout);
    else
      k_tail_vcycle<3, float, true, false><<<1, threads, lds, s>>>(d, lv, nlev, nsmooth, use, cin,
                                                                  nin, cout, nout);
  }
  return hipGetLastError();
}

int tail_max_levels() { return TAIL_MAXLEV; }
size_t tail_static_lds() { return 0; }

hipError_t launch_tail_filldown(int nd, const Dev &d, const TailLevel *lv, int nlev, int field,
                                int threads, hipStream_t s) {
  if (nlev <= 1) return hipSuccess;
  if (nd == 2) k_tail_filldown<2><<<1, threads, 0, s>>>(d, lv, nlev, field);
  else k_tail_filldown<3><<<1, threads, 0, s>>>(d, lv, nlev, field);
  return hipGetLastError();
}

hipError_t launch_smooth_iso(int nd, int prec, const Dev &d, const IsoPatch *iso, int n,
                             const int *stage, const IsoRec *rec,
                             const unsigned long long *srec, int npasses, int ghost0_first,
                             int final2, size_t lds, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 grid((unsigned)n);
  if (prec == PREC_SOLUTION_FP64) {
    if (nd == 2)
      k_smooth_iso<2, double, false><<<grid, ISO_THREADS, lds, s>>>(d, iso, stage, rec, srec, npasses,
                                                                    ghost0_first, final2);
    else
      k_smooth_iso<3, double, false><<<grid, ISO_THREADS, lds, s>>>(d, iso, stage, rec, srec, npasses,
                                                                    ghost0_first, final2);
  } else {
    if (nd == 2)
      k_smooth_iso<2, float, true><<<grid, ISO_THREADS, lds, s>>>(d, iso, stage, rec, srec, npasses,
                                                                  ghost0_first, final2);
    else
      k_smooth_iso<3, float, true><<<grid, ISO_THREADS, lds, s>>>(d, iso, stage, rec, srec, npasses,
                                                                  ghost0_first, final2);
  }
  return hipGetLastError();
}

hipError_t iso_allow_lds(size_t lds) {
  const void *fns[] = {
      (const void *)k_smooth_iso<2, double, false>, (const void *)k_smooth_iso<3, double, false>,
      (const void *)k_smooth_iso<2, float, true>, (const void *)k_smooth_iso<3, float, true>};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t tail_allow_lds(size_t lds) {
  const void *fns[] = {
      (const void *)k_tail_vcycle<2, double, false, true>,
      (const void *)k_tail_vcycle<3, double, false, true>,
      (const void *)k_tail_vcycle<2, double, false, false>,
      (const void *)k_tail_vcycle<3, double, false, false>,
      (const void *)k_tail_vcycle<2, float, true, false>,
      (const void *)k_tail_vcycle<3, float, true, false>};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}


}  // namespace bsam
