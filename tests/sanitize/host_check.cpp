// Host-side check of the engine library under AddressSanitizer and
// UndefinedBehaviorSanitizer (test infrastructure): host-only contexts
// (device -1) commit random hierarchies -- 2D and 3D, Neumann, user and
// periodic boundaries, siblings touching and not -- which runs every host
// planner (tiles, ghost-fill programs, fusion and riding plans, tail
// placement); the host ghost programs are applied to host arrays; the
// regrid clustering runs on random flag fields; the frame formatter on
// random values. Exits 0 if everything returned normally.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "bsam_gpu.h"

namespace {

struct P {
  int level, parent, mx[3], mb[6], mg[6], bc[6];
  double dx, xl[3];
};

std::mt19937 rng(12345);
int rnd(int lo, int hi) { return std::uniform_int_distribution<int>(lo, hi)(rng); }

// a root of mx, its below-root chain to 1 cell, then random children on
// levels 1..2 (non-overlapping within a level, by construction on a grid of
// candidate blocks)
std::vector<P> hierarchy(int nd, int root, const int bc[6], int &minlevel) {
  std::vector<P> ps;
  int m = root, lev = 0;
  while (m % 2 == 0 && lev < 10) { m /= 2; lev++; }
  minlevel = -lev;
  for (int l = minlevel; l <= 0; l++) {
    P p{};
    int s = root >> (0 - l);
    p.level = l;
    p.parent = l == minlevel ? -1 : (int)ps.size() - 1;
    for (int d = 0; d < 3; d++) {
      p.mx[d] = d < nd ? s : 1;
      p.mb[2 * d] = 1;
      p.mb[2 * d + 1] = d < nd ? (l == minlevel ? s : s / 2) : 1;
      p.mg[2 * d] = 1;
      p.mg[2 * d + 1] = d < nd ? s : 1;
      p.xl[d] = 0.0;
    }
    for (int f = 0; f < 6; f++) p.bc[f] = bc[f];
    p.dx = 1.0 / s;
    ps.push_back(p);
  }
  // level-0 patch is ps.back(); children on a coarse 4-cell block lattice
  for (int l = 1; l <= 2; l++) {
    std::vector<int> parents;
    for (int i = 0; i < (int)ps.size(); i++)
      if (ps[i].level == l - 1) parents.push_back(i);
    for (int pi : parents) {
      const P pa = ps[pi];   // a copy: ps grows below
      int nb[3] = {pa.mx[0] / 4, pa.mx[1] / 4, nd == 3 ? pa.mx[2] / 4 : 1};
      for (int bk = 0; bk < nb[2]; bk++)
        for (int bj = 0; bj < nb[1]; bj++)
          for (int bi = 0; bi < nb[0]; bi++) {
            if (rnd(0, 3) != 0) continue;
            P c{};
            c.level = l;
            c.parent = pi;
            int b[3] = {bi, bj, bk};
            for (int d = 0; d < 3; d++) {
              int lo = d < nd ? 4 * b[d] + 1 : 1, hi = d < nd ? 4 * b[d] + 4 : 1;
              c.mb[2 * d] = lo;
              c.mb[2 * d + 1] = hi;
              c.mx[d] = d < nd ? 2 * (hi - lo + 1) : 1;
              int g1 = pa.mg[2 * d] + lo - 1, g2 = g1 + hi - lo;
              c.mg[2 * d] = d < nd ? (g1 - 1) * 2 + 1 : 1;
              c.mg[2 * d + 1] = d < nd ? g2 * 2 : 1;
              c.xl[d] = d < nd ? (lo - 1) * pa.dx + pa.xl[d] : 0.0;
            }
            for (int d = 0; d < nd; d++) {
              c.bc[2 * d] = (pa.bc[2 * d] < 999 && c.mb[2 * d] == 1) ? pa.bc[2 * d] : 999;
              c.bc[2 * d + 1] =
                  (pa.bc[2 * d + 1] < 999 && c.mb[2 * d + 1] == pa.mx[d]) ? pa.bc[2 * d + 1] : 999;
            }
            for (int f = 2 * nd; f < 6; f++) c.bc[f] = 1;
            c.dx = pa.dx / 2;
            ps.push_back(c);
          }
    }
  }
  return ps;
}

int check_hierarchy(int nd, int root, const int bc[6]) {
  int minlevel = 0;
  std::vector<P> ps = hierarchy(nd, root, bc, minlevel);
  bsam_cfg cfg{};
  cfg.ndims = nd; cfg.nrvars = 2; cfg.mbc = 1; cfg.minlevel = minlevel;
  cfg.dt = 1e-3; cfg.eps2 = 9e-4; cfg.fc_alpha = 1; cfg.fc_shift = 0.5;
  cfg.fe_num = 1; cfg.fe_den = 4; cfg.nsmoothingpasses = 2; cfg.maxvcycles = 10;
  cfg.qerrortol = 1e-8; cfg.precision = 0; cfg.device = -1; cfg.errortype = 1;
  bsam_ctx *c = bsam_ctx_create(&cfg);
  if (!c) return 1;
  for (int rep = 0; rep < 2; rep++) {   // the second commit reuses the first
    if (bsam_hierarchy_begin(c)) return 2;
    for (const P &p : ps)
      if (bsam_add_patch(c, 0, p.level, p.parent, p.mx, p.mb, p.mg, p.bc, p.dx, p.xl)) return 3;
    if (bsam_hierarchy_commit(c)) {
      fprintf(stderr, "commit: %s\n", bsam_last_error(c));
      return 4;
    }
  }
  // the ghost programs applied to host arrays of every level and pass
  std::vector<std::vector<double>> q(ps.size());
  std::vector<double *> qp(ps.size());
  for (size_t i = 0; i < ps.size(); i++) {
    long n = 2L * (ps[i].mx[0] + 2) * (ps[i].mx[1] + 2) * (nd == 3 ? ps[i].mx[2] + 2 : 1);
    q[i].resize(n);
    for (double &v : q[i]) v = std::uniform_real_distribution<double>(-1, 1)(rng);
    qp[i] = q[i].data();
  }
  for (int l = minlevel; l <= 2; l++)
    for (int ip = 0; ip < 3; ip++)
      if (bsam_ghost_program_apply_host(c, l, ip, qp.data()) && l <= 0) return 5;
  bsam_ctx_destroy(c);
  return 0;
}

int check_regrid(int nd, int level) {
  for (int t = 0; t < 20; t++) {
    int np = rnd(1, 4);
    std::vector<int> mx(3 * np), mg(6 * np);
    std::vector<std::vector<int>> fl(np);
    std::vector<int *> fp(np);
    int x0 = 1;
    for (int k = 0; k < np; k++) {
      for (int d = 0; d < 3; d++) {
        int m = d < nd ? rnd(2, 24) : 1;
        mx[3 * k + d] = m;
        int lo = d == 0 ? x0 : 1;
        mg[6 * k + 2 * d] = lo;
        mg[6 * k + 2 * d + 1] = lo + m - 1;
      }
      x0 += mx[3 * k] + rnd(0, 3);
      fl[k].resize((size_t)mx[3 * k] * mx[3 * k + 1] * mx[3 * k + 2]);
      for (int &v : fl[k]) v = rnd(0, 9) == 0;
      fp[k] = fl[k].data();
    }
    bsam_regrid_params rp{};
    rp.ndims = nd; rp.level = level; rp.ibuffer = rnd(0, 2); rp.minimumgridpoints = rnd(1, 4);
    rp.desiredfillratio = 0.5 + 0.1 * rnd(0, 4);
    for (int d = 0; d < 3; d++) rp.root_mx[d] = d < nd ? 64 : 1;
    for (int f = 0; f < 6; f++) rp.root_mthbc[f] = f < 2 * nd ? (rnd(0, 1) ? 2 : 1) : 1;
    for (int d = 0; d < nd; d++) rp.root_mthbc[2 * d + 1] = rp.root_mthbc[2 * d];
    int cap = 4096, ng = 0;
    std::vector<int> par(cap), mb(6 * cap);
    int rc = bsam_regrid_level(&rp, np, mx.data(), mg.data(), fp.data(), cap, &ng, par.data(),
                               mb.data());
    if (rc && rp.minimumgridpoints <= 24) {
      // a grid below minimumgridpoints is an error the reference stops on
      continue;
    }
  }
  return 0;
}

}  // namespace

int main() {
  const int neumann[6] = {1, 1, 1, 1, 1, 1}, user[6] = {10, 10, 10, 10, 10, 10},
            periodic[6] = {2, 2, 2, 2, 2, 2}, mixed[6] = {2, 2, 10, 10, 1, 1};
  int rc = 0;
  for (int nd = 2; nd <= 3; nd++)
    for (const int *bc : {neumann, user, periodic, mixed}) {
      int r = check_hierarchy(nd, nd == 2 ? 32 : 16, bc);
      if (r) {
        fprintf(stderr, "hierarchy nd=%d bc=%d: %d\n", nd, bc[0], r);
        rc = 1;
      }
    }
  for (int nd = 2; nd <= 3; nd++)
    for (int level = 0; level <= 2; level++) rc |= check_regrid(nd, level);
  std::vector<double> v(10000);
  for (double &x : v) x = std::uniform_real_distribution<double>(-1e3, 1e3)(rng);
  std::vector<char> out(10000 * 60);
  if (bsam_format_rows(v.data(), 5000, 2, 2, 1, 0, out.data(), (long)out.size()) < 0) rc = 1;
  if (bsam_format_rows(v.data(), 5000, 2, 1, 5000, 1, out.data(), (long)out.size()) < 0) rc = 1;
  printf("host_check %s\n", rc ? "FAILED" : "ok");
  return rc;
}
