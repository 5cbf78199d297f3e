// ghost.cpp -- ghost-fill programs.
//
// SetGhost(level, ipass) (boundary.f90:58-89) runs four ordered phases per
// level: coarse-fine interpolation of the colour-selected ghosts
// (Interpolate2D/3D), same-level sibling copies (GhostOverlap over
// ApplyOnLevelPairs), periodic copies, then the physical boundary faces in
// order 1..2*ndims. Later phases overwrite earlier ones and boundary faces read
// ghosts written before them, so the order decides the result.
//
// Here the phases are executed symbolically, once per hierarchy, on the host:
// each ghost cell ends with an expression over values that the program never
// writes (interior cells, parent cells, sibling interiors, or ghosts this pass
// leaves untouched). The device then applies all updates of a program in one
// launch with no ordering between them; the builder checks the no-overlap
// property and refuses programs that violate it.
#include "ghost.hpp"

#include <algorithm>
#include <array>
#include <climits>
#include <cstring>
#include <initializer_list>
#include <map>
#include <unordered_map>
#include <unordered_set>

namespace bsam {
namespace {

struct Cell {
  int p = -1, i = 0, j = 0, k = 0;
};

struct Expr {
  int kind = 0;   // GhostOp kinds
  Cell f0, f1;
  Cell c[4];
  int nflip = 0;
  int qcode = 0;  // periodic offset added before the maps (flip_byte's code)
};

// A Fortran array-section subscript: start, stop, stride (scalars have
// start == stop, stride 1).
struct R {
  int s, e, st;
  int count() const { return e < s ? 0 : (e - s) / st + 1; }
};
inline R S(int x) { return R{x, x, 1}; }
inline R G(int s, int e, int st) { return R{s, e, st}; }
using Sec = std::array<R, 3>;

// Ghost-cell expressions of one patch, keyed by ext-linear index: a dense
// slot table over the ghost-extended box and the entries in insertion order
// (sorted by key once the program is complete).
struct GhostMap {
  std::vector<int> slot;
  std::vector<std::pair<long long, Expr>> ent;
  void init(long long extvol) { slot.assign((size_t)extvol, -1); }
  Expr &operator[](long long k) {
    int &s = slot[(size_t)k];
    if (s < 0) {
      s = (int)ent.size();
      ent.emplace_back(k, Expr());
    }
    return ent[(size_t)s].second;
  }
  const Expr *find(long long k) const {
    int s = slot[(size_t)k];
    return s < 0 ? nullptr : &ent[(size_t)s].second;
  }
  // entry numbers by ascending key
  std::vector<int> order() const {
    std::vector<int> o(ent.size());
    for (size_t n = 0; n < o.size(); n++) o[n] = (int)n;
    std::sort(o.begin(), o.end(), [&](int a, int b) { return ent[(size_t)a].first < ent[(size_t)b].first; });
    return o;
  }
};

struct PatchMap {
  int p = -1;
  GhostMap m;  // ext-linear index -> expression
};

struct Builder {
  const std::vector<HPatch> &hp;
  int nd;
  std::string &msg;
  bool ok = true;

  Builder(const std::vector<HPatch> &h, int n, std::string &m) : hp(h), nd(n), msg(m) {}

  long long key(const HPatch &h, int i, int j, int k) const {
    long long ex = h.mx[0] + 2, ey = h.mx[1] + 2;
    return i + ex * (j + ey * (long long)k);
  }
  Cell cell_of(const HPatch &h, int p, long long key) const {
    long long ex = h.mx[0] + 2, ey = h.mx[1] + 2;
    Cell c;
    c.p = p;
    c.i = (int)(key % ex);
    c.j = (int)((key / ex) % ey);
    c.k = (int)(key / (ex * ey));
    return c;
  }
  bool interior(const HPatch &h, int i, int j, int k) const {
    return i >= 1 && i <= h.mx[0] && j >= 1 && j <= h.mx[1] &&
           (nd == 2 || (k >= 1 && k <= h.mx[2]));
  }

  // One array-section assignment of Interpolate2D/3D: dst(d) = f(a, b, cs...)
  // with the subscripts of every operand advancing together.
  void put(PatchMap &pm, int kind, const Sec &d, const Sec &a, const Sec &b,
           std::initializer_list<Sec> cs) {
    const HPatch &h = hp[pm.p];
    int par = h.parent;
    int n[3];
    for (int x = 0; x < 3; x++) {
      n[x] = d[x].count();
      auto chk = [&](const Sec &s) {
        if (s[x].count() != n[x]) {
          ok = false;
          msg = "non-conforming interpolation section";
        }
      };
      chk(a);
      chk(b);
      for (const Sec &s : cs) chk(s);
    }
    if (!ok) return;
    for (int z = 0; z < n[2]; z++)
      for (int y = 0; y < n[1]; y++)
        for (int x = 0; x < n[0]; x++) {
          int pos[3] = {x, y, z};
          auto at = [&](const Sec &s, int p) {
            Cell c;
            c.p = p;
            c.i = s[0].s + s[0].st * pos[0];
            c.j = s[1].s + s[1].st * pos[1];
            c.k = s[2].s + s[2].st * pos[2];
            return c;
          };
          Cell dc = at(d, pm.p);
          Expr e;
          e.kind = kind;
          e.f0 = at(a, pm.p);
          e.f1 = at(b, pm.p);
          int q = 0;
          for (const Sec &s : cs) e.c[q++] = at(s, par);
          pm.m[key(h, dc.i, dc.j, dc.k)] = e;
        }
  }

  // Interpolate2D, boundary.f90:135-323.
  void interp2d(PatchMap &pm, int ipass) {
    const HPatch &h = hp[pm.p];
    int mx = h.mx[0], my = h.mx[1];
    int mb11 = h.mb[0][0], mb12 = h.mb[0][1], mb21 = h.mb[1][0], mb22 = h.mb[1][1];
    bool red = ipass == 1 || ipass == 0, black = ipass == 2 || ipass == 0;
    auto P2 = [](R a, R b) { return Sec{a, b, S(0)}; };
    if (red) {
      put(pm, 1, P2(S(mx + 1), S(my + 1)), P2(S(mx - 1), S(my - 1)), P2(S(mx), S(my)),
          {P2(S(mb12 + 1), S(mb22 + 1))});
      put(pm, 1, P2(S(0), S(0)), P2(S(2), S(2)), P2(S(1), S(1)),
          {P2(S(mb11 - 1), S(mb21 - 1))});
    }
    if (black) {
      put(pm, 1, P2(S(mx + 1), S(0)), P2(S(mx - 1), S(2)), P2(S(mx), S(1)),
          {P2(S(mb12 + 1), S(mb21 - 1))});
      put(pm, 1, P2(S(0), S(my + 1)), P2(S(2), S(my - 1)), P2(S(1), S(my)),
          {P2(S(mb11 - 1), S(mb22 + 1))});
    }
    auto internal = [&](int f) { return h.mthbc[f] == 2 || h.mthbc[f] == 999; };
    if (internal(0)) {
      if (red) {
        put(pm, 1, P2(S(0), G(2, my - 2, 2)), P2(S(2), G(4, my, 2)), P2(S(1), G(3, my - 1, 2)),
            {P2(S(mb11 - 1), G(mb21, mb22 - 1, 1))});
        put(pm, 2, P2(S(0), S(my)), P2(S(2), S(my)), P2(S(1), S(my)),
            {P2(S(mb11 - 1), S(mb22 + 1)), P2(S(mb11 - 1), S(mb22)),
             P2(S(mb11 - 1), S(mb22 - 1))});
      }
      if (black) {
        put(pm, 1, P2(S(0), G(3, my - 1, 2)), P2(S(2), G(1, my - 3, 2)),
            P2(S(1), G(2, my - 2, 2)), {P2(S(mb11 - 1), G(mb21 + 1, mb22, 1))});
        put(pm, 2, P2(S(0), S(1)), P2(S(2), S(1)), P2(S(1), S(1)),
            {P2(S(mb11 - 1), S(mb21 - 1)), P2(S(mb11 - 1), S(mb21)),
             P2(S(mb11 - 1), S(mb21 + 1))});
      }
    }
    if (internal(1)) {
      if (red) {
        put(pm, 1, P2(S(mx + 1), G(3, my - 1, 2)), P2(S(mx - 1), G(1, my - 3, 2)),
            P2(S(mx), G(2, my - 2, 2)), {P2(S(mb12 + 1), G(mb21 + 1, mb22, 1))});
        put(pm, 2, P2(S(mx + 1), S(1)), P2(S(mx - 1), S(1)), P2(S(mx), S(1)),
            {P2(S(mb12 + 1), S(mb21 - 1)), P2(S(mb12 + 1), S(mb21)),
             P2(S(mb12 + 1), S(mb21 + 1))});
      }
      if (black) {
        put(pm, 1, P2(S(mx + 1), G(2, my - 2, 2)), P2(S(mx - 1), G(4, my, 2)),
            P2(S(mx), G(3, my - 1, 2)), {P2(S(mb12 + 1), G(mb21, mb22 - 1, 1))});
        put(pm, 2, P2(S(mx + 1), S(my)), P2(S(mx - 1), S(my)), P2(S(mx), S(my)),
            {P2(S(mb12 + 1), S(mb22 + 1)), P2(S(mb12 + 1), S(mb22)),
             P2(S(mb12 + 1), S(mb22 - 1))});
      }
    }
    if (internal(2)) {
      if (red) {
        put(pm, 1, P2(G(2, mx - 2, 2), S(0)), P2(G(4, mx, 2), S(2)), P2(G(3, mx - 1, 2), S(1)),
            {P2(G(mb11, mb12 - 1, 1), S(mb21 - 1))});
        put(pm, 2, P2(S(mx), S(0)), P2(S(mx), S(2)), P2(S(mx), S(1)),
            {P2(S(mb12 + 1), S(mb21 - 1)), P2(S(mb12), S(mb21 - 1)),
             P2(S(mb12 - 1), S(mb21 - 1))});
      }
      if (black) {
        put(pm, 1, P2(G(3, mx - 1, 2), S(0)), P2(G(1, mx - 3, 2), S(2)),
            P2(G(2, mx - 2, 2), S(1)), {P2(G(mb11 + 1, mb12, 1), S(mb21 - 1))});
        put(pm, 2, P2(S(1), S(0)), P2(S(1), S(2)), P2(S(1), S(1)),
            {P2(S(mb11 - 1), S(mb21 - 1)), P2(S(mb11), S(mb21 - 1)),
             P2(S(mb11 + 1), S(mb21 - 1))});
      }
    }
    if (internal(3)) {
      if (red) {
        put(pm, 1, P2(G(3, mx - 1, 2), S(my + 1)), P2(G(1, mx - 3, 2), S(my - 1)),
            P2(G(2, mx - 2, 2), S(my)), {P2(G(mb11 + 1, mb12, 1), S(mb22 + 1))});
        put(pm, 2, P2(S(1), S(my + 1)), P2(S(1), S(my - 1)), P2(S(1), S(my)),
            {P2(S(mb11 - 1), S(mb22 + 1)), P2(S(mb11), S(mb22 + 1)),
             P2(S(mb11 + 1), S(mb22 + 1))});
      }
      if (black) {
        put(pm, 1, P2(G(2, mx - 2, 2), S(my + 1)), P2(G(4, mx, 2), S(my - 1)),
            P2(G(3, mx - 1, 2), S(my)), {P2(G(mb11, mb12 - 1, 1), S(mb22 + 1))});
        put(pm, 2, P2(S(mx), S(my + 1)), P2(S(mx), S(my - 1)), P2(S(mx), S(my)),
            {P2(S(mb12 + 1), S(mb22 + 1)), P2(S(mb12), S(mb22 + 1)),
             P2(S(mb12 - 1), S(mb22 + 1))});
      }
    }
  }

  // GetFaceIndex, boundary.f90:1271-1297: {ghost, 2nd fine, 1st fine,
  // coarse outside, coarse inside} along dimension d for side parity.
  std::array<int, 5> face_index(const HPatch &h, int d, int parity) const {
    if (parity == 1) return {0, 2, 1, h.mb[d][0] - 1, h.mb[d][0]};
    return {h.mx[d] + 1, h.mx[d] - 1, h.mx[d], h.mb[d][1] + 1, h.mb[d][1]};
  }

  // Interpolate3D, boundary.f90:1299-1799, including the edge test of :1354
  // that compares the z-face code with 99 rather than 999.
  void interp3d(PatchMap &pm, int ipass) {
    const HPatch &h = hp[pm.p];
    int mx = h.mx[0], my = h.mx[1], mz = h.mx[2];
    int mb11 = h.mb[0][0], mb12 = h.mb[0][1], mb21 = h.mb[1][0], mb22 = h.mb[1][1],
        mb31 = h.mb[2][0], mb32 = h.mb[2][1];
    auto phys = [&](int f) { return h.mthbc[f] != 2 && h.mthbc[f] != 999; };
    // (1) corners
    for (int p1 = 1; p1 <= 2; p1++) {
      if (phys(p1 - 1) && ipass == 1) continue;
      auto f1 = face_index(h, 0, p1);
      for (int p2 = 1; p2 <= 2; p2++) {
        if (phys(p2 + 1) && ipass == 1) continue;
        auto f2 = face_index(h, 1, p2);
        for (int p3 = 1; p3 <= 2; p3++) {
          if (phys(p3 + 3) && ipass == 1) continue;
          auto f3 = face_index(h, 2, p3);
          put(pm, 1, Sec{S(f1[0]), S(f2[0]), S(f3[0])}, Sec{S(f1[1]), S(f2[1]), S(f3[1])},
              Sec{S(f1[2]), S(f2[2]), S(f3[2])}, {Sec{S(f1[3]), S(f2[3]), S(f3[3])}});
        }
      }
    }
    // (2) edges 1-4: coordinates 1 and 3 fixed
    for (int p1 = 1; p1 <= 2; p1++) {
      if (phys(p1 - 1) && ipass == 1) continue;
      auto f1 = face_index(h, 0, p1);
      for (int p3 = 1; p3 <= 2; p3++) {
        if (h.mthbc[p3 + 3] != 2 && h.mthbc[p3 + 3] != 99 && ipass == 1) continue;
        auto f3 = face_index(h, 2, p3);
        put(pm, 1, Sec{S(f1[0]), G(3, my - 1, 2), S(f3[0])},
            Sec{S(f1[1]), G(1, my - 3, 2), S(f3[1])}, Sec{S(f1[2]), G(2, my - 2, 2), S(f3[2])},
            {Sec{S(f1[3]), G(mb21 + 1, mb22, 1), S(f3[3])}});
        put(pm, 2, Sec{S(f1[0]), S(1), S(f3[0])}, Sec{S(f1[1]), S(1), S(f3[1])},
            Sec{S(f1[2]), S(1), S(f3[2])},
            {Sec{S(f1[3]), S(mb21 - 1), S(f3[3])}, Sec{S(f1[3]), S(mb21), S(f3[3])},
             Sec{S(f1[3]), S(mb21 + 1), S(f3[3])}});
        put(pm, 1, Sec{S(f1[0]), G(2, my - 2, 2), S(f3[0])},
            Sec{S(f1[1]), G(4, my, 2), S(f3[1])}, Sec{S(f1[2]), G(3, my - 1, 2), S(f3[2])},
            {Sec{S(f1[3]), G(mb21, mb22 - 1, 1), S(f3[3])}});
        put(pm, 2, Sec{S(f1[0]), S(my), S(f3[0])}, Sec{S(f1[1]), S(my), S(f3[1])},
            Sec{S(f1[2]), S(my), S(f3[2])},
            {Sec{S(f1[3]), S(mb22 + 1), S(f3[3])}, Sec{S(f1[3]), S(mb22), S(f3[3])},
             Sec{S(f1[3]), S(mb22 - 1), S(f3[3])}});
      }
    }
    // edges 5-8: coordinates 1 and 2 fixed
    for (int p1 = 1; p1 <= 2; p1++) {
      if (phys(p1 - 1) && ipass == 1) continue;
      auto f1 = face_index(h, 0, p1);
      for (int p2 = 1; p2 <= 2; p2++) {
        if (phys(p2 + 1) && ipass == 1) continue;
        auto f2 = face_index(h, 1, p2);
        put(pm, 1, Sec{S(f1[0]), S(f2[0]), G(3, mz - 1, 2)},
            Sec{S(f1[1]), S(f2[1]), G(1, mz - 3, 2)}, Sec{S(f1[2]), S(f2[2]), G(2, mz - 2, 2)},
            {Sec{S(f1[3]), S(f2[3]), G(mb31 + 1, mb32, 1)}});
        put(pm, 2, Sec{S(f1[0]), S(f2[0]), S(1)}, Sec{S(f1[1]), S(f2[1]), S(1)},
            Sec{S(f1[2]), S(f2[2]), S(1)},
            {Sec{S(f1[3]), S(f2[3]), S(mb31 - 1)}, Sec{S(f1[3]), S(f2[3]), S(mb31)},
             Sec{S(f1[3]), S(f2[3]), S(mb31 + 1)}});
        put(pm, 1, Sec{S(f1[0]), S(f2[0]), G(2, mz - 2, 2)},
            Sec{S(f1[1]), S(f2[1]), G(4, mz, 2)}, Sec{S(f1[2]), S(f2[2]), G(3, mz - 1, 2)},
            {Sec{S(f1[3]), S(f2[3]), G(mb31, mb32 - 1, 1)}});
        put(pm, 2, Sec{S(f1[0]), S(f2[0]), S(mz)}, Sec{S(f1[1]), S(f2[1]), S(mz)},
            Sec{S(f1[2]), S(f2[2]), S(mz)},
            {Sec{S(f1[3]), S(f2[3]), S(mb32 + 1)}, Sec{S(f1[3]), S(f2[3]), S(mb32)},
             Sec{S(f1[3]), S(f2[3]), S(mb32 - 1)}});
      }
    }
    // edges 9-12: coordinates 2 and 3 fixed
    for (int p2 = 1; p2 <= 2; p2++) {
      if (phys(p2 + 1) && ipass == 1) continue;
      auto f2 = face_index(h, 1, p2);
      for (int p3 = 1; p3 <= 2; p3++) {
        if (phys(p3 + 3) && ipass == 1) continue;
        auto f3 = face_index(h, 2, p3);
        put(pm, 1, Sec{G(3, mx - 1, 2), S(f2[0]), S(f3[0])},
            Sec{G(1, mx - 3, 2), S(f2[1]), S(f3[1])}, Sec{G(2, mx - 2, 2), S(f2[2]), S(f3[2])},
            {Sec{G(mb11 + 1, mb12, 1), S(f2[3]), S(f3[3])}});
        put(pm, 2, Sec{S(1), S(f2[0]), S(f3[0])}, Sec{S(1), S(f2[1]), S(f3[1])},
            Sec{S(1), S(f2[2]), S(f3[2])},
            {Sec{S(mb11 - 1), S(f2[3]), S(f3[3])}, Sec{S(mb11), S(f2[3]), S(f3[3])},
             Sec{S(mb11 + 1), S(f2[3]), S(f3[3])}});
        put(pm, 1, Sec{G(2, mx - 2, 2), S(f2[0]), S(f3[0])},
            Sec{G(4, mx, 2), S(f2[1]), S(f3[1])}, Sec{G(3, mx - 1, 2), S(f2[2]), S(f3[2])},
            {Sec{G(mb11, mb12 - 1, 1), S(f2[3]), S(f3[3])}});
        put(pm, 2, Sec{S(mx), S(f2[0]), S(f3[0])}, Sec{S(mx), S(f2[1]), S(f3[1])},
            Sec{S(mx), S(f2[2]), S(f3[2])},
            {Sec{S(mb12 + 1), S(f2[3]), S(f3[3])}, Sec{S(mb12), S(f2[3]), S(f3[3])},
             Sec{S(mb12 - 1), S(f2[3]), S(f3[3])}});
      }
    }
    // (3) faces, periodic or internal only
    const R Y0 = G(2, my - 2, 2), Y1 = G(3, my - 1, 2), YA = G(4, my, 2), YB = G(1, my - 3, 2);
    const R Z0 = G(2, mz - 2, 2), Z1 = G(3, mz - 1, 2), ZA = G(4, mz, 2), ZB = G(1, mz - 3, 2);
    const R X0 = G(2, mx - 2, 2), X1 = G(3, mx - 1, 2), XA = G(4, mx, 2), XB = G(1, mx - 3, 2);
    const R CY0 = G(mb21, mb22 - 1, 1), CY1 = G(mb21 + 1, mb22, 1);
    const R CZ0 = G(mb31, mb32 - 1, 1), CZ1 = G(mb31 + 1, mb32, 1);
    const R CX0 = G(mb11, mb12 - 1, 1), CX1 = G(mb11 + 1, mb12, 1);
    // faces 1 and 2 (coordinate 1 fixed)
    for (int p1 = 1; p1 <= 2; p1++) {
      if (phys(p1 - 1)) continue;
      auto f = face_index(h, 0, p1);
      R g = S(f[0]), a = S(f[1]), b = S(f[2]), cc = S(f[3]);
      if (ipass == 0 || (p1 + ipass) % 2 == 0) {
        // yellow
        put(pm, 1, Sec{g, Y0, Z0}, Sec{a, YA, ZA}, Sec{b, Y1, Z1}, {Sec{cc, CY0, CZ0}});
        put(pm, 2, Sec{g, S(my), Z0}, Sec{a, S(my), ZA}, Sec{b, S(my), Z1},
            {Sec{cc, S(mb22 + 1), CZ0}, Sec{cc, S(mb22), CZ0}, Sec{cc, S(mb22 - 1), CZ0}});
        put(pm, 2, Sec{g, Y0, S(mz)}, Sec{a, YA, S(mz)}, Sec{b, Y1, S(mz)},
            {Sec{cc, CY0, S(mb32 + 1)}, Sec{cc, CY0, S(mb32)}, Sec{cc, CY0, S(mb32 - 1)}});
        // blue
        put(pm, 1, Sec{g, Y1, Z1}, Sec{a, YB, ZB}, Sec{b, Y0, Z0}, {Sec{cc, CY1, CZ1}});
        put(pm, 2, Sec{g, Y1, S(1)}, Sec{a, YB, S(1)}, Sec{b, Y0, S(1)},
            {Sec{cc, CY1, S(mb31 - 1)}, Sec{cc, CY1, S(mb31)}, Sec{cc, CY1, S(mb31 + 1)}});
        put(pm, 2, Sec{g, S(1), Z1}, Sec{a, S(1), ZB}, Sec{b, S(1), Z0},
            {Sec{cc, S(mb21 - 1), CZ1}, Sec{cc, S(mb21), CZ1}, Sec{cc, S(mb21 + 1), CZ1}});
      }
      if (ipass == 0 || (p1 + ipass) % 2 == 1) {
        // green
        put(pm, 1, Sec{g, Y1, Z0}, Sec{a, YB, ZA}, Sec{b, Y0, Z1}, {Sec{cc, CY1, CZ0}});
        put(pm, 2, Sec{g, S(1), Z0}, Sec{a, S(1), ZA}, Sec{b, S(1), Z1},
            {Sec{cc, S(mb21 - 1), CZ0}, Sec{cc, S(mb21), CZ0}, Sec{cc, S(mb21 + 1), CZ0}});
        put(pm, 2, Sec{g, Y1, S(mz)}, Sec{a, YB, S(mz)}, Sec{b, Y0, S(mz)},
            {Sec{cc, CY1, S(mb32 + 1)}, Sec{cc, CY1, S(mb32)}, Sec{cc, CY1, S(mb32 - 1)}});
        // red
        put(pm, 1, Sec{g, Y0, Z1}, Sec{a, YA, ZB}, Sec{b, Y1, Z0}, {Sec{cc, CY0, CZ1}});
        put(pm, 2, Sec{g, Y0, S(1)}, Sec{a, YA, S(1)}, Sec{b, Y1, S(1)},
            {Sec{cc, CY0, S(mb31 - 1)}, Sec{cc, CY0, S(mb31)}, Sec{cc, CY0, S(mb31 + 1)}});
        put(pm, 2, Sec{g, S(my), Z1}, Sec{a, S(my), ZB}, Sec{b, S(my), Z0},
            {Sec{cc, S(mb22 + 1), CZ1}, Sec{cc, S(mb22), CZ1}, Sec{cc, S(mb22 - 1), CZ1}});
      }
      for (int p2 = 1; p2 <= 2; p2++) {
        auto f2 = face_index(h, 1, p2);
        for (int p3 = 1; p3 <= 2; p3++) {
          auto f3 = face_index(h, 2, p3);
          put(pm, 3, Sec{g, S(f2[2]), S(f3[2])}, Sec{a, S(f2[2]), S(f3[2])},
              Sec{b, S(f2[2]), S(f3[2])},
              {Sec{cc, S(f2[4]), S(f3[4])}, Sec{cc, S(f2[3]), S(f3[4])},
               Sec{cc, S(f2[4]), S(f3[3])}, Sec{cc, S(f2[3]), S(f3[3])}});
        }
      }
    }
    // faces 3 and 4 (coordinate 2 fixed)
    for (int p2 = 1; p2 <= 2; p2++) {
      if (phys(p2 + 1)) continue;
      auto f = face_index(h, 1, p2);
      R g = S(f[0]), a = S(f[1]), b = S(f[2]), cc = S(f[3]);
      if (ipass == 0 || (p2 + ipass) % 2 == 0) {
        // yellow
        put(pm, 1, Sec{X0, g, Z0}, Sec{XA, a, ZA}, Sec{X1, b, Z1}, {Sec{CX0, cc, CZ0}});
        put(pm, 2, Sec{X0, g, S(mz)}, Sec{XA, a, S(mz)}, Sec{X1, b, S(mz)},
            {Sec{CX0, cc, S(mb32 + 1)}, Sec{CX0, cc, S(mb32)}, Sec{CX0, cc, S(mb32 - 1)}});
        put(pm, 2, Sec{S(mx), g, Z0}, Sec{S(mx), a, ZA}, Sec{S(mx), b, Z1},
            {Sec{S(mb12 + 1), cc, CZ0}, Sec{S(mb12), cc, CZ0}, Sec{S(mb12 - 1), cc, CZ0}});
        // blue
        put(pm, 1, Sec{X1, g, Z1}, Sec{XB, a, ZB}, Sec{X0, b, Z0}, {Sec{CX1, cc, CZ1}});
        put(pm, 2, Sec{S(1), g, Z1}, Sec{S(1), a, ZB}, Sec{S(1), b, Z0},
            {Sec{S(mb11 - 1), cc, CZ1}, Sec{S(mb11), cc, CZ1}, Sec{S(mb11 + 1), cc, CZ1}});
        put(pm, 2, Sec{X1, g, S(1)}, Sec{XB, a, S(1)}, Sec{X0, b, S(1)},
            {Sec{CX1, cc, S(mb31 - 1)}, Sec{CX1, cc, S(mb31)}, Sec{CX1, cc, S(mb31 + 1)}});
      }
      if (ipass == 0 || (p2 + ipass) % 2 == 1) {
        // green
        put(pm, 1, Sec{X1, g, Z0}, Sec{XB, a, ZA}, Sec{X0, b, Z1}, {Sec{CX1, cc, CZ0}});
        put(pm, 2, Sec{X0, g, S(1)}, Sec{XA, a, S(1)}, Sec{X1, b, S(1)},
            {Sec{CX0, cc, S(mb31 - 1)}, Sec{CX0, cc, S(mb31)}, Sec{CX0, cc, S(mb31 + 1)}});
        put(pm, 2, Sec{S(mx), g, Z1}, Sec{S(mx), a, ZB}, Sec{S(mx), b, Z0},
            {Sec{S(mb12 + 1), cc, CZ1}, Sec{S(mb12), cc, CZ1}, Sec{S(mb12 - 1), cc, CZ1}});
        // red
        put(pm, 1, Sec{X0, g, Z1}, Sec{XA, a, ZB}, Sec{X1, b, Z0}, {Sec{CX0, cc, CZ1}});
        put(pm, 2, Sec{S(1), g, Z0}, Sec{S(1), a, ZA}, Sec{S(1), b, Z1},
            {Sec{S(mb11 - 1), cc, CZ0}, Sec{S(mb11), cc, CZ0}, Sec{S(mb11 + 1), cc, CZ0}});
        put(pm, 2, Sec{X1, g, S(mz)}, Sec{XB, a, S(mz)}, Sec{X0, b, S(mz)},
            {Sec{CX1, cc, S(mb32 + 1)}, Sec{CX1, cc, S(mb32)}, Sec{CX1, cc, S(mb32 - 1)}});
      }
      for (int p1 = 1; p1 <= 2; p1++) {
        auto f1 = face_index(h, 0, p1);
        for (int p3 = 1; p3 <= 2; p3++) {
          auto f3 = face_index(h, 2, p3);
          put(pm, 3, Sec{S(f1[2]), g, S(f3[2])}, Sec{S(f1[2]), a, S(f3[2])},
              Sec{S(f1[2]), b, S(f3[2])},
              {Sec{S(f1[4]), cc, S(f3[4])}, Sec{S(f1[3]), cc, S(f3[4])},
               Sec{S(f1[4]), cc, S(f3[3])}, Sec{S(f1[3]), cc, S(f3[3])}});
        }
      }
    }
    // faces 5 and 6 (coordinate 3 fixed)
    for (int p3 = 1; p3 <= 2; p3++) {
      if (phys(p3 + 3)) continue;
      auto f = face_index(h, 2, p3);
      R g = S(f[0]), a = S(f[1]), b = S(f[2]), cc = S(f[3]);
      if (ipass == 0 || (p3 + ipass) % 2 == 0) {
        // yellow
        put(pm, 1, Sec{X0, Y0, g}, Sec{XA, YA, a}, Sec{X1, Y1, b}, {Sec{CX0, CY0, cc}});
        put(pm, 2, Sec{S(mx), Y0, g}, Sec{S(mx), YA, a}, Sec{S(mx), Y1, b},
            {Sec{S(mb12 + 1), CY0, cc}, Sec{S(mb12), CY0, cc}, Sec{S(mb12 - 1), CY0, cc}});
        put(pm, 2, Sec{X0, S(my), g}, Sec{XA, S(my), a}, Sec{X1, S(my), b},
            {Sec{CX0, S(mb22 + 1), cc}, Sec{CX0, S(mb22), cc}, Sec{CX0, S(mb22 - 1), cc}});
        // blue
        put(pm, 1, Sec{X1, Y1, g}, Sec{XB, YB, a}, Sec{X0, Y0, b}, {Sec{CX1, CY1, cc}});
        put(pm, 2, Sec{X1, S(1), g}, Sec{XB, S(1), a}, Sec{X0, S(1), b},
            {Sec{CX1, S(mb21 - 1), cc}, Sec{CX1, S(mb21), cc}, Sec{CX1, S(mb21 + 1), cc}});
        put(pm, 2, Sec{S(1), Y1, g}, Sec{S(1), YB, a}, Sec{S(1), Y0, b},
            {Sec{S(mb11 - 1), CY1, cc}, Sec{S(mb11), CY1, cc}, Sec{S(mb11 + 1), CY1, cc}});
      }
      if (ipass == 0 || (p3 + ipass) % 2 == 1) {
        // green
        put(pm, 1, Sec{X1, Y0, g}, Sec{XB, YA, a}, Sec{X0, Y1, b}, {Sec{CX1, CY0, cc}});
        put(pm, 2, Sec{S(1), Y0, g}, Sec{S(1), YA, a}, Sec{S(1), Y1, b},
            {Sec{S(mb11 - 1), CY0, cc}, Sec{S(mb11), CY0, cc}, Sec{S(mb11 + 1), CY0, cc}});
        put(pm, 2, Sec{X1, S(my), g}, Sec{XB, S(my), a}, Sec{X0, S(my), b},
            {Sec{CX1, S(mb22 + 1), cc}, Sec{CX1, S(mb22), cc}, Sec{CX1, S(mb22 - 1), cc}});
        // red
        put(pm, 1, Sec{X0, Y1, g}, Sec{XA, YB, a}, Sec{X1, Y0, b}, {Sec{CX0, CY1, cc}});
        put(pm, 2, Sec{X0, S(1), g}, Sec{XA, S(1), a}, Sec{X1, S(1), b},
            {Sec{CX0, S(mb21 - 1), cc}, Sec{CX0, S(mb21), cc}, Sec{CX0, S(mb21 + 1), cc}});
        put(pm, 2, Sec{S(mx), Y1, g}, Sec{S(mx), YB, a}, Sec{S(mx), Y0, b},
            {Sec{S(mb12 + 1), CY1, cc}, Sec{S(mb12), CY1, cc}, Sec{S(mb12 - 1), CY1, cc}});
      }
      for (int p1 = 1; p1 <= 2; p1++) {
        auto f1 = face_index(h, 0, p1);
        for (int p2 = 1; p2 <= 2; p2++) {
          auto f2 = face_index(h, 1, p2);
          put(pm, 3, Sec{S(f1[2]), S(f2[2]), g}, Sec{S(f1[2]), S(f2[2]), a},
              Sec{S(f1[2]), S(f2[2]), b},
              {Sec{S(f1[4]), S(f2[4]), cc}, Sec{S(f1[3]), S(f2[4]), cc},
               Sec{S(f1[4]), S(f2[3]), cc}, Sec{S(f1[3]), S(f2[3]), cc}});
        }
      }
    }
  }

  // GhostOverlap, boundary.f90:346-415, for one ordered pair: cells of the
  // ghost-extended overlap that lie in g2's interior are copied into g1.
  void overlap_into(PatchMap &m1, int p2) {
    const HPatch &g1 = hp[m1.p], &g2 = hp[p2];
    int lo[3], hi[3];
    for (int n = 0; n < 3; n++) {
      int a1 = n < nd ? g1.mg[n][0] : 1, a2 = n < nd ? g1.mg[n][1] : 1;
      int b1 = n < nd ? g2.mg[n][0] : 1, b2 = n < nd ? g2.mg[n][1] : 1;
      int o0 = std::max(a1 - 1, b1 - 1), o1 = std::min(a2 + 1, b2 + 1);
      if (n < nd && o1 - o0 < 0) return;
      lo[n] = std::max(o0, b1);
      hi[n] = std::min(o1, b2);
    }
    int g1lo[3], g2lo[3];
    for (int n = 0; n < 3; n++) {
      g1lo[n] = n < nd ? g1.mg[n][0] : 1;
      g2lo[n] = n < nd ? g2.mg[n][0] : 1;
    }
    for (int k = lo[2]; k <= hi[2]; k++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          int di = i - g1lo[0] + 1, dj = j - g1lo[1] + 1;
          int dk = nd == 3 ? k - g1lo[2] + 1 : 0;
          Expr e;
          e.kind = 0;
          e.f0 = Cell{p2, i - g2lo[0] + 1, j - g2lo[1] + 1, nd == 3 ? k - g2lo[2] + 1 : 0};
          m1.m[key(g1, di, dj, dk)] = e;
        }
  }

  // Resolves the current expression of a cell read by a later phase: an
  // interior cell or an untouched ghost is a plain copy of its old value.
  Expr source(PatchMap &pm, const HPatch &h, int i, int j, int k) {
    if (!interior(h, i, j, k)) {
      if (const Expr *e = pm.m.find(key(h, i, j, k))) return *e;
    }
    Expr e;
    e.kind = 0;
    e.f0 = Cell{pm.p, i, j, k};
    return e;
  }

  // PeriodicBC (boundary.f90:417-486): a patch periodic on both sides of a
  // dimension wraps itself, dimensions in order, over the full
  // ghost-extended tangential range; the wrap in x adds +qpo to the high
  // ghosts and -qpo to the low ones.
  void periodic_self(PatchMap &pm) {
    const HPatch &h = hp[pm.p];
    for (int d = 0; d < nd; d++) {
      if (!(h.mthbc[2 * d] == 2 && h.mthbc[2 * d + 1] == 2)) continue;
      int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
      int ahi = ta < nd ? h.mx[ta] + 1 : 0, bhi = tb < nd ? h.mx[tb] + 1 : 0;
      std::vector<std::pair<long long, Expr>> upd;
      for (int b = 0; b <= bhi; b++)
        for (int a = 0; a <= ahi; a++) {
          int hi[3], shi[3], lo[3], slo[3];
          hi[d] = h.mx[d] + 1; shi[d] = 1;
          lo[d] = 0; slo[d] = h.mx[d];
          hi[ta] = shi[ta] = lo[ta] = slo[ta] = a;
          hi[tb] = shi[tb] = lo[tb] = slo[tb] = b;
          Expr ehi = source(pm, h, shi[0], shi[1], shi[2]);
          Expr elo = source(pm, h, slo[0], slo[1], slo[2]);
          if (d == 0) {
            // the sources are interior cells or ghosts set by the
            // interpolation phase: no offset or map yet
            if (ehi.qcode || elo.qcode || ehi.nflip || elo.nflip) {
              ok = false;
              msg = "ghost program: a periodic offset over another offset or map";
            }
            ehi.qcode = 1;   // + qpo(n)
            elo.qcode = 2;   // - qpo(n)
          }
          upd.push_back({key(h, hi[0], hi[1], hi[2]), ehi});
          upd.push_back({key(h, lo[0], lo[1], lo[2]), elo});
        }
      for (auto &u : upd) pm.m[u.first] = u.second;
    }
  }

  // PeriodicGhostOverlap (boundary.f90:510-596) for the ordered pair
  // (g1, g2): every periodic image of g2 exchanges its ghost-extended overlap
  // with g1 both ways; all sources are interior cells. The copies into g1
  // add qoffset (GetPeriodicOffsets, :965-1044: +qpo for an image shifted by
  // +x, -qpo for -x, none otherwise), those into g2 subtract it.
  void periodic_pair(PatchMap &m1, PatchMap &m2, int level, const Periodic &per) {
    const HPatch &g1 = hp[m1.p], &g2 = hp[m2.p];
    static const int tab[13][3] = {{1, 0, 0},  {0, 1, 0},  {0, 0, 1},  {1, 1, 0},  {-1, 1, 0},
                                   {0, 1, 1},  {0, 1, -1}, {1, 0, 1},  {-1, 0, 1}, {1, 1, 1},
                                   {1, 1, -1}, {-1, 1, 1}, {1, -1, 1}};
    for (int o = 0; o < per.noff; o++) {
      int off[3];
      for (int d = 0; d < 3; d++) off[d] = tab[per.idx[o] - 1][d] * (per.rootmx[d] << level);
      int a1[3], a2[3], b1[3], b2[3], lo1[3], hi1[3], lo2[3], hi2[3];
      bool empty = false;
      for (int n = 0; n < 3; n++) {
        a1[n] = n < nd ? g1.mg[n][0] : 1; a2[n] = n < nd ? g1.mg[n][1] : 1;
        b1[n] = n < nd ? g2.mg[n][0] + off[n] : 1; b2[n] = n < nd ? g2.mg[n][1] + off[n] : 1;
        int o0 = std::max(a1[n] - 1, b1[n] - 1), o1 = std::min(a2[n] + 1, b2[n] + 1);
        if (n < nd && o1 - o0 < 0) empty = true;
        lo1[n] = std::max(o0, b1[n]); hi1[n] = std::min(o1, b2[n]);
        lo2[n] = std::max(o0, a1[n]); hi2[n] = std::min(o1, a2[n]);
      }
      if (empty) continue;
      // into g1: + qoffset (+qpo, -qpo, or +0.0 for an image not shifted in
      // x); into g2: - qoffset (-qpo, +qpo, or - 0.0, which leaves every
      // value as it is)
      const int sx = tab[per.idx[o] - 1][0];
      const int code1 = sx > 0 ? 1 : sx < 0 ? 2 : 3, code2 = sx > 0 ? 2 : sx < 0 ? 1 : 0;
      auto put_copy = [&](PatchMap &dm, const HPatch &dh, int di, int dj, int dk, int sp,
                          int si, int sj, int sk, int code) {
        Expr e;
        e.kind = 0;
        e.f0 = Cell{sp, si, sj, sk};
        e.qcode = code;
        dm.m[key(dh, di, dj, dk)] = e;
      };
      for (int k = lo1[2]; k <= hi1[2]; k++)
        for (int j = lo1[1]; j <= hi1[1]; j++)
          for (int i = lo1[0]; i <= hi1[0]; i++)
            put_copy(m1, g1, i - a1[0] + 1, j - a1[1] + 1, nd == 3 ? k - a1[2] + 1 : 0, m2.p,
                     i - b1[0] + 1, j - b1[1] + 1, nd == 3 ? k - b1[2] + 1 : 0, code1);
      for (int k = lo2[2]; k <= hi2[2]; k++)
        for (int j = lo2[1]; j <= hi2[1]; j++)
          for (int i = lo2[0]; i <= hi2[0]; i++)
            put_copy(m2, g2, i - b1[0] + 1, j - b1[1] + 1, nd == 3 ? k - b1[2] + 1 : 0, m1.p,
                     i - a1[0] + 1, j - a1[1] + 1, nd == 3 ? k - a1[2] + 1 : 0, code2);
    }
  }

  // SetBC2D/SetBC3D (boundary.f90:643-891) with the CH module's UserBC2D/3D
  // (problem.f90:646-756): faces in order over the full tangential range.
  void physical_bc(PatchMap &pm) {
    const HPatch &h = hp[pm.p];
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
      lo[d] = d < nd ? 0 : 0;
      hi[d] = d < nd ? h.mx[d] + 1 : 0;
    }
    for (int face = 0; face < 2 * nd; face++) {
      int t = h.mthbc[face];
      if (t == 2 || t == 999) continue;
      int flip = (t == 1) ? 0 : ((nd == 3 || face < 2) ? 1 : 0);
      int d = face / 2, right = face % 2;
      int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
      std::vector<std::pair<long long, Expr>> upd;
      for (int b = lo[tb]; b <= hi[tb]; b++)
        for (int a = lo[ta]; a <= hi[ta]; a++) {
          int di[3], si[3];
          di[d] = right ? h.mx[d] + 1 : 0;
          si[d] = right ? h.mx[d] : 1;
          di[ta] = si[ta] = a;
          di[tb] = si[tb] = b;
          Expr e;
          if (interior(h, si[0], si[1], si[2])) {
            e.kind = 0;
            e.f0 = Cell{pm.p, si[0], si[1], si[2]};
            e.nflip = flip;
          } else {
            const Expr *it = pm.m.find(key(h, si[0], si[1], si[2]));
            if (it) {
              e = *it;
              e.nflip += flip;
            } else {
              e.kind = 0;
              e.f0 = Cell{pm.p, si[0], si[1], si[2]};
              e.nflip = flip;
            }
          }
          upd.push_back({key(h, di[0], di[1], di[2]), e});
        }
      for (auto &u : upd) pm.m[u.first] = u.second;
    }
  }
};

long long qoff(const std::vector<HPatch> &hp, const Cell &c) {
  const HPatch &h = hp[c.p];
  return h.qoff + qcell(h.px, h.py, c.i, c.j, c.k);
}

}  // namespace

Periodic periodic_setup(const HPatch &root, int ndims) {
  Periodic per;
  bool px = root.mthbc[0] == 2, py = root.mthbc[2] == 2, pz = ndims == 3 && root.mthbc[4] == 2;
  std::vector<int> list;
  if (px && py && pz) list = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13};
  else if (px && py) list = {1, 2, 4, 5};
  else if (px && pz) list = {1, 3, 8, 9};
  else if (px) list = {1};
  else if (py && pz) list = {2, 3, 6, 7};
  else if (py) list = {2};
  else if (pz) list = {3};
  per.noff = (int)list.size();
  for (int k = 0; k < per.noff; k++) per.idx[k] = list[k];
  for (int d = 0; d < 3; d++) per.rootmx[d] = root.mx[d];
  return per;
}

int build_ghost_program(const std::vector<HPatch> &hp, const std::vector<int> &patches,
                        int ndims, int level, int ipass, const Periodic &per,
                        std::vector<GhostOp> &out, std::string &msg) {
  out.clear();
  Builder B(hp, ndims, msg);
  std::vector<PatchMap> maps(patches.size());
  for (size_t a = 0; a < patches.size(); a++) {
    maps[a].p = patches[a];
    const HPatch &h = hp[patches[a]];
    maps[a].m.init((long long)(h.mx[0] + 2) * (h.mx[1] + 2) * (ndims == 3 ? h.mx[2] + 2 : 1));
  }
  if (level > 0) {
    for (auto &pm : maps) {
      if (ndims == 2) B.interp2d(pm, ipass);
      else B.interp3d(pm, ipass);
      if (!B.ok) return -1;
    }
    // ApplyOnLevelPairs order (treeops.f90:574-610); both directions per pair
    for (size_t a = 0; a < maps.size(); a++)
      for (size_t b = a + 1; b < maps.size(); b++) {
        B.overlap_into(maps[a], maps[b].p);
        B.overlap_into(maps[b], maps[a].p);
      }
  }
  if (per.noff > 0) {
    for (auto &pm : maps) B.periodic_self(pm);
    if (!B.ok) return -1;
    if (level > 0)
      for (size_t a = 0; a < maps.size(); a++)
        for (size_t b = a + 1; b < maps.size(); b++) {
          B.periodic_pair(maps[a], maps[b], level, per);
          B.periodic_pair(maps[b], maps[a], level, per);
        }
  }
  for (auto &pm : maps) B.physical_bc(pm);
  for (auto &pm : maps) {
    const HPatch &h = hp[pm.p];
    for (int n : pm.m.order()) {
      const auto &kv = pm.m.ent[(size_t)n];
      const Expr &e = kv.second;
      Cell dc = B.cell_of(h, pm.p, kv.first);
      if (B.interior(h, dc.i, dc.j, dc.k)) {
        msg = "ghost program writes an interior cell";
        return -1;
      }
      GhostOp g;
      memset(&g, 0, sizeof g);
      if (e.nflip < 0 || e.nflip > 15) {
        msg = "ghost program: too many boundary maps on one ghost";
        return -1;
      }
      g.dst = (int)qoff(hp, dc);
      g.kind = (unsigned char)e.kind;
      g.nflip = (unsigned char)flip_byte(e.nflip, e.qcode);
      g.f0 = (int)qoff(hp, e.f0);
      if (e.kind > 0) {
        g.f1 = (int)qoff(hp, e.f1);
        int nc = e.kind == 1 ? 1 : (e.kind == 2 ? 3 : 4);
        for (int q = 0; q < nc; q++) g.c[q] = (int)qoff(hp, e.c[q]);
      }
      out.push_back(g);
    }
  }
  // every source must be a value the program leaves untouched
  std::vector<char> written;
  {
    size_t n = 0;
    for (auto &g : out) n = std::max(n, (size_t)g.dst + 1);
    written.assign(n, 0);
    for (auto &g : out) written[(size_t)g.dst] = 1;
  }
  auto wr = [&](int off) { return off >= 0 && (size_t)off < written.size() && written[(size_t)off]; };
  for (auto &g : out) {
    bool bad = wr(g.f0);
    if (g.kind > 0) {
      bad = bad || wr(g.f1);
      int nc = g.kind == 1 ? 1 : (g.kind == 2 ? 3 : 4);
      for (int q = 0; q < nc; q++) bad = bad || wr(g.c[q]);
    }
    if (bad) {
      msg = "ghost program reads a cell it writes";
      return -1;
    }
  }
  return 0;
}

namespace {
// Cells of a level's relaxation launch by geometry: the patch and local index
// of a q-pool offset (pools are laid out in patch order), whether a cell is
// interior, the interior tile holding it (tiles of a patch are k, j, i
// nested, from the patch's first tile) and whether colour `color` relaxes it.
struct LaunchGeometry {
  const std::vector<HPatch> &hp;
  int nd, tx, ty, tz, color;
  std::vector<long long> qoffs;
  std::vector<int> first;
  LaunchGeometry(const std::vector<HPatch> &h, const std::vector<Tile> &tiles, int nd_, int tx_,
                 int ty_, int tz_, int color_)
      : hp(h), nd(nd_), tx(tx_), ty(ty_), tz(tz_), color(color_), qoffs(h.size()),
        first(h.size(), -1) {
    for (size_t n = 0; n < hp.size(); n++) qoffs[n] = hp[n].qoff;
    for (size_t t = 0; t < tiles.size(); t++)
      if (first[tiles[t].patch] < 0) first[tiles[t].patch] = (int)t;
  }
  void locate(long long off, int &p, int &i, int &j, int &k) const {
    auto it = std::upper_bound(qoffs.begin(), qoffs.end(), off);
    p = (int)(it - qoffs.begin()) - 1;
    if (p < 0 || off >= hp[p].qoff + hp[p].qn) {
      p = -1;
      return;
    }
    const HPatch &h = hp[p];
    qcell_index(off - h.qoff, h.px, h.py, i, j, k);
  }
  bool inside(int p, int i, int j, int k) const {
    const HPatch &h = hp[p];
    return i >= 1 && i <= h.mx[0] && j >= 1 && j <= h.mx[1] &&
           (nd == 2 || (k >= 1 && k <= h.mx[2]));
  }
  int tile_of(int p, int i, int j, int k) const {
    const HPatch &h = hp[p];
    int nx = (h.mx[0] + tx - 1) / tx, ny = (h.mx[1] + ty - 1) / ty;
    int a = (i - 1) / tx, b = (j - 1) / ty, cz = nd == 3 ? (k - 1) / tz : 0;
    return first[p] + (cz * ny + b) * nx + a;
  }
  bool relaxed(int i, int j, int k) const {
    return ((i + j + (nd == 3 ? k : 0) + color) & 1) == 1;
  }
  // the tile of this launch that relaxes the cell at `off`, -1 if none
  int relaxing_tile(long long off) const {
    int p, i, j, k;
    locate(off, p, i, j, k);
    if (p < 0 || first[p] < 0 || !inside(p, i, j, k) || !relaxed(i, j, k)) return -1;
    return tile_of(p, i, j, k);
  }
  // for a ghost cell: the patch, the interior cell it borders (clamped
  // index) and how many dimensions it lies outside in; p < 0 if the offset
  // is not a cell of a tiled patch
  void border(long long off, int &p, int c[3], int &outside) const {
    int i, j, k;
    locate(off, p, i, j, k);
    outside = 0;
    if (p < 0 || first[p] < 0) {
      p = -1;
      return;
    }
    const HPatch &h = hp[p];
    int m[3] = {h.mx[0], h.mx[1], nd == 3 ? h.mx[2] : 1};
    c[0] = i;
    c[1] = j;
    c[2] = nd == 3 ? k : 1;
    for (int d = 0; d < 3; d++) {
      if (c[d] < 1 || c[d] > m[d]) outside++;
      c[d] = std::min(std::max(c[d], 1), m[d]);
    }
  }
};
}  // namespace

FusedGhost plan_fused_ghost(const std::vector<HPatch> &hp, const std::vector<Tile> &tiles,
                            int nd, int tx, int ty, int tz,
                            const std::vector<GhostOp> &prog, const std::vector<GhostOp> &next,
                            int color, std::string &why) {
  FusedGhost out;
  const LaunchGeometry G(hp, tiles, nd, tx, ty, tz, color);
  auto locate = [&](long long off, int &p, int &i, int &j, int &k) { G.locate(off, p, i, j, k); };
  const std::vector<int> &first = G.first;
  auto inside = [&](int p, int i, int j, int k) { return G.inside(p, i, j, k); };
  auto tile_of = [&](int p, int i, int j, int k) { return G.tile_of(p, i, j, k); };
  auto relaxed = [&](int i, int j, int k) { return G.relaxed(i, j, k); };
  std::vector<char> next_dst;   // destinations of the next program
  {
    size_t n = 0;
    for (const GhostOp &g : next) n = std::max(n, (size_t)g.dst + 1);
    next_dst.assign(n, 0);
    for (const GhostOp &g : next) next_dst[(size_t)g.dst] = 1;
  }
  std::vector<std::vector<GhostOp>> per(tiles.size());
  for (const GhostOp &g : prog) {
    int p, i, j, k;
    locate(g.dst, p, i, j, k);
    if (p < 0 || first[p] < 0) {
      why = "ghost destination outside the level's patches";
      return out;
    }
    const HPatch &h = hp[p];
    int m[3] = {h.mx[0], h.mx[1], nd == 3 ? h.mx[2] : 1};
    int c[3] = {i, j, nd == 3 ? k : 1};
    int outside = 0;
    for (int d = 0; d < 3; d++) {
      if (c[d] < 1 || c[d] > m[d]) outside++;
      c[d] = std::min(std::max(c[d], 1), m[d]);
    }
    // the interior cell a face ghost borders is the only reader of the ghost
    bool read = outside == 1 && relaxed(c[0], c[1], nd == 3 ? c[2] : 0);
    if (!read && (size_t)g.dst < next_dst.size() && next_dst[(size_t)g.dst]) continue;
    // the tiles of this launch that relax a source of the update (coarse
    // sources belong to another level; f0/f1 are the level's cells)
    int src_tile = -1;
    bool several = false;
    long long srcs[2] = {g.f0, g.f1};
    int nsrc = g.kind == 0 ? 1 : 2;
    for (int s2 = 0; s2 < nsrc; s2++) {
      int sp, si, sj, sk;
      locate(srcs[s2], sp, si, sj, sk);
      if (sp < 0) {
        why = "ghost source outside the q pool";
        return out;
      }
      if (first[sp] < 0 || !inside(sp, si, sj, sk)) continue;   // not relaxed here
      if (!relaxed(si, sj, nd == 3 ? sk : 0)) continue;
      int t = tile_of(sp, si, sj, sk);
      if (src_tile >= 0 && t != src_tile) several = true;
      src_tile = t;
    }
    // The update must read its sources before they are relaxed and, when the
    // relaxation reads its destination, be written before that read. The
    // tile that relaxes the sources applies it before relaxing; an update
    // whose destination is read goes to the reading tile, and to `pre` if a
    // source is relaxed by another tile (or by several tiles).
    int owner;
    if (read) {
      owner = tile_of(p, c[0], c[1], c[2]);
      if (src_tile >= 0 && src_tile != owner) owner = -1;
    } else {
      owner = src_tile >= 0 ? src_tile : tile_of(p, c[0], c[1], c[2]);
    }
    if (several) owner = -1;
    if (owner < 0) out.pre.push_back(g);
    else per[owner].push_back(g);
  }
  out.start.assign(tiles.size() + 1, 0);
  for (size_t t = 0; t < tiles.size(); t++) {
    out.start[t + 1] = out.start[t] + (int)per[t].size();
    out.ops.insert(out.ops.end(), per[t].begin(), per[t].end());
  }
  out.kept = (int)(out.ops.size() + out.pre.size());
  if (!fused_plan_race_free(hp, tiles, nd, tx, ty, tz, color, out, why)) return out;
  out.ok = true;
  return out;
}

namespace {
// The cells of a relaxation launch from the cells' side: every cell colour
// `color` relaxes marked with its tile, every ghost cell a relaxed cell's
// 2*nd-point stencil reads marked with the reading tile (-2 if several), over
// the q-pool range [lo, hi) of the launch's patches.
struct LaunchCells {
  long long lo = LLONG_MAX, hi = 0;
  std::vector<int> relaxed_by, read_by;
  int at(const std::vector<int> &m, long long off) const {
    return off >= lo && off < hi ? m[(size_t)(off - lo)] : -1;
  }
};
LaunchCells launch_cells(const std::vector<HPatch> &hp, const std::vector<Tile> &tiles, int nd,
                         int tx, int ty, int tz, int color) {
  LaunchCells lc;
  for (const Tile &t : tiles) {
    const HPatch &h = hp[t.patch];
    lc.lo = std::min(lc.lo, h.qoff);
    lc.hi = std::max(lc.hi, h.qoff + h.qn);
  }
  if (tiles.empty()) lc.lo = 0;
  lc.relaxed_by.assign((size_t)(lc.hi - lc.lo), -1);
  lc.read_by.assign((size_t)(lc.hi - lc.lo), -1);
  const int nz = nd == 3 ? tz : 1;
  for (size_t t = 0; t < tiles.size(); t++) {
    const Tile &tl = tiles[t];
    const HPatch &h = hp[tl.patch];
    const int mz = nd == 3 ? h.mx[2] : 1;
    for (int k = tl.k0; k < tl.k0 + nz; k++)
      for (int j = tl.j0; j < tl.j0 + ty; j++)
        for (int i = tl.i0; i < tl.i0 + tx; i++) {
          const int kk = nd == 3 ? k : 0;
          if (i > h.mx[0] || j > h.mx[1] || (nd == 3 && k > mz)) continue;
          if (((i + j + kk + color) & 1) != 1) continue;
          lc.relaxed_by[(size_t)(h.qoff + qcell(h.px, h.py, i, j, kk) - lc.lo)] = (int)t;
          const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0},
                    dk[6] = {0, 0, 0, 0, 1, -1};
          for (int n = 0; n < 2 * nd; n++) {
            const int a = i + di[n], b = j + dj[n], c = kk + dk[n];
            const bool ghost = a < 1 || a > h.mx[0] || b < 1 || b > h.mx[1] ||
                               (nd == 3 && (c < 1 || c > mz));
            if (!ghost) continue;
            int &r = lc.read_by[(size_t)(h.qoff + qcell(h.px, h.py, a, b, c) - lc.lo)];
            r = (r == -1 || r == (int)t) ? (int)t : -2;
          }
        }
  }
  return lc;
}
// Sources of an update: f0, f1 (fine) and the coarse c[] by kind.
int op_sources(const GhostOp &g, long long out[6]) {
  int n = 0;
  out[n++] = g.f0;
  if (g.kind > 0) {
    out[n++] = g.f1;
    const int nc = g.kind == 1 ? 1 : (g.kind == 2 ? 3 : 4);
    for (int q = 0; q < nc; q++) out[n++] = g.c[q];
  }
  return n;
}
}  // namespace

// Checks a fused plan from the cells' side: each tile's updates may read
// relaxed cells of that tile only and write ghost cells that no other tile
// reads.
bool fused_plan_race_free(const std::vector<HPatch> &hp, const std::vector<Tile> &tiles, int nd,
                          int tx, int ty, int tz, int color, const FusedGhost &plan,
                          std::string &why) {
  if (tiles.empty()) return true;
  const LaunchCells lc = launch_cells(hp, tiles, nd, tx, ty, tz, color);
  for (size_t t = 0; t < tiles.size(); t++)
    for (int e = plan.start[t]; e < plan.start[t + 1]; e++) {
      const GhostOp &g = plan.ops[(size_t)e];
      long long src[6];
      const int ns = op_sources(g, src);
      bool bad = lc.at(lc.relaxed_by, g.dst) >= 0;
      for (int q = 0; q < ns; q++) {
        const int r = lc.at(lc.relaxed_by, src[q]);
        bad = bad || (r >= 0 && r != (int)t);
      }
      const int w = lc.at(lc.read_by, g.dst);
      if (bad || (w != -1 && w != (int)t)) {
        why = "fused ghost update shared between tiles";
        return false;
      }
    }
  return true;
}

PostGhost plan_post_ghost(const std::vector<HPatch> &hp, const std::vector<Tile> &tiles, int nd,
                          int tx, int ty, int tz, int color, const FusedGhost *before,
                          const std::vector<GhostOp> &cand, std::string &why) {
  PostGhost out;
  const LaunchGeometry G(hp, tiles, nd, tx, ty, tz, color);
  // the tiles whose in-launch updates (applied before their relaxation)
  // write or read each cell: -2 for several
  std::unordered_map<long long, int> bdst, bsrc;
  auto mark = [](std::unordered_map<long long, int> &m, long long off, int t) {
    auto it = m.find(off);
    if (it == m.end()) m.emplace(off, t);
    else if (it->second != t) it->second = -2;
  };
  if (before && before->ok)
    for (size_t t = 0; t < tiles.size(); t++)
      for (int e = before->start[t]; e < before->start[t + 1]; e++) {
        const GhostOp &g = before->ops[(size_t)e];
        mark(bdst, g.dst, (int)t);
        long long src[6];
        const int ns = op_sources(g, src);
        for (int q = 0; q < ns; q++) mark(bsrc, src[q], (int)t);
      }
  auto other = [](const std::unordered_map<long long, int> &m, long long off, int t) {
    auto it = m.find(off);
    return it != m.end() && it->second != t;
  };
  std::vector<std::vector<GhostOp>> per(tiles.size());
  for (const GhostOp &g : cand) {
    int p, c[3], outside;
    G.border(g.dst, p, c, outside);
    if (p < 0) {
      why = "ghost destination outside the level's patches";
      return out;
    }
    // the tile relaxing the update's fine sources of this colour; it applies
    // the update after relaxing them
    int owner = -1;
    bool several = false;
    const int nf = g.kind == 0 ? 1 : 2;
    const long long fs[2] = {g.f0, g.f1};
    for (int q = 0; q < nf; q++) {
      const int t = G.relaxing_tile(fs[q]);
      if (t < 0) continue;
      if (owner >= 0 && t != owner) several = true;
      owner = t;
    }
    const int near = G.tile_of(p, c[0], c[1], c[2]);
    if (owner < 0) owner = near;
    // the relaxation reads a face ghost through the one interior cell it
    // borders, in that cell's tile
    const bool read = outside == 1 && G.relaxed(c[0], c[1], c[2]);
    bool keep = several || (read && near != owner) || other(bdst, g.dst, owner) ||
                other(bsrc, g.dst, owner);
    long long src[6];
    const int ns = op_sources(g, src);
    for (int q = 0; q < ns && !keep; q++) keep = other(bdst, src[q], owner);
    if (keep) out.rest.push_back(g);
    else per[(size_t)owner].push_back(g);
  }
  out.start.assign(tiles.size() + 1, 0);
  for (size_t t = 0; t < tiles.size(); t++) {
    out.start[t + 1] = out.start[t] + (int)per[t].size();
    out.ops.insert(out.ops.end(), per[t].begin(), per[t].end());
  }
  if (!post_plan_race_free(hp, tiles, nd, tx, ty, tz, color, before, out, why)) return out;
  out.ok = true;
  return out;
}

// Checks a post plan from the cells' side: a tile's updates run after its
// relaxation, so they may read cells relaxed by that tile only, cells no
// in-launch update of another tile writes, and write ghost cells that no
// other tile's stencil or in-launch update reads or writes.
bool post_plan_race_free(const std::vector<HPatch> &hp, const std::vector<Tile> &tiles, int nd,
                         int tx, int ty, int tz, int color, const FusedGhost *before,
                         const PostGhost &plan, std::string &why) {
  if (tiles.empty()) return true;
  const LaunchCells lc = launch_cells(hp, tiles, nd, tx, ty, tz, color);
  std::vector<int> wrote((size_t)(lc.hi - lc.lo), -1), used((size_t)(lc.hi - lc.lo), -1);
  auto put = [&](std::vector<int> &m, long long off, int t) {
    if (off < lc.lo || off >= lc.hi) return;
    int &r = m[(size_t)(off - lc.lo)];
    r = (r == -1 || r == t) ? t : -2;
  };
  if (before && before->ok)
    for (size_t t = 0; t < tiles.size(); t++)
      for (int e = before->start[t]; e < before->start[t + 1]; e++) {
        const GhostOp &g = before->ops[(size_t)e];
        put(wrote, g.dst, (int)t);
        long long src[6];
        const int ns = op_sources(g, src);
        for (int q = 0; q < ns; q++) put(used, src[q], (int)t);
      }
  auto foreign = [&](const std::vector<int> &m, long long off, int t) {
    const int r = lc.at(m, off);
    return r != -1 && r != t;
  };
  for (size_t t = 0; t < tiles.size(); t++)
    for (int e = plan.start[t]; e < plan.start[t + 1]; e++) {
      const GhostOp &g = plan.ops[(size_t)e];
      const int tt = (int)t;
      bool bad = lc.at(lc.relaxed_by, g.dst) >= 0 || foreign(lc.read_by, g.dst, tt) ||
                 foreign(wrote, g.dst, tt) || foreign(used, g.dst, tt);
      long long src[6];
      const int ns = op_sources(g, src);
      for (int q = 0; q < ns; q++)
        bad = bad || foreign(lc.relaxed_by, src[q], tt) || foreign(wrote, src[q], tt);
      if (bad) {
        why = "post ghost update shared between tiles";
        return false;
      }
    }
  return true;
}

// Parent cell (ci, cj, ck) of the child's coarse index space (0..cmx+1; the
// footprint is 1..cmx) as a q-pool offset of the parent patch.
static long long parent_offset(const std::vector<HPatch> &hp, const HPatch &h, int nd, int ci,
                               int cj, int ck) {
  const HPatch &pa = hp[h.parent];
  long long pi = h.mb[0][0] - 1 + ci, pj = h.mb[1][0] - 1 + cj;
  long long pk = nd == 3 ? h.mb[2][0] - 1 + ck : 0;
  return pa.qoff + qcell(pa.px, pa.py, (int)pi, (int)pj, (int)pk);
}

bool destination_index(const std::vector<GhostOp> &prog, std::vector<int> &at) {
  long long n = 0;
  for (const GhostOp &g : prog) n = std::max(n, (long long)g.dst + 1);
  at.assign((size_t)n, -1);
  for (size_t u = 0; u < prog.size(); u++) {
    int &a = at[(size_t)prog[u].dst];
    if (a >= 0) return false;
    a = (int)u;
  }
  return true;
}

static inline bool listed(const std::vector<int> *only, long long off) {
  return !only || (off >= 0 && off < (long long)only->size() && (*only)[(size_t)off] >= 0);
}

// True if the parent cells (child coarse indices lo..hi per dimension, as
// parent_offset takes them) are all interior cells of the parent: a SetGhost
// program of the parent level writes none of them.
static bool parent_interior(const std::vector<HPatch> &hp, const HPatch &h, int nd,
                            const int lo[3], const int hi[3]) {
  const HPatch &pa = hp[h.parent];
  for (int d = 0; d < nd; d++)
    if (h.mb[d][0] - 1 + lo[d] < 1 || h.mb[d][0] - 1 + hi[d] > pa.mx[d]) return false;
  return true;
}

// The cells (i, j, k) of linear tile `tl` over a box of extents e[0..2]
// (e[2] = 1 in 2D), in order.
template <class F>
static void for_lin_cells(const Tile &tl, const int e[3], F f) {
  const long long n = (long long)e[0] * e[1] * e[2];
  for (long long x = tl.i0; x < std::min<long long>(tl.i0 + LIN_TILE, n); x++)
    f((int)(x % e[0]), (int)((x / e[0]) % e[1]), (int)(x / ((long long)e[0] * e[1])));
}

// The coarse cells tile `tl` of the coarse ghost + loading launch works on:
// a fine tile (res) the interior coarse cells whose children it holds, a ring
// tile its ring cells (ring_cell).
template <class F>
static void for_cgl_cells(const HPatch &h, const Tile &tl, int kind, int nd, F f) {
  const int cz = nd == 3 ? h.cmx[2] : 1;
  if (kind < 0) {   // a linear tile over the ghost-extended coarse box
    const int box[3] = {h.cmx[0] + 2, h.cmx[1] + 2, nd == 3 ? cz + 2 : 1};
    for_lin_cells(tl, box, f);
    return;
  }
  if (kind == 1) {
    const int hx = (1 << cl_lx(nd)) / 2, hy = (1 << cl_ly(nd)) / 2,
              hz = nd == 3 ? (1 << cl_lz(nd)) / 2 : 1;
    for (int c = 0; c < hz; c++)
      for (int b = 0; b < hy; b++)
        for (int a = 0; a < hx; a++) {
          const int i = (tl.i0 + 1) / 2 + a, j = (tl.j0 + 1) / 2 + b,
                    k = nd == 3 ? (tl.k0 + 1) / 2 + c : 0;
          if (i <= h.cmx[0] && j <= h.cmx[1] && (nd == 2 || k <= cz)) f(i, j, k);
        }
    return;
  }
  const long long n = ring_count(nd, h.cmx[0], h.cmx[1], cz);
  for (long long e = tl.i0; e < std::min<long long>(tl.i0 + LIN_TILE, n); e++) {
    int i, j, k;
    ring_cell(nd, h.cmx[0], h.cmx[1], cz, e, i, j, k);
    f(i, j, k);
  }
}

std::vector<std::vector<long long>> coarse_ghost_load_reads(const std::vector<HPatch> &hp,
                                                            const std::vector<Tile> &tiles,
                                                            int nres, int nd,
                                                            const std::vector<int> *only) {
  std::vector<std::vector<long long>> out(tiles.size());
  for (size_t t = 0; t < tiles.size(); t++) {
    const Tile &tl = tiles[t];
    const HPatch &h = hp[tl.patch];
    // -1: linear tile over the coarse box, 1: fine tile, 0: ring tile
    const int kind = nres < 0 ? -1 : (int)t < nres ? 1 : 0;
    const int e[3] = {h.cmx[0] + 1, h.cmx[1] + 1, nd == 3 ? h.cmx[2] + 1 : 0};
    auto ring = [&](int a, int b, int c) {
      return a == 0 || a == e[0] || b == 0 || b == e[1] || (nd == 3 && (c == 0 || c == e[2]));
    };
    std::vector<long long> &r = out[t];
    if (only) {
      // every cell read lies within one cell of the tile's cells
      int lo[3] = {INT_MAX, INT_MAX, nd == 3 ? INT_MAX : 0}, hi[3] = {INT_MIN, INT_MIN, 0};
      for_cgl_cells(h, tl, kind, nd, [&](int i, int j, int k) {
        const int v[3] = {i, j, k};
        for (int d = 0; d < nd; d++) {
          lo[d] = std::min(lo[d], std::max(v[d] - 1, 0));
          hi[d] = std::max(hi[d], std::min(v[d] + 1, e[d]));
        }
      });
      if (lo[0] == INT_MAX || parent_interior(hp, h, nd, lo, hi)) continue;
    }
    for_cgl_cells(h, tl, kind, nd, [&](int i, int j, int k) {
      // a ring cell copies the parent; an interior cell reads its stencil
      // neighbours on the ring from the parent
      const bool inner = !ring(i, j, k);
      for (int s = 0; s < (inner ? 2 * nd + 1 : 1); s++) {
        const int d[7][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0},
                             {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        int a = i + d[s][0], b = j + d[s][1], c = k + d[s][2];
        if (a < 0 || a > e[0] || b < 0 || b > e[1] || c < 0 || c > e[2]) continue;
        if (!ring(a, b, c)) continue;
        long long off = parent_offset(hp, h, nd, a, b, c);
        if (listed(only, off)) r.push_back(off);
      }
    });
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
  }
  return out;
}

std::vector<std::vector<long long>> correct_reads(const std::vector<HPatch> &hp,
                                                  const std::vector<Tile> &tiles, int nd,
                                                  const std::vector<int> *only) {
  std::vector<std::vector<long long>> out(tiles.size());
  for (size_t t = 0; t < tiles.size(); t++) {
    const Tile &tl = tiles[t];
    const HPatch &h = hp[tl.patch];
    const int box[3] = {h.cmx[0] + 1, h.cmx[1] + 1, nd == 3 ? h.cmx[2] + 1 : 1};
    std::vector<long long> &r = out[t];
    if (only) {
      int lo[3] = {INT_MAX, INT_MAX, nd == 3 ? INT_MAX : 0}, hi[3] = {INT_MIN, INT_MIN, 0};
      for_lin_cells(tl, box, [&](int m, int n, int o) {
        const int v[3] = {m, n, o};
        for (int d = 0; d < nd; d++) {
          lo[d] = std::min(lo[d], v[d]);
          hi[d] = std::max(hi[d], v[d] + 1);
        }
      });
      if (parent_interior(hp, h, nd, lo, hi)) continue;
    }
    // block (m, n, o) reads parent cells m + {0, 1} per dimension
    for_lin_cells(tl, box, [&](int m, int n, int o) {
      for (int c = 0; c < (nd == 3 ? 2 : 1); c++)
        for (int b = 0; b < 2; b++)
          for (int a = 0; a < 2; a++) {
            long long off = parent_offset(hp, h, nd, m + a, n + b, nd == 3 ? o + c : 0);
            if (listed(only, off)) r.push_back(off);
          }
    });
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
  }
  return out;
}

RideGhost plan_ride(const std::vector<GhostOp> &prog,
                    const std::vector<std::vector<long long>> &reads) {
  RideGhost out;
  std::vector<int> at;   // destination -> update
  if (!destination_index(prog, at)) return out;   // one update per destination
  std::vector<char> read_somewhere(prog.size(), 0);
  out.start.assign(reads.size() + 1, 0);
  for (size_t t = 0; t < reads.size(); t++) {
    for (long long off : reads[t]) {
      if (off < 0 || off >= (long long)at.size() || at[(size_t)off] < 0) continue;
      int u = at[(size_t)off];
      out.ops.push_back(prog[u]);
      read_somewhere[u] = 1;
    }
    out.start[t + 1] = (int)out.ops.size();
  }
  for (size_t n = 0; n < prog.size(); n++)
    if (!read_somewhere[n]) out.rest.push_back(prog[n]);
  out.ok = true;
  return out;
}

bool program_reads_avoid(const std::vector<GhostOp> &prog,
                         const std::vector<unsigned char> &written) {
  auto hit = [&](long long off) {
    return off >= 0 && off < (long long)written.size() && written[off];
  };
  for (const GhostOp &g : prog) {
    if (hit(g.f0)) return false;
    if (g.kind == 0) continue;
    if (hit(g.f1)) return false;
    int nc = g.kind == 1 ? 1 : g.kind == 2 ? 3 : 4;
    for (int k = 0; k < nc; k++)
      if (hit(g.c[k])) return false;
  }
  return true;
}

bool short_ghost_record(const GhostOp &g, GhostS &out) {
  auto near = [&](int src) {
    const long long d = (long long)src - g.dst;
    return d >= -32768 && d <= 32767;
  };
  if (g.kind == 0) {
    out = GhostS{g.dst, g.f0, 0, 0, g.kind, g.nflip, g.group};
    return true;
  }
  if (g.kind == 1 && near(g.f0) && near(g.f1)) {
    out = GhostS{g.dst, g.c[0], (short)(g.f0 - g.dst), (short)(g.f1 - g.dst), g.kind, g.nflip,
                 g.group};
    return true;
  }
  return false;
}

GhostRecords ghost_records(const std::vector<GhostOp> &ops, const std::vector<int> *start,
                           bool use_short) {
  GhostRecords r;
  auto add = [&](const GhostOp &g) {
    GhostS t;
    if (use_short && short_ghost_record(g, t)) r.s.push_back(t);
    else r.l.push_back(g);
  };
  if (!start) {
    for (const GhostOp &g : ops) add(g);
    return r;
  }
  const size_t nt = start->empty() ? 0 : start->size() - 1;
  r.ss.assign(nt + 1, 0);
  r.sl.assign(nt + 1, 0);
  for (size_t b = 0; b < nt; b++) {
    r.ss[b] = (int)r.s.size();
    r.sl[b] = (int)r.l.size();
    for (int e = (*start)[b]; e < (*start)[b + 1]; e++) add(ops[e]);
  }
  r.ss[nt] = (int)r.s.size();
  r.sl[nt] = (int)r.l.size();
  return r;
}

void set_op_groups(const std::vector<HPatch> &hp, std::vector<GhostOp> &ops) {
  std::vector<long long> qoffs(hp.size());
  for (size_t n = 0; n < hp.size(); n++) qoffs[n] = hp[n].qoff;
  for (GhostOp &g : ops) {
    auto it = std::upper_bound(qoffs.begin(), qoffs.end(), (long long)g.dst);
    const int p = (int)(it - qoffs.begin()) - 1;
    g.group = (short)(p >= 0 ? hp[p].group : 0);
  }
}

void locate_q_cell(const std::vector<HPatch> &hp, long long off, int &p, int &i, int &j,
                   int &k) {
  p = -1;
  for (size_t n = 0; n < hp.size(); n++)
    if (off >= hp[n].qoff && off < hp[n].qoff + hp[n].qn) {
      p = (int)n;
      break;
    }
  if (p < 0) {
    i = j = k = -1;
    return;
  }
  const HPatch &h = hp[p];
  qcell_index(off - h.qoff, h.px, h.py, i, j, k);
}

void apply_ghost_program_host(const std::vector<HPatch> &hp, const std::vector<GhostOp> &ops,
                              int nrvars, double *const *q_patches, const double *qpo) {
  // a device offset's element in the patch's reference-layout host array,
  // and that array's variable stride (the device's may hold padding)
  auto host_index = [&](long long off, int &p) {
    int i, j, k;
    locate_q_cell(hp, off, p, i, j, k);
    const HPatch &h = hp[p];
    return i + h.px * (j + h.py * (long long)k);
  };
  auto host_vs = [&](int p) {
    const HPatch &h = hp[p];
    return h.px * h.py * (h.qn / split_of(h.px, h.py).pl);
  };
  auto val = [&](long long off, int v) {
    int p;
    long long at = host_index(off, p);
    return q_patches[p][at + (long long)v * host_vs(p)];
  };
  std::vector<double> res(ops.size() * nrvars);
  for (size_t e = 0; e < ops.size(); e++) {
    const GhostOp &g = ops[e];
    for (int v = 0; v < nrvars; v++) {
      double x;
      switch (g.kind) {
        case 0: x = val(g.f0, v); break;
        case 1:
          x = BSAM_A1 * val(g.f0, v) + BSAM_A2 * val(g.f1, v) +
              BSAM_A3 * val(g.c[0], v);
          break;
        case 2:
          x = BSAM_A1 * val(g.f0, v) + BSAM_A2 * val(g.f1, v) +
              BSAM_A3 * (BSAM_B1 * val(g.c[0], v) + BSAM_B2 * val(g.c[1], v) +
                         BSAM_B3 * val(g.c[2], v));
          break;
        default:
          x = BSAM_A1 * val(g.f0, v) + BSAM_A2 * val(g.f1, v) +
              BSAM_A3 * (BSAM_C1 * val(g.c[0], v) + BSAM_C2 * val(g.c[1], v) +
                         BSAM_C3 * val(g.c[2], v) + BSAM_C4 * val(g.c[3], v));
          break;
      }
      x = add_qpo<double>(x, qpo_code(g.nflip), Qpo{qpo[0], qpo[1]}, v);
      for (int r = 0; r < flip_count(g.nflip); r++) x = 2.0 - x;
      res[e * nrvars + v] = x;
    }
  }
  for (size_t e = 0; e < ops.size(); e++) {
    const GhostOp &g = ops[e];
    int p;
    long long at = host_index(g.dst, p);
    for (int v = 0; v < nrvars; v++)
      q_patches[p][at + (long long)v * host_vs(p)] = res[e * nrvars + v];
  }
}

}  // namespace bsam
