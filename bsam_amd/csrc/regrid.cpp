// regrid.cpp -- buffering, edge-tag inflation and Berger-Rigoutsos
// clustering of one level's error flags into new subgrids; every step keeps
// the reference's loop order, so the patch lists come out identical.
#include "regrid.hpp"

#include <algorithm>
#include <array>
#include <cstdlib>

namespace bsam {
namespace {

constexpr int kMaxSubgrids = 1024;     // maxsubgrids (nodeinfodef.f90:59)
constexpr int kMaxSplitPasses = 15;    // NewSubGrids' maxsplitpasses

using Box = std::array<std::array<int, 2>, 3>;   // [dim][lo, hi], 1-based

inline long long fidx(const RegridPatch &p, int i, int j, int k) {
  return (i - 1) + (long long)p.mx[0] * ((j - 1) + (long long)p.mx[1] * (k - 1));
}

// BufferTaggedCells (bsamroutines.f90:1240-1280): flags the cells of every
// patch inside the buffer box of the global cell c.
void buffer_tag(const RegridParams &rp, std::vector<RegridPatch> &ps, const int c[3]) {
  const int nd = rp.ndims, ib = rp.ibuffer;
  for (RegridPatch &p : ps) {
    int lo[3] = {1, 1, 1}, hi[3] = {1, 1, 1};
    bool empty = false;
    for (int d = 0; d < nd; d++) {
      lo[d] = std::max(c[d] - ib, p.mg[d][0]);
      hi[d] = std::min(c[d] + ib, p.mg[d][1]);
      if (hi[d] < lo[d]) empty = true;
    }
    if (empty) continue;
    int l[3] = {1, 1, 1}, h[3] = {1, 1, 1};
    for (int d = 0; d < nd; d++) {
      l[d] = lo[d] - p.mg[d][0] + 1;
      h[d] = hi[d] - p.mg[d][0] + 1;
    }
    for (int k = l[2]; k <= h[2]; k++)
      for (int j = l[1]; j <= h[1]; j++)
        for (int i = l[0]; i <= h[0]; i++) p.flags[fidx(p, i, j, k)] = 1;
  }
}

// GetPeriodicTagOffset (boundary.f90:1046-1201): the periodic image of a tag
// whose buffer box crosses the domain edge along offset vector `off`
// (numbering of GetPeriodicOffsets); false if it does not cross there.
bool periodic_tag_offset(const RegridParams &rp, int c[3], int off) {
  const int nd = rp.ndims, ib = rp.ibuffer;
  int m[3][2] = {{1, 1}, {1, 1}, {1, 1}};
  int mm[3] = {1, 1, 1};   // mxmax(level, :)
  for (int d = 0; d < nd; d++) {
    m[d][0] = c[d] - ib;
    m[d][1] = c[d] + ib;
    mm[d] = rp.per.rootmx[d] << rp.level;
  }
  // each case tests the upper crossing first, then the lower one; signs
  // s[d] = +1 (upper edge, shift down), -1 (lower edge, shift up), 0 unused
  auto hi = [&](int d) { return m[d][1] > mm[d]; };
  auto lo = [&](int d) { return m[d][0] < 1; };
  auto shift = [&](int d, int s) { c[d] += s > 0 ? -mm[d] : mm[d]; };
  switch (off) {
    case 1: if (hi(0)) shift(0, 1); else if (lo(0)) shift(0, -1); else return false; break;
    case 2: if (hi(1)) shift(1, 1); else if (lo(1)) shift(1, -1); else return false; break;
    case 3: if (hi(2)) shift(2, 1); else if (lo(2)) shift(2, -1); else return false; break;
    case 4:
      if (hi(0) && hi(1)) { shift(0, 1); shift(1, 1); }
      else if (lo(0) && lo(1)) { shift(0, -1); shift(1, -1); }
      else return false;
      break;
    case 5:
      if (hi(0) && lo(1)) { shift(0, 1); shift(1, -1); }
      else if (lo(0) && hi(1)) { shift(0, -1); shift(1, 1); }
      else return false;
      break;
    case 6:
      if (hi(1) && hi(2)) { shift(1, 1); shift(2, 1); }
      else if (lo(1) && lo(2)) { shift(1, -1); shift(2, -1); }
      else return false;
      break;
    case 7:
      if (hi(1) && lo(2)) { shift(1, 1); shift(2, -1); }
      else if (lo(1) && hi(2)) { shift(1, -1); shift(2, 1); }
      else return false;
      break;
    case 8:
      if (hi(0) && hi(2)) { shift(0, 1); shift(2, 1); }
      else if (lo(0) && lo(2)) { shift(0, -1); shift(2, -1); }
      else return false;
      break;
    case 9:
      // the reference compares the x3 upper edge with mxmax(level, 2) here
      if (hi(0) && lo(2)) { shift(0, 1); shift(2, -1); }
      else if (lo(0) && m[2][1] > mm[1]) { shift(0, -1); shift(2, 1); }
      else return false;
      break;
    case 10:
      if (hi(0) && hi(1) && hi(2)) { shift(0, 1); shift(1, 1); shift(2, 1); }
      else if (lo(0) && lo(1) && lo(2)) { shift(0, -1); shift(1, -1); shift(2, -1); }
      else return false;
      break;
    case 11:
      if (hi(0) && hi(1) && lo(2)) { shift(0, 1); shift(1, 1); shift(2, -1); }
      else if (lo(0) && lo(1) && hi(2)) { shift(0, -1); shift(1, -1); shift(2, 1); }
      else return false;
      break;
    case 12:
      if (lo(0) && hi(1) && hi(2)) { shift(0, -1); shift(1, 1); shift(2, 1); }
      else if (hi(0) && lo(1) && lo(2)) { shift(0, 1); shift(1, -1); shift(2, -1); }
      else return false;
      break;
    case 13:
      if (hi(0) && lo(1) && hi(2)) { shift(0, 1); shift(1, -1); shift(2, 1); }
      else if (lo(0) && hi(1) && lo(2)) { shift(0, -1); shift(1, 1); shift(2, -1); }
      else return false;
      break;
    default: return false;
  }
  return true;
}

// GridFlagRatio (bsamroutines.f90:1614-1643): flagged cells / cells of the box.
double flag_ratio(const RegridPatch &p, int nd, const Box &b) {
  long long total = 1, flagged = 0;
  for (int d = 0; d < nd; d++) total *= b[d][1] - b[d][0] + 1;
  int k0 = nd == 3 ? b[2][0] : 1, k1 = nd == 3 ? b[2][1] : 1;
  for (int k = k0; k <= k1; k++)
    for (int j = b[1][0]; j <= b[1][1]; j++)
      for (int i = b[0][0]; i <= b[0][1]; i++) flagged += p.flags[fidx(p, i, j, k)];
  return (double)flagged / (double)total;
}

// GetSignatures (bsamroutines.f90:1581-1612): per dimension n and index i of
// the box, the number of flags in the box's slab at i.
void signatures(const RegridPatch &p, int nd, const Box &b, std::vector<long long> sig[3]) {
  int k0 = nd == 3 ? b[2][0] : 1, k1 = nd == 3 ? b[2][1] : 1;
  for (int n = 0; n < nd; n++)
    for (int i = b[n][0]; i <= b[n][1]; i++) sig[n][(size_t)i] = 0;
  for (int k = k0; k <= k1; k++)
    for (int j = b[1][0]; j <= b[1][1]; j++)
      for (int i = b[0][0]; i <= b[0][1]; i++) {
        int f = p.flags[fidx(p, i, j, k)];
        if (!f) continue;
        sig[0][(size_t)i] += f;
        sig[1][(size_t)j] += f;
        if (nd == 3) sig[2][(size_t)k] += f;
      }
}

// NewSubGrids (bsamroutines.f90:1352-1579), Berger-Rigoutsos on one patch's
// flags over its whole interior.
int new_subgrids(const RegridParams &rp, const RegridPatch &p, std::vector<Box> &grids,
                 std::string &msg) {
  const int nd = rp.ndims, mgp = rp.mgp;
  Box mb;
  for (int d = 0; d < 3; d++) mb[d] = {1, d < nd ? p.mx[d] : 1};
  grids.clear();
  if (flag_ratio(p, nd, mb) < 1.0e-10) return 0;
  int maxm = 0;
  for (int d = 0; d < nd; d++) maxm = std::max(maxm, p.mx[d]);
  std::vector<long long> sig[3], dd[3];
  for (int n = 0; n < nd; n++) {
    sig[n].assign((size_t)maxm + 2, 0);
    dd[n].assign((size_t)maxm + 2, 0);
  }
  std::vector<Box> sub{mb};
  std::vector<char> cansplit{1};
  size_t igrid = 0;
  while ((int)sub.size() < kMaxSubgrids && igrid < sub.size()) {
    int npass = 0;
    while (cansplit[igrid] && npass < kMaxSplitPasses) {
      npass++;
      signatures(p, nd, sub[igrid], sig);
      // trim unflagged slabs on the box's edges
      for (int n = 0; n < nd; n++) {
        const int lo = sub[igrid][n][0], hi = sub[igrid][n][1];
        int i1 = lo, i2 = hi;
        while (sig[n][(size_t)i1] == 0 && i1 < hi && i2 - i1 + 1 > mgp) i1++;
        while (sig[n][(size_t)i2] == 0 && i2 > lo && i2 - i1 + 1 > mgp) i2--;
        sub[igrid][n] = {i1, i2};
      }
      const double fill = flag_ratio(p, nd, sub[igrid]);
      if (!(fill < rp.desfill)) {
        cansplit[igrid] = 0;
        continue;
      }
      signatures(p, nd, sub[igrid], sig);
      int isplit[3] = {0, 0, 0};
      bool have = false;
      // a hole along which to split
      for (int n = 0; n < nd && !have; n++) {
        const int i1 = sub[igrid][n][0], i2 = sub[igrid][n][1];
        for (int i = i1 + mgp; i <= i2 - mgp; i++)
          if (sig[n][(size_t)i] == 0 && std::min(i - i1 + 1, i2 - i) >= mgp) {
            isplit[n] = i;
            have = true;
            break;
          }
      }
      if (!have) {
        // the strongest inflection of the signature
        for (int n = 0; n < nd; n++) {
          const int i1 = sub[igrid][n][0], i2 = sub[igrid][n][1];
          for (int i = i1 + 1; i <= i2 - 1; i++)
            dd[n][(size_t)i] = sig[n][(size_t)i - 1] - 2 * sig[n][(size_t)i] + sig[n][(size_t)i + 1];
        }
        long long inflect = 0;
        int dist0 = 0;
        for (int n = 0; n < nd; n++) {
          const int i1 = sub[igrid][n][0], i2 = sub[igrid][n][1];
          for (int i = i1 + std::max(2, mgp); i <= i2 - std::max(1, mgp - 1); i++) {
            long long del = std::llabs(dd[n][(size_t)i] - dd[n][(size_t)i - 1]);
            if (del > inflect) {
              inflect = del;
              isplit[0] = isplit[1] = isplit[2] = 0;
              isplit[n] = i - 1;
              have = true;
              dist0 = std::min(i - i1, i2 - i + 1);
            } else if (del == inflect && inflect > 0) {
              int dist1 = std::min(i - i1, i2 - i + 1);
              if (dist1 > dist0) {
                isplit[0] = isplit[1] = isplit[2] = 0;
                isplit[n] = i - 1;
                have = true;
                dist0 = dist1;
              }
            }
          }
        }
      }
      if (!have) {
        cansplit[igrid] = 0;
        continue;
      }
      for (int n = 0; n < nd; n++) {
        const int s = isplit[n];
        if (s > 0 && std::min(sub[igrid][n][1] - s, s - sub[igrid][n][0] + 1) >= mgp) {
          if ((int)sub.size() >= kMaxSubgrids) {
            msg = "NewSubGrids: more than maxsubgrids subgrids";
            return -1;
          }
          Box nb = sub[igrid];
          nb[n][0] = s + 1;
          sub[igrid][n][1] = s;
          sub.push_back(nb);
          cansplit.push_back(1);
          break;
        }
      }
    }
    igrid++;
  }
  for (const Box &b : sub) {
    int minside = 1 << 30;
    for (int d = 0; d < nd; d++) minside = std::min(minside, b[d][1] - b[d][0] + 1);
    if (minside < mgp) {
      msg = "NewSubGrids: grid smaller than minimumgridpoints";
      return -1;
    }
  }
  grids = sub;
  return 0;
}

}  // namespace

int regrid_level(const RegridParams &rp, std::vector<RegridPatch> &ps, std::vector<NewGrid> &out,
                 std::string &msg) {
  const int nd = rp.ndims, ib = rp.ibuffer;
  out.clear();
  if (nd < 2 || nd > 3) {
    msg = "regrid: ndims must be 2 or 3";
    return -1;
  }
  if (rp.mgp < 1 || ib < 0) {
    msg = "regrid: minimumgridpoints >= 1 and ibuffer >= 0 required";
    return -1;
  }
  // BufferAndList per patch, in level-list order: tags whose buffer box
  // leaves the patch go to the list (global coordinates), the others are
  // buffered in place, both read from the flags before buffering
  std::vector<std::array<int, 3>> tags;
  if (ib > 0) {
    for (RegridPatch &p : ps) {
      const long long n = (long long)p.mx[0] * p.mx[1] * p.mx[2];
      std::vector<int> tmp(p.flags, p.flags + n);
      for (int k = 1; k <= p.mx[2]; k++)
        for (int j = 1; j <= p.mx[1]; j++)
          for (int i = 1; i <= p.mx[0]; i++) {
            if (tmp[(size_t)fidx(p, i, j, k)] != 1) continue;
            const int idx[3] = {i, j, k};
            int lo[3] = {1, 1, 1}, hi[3] = {1, 1, 1};
            bool edge = false;
            for (int d = 0; d < nd; d++) {
              lo[d] = idx[d] - ib;
              hi[d] = idx[d] + ib;
            }
            for (int d = 0; d < 3; d++) edge = edge || lo[d] < 1 || hi[d] > p.mx[d];
            if (edge) {
              std::array<int, 3> c{1, 1, 1};
              for (int d = 0; d < nd; d++) c[d] = idx[d] + p.mg[d][0] - 1;
              tags.push_back(c);
            } else {
              for (int kk = lo[2]; kk <= hi[2]; kk++)
                for (int jj = lo[1]; jj <= hi[1]; jj++)
                  for (int ii = lo[0]; ii <= hi[0]; ii++) p.flags[fidx(p, ii, jj, kk)] = 1;
            }
          }
    }
  }
  // InflateEdgeTags: the list from its newest tag to its oldest
  for (size_t t = tags.size(); t-- > 0;) {
    int c[3] = {tags[t][0], tags[t][1], tags[t][2]};
    buffer_tag(rp, ps, c);
    for (int o = 0; o < rp.per.noff; o++) {
      int cc[3] = {tags[t][0], tags[t][1], tags[t][2]};
      if (periodic_tag_offset(rp, cc, rp.per.idx[o])) buffer_tag(rp, ps, cc);
    }
  }
  // RefineGrid on every patch of the level, in level-list order
  std::vector<Box> grids;
  for (size_t k = 0; k < ps.size(); k++) {
    if (new_subgrids(rp, ps[k], grids, msg)) return -1;
    for (const Box &b : grids) {
      NewGrid g;
      g.parent = (int)k;
      for (int d = 0; d < 3; d++) {
        g.mb[d][0] = d < nd ? b[d][0] : 1;
        g.mb[d][1] = d < nd ? b[d][1] : 1;
      }
      out.push_back(g);
    }
  }
  return 0;
}

}  // namespace bsam
