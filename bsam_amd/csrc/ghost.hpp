// ghost.hpp -- host-side integer work of the ghost fill: builds, per level and
// colour pass, the list of ghost-cell updates that SetGhost performs.
#pragma once
#include <string>
#include <vector>

#include "common.hpp"

namespace bsam {

// Host record of one patch (the geometry of nodeinfo, nodeinfodef.f90:124-219)
// plus its q-pool placement, filled at commit.
struct HPatch {
  int id = 0, level = 0, parent = -1;
  int mx[3] = {1, 1, 1}, cmx[3] = {1, 1, 1};
  int mb[3][2] = {{1, 1}, {1, 1}, {1, 1}};
  int mg[3][2] = {{1, 1}, {1, 1}, {1, 1}};
  int mthbc[6] = {1, 1, 1, 1, 1, 1};
  double dx = 0.0, xlower[3] = {0.0, 0.0, 0.0};
  // q-pool offset of the first variable, cells per variable, row pitch and
  // rows per plane of the ghost-extended q
  long long qoff = 0, qn = 0, px = 0, py = 0;
  int group = 0;   // independent sub-problem the patch belongs to
};

// Periodic set-up of the hierarchy (PeriodicSetup, boundary.f90:893-963):
// the offset vectors in use (GetPeriodicOffsets numbering 1..13) and the root
// extents that scale them (mxmax = root mx * 2^level).
struct Periodic {
  int noff = 0;
  int idx[13] = {0};
  int rootmx[3] = {1, 1, 1};
};
Periodic periodic_setup(const HPatch &root, int ndims);

// Builds the ghost-fill program of SetGhost(level, ipass) (boundary.f90:58-89)
// for the patches of one level. Returns 0 on success, else fills msg.
int build_ghost_program(const std::vector<HPatch> &hp, const std::vector<int> &patches,
                        int ndims, int level, int ipass, const Periodic &per,
                        std::vector<GhostOp> &out, std::string &msg);

// A SetGhost program of the parent level applied in full inside a level
// kernel of the child level that reads some of its destinations and writes
// none of its sources (the child's coarse ghost + loading launch takes the
// parent's SetGhost(0), its correction launch the parent's last SetGhost of
// post-smoothing). Every update whose destination a tile reads is listed for
// that tile (start/ops; an update may be listed for several tiles, each
// writing the same value) and applied by it before its own work; the others
// (rest) are applied grid-stride by all the launch's workgroups.
struct RideGhost {
  bool ok = false;
  std::vector<int> start;
  std::vector<GhostOp> ops;
  std::vector<GhostOp> rest;
};
// Linear tiles: LIN_TILE consecutive cells (x fastest) of a box of one
// patch, the first at cell Tile::i0 of the box.
constexpr int LIN_TILE = 256;
// Parent q-pool offsets (variable 1) that each tile of the child level's
// coarse ghost + loading launch (fine tiles tiles[0..nres), then linear tiles
// over the qc ring cells; with nres < 0, linear tiles over the ghost-extended
// coarse box) or correction launch (linear tiles over the
// 2x2(x2) fine blocks 0..cmx) may read, ascending; with `only`, just the
// offsets off with only[off] >= 0.
std::vector<std::vector<long long>> coarse_ghost_load_reads(const std::vector<HPatch> &hp,
                                                            const std::vector<Tile> &tiles,
                                                            int nres, int ndims,
                                                            const std::vector<int> *only = nullptr);
std::vector<std::vector<long long>> correct_reads(const std::vector<HPatch> &hp,
                                                  const std::vector<Tile> &tiles, int ndims,
                                                  const std::vector<int> *only = nullptr);
// at[off] = the update of `prog` that writes q-pool offset off, -1 where
// none; false if two updates share a destination.
bool destination_index(const std::vector<GhostOp> &prog, std::vector<int> &at);
RideGhost plan_ride(const std::vector<GhostOp> &prog,
                    const std::vector<std::vector<long long>> &reads);

// The ghost-cell updates of SetGhost(level, ipass) that a relaxation kernel
// for colour `color` applies before relaxing, grouped by the tile that
// applies them (ops[start[t] .. start[t+1]) for tile t of the level's
// interior tile list). Updates whose destination the relaxation does not read
// and that the next program, SetGhost(level, color), rewrites are dropped
// (their value is never observed). A kept update goes to the tile that
// relaxes its sources of colour `color` (it reads them before they are
// relaxed, as SetGhost before RelaxPatch does), or, if it has none, to the
// tile of the interior cell its destination borders; an update whose
// destination the relaxation reads goes to the reading tile. An update that
// fits none of these (a source relaxed by another tile than the one reading
// the destination, or sources in several tiles) goes to `pre`, a program of
// its own the engine launches just before the fused relaxation. The updates
// in `ops` read nothing another tile of the launch writes and write nothing
// another tile reads, so the tiles never race.
struct FusedGhost {
  bool ok = false;
  int kept = 0;                // ops + pre
  std::vector<int> start;
  std::vector<GhostOp> ops;
  std::vector<GhostOp> pre;
};
// tiles: interior tiles of the level (origin i0/j0/k0 = 1 + multiples of the
// tile extents tx, ty, tz; k0 = 0 in 2D), in patch order, k, j, i nested.
FusedGhost plan_fused_ghost(const std::vector<HPatch> &hp, const std::vector<Tile> &tiles,
                            int ndims, int tx, int ty, int tz,
                            const std::vector<GhostOp> &prog, const std::vector<GhostOp> &next,
                            int color, std::string &why);
// Updates of SetGhost(level, color) that the relaxation launch of `color`
// just before the one they are read in can apply at its end, each in a tile
// after that tile's relaxation and a workgroup barrier (`cand`: the `pre`
// updates of the next launch's fused plan; `before`: the fused plan applied
// inside this launch, or null). An update goes to the tile that relaxes its
// fine sources of colour `color` (it must read them relaxed), or, with none,
// to the tile of the interior cell its destination borders. It stays in
// `rest` (still launched before the next relaxation) if its sources of that
// colour lie in several tiles, if this launch's relaxation reads its
// destination in another tile, or if an in-launch update of another tile
// writes one of its sources or writes or reads its destination. Between this
// launch and the next nothing else touches the level, and the next launch
// relaxes the other colour, so the updates read what SetGhost(level, color)
// would.
struct PostGhost {
  bool ok = false;
  std::vector<int> start;
  std::vector<GhostOp> ops;
  std::vector<GhostOp> rest;
};
PostGhost plan_post_ghost(const std::vector<HPatch> &hp, const std::vector<Tile> &tiles, int nd,
                          int tx, int ty, int tz, int color, const FusedGhost *before,
                          const std::vector<GhostOp> &cand, std::string &why);
// Race check of a post plan (run by plan_post_ghost before it accepts one),
// from the cells' side: each tile's updates read cells relaxed by that tile
// or by none, and cells no other tile's in-launch update writes; they write
// ghost cells no other tile's stencil or in-launch update reads or writes.
bool post_plan_race_free(const std::vector<HPatch> &hp, const std::vector<Tile> &tiles, int nd,
                         int tx, int ty, int tz, int color, const FusedGhost *before,
                         const PostGhost &plan, std::string &why);

// Race check of a plan (run by plan_fused_ghost before it accepts one): each
// tile's updates read only cells that tile relaxes (or cells the launch does
// not relax) and write only ghost cells that no other tile's stencil reads.
bool fused_plan_race_free(const std::vector<HPatch> &hp, const std::vector<Tile> &tiles, int nd,
                          int tx, int ty, int tz, int color, const FusedGhost &plan,
                          std::string &why);

// True if no update of `prog` reads a q-pool cell (variable-1 offset) in
// `written`: the program may then run in the same launch as a kernel that
// writes exactly those cells.
bool program_reads_avoid(const std::vector<GhostOp> &prog,
                         const std::vector<unsigned char> &written);

// The device records of a list of updates: the 16-byte form (GhostS) of each
// update that has one (use_short) and the 32-byte form of the others, each in
// the list's order. With per-tile groups (start: nt + 1 entries), ss and sl
// are the per-tile starts into s and l, every update staying with its tile.
struct GhostRecords {
  std::vector<GhostS> s;
  std::vector<GhostOp> l;
  std::vector<int> ss, sl;
};
// The 16-byte form of g, if it has one: a copy, or a kind-1 interpolation
// whose fine sources lie within 16-bit offsets of the destination.
bool short_ghost_record(const GhostOp &g, GhostS &out);
GhostRecords ghost_records(const std::vector<GhostOp> &ops, const std::vector<int> *start,
                           bool use_short);

// Sets every update's group to its destination patch's group.
void set_op_groups(const std::vector<HPatch> &hp, std::vector<GhostOp> &ops);

// Patch and (i,j,k) (lower bound 0; k = 0 in 2D) of a q-pool offset.
void locate_q_cell(const std::vector<HPatch> &hp, long long off, int &p, int &i, int &j,
                   int &k);

// Evaluates a program on host arrays (one reference-layout q array per patch),
// reading the state before the program as the device kernel does.
void apply_ghost_program_host(const std::vector<HPatch> &hp, const std::vector<GhostOp> &ops,
                              int nrvars, double *const *q_patches, const double *qpo);

}  // namespace bsam
