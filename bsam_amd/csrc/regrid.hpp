// regrid.hpp -- host-side integer work of the regrid of one level: the
// buffering of tagged cells and their list (BufferAndList), the inflation of
// edge tags into neighbouring and periodic-image patches (InflateEdgeTags)
// and Berger-Rigoutsos clustering of each patch's flags into new subgrids
// (NewSubGrids).
#pragma once
#include <string>
#include <vector>

#include "ghost.hpp"

namespace bsam {

struct RegridPatch {
  int mx[3] = {1, 1, 1};
  int mg[3][2] = {{1, 1}, {1, 1}, {1, 1}};
  int *flags = nullptr;   // errorflags(1:mx(1), 1:mx(2), 1:mx(3)), Fortran order
};

struct RegridParams {
  int ndims = 2, level = 0;
  int ibuffer = 0;              // ibuffer(level)
  int mgp = 2;                  // minimumgridpoints(level)
  double desfill = 0.8;         // desiredfillratios(level)
  Periodic per;                 // PeriodicSetup of the root
};

struct NewGrid {
  int parent = 0;               // index into the level's patch list
  int mb[3][2] = {{1, 1}, {1, 1}, {1, 1}};
};

// Buffers every patch's flags and inflates the edge tags (ErrFlag's
// BufferAndList, lib/bsamroutines.f90:1159-1205, and InflateEdgeTags
// :1207-1280), then clusters each patch's flags into subgrids (RefineGrid /
// NewSubGrids :1334-1579) in level-list order. Returns 0, or -1 with msg.
int regrid_level(const RegridParams &rp, std::vector<RegridPatch> &patches,
                 std::vector<NewGrid> &out, std::string &msg);

}  // namespace bsam
