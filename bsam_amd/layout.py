"""Patch hierarchy geometry: the host-side integer work of the BSAM tree.

A :class:`Layout` records the patches of a block-structured AMR hierarchy the
way the reference tree builds them, and yields them in the order the
reference traverses them:

* the root patch (level 0) is read from grid data as in ``RootInit``
  (lib/bsamroutines.f90:205-424), including the automatic choice of
  ``minlevel`` when it is given as 1 (:287-301);
* each below-root level is one patch half the size of the level above, as
  built by ``InitSeed`` (lib/bsamroutines.f90:426-494);
* refined patches follow ``MakeNewGrid`` (lib/bsamroutines.f90:1581-1683):
  ``mx = 2*(mb2-mb1+1)``, ``mglobal`` from the parent's, and a physical
  boundary code inherited only where the child touches that parent edge
  (internal code 999 elsewhere);
* every level is traversed youngest patch first, because ``CreateChild``
  prepends to the level list (lib/treeops.f90:339-430, ApplyOnLevel
  :536-572).

All of it is integer (plus the derived dx/xlower) and must agree exactly with
the reference; ``tests/test_layout.py`` checks it against the reference
tree's own listing.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

INTERNAL_BC = 999
MAX_DEPTH = 10


@dataclass
class Patch:
    id: int
    level: int
    parent_id: Optional[int]
    mx: Tuple[int, int, int]
    mbounds: Tuple[Tuple[int, int], ...]
    mglobal: Tuple[Tuple[int, int], ...]
    mthbc: Tuple[int, ...]
    dx: float
    xlower: Tuple[float, float, float]
    children: List[int] = field(default_factory=list)

    def cells(self, ndims: int) -> int:
        n = 1
        for d in range(ndims):
            n *= self.mx[d]
        return n


def auto_minlevel(mx: Sequence[int], ndims: int) -> int:
    """Depth of the below-root chain when minlevel = 1 (lib/bsamroutines.f90:287-301)."""
    m = list(mx[:ndims])
    lev = 0
    while all(v % 2 == 0 for v in m) and lev < MAX_DEPTH:
        lev += 1
        m = [v // 2 for v in m]
    return -lev


class Layout:
    """A block-structured hierarchy of patches over a rectangular domain."""

    def __init__(self, ndims: int, mx: Sequence[int], dx: float,
                 xlower: Sequence[float] = (0.0, 0.0, 0.0),
                 mthbc: Sequence[int] = (1, 1, 1, 1, 1, 1),
                 minlevel: int = 1, nrvars: int = 2, mbc: int = 1):
        if ndims not in (2, 3):
            raise ValueError("ndims must be 2 or 3")
        if mbc != 1:
            raise ValueError("only mbc = 1 is supported")
        mx = tuple(int(v) for v in mx[:ndims]) + (1,) * (3 - ndims)
        for d in range(ndims):
            if mx[d] <= 0 or mx[d] % 2:
                raise ValueError("root mx must be positive and even")
        if minlevel > 1:
            raise ValueError("minlevel must be <= 1")
        if minlevel == 1:
            minlevel = auto_minlevel(mx, ndims)
        self.ndims, self.nrvars, self.mbc = ndims, nrvars, mbc
        self.minlevel = minlevel
        xl = tuple(float(v) for v in xlower[:ndims]) + (0.0,) * (3 - ndims)
        bc = tuple(int(v) for v in mthbc[:2 * ndims]) + (1,) * (6 - 2 * ndims)
        self.patches: Dict[int, Patch] = {}
        self.creation: Dict[int, List[int]] = {}
        root = Patch(0, 0, None, mx,
                     tuple((1, mx[d]) for d in range(3)),
                     tuple((1, mx[d]) if d < ndims else (1, 1) for d in range(3)),
                     bc, float(dx), xl)
        self.patches[0] = root
        self.creation[0] = [0]
        child = root
        for lev in range(-1, minlevel - 1, -1):
            if any(child.mx[d] % 2 for d in range(ndims)):
                raise ValueError(f"grid on level {child.level} will not coarsen")
            smx = tuple(child.mx[d] // 2 if d < ndims else 1 for d in range(3))
            seed = Patch(lev, lev, None, smx,
                         tuple((1, smx[d]) if d < ndims else (1, 1) for d in range(3)),
                         tuple((1, child.mglobal[d][1] // 2) if d < ndims else (1, 1)
                               for d in range(3)),
                         child.mthbc, child.dx * 2.0, child.xlower)
            child.mbounds = tuple((1, smx[d]) if d < ndims else (1, 1) for d in range(3))
            child.parent_id = lev
            seed.children.append(child.id)
            self.patches[lev] = seed
            self.creation[lev] = [lev]
            child = seed
        self._next_id = 1

    # -- construction -----------------------------------------------------
    def add_child(self, parent_id: int, mbounds: Sequence[Tuple[int, int]]) -> int:
        """Create a refined patch over ``mbounds`` (parent-local coarse cells)."""
        par = self.patches[parent_id]
        nd = self.ndims
        if par.level < 0:
            raise ValueError("children are created on levels >= 0 only")
        mb = tuple((int(mbounds[d][0]), int(mbounds[d][1])) if d < nd else (1, 1)
                   for d in range(3))
        for d in range(nd):
            if not (1 <= mb[d][0] <= mb[d][1] <= par.mx[d]):
                raise ValueError(f"mbounds {mb} outside parent {parent_id}")
        cmx = tuple(2 * (mb[d][1] - mb[d][0] + 1) if d < nd else 1 for d in range(3))
        mg = []
        for d in range(3):
            if d >= nd:
                mg.append((1, 1))
                continue
            g1 = par.mglobal[d][0] + mb[d][0] - 1
            g2 = g1 + mb[d][1] - mb[d][0]
            mg.append(((g1 - 1) * 2 + 1, g2 * 2))
        # all six codes start internal, as in MakeNewGrid; only the first
        # 2*ndims can inherit a physical code
        bc = [INTERNAL_BC] * 6
        for d in range(nd):
            if par.mthbc[2 * d] < INTERNAL_BC and mb[d][0] == 1:
                bc[2 * d] = par.mthbc[2 * d]
            if par.mthbc[2 * d + 1] < INTERNAL_BC and mb[d][1] == par.mx[d]:
                bc[2 * d + 1] = par.mthbc[2 * d + 1]
        xl = tuple(float(mb[d][0] - 1) * par.dx + par.xlower[d] if d < nd else 0.0
                   for d in range(3))
        pid = self._next_id
        self._next_id += 1
        p = Patch(pid, par.level + 1, parent_id, cmx, mb, tuple(mg), tuple(bc),
                  par.dx / 2.0, xl)
        self.patches[pid] = p
        par.children.append(pid)
        self.creation.setdefault(p.level, []).append(pid)
        return pid

    # -- traversal --------------------------------------------------------
    @property
    def finestlevel(self) -> int:
        return max(self.creation)

    def level_ids(self, level: int) -> List[int]:
        """Patch ids of a level in level-list (ApplyOnLevel) order: youngest first."""
        return list(reversed(self.creation.get(level, [])))

    def ordered(self) -> List[Patch]:
        """All patches, levels ascending, each level in level-list order."""
        out = []
        for lev in range(self.minlevel, self.finestlevel + 1):
            out.extend(self.patches[i] for i in self.level_ids(lev))
        return out

    def index_map(self) -> Dict[int, int]:
        return {p.id: n for n, p in enumerate(self.ordered())}

    def refined(self) -> List[Patch]:
        """Refined patches (level >= 1) in creation order."""
        out = []
        for lev in range(1, self.finestlevel + 1):
            out.extend(self.patches[i] for i in self.creation[lev])
        return out

    def cells_on_level(self, level: int) -> int:
        return sum(self.patches[i].cells(self.ndims) for i in self.level_ids(level))

    def field_shape(self, p: Patch, field: str) -> Tuple[int, int, int, int]:
        """Fortran-order array shape of a patch field (AllocFields)."""
        nd, mbc, nv = self.ndims, self.mbc, self.nrvars
        if field in ("q", "qold"):
            return tuple(p.mx[d] + 2 * mbc if d < nd else 1 for d in range(3)) + (nv,)
        if field == "qc":
            return tuple((1 if p.mx[d] % 2 else p.mx[d] // 2) + 2 * mbc if d < nd else 1
                         for d in range(3)) + (nv,)
        if field in ("f", "rf", "ftmp"):
            return tuple(p.mx[d] if d < nd else 1 for d in range(3)) + (nv,)
        if field == "qrte":
            return tuple((1 if p.mx[d] % 2 else p.mx[d] // 2) if d < nd else 1
                         for d in range(3)) + (nv,)
        if field == "errorflags":
            return tuple(p.mx[d] if d < nd else 1 for d in range(3))
        raise KeyError(field)

    def relax_cell_updates_per_vcycle(self, nsmoothingpasses: int = 2) -> int:
        """Sum over levels of relaxed cells: 2*nu*N per level above minlevel, nu*N at it."""
        tot = 0
        for lev in range(self.minlevel, self.finestlevel + 1):
            n = self.cells_on_level(lev)
            tot += (nsmoothingpasses if lev == self.minlevel else 2 * nsmoothingpasses) * n
        return tot
