"""The loop's reduced SetGhost(level - 1, 0) ride, checked on the CPU.

Inside the FP64 multigrid loop the engine's coarse ghost + loading launch of
level L carries only the updates of SetGhost(L - 1, 0) that read a cell
changed since the loop's FillDown applied that program in full: level L - 1
cells under level L (the V-cycle's restriction), level L - 2 cells under
level L - 1 and level L - 2 ghost cells (FillDown's next level and its
SetGhost(L - 2, 0)). Here the oracle runs FillDown(q) and a
V-cycle's descent (smoothing, RestrictSolution, SetGhost(L - 1, 0),
GetCoarseGhostPoints, CoarseLoadingFunction per level, the reference order);
at each SetGhost(L - 1, 0) the reduced program, applied to a copy of the
state instead of the full one, must leave every array bit-identical.
"""

import numpy as np
import pytest

import orc
from bsam_amd.engine import Engine
from bsam_amd.workloads import nested_block_layout_3d
from test_ghost_program_host import LAYOUTS

KW = dict(dt=1e-3, eps=0.03)
# layouts with a refined level (a level L >= 1 whose parent FillDown fills)
CASES = {k: v for k, v in LAYOUTS.items() if v().finestlevel >= 1}
CASES["blocks3d"] = lambda: nested_block_layout_3d(32, 2, (0.25, 0.25), seed=5, sides=(8, 16))


def smooth(ora, level, g0, nsmooth=2):
    pending = 0 if g0 else -1
    for _ in range(nsmooth):
        for color in (1, 2):
            if pending >= 0:
                ora.set_ghost(level, pending)
            ora.relax(level, color)
            pending = color
    ora.set_ghost(level, pending)


@pytest.mark.parametrize("name", sorted(CASES))
def test_loop_ride_matches_full_set_ghost(name):
    lay = CASES[name]()
    eng = Engine(lay, device=-1, **KW)
    ora = orc.Oracle(lay, **KW)
    rng = np.random.default_rng(17)
    for n, p in enumerate(ora.patches):
        for f in ("q", "f", "qc"):
            ora.field(n, f)[...] = rng.uniform(-1, 1, lay.field_shape(p, f))
    ora.fill_down(2)
    checked = reduced = 0
    for level in range(lay.finestlevel, max(lay.minlevel, 0), -1):
        smooth(ora, level, g0=False)
        ora.restrict(level)
        size = eng.ride_loop_size(level)
        assert size >= 0, (name, level)
        full = len(eng.ghost_program_dst(level - 1, 0))
        assert size <= full
        reduced += full - size
        qs = [np.asfortranarray(ora.field(n, "q").copy()) for n in range(len(ora.patches))]
        eng.ride_loop_apply_host(level, qs)
        ora.set_ghost(level - 1, 0)
        for n in range(len(ora.patches)):
            assert np.array_equal(qs[n], ora.field(n, "q")), (name, level, n)
        checked += 1
        ora.coarse_ghost(level)
        ora.coarse_load(level)
    assert checked > 0
    if lay.finestlevel >= 2:
        assert reduced > 0


def test_loop_ride_needs_the_fill_down():
    """Without the FillDown the reduced program is not the full one: the
    planted omission is seen (the check above has teeth)."""
    lay = CASES["blocks3d"]()
    eng = Engine(lay, device=-1, **KW)
    ora = orc.Oracle(lay, **KW)
    rng = np.random.default_rng(3)
    for n, p in enumerate(ora.patches):
        for f in ("q", "f", "qc"):
            ora.field(n, f)[...] = rng.uniform(-1, 1, lay.field_shape(p, f))
    level = lay.finestlevel
    smooth(ora, level, g0=False)
    ora.restrict(level)
    qs = [np.asfortranarray(ora.field(n, "q").copy()) for n in range(len(ora.patches))]
    eng.ride_loop_apply_host(level, qs)
    ora.set_ghost(level - 1, 0)
    assert any(not np.array_equal(qs[n], ora.field(n, "q")) for n in range(len(ora.patches)))
