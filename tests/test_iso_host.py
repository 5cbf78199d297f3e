"""Isolated patches smoothed out of LDS, checked on the CPU.

A level's isolated patches (whose SetGhost programs read no other patch of
the level and which no other patch's programs read) run their whole
smoothing in one workgroup; the level's other patches take the per-colour
launches, whose fused and post plans then leave the isolated patches'
updates out. Here the engine's schedule is replayed on the CPU -- per
half-sweep the isolated patches' program decoded from the kernel's LDS
records, the other patches' fused and post plans, and the oracle's
relaxation of every patch -- and must leave every array bit-identical to the
reference sequence (relax, SetGhost, relax, SetGhost, ... with the oracle's
full SetGhost). The two sets share no cell, so replaying them interleaved
equals running the isolated patches' smoothing first.
"""

import numpy as np
import pytest

import orc
from bsam_amd.engine import Engine
from bsam_amd.workloads import nested_block_layout_3d
from test_fused_ghost_host import KW, run
from test_ghost_program_host import LAYOUTS


def layout_blocks3d():
    # 3D blocks at random places in a 32^3 root: some touch a sibling, some not
    return nested_block_layout_3d(32, 2, (0.25, 0.25), seed=5, sides=(8, 16))


CASES = {"blocks3d": layout_blocks3d, "amr3d": LAYOUTS["amr3d"], "amr2d": LAYOUTS["amr2d"],
         "amr3d_user_bc": LAYOUTS["amr3d_user_bc"], "periodic3d_xz": LAYOUTS["periodic3d_xz"],
         "amr2d_tiles": LAYOUTS["amr2d_tiles"]}


def run_iso(lay, eng, level, ghost0_first, seed, nsmooth=2):
    """The engine's level_relax with isolated patches, on the oracle's arrays,
    in the engine's order: the isolated patches' whole smoothing first (the
    smoothing kernel's launch), then the other patches' launches."""
    ora = orc.Oracle(lay, nsmoothingpasses=nsmooth, **KW)
    rng = np.random.default_rng(seed)
    for n, p in enumerate(ora.patches):
        for name in ("q", "f"):
            ora.field(n, name)[...] = rng.uniform(-1, 1, lay.field_shape(p, name))
    qs = [ora.field(n, "q") for n in range(len(ora.patches))]
    isos = set(eng.iso_patches(level))
    others = [n for n, p in enumerate(ora.patches) if p.level == level and n not in isos]

    def relax_only(color, which):
        # the oracle relaxes the whole level: keep the other set's q as it was
        keep = {n: qs[n].copy() for n in range(len(qs)) if n not in which}
        ora.relax(level, color)
        for n, a in keep.items():
            qs[n][...] = a

    # each isolated patch's whole smoothing on its own (the workgroups of the
    # smoothing launch run in any order): a patch that read another, or was
    # read by another, would see it before or after its smoothing, never in
    # between as the reference sequence does
    for n in sorted(isos):
        pending = 0 if ghost0_first else -1
        for s in range(nsmooth):
            for color in (1, 2):
                if pending >= 0:
                    keep = {m: qs[m].copy() for m in range(len(qs)) if m != n}
                    eng.iso_program_apply_host(level, pending, qs)
                    for m, a in keep.items():
                        qs[m][...] = a
                relax_only(color, {n})
                pending = color
        keep = {m: qs[m].copy() for m in range(len(qs)) if m != n}
        eng.iso_program_apply_host(level, 2, qs)
        for m, a in keep.items():
            qs[m][...] = a
    if not others:
        return [ora.field(n, "q").copy() for n in range(len(ora.patches))]
    pending = 0 if ghost0_first else -1
    moved = None
    for s in range(nsmooth):
        for color in (1, 2):
            last = s == nsmooth - 1 and color == 2
            if pending >= 0:
                assert eng.fused_ghost_size(level, pending, color) >= 0
                if moved is not None:
                    eng.post_ghost_apply_host(level, moved[0], moved[1], 1, qs)
                else:
                    eng.fused_ghost_part_apply_host(level, pending, color, 1, qs)
                eng.fused_ghost_part_apply_host(level, pending, color, 0, qs)
            relax_only(color, set(others))
            post = not last and eng.post_ghost_size(level, pending, color, 0) >= 0
            if post:
                eng.post_ghost_apply_host(level, pending, color, 0, qs)
            moved = (pending, color) if post else None
            pending = color
    # SetGhost(2) after the smoothing for the other patches: through the last
    # launch's final plan and its rest, or the full program (which rewrites
    # the isolated patches' ghosts with the values they already hold)
    if eng.post_ghost_size(level, 1, 2, 2) >= 0:
        eng.post_ghost_apply_host(level, 1, 2, 2, qs)
        eng.post_ghost_apply_host(level, 1, 2, 3, qs)
    else:
        eng.ghost_apply_host(level, 2, qs)
    return [ora.field(n, "q").copy() for n in range(len(ora.patches))]


@pytest.mark.parametrize("name", sorted(CASES))
def test_iso_smoothing_matches_reference_sequence(name):
    lay = CASES[name]()
    eng = Engine(lay, device=-1, tune={"iso_min_cells": 1}, **KW)
    niso = 0
    for level in range(lay.minlevel, lay.finestlevel + 1):
        niso += len(eng.iso_patches(level))
        for g0 in (False, True):
            seed = 13 * (level - lay.minlevel) + g0
            want, _ = run(lay, eng, level, g0, False, seed)
            got = run_iso(lay, eng, level, g0, seed)
            for n, (a, b) in enumerate(zip(got, want)):
                assert np.array_equal(a, b), (level, g0, n)
    assert niso > 0


def test_iso_patches_touch_no_sibling():
    """The isolated patches of the 3D block layout are those whose
    ghost-extended box meets no other patch of the level (no periodic or
    other coupling there), and both kinds occur."""
    lay = layout_blocks3d()
    eng = Engine(lay, device=-1, tune={"iso_min_cells": 1}, **KW)
    pats = lay.ordered()
    seen = {True: 0, False: 0}
    for level in range(1, lay.finestlevel + 1):
        iso = set(eng.iso_patches(level))
        ids = [n for n, p in enumerate(pats) if p.level == level]
        for a in ids:
            lo = np.array(pats[a].mglobal)[:, 0]
            hi = np.array(pats[a].mglobal)[:, 1]
            touch = False
            for b in ids:
                if b == a:
                    continue
                lo2 = np.array(pats[b].mglobal)[:, 0]
                hi2 = np.array(pats[b].mglobal)[:, 1]
                touch |= bool(np.all(lo - 1 <= hi2) and np.all(lo2 <= hi + 1))
            # a patch of side 16 and its staged parent cells fit the LDS
            assert (a in iso) == (not touch), (level, a)
            seen[a in iso] += 1
    assert seen[True] > 0 and seen[False] > 0


def test_iso_threshold_and_fusion_gate():
    lay = layout_blocks3d()
    off = Engine(lay, device=-1, tune={"iso_min_cells": 0}, **KW)
    big = Engine(lay, device=-1, tune={"iso_min_cells": 10 ** 9}, **KW)
    nofuse = Engine(lay, device=-1, tune={"iso_min_cells": 1, "fuse_ghost": 0}, **KW)
    on = Engine(lay, device=-1, tune={"iso_min_cells": 1}, **KW)
    for level in range(lay.minlevel, lay.finestlevel + 1):
        assert off.iso_patches(level) == []
        assert big.iso_patches(level) == []
        assert nofuse.iso_patches(level) == []
    assert on.tune_get("iso_patches") > 0
    assert on.tune_get("iso_cells") > 0
    # the fused plans of an isolated level cover fewer updates
    lvl = lay.finestlevel
    assert on.fused_ghost_size(lvl, 1, 2) < off.fused_ghost_size(lvl, 1, 2)
