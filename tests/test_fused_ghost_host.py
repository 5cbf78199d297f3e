"""SetGhost fused into the following relaxation, checked on the CPU.

The engine applies a SetGhost program inside the next relaxation launch:
each tile first applies the ghost updates its cells read, and updates whose
value nobody reads before the next SetGhost rewrites them are left out. Here
the fused sequence is replayed on the CPU -- kept updates applied to the
oracle's arrays, then the oracle's relaxation -- and must leave every array
bit-identical to the reference sequence (relax, SetGhost, relax, SetGhost,
... with the oracle's full SetGhost), for every level, for smoothing with and
without a leading SetGhost(level, 0).
"""

import numpy as np
import pytest

import orc
from bsam_amd.engine import Engine
from test_ghost_program_host import LAYOUTS

KW = dict(dt=1e-3, eps=0.03)


def run(lay, eng, level, ghost0_first, fused, seed, nsmooth=2):
    ora = orc.Oracle(lay, nsmoothingpasses=nsmooth, **KW)
    rng = np.random.default_rng(seed)
    for n, p in enumerate(ora.patches):
        for name in ("q", "f"):
            ora.field(n, name)[...] = rng.uniform(-1, 1, lay.field_shape(p, name))
    qs = [ora.field(n, "q") for n in range(len(ora.patches))]
    pending = 0 if ghost0_first else -1
    nfused = 0
    for _ in range(nsmooth):
        for color in (1, 2):
            if pending >= 0 and fused and eng.fused_ghost_size(level, pending, color) >= 0:
                eng.fused_ghost_apply_host(level, pending, color, qs)
                nfused += 1
            elif pending >= 0:
                ora.set_ghost(level, pending)
            ora.relax(level, color)
            pending = color
    ora.set_ghost(level, pending)
    return [ora.field(n, "q").copy() for n in range(len(ora.patches))], nfused


def run_post(lay, eng, level, ghost0_first, seed, nsmooth=2, final=False):
    """The engine's smoothing schedule with post plans: a launch with a post
    plan ends with the next launch's moved updates, and the next launch runs
    only the rest before its relaxation (the engine's level_relax)."""
    ora = orc.Oracle(lay, nsmoothingpasses=nsmooth, **KW)
    rng = np.random.default_rng(seed)
    for n, p in enumerate(ora.patches):
        for name in ("q", "f"):
            ora.field(n, name)[...] = rng.uniform(-1, 1, lay.field_shape(p, name))
    qs = [ora.field(n, "q") for n in range(len(ora.patches))]
    pending = 0 if ghost0_first else -1
    moved, nposts = None, 0
    for s in range(nsmooth):
        for color in (1, 2):
            last = s == nsmooth - 1 and color == 2
            fused = pending >= 0 and eng.fused_ghost_size(level, pending, color) >= 0
            post = not last and eng.post_ghost_size(level, pending, color, 0) >= 0
            if fused:
                if moved is not None:
                    eng.post_ghost_apply_host(level, moved[0], moved[1], 1, qs)
                else:
                    eng.fused_ghost_part_apply_host(level, pending, color, 1, qs)
                eng.fused_ghost_part_apply_host(level, pending, color, 0, qs)
            elif pending >= 0:
                ora.set_ghost(level, pending)
            ora.relax(level, color)
            if post:
                eng.post_ghost_apply_host(level, pending, color, 0, qs)
                nposts += 1
            moved = (pending, color) if post else None
            pending = color
    # the program after the smoothing: its final plan's moved updates at the
    # end of the last launch, then the rest (the engine's op_apply_pending)
    if final and eng.post_ghost_size(level, 1, 2, 2) >= 0:
        eng.post_ghost_apply_host(level, 1, 2, 2, qs)
        eng.post_ghost_apply_host(level, 1, 2, 3, qs)
        nposts += 1
    else:
        ora.set_ghost(level, pending)
    return [ora.field(n, "q").copy() for n in range(len(ora.patches))], nposts


@pytest.mark.parametrize("name", ["amr3d", "amr3d_user_bc", "amr3d_mixed_bc", "periodic3d_xz"])
def test_post_plans_match_reference_sequence(name):
    """Moved and remaining updates together are exactly the next launch's
    pre updates: the smoothing with post plans leaves every array
    bit-identical to the reference sequence."""
    lay = LAYOUTS[name]()
    eng = Engine(lay, device=-1, **KW)
    total = 0
    for level in range(lay.minlevel, lay.finestlevel + 1):
        for g0 in (False, True):
            seed = 11 * (level - lay.minlevel) + g0
            want, _ = run(lay, eng, level, g0, False, seed)
            for final in (False, True):
                got, nposts = run_post(lay, eng, level, g0, seed, final=final)
                total += nposts
                for n, (a, b) in enumerate(zip(got, want)):
                    assert np.array_equal(a, b), (level, g0, final, n)
    assert total > 0


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_fused_smoothing_matches_reference_sequence(name):
    lay = LAYOUTS[name]()
    eng = Engine(lay, device=-1, **KW)
    total = 0
    for level in range(lay.minlevel, lay.finestlevel + 1):
        for g0 in (False, True):
            seed = 7 * (level - lay.minlevel) + g0
            want, _ = run(lay, eng, level, g0, False, seed)
            got, nf = run(lay, eng, level, g0, True, seed)
            total += nf
            for n, (a, b) in enumerate(zip(got, want)):
                assert np.array_equal(a, b), (level, g0, n)
    assert total > 0   # the fused path was exercised


def test_fused_plans_cover_the_2d_levels():
    """On the 2D layouts every transition of every level fuses, and each
    level drops part of its programs. (On refined 3D levels the face
    interpolation reads tangentially diagonal fine cells that can sit in the
    neighbouring tile, so those transitions stay two launches.)"""
    for name in ("amr2d_tiles", "amr2d", "periodic_xy", "amr2d_user_bc"):
        lay = LAYOUTS[name]()
        eng = Engine(lay, device=-1, **KW)
        for level in range(lay.minlevel, lay.finestlevel + 1):
            dropped = 0
            for ipass, color in ((0, 1), (1, 2), (2, 1)):
                n = eng.fused_ghost_size(level, ipass, color)
                full = len(eng.ghost_program_dst(level, ipass))
                assert 0 <= n <= full, (name, level, ipass, color, n, full)
                dropped += full - n
            assert dropped > 0, (name, level)


def test_fused_plans_pass_the_race_check_in_3d():
    """Every transition of every 3D level has an accepted fused plan: the
    engine checks each plan from the cells' side (every cell the launch
    relaxes marked with its tile, every ghost a relaxed stencil reads marked
    with the reading tile) and refuses one whose tile updates read another
    tile's relaxed cells or write a ghost another tile reads. The updates
    launched ahead of the relaxation (tangentially diagonal fine sources in
    a neighbouring tile) stay a minority of the kept ones."""
    from bsam_amd import engine as E
    for name in ("amr3d", "amr3d_user_bc", "amr3d_mixed_bc", "periodic3d_xz"):
        lay = LAYOUTS[name]()
        eng = Engine(lay, device=-1, **KW)
        for level in range(lay.minlevel, lay.finestlevel + 1):
            for ipass, color in ((0, 1), (1, 2), (2, 1)):
                n = eng.fused_ghost_size(level, ipass, color)
                assert n >= 0, (name, level, ipass, color)
                pre = E.lib().bsam_fused_ghost_pre_size(eng.ctx, level, ipass, color)
                assert 0 <= pre <= n // 2 + 1, (name, level, ipass, color, pre, n)
            # the launches before a transition with pre updates take them
            # over at their end, each plan race-checked the same way
            for pending, color in ((-1, 1), (0, 1), (2, 1), (1, 2)):
                nxt = (color, 3 - color)
                pre = E.lib().bsam_fused_ghost_pre_size(eng.ctx, level, *nxt)
                if pre <= 0:
                    continue
                moved = eng.post_ghost_size(level, pending, color, 0)
                rest = eng.post_ghost_size(level, pending, color, 1)
                assert moved + rest == pre and moved >= pre // 2, \
                    (name, level, pending, color, moved, rest, pre)
            # the program after a smoothing, taken over by its last launch
            full = len(eng.ghost_program_dst(level, 2))
            if full:
                moved = eng.post_ghost_size(level, 1, 2, 2)
                rest = eng.post_ghost_size(level, 1, 2, 3)
                assert moved + rest == full and moved >= full // 2, \
                    (name, level, moved, rest, full)


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_ghost_riding_restriction_commutes(name):
    """Where the engine runs SetGhost(level, p) in the same launch as the
    level's restriction, the two must commute: SetGhost then RestrictSolution
    and RestrictSolution then SetGhost leave every array bit-identical (the
    program reads none of the parent cells the restriction writes, and the
    restriction reads no ghost)."""
    lay = LAYOUTS[name]()
    eng = Engine(lay, device=-1, **KW)
    checked = 0
    for level in range(lay.minlevel + 1, lay.finestlevel + 1):
        for p in (0, 1, 2):
            if eng.ghost_rides_down(level, p) != 1:
                continue
            outs = []
            for order in ("ghost_first", "restrict_first"):
                ora = orc.Oracle(lay, **KW)
                rng = np.random.default_rng((level - lay.minlevel) * 3 + p)
                for n, pt in enumerate(ora.patches):
                    for f in ("q", "qc"):
                        ora.field(n, f)[...] = rng.uniform(-1, 1, lay.field_shape(pt, f))
                if order == "ghost_first":
                    ora.set_ghost(level, p)
                    ora.restrict(level)
                else:
                    ora.restrict(level)
                    ora.set_ghost(level, p)
                outs.append([(ora.field(n, "q").copy(), ora.field(n, "qc").copy())
                             for n in range(len(ora.patches))])
            for (qa, ca), (qb, cb) in zip(*outs):
                assert np.array_equal(qa, qb) and np.array_equal(ca, cb), (level, p)
            checked += 1
    if lay.finestlevel >= 1 and lay.ndims == 2:
        assert checked > 0
