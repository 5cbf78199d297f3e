"""The regrid's host clustering (bsam_regrid_level: BufferAndList,
InflateEdgeTags, Berger-Rigoutsos NewSubGrids) against the oracle
restatement (oracle/regrid.py) and against the reference program itself.

The reference keeps these routines private to BSAMRoutines, so the anchor
is the reference program's first frame: its hierarchy comes from QInit2D on
the root, SetGhost, the problem's user tagging and the clustering, level by
level, with no solve in between. Rebuilding it from the same steps (the
oracle's ghost fill and tagging, the engine's clustering) must give the
frame's patch lists exactly, in level-list order.
"""

import os
import zlib

import numpy as np
import pytest

import bsam_program as bp
import orc
import regrid as oregrid
from bsam_amd import frames
from bsam_amd.engine import regrid_level
from bsam_amd.layout import Layout
from helpers import layout_amr2d, layout_amr3d, layout_periodic2d, layout_periodic3d


def random_flags(patches, ndims, rng, density):
    out = []
    for mx, _ in patches:
        shape = tuple(mx[d] if d < ndims else 1 for d in range(3))
        f = (rng.random(shape) < density).astype(np.int32)
        # a few solid blobs so that clustering has structure to find
        for _ in range(2):
            c = [rng.integers(0, shape[d]) for d in range(3)]
            r = [rng.integers(1, max(2, shape[d] // 3)) for d in range(3)]
            f[c[0]:c[0] + r[0], c[1]:c[1] + r[1], c[2]:c[2] + r[2]] = 1
        out.append(np.asfortranarray(f))
    return out


LAYOUTS = {
    "amr2d": layout_amr2d,
    "amr2d_periodic": lambda: layout_periodic2d((2, 2, 2, 2)),
    "amr2d_periodic_x": lambda: layout_periodic2d((2, 2, 10, 10)),
    "amr3d": layout_amr3d,
    "amr3d_periodic": layout_periodic3d,
}


@pytest.mark.parametrize("name", sorted(LAYOUTS))
@pytest.mark.parametrize("ib,mgp,fill", [(1, 4, 0.9), (2, 2, 0.7), (0, 3, 0.8), (1, 1, 0.5)])
def test_clustering_matches_oracle(name, ib, mgp, fill):
    lay = LAYOUTS[name]()
    nd = lay.ndims
    root = lay.patches[0]
    rng = np.random.default_rng(zlib.crc32(f"{name}/{ib}/{mgp}".encode()))
    for level in range(0, lay.finestlevel + 1):
        pats = [lay.patches[i] for i in lay.level_ids(level)]
        geo = [(p.mx, p.mglobal) for p in pats]
        for density in (0.02, 0.15):
            f1 = random_flags(geo, nd, rng, density)
            f2 = [f.copy() for f in f1]
            got = regrid_level(nd, level, ib, mgp, fill, root.mx, root.mthbc, geo, f1)
            want = oregrid.regrid_level(nd, level, ib, mgp, fill, geo, f2, root.mx, root.mthbc)
            assert [(k, tuple(tuple(r) for r in b)) for k, b in want] == got, (level, density)
            for a, b in zip(f1, f2):
                assert np.array_equal(a, b)
            # every flagged cell lies in a subgrid of its patch (a flag may
            # be left out only where a subgrid could not shrink below mgp)
            for k, fl in enumerate(f1):
                cover = np.zeros_like(fl)
                for kk, mb in got:
                    if kk == k:
                        cover[mb[0][0] - 1:mb[0][1], mb[1][0] - 1:mb[1][1],
                              mb[2][0] - 1:mb[2][1]] = 1
                assert not np.any((fl == 1) & (cover == 0)), (level, k)


def test_periodic_inflation_wraps():
    """An edge tag of a periodic root buffers its image across the domain."""
    lay = layout_periodic2d((2, 2, 2, 2))
    root = lay.patches[0]
    geo = [(root.mx, root.mglobal)]
    f = np.zeros((root.mx[0], root.mx[1], 1), dtype=np.int32, order="F")
    f[0, 5, 0] = 1
    regrid_level(2, 0, 1, 2, 0.9, root.mx, root.mthbc, geo, [f])
    assert f[root.mx[0] - 1, 4:7, 0].tolist() == [1, 1, 1]
    assert f[0:2, 4:7, 0].sum() == 6


def _frame0_hierarchy(tmp_path):
    if not os.path.exists(bp.REF_PROGRAM):
        pytest.skip("reference program not built (make -C oracle run)")
    deck = bp.shipped_deck(1)
    ref, _, _ = bp.run(bp.REF_PROGRAM, deck, str(tmp_path / "ref"))
    _, finest, ps = frames.read_frame(str(tmp_path / "ref" / "out" / "m00000.dat"))
    return deck, finest, ps


def test_initial_hierarchy_matches_reference(tmp_path):
    """The shipped deck's first frame: QInit2D on the root, then per level
    SetGhost(level, 0), SetErrFlagsUser2D with qtolerance(level) and the
    clustering, each new patch initialised by QInit2D (initial grids)."""
    deck, finest, ps = _frame0_hierarchy(tmp_path)
    g = deck["griddata"]
    nd = g["ndims"]
    mthbc = tuple(g["mthbc"][i] for i in range(1, 2 * nd + 1))
    mx = tuple(g["mx"][d + 1] for d in range(nd))
    h = (g["xupper"][1] - g["xlower"][1]) / mx[0]
    lay = Layout(nd, mx, h, xlower=(g["xlower"][1], g["xlower"][2], 0.0), mthbc=mthbc,
                 minlevel=g["minlevel"])
    for level in range(0, g["maxlevel"]):
        o = orc.Oracle(lay, dt=deck["rundata"]["dt"], eps=deck["problemdata"]["eps"])
        for n, p in enumerate(o.patches):
            if p.level >= 0:
                o.field(n, "q")[...] = oregrid.qinit2d(p.mx, p.dx, p.xlower)
        for lv in range(0, level + 1):   # AMR(0..level)'s SetGhost(lv, 0)
            o.set_ghost(lv, 0)
        o.error_flags_user(level, g["qtolerance"][level])
        pats = [(n, p) for n, p in enumerate(o.patches) if p.level == level]
        geo = [(p.mx, p.mglobal) for _, p in pats]
        fl = [np.asfortranarray(o.flags(n).astype(np.int32)) for n, _ in pats]
        fo = [f.copy() for f in fl]
        new = regrid_level(nd, level, g["ibuffer"][level], g["minimumgridpoints"][level],
                           g["desiredfillratios"][level], lay.patches[0].mx, mthbc, geo, fl)
        want = oregrid.regrid_level(nd, level, g["ibuffer"][level],
                                    g["minimumgridpoints"][level],
                                    g["desiredfillratios"][level], geo, fo,
                                    lay.patches[0].mx, mthbc)
        assert [(k, tuple(tuple(r) for r in b)) for k, b in want] == new
        if not new:
            break
        for k, mb in new:
            lay.add_child(pats[k][1].id, mb[:nd])
    # the rebuilt hierarchy's patch lists equal the frame's, level by level
    got = []
    for level in range(0, lay.finestlevel + 1):
        for pid in lay.level_ids(level):
            p = lay.patches[pid]
            got.append((p.level, p.mx[:nd], tuple(p.mglobal[:nd])))
    want = [(p.level, p.mx, p.mglobal) for p in ps]
    assert lay.finestlevel == finest
    assert got == want


def test_regrid_rejects_bad_parameters():
    from bsam_amd.engine import BsamError
    f = np.zeros((8, 8, 1), dtype=np.int32, order="F")
    with pytest.raises(BsamError):
        regrid_level(2, 0, 1, 0, 0.9, (8, 8, 1), (1, 1, 1, 1), [((8, 8, 1), ((1, 8), (1, 8), (1, 1)))], [f])
    with pytest.raises(BsamError):
        regrid_level(4, 0, 1, 2, 0.9, (8, 8, 1), (1, 1, 1, 1), [((8, 8, 1), ((1, 8), (1, 8), (1, 1)))], [f])
    f[:] = 1
    assert regrid_level(2, 0, 1, 2, 0.9, (8, 8, 1), (1, 1, 1, 1),
                        [((8, 8, 1), ((1, 8), (1, 8), (1, 1)))], [f]) == [(0, ((1, 8), (1, 8), (1, 1)))]
