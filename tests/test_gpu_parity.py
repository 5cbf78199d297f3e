"""GPU engine vs the FP64 oracle, through the C ABI.

Field results must be bit-identical: the kernels keep the reference's
floating-point evaluation order and are built without multiply-add
contraction. Residual norms are reduced in a different (parallel, fixed)
order, so they are compared to 1e-12 relative.
"""

import numpy as np
import pytest

import orc
from bsam_amd.engine import Engine
from bsam_amd.layout import Layout
from bsam_amd.workloads import nested_block_layout_3d
from helpers import (layout_amr2d, layout_amr2d_tiles, layout_amr3d, layout_periodic2d,
                     layout_periodic3d, layout_single, noise_q, random_fields)

pytestmark = pytest.mark.gpu

KW = dict(dt=1e-3, eps=0.03, alpha=1.0, maxvcycles=30, qerrortol=1e-8)
LAYOUTS = {
    "single32": lambda: layout_single(32),
    "amr2d": layout_amr2d,
    "amr2d_tiles": layout_amr2d_tiles,
    "amr2d_user_bc": lambda: layout_amr2d(mthbc=(10, 10, 10, 10)),
    "single8_3d": lambda: layout_single(8, ndims=3),
    "amr3d": layout_amr3d,
    "amr3d_mixed_bc": lambda: layout_amr3d(mthbc=(1, 10, 1, 10, 10, 1)),
    "periodic2d_xy": lambda: layout_periodic2d((2, 2, 2, 2)),
    "periodic2d_x_user_y": lambda: layout_periodic2d((2, 2, 10, 10)),
    "periodic3d_xyz": layout_periodic3d,
    # 8^3 and 16^3 blocks at random places: some touch a sibling, some not
    "blocks3d": lambda: nested_block_layout_3d(32, 2, (0.25, 0.25), seed=5, sides=(8, 16)),
}


# engine schedules: per-level launches only; default (single-workgroup tail
# for small coarse levels with its arrays in LDS, SetGhost applied inside the
# next relaxation launch, hipGraph replay); every single-patch level in the
# tail, in LDS as far as it fits or all in global memory; SetGhost and
# relaxation as separate launches; other tile shapes (in 3D the third extent
# fills the 256 threads); relaxation stencils read through an LDS copy of the
# tile and its halo (fused, unfused, and with another tile shape); no ghost
# updates moved to the end of the previous launch; every ghost update in the
# 32-byte device record; isolated patches smoothed out of LDS on every level
# that has one (with the launches of the other patches fused, moved or
# staged, with every single-patch coarse level run by level launches, and on
# the engine's stream instead of beside the other patches' launches); the
# coarse loading over fine tiles on every level, or over linear coarse tiles
# on every level
SCHEDULES = {
    "launch": {"tail_cells": 0, "graph": 0},
    "default": {},
    "all_tail": {"tail_cells": 1 << 30, "graph": 1},
    "all_tail_global": {"tail_cells": 1 << 30, "tail_lds": 0, "graph": 0},
    "unfused": {"fuse_ghost": 0, "graph": 0},
    "tiles_64x4": {"tile_x": 64, "tile_y": 4},
    "tiles_16x16": {"tile_x": 16, "tile_y": 16, "fuse_ghost": 1},
    "staged": {"relax_lds": 1},
    "staged_unfused": {"relax_lds": 1, "fuse_ghost": 0, "graph": 0},
    "staged_64x4": {"relax_lds": 1, "tile_x": 64, "tile_y": 4},
    "no_post": {"post_ghost": 0},
    "long_records": {"short_ghost": 0},
    "iso": {"iso_min_cells": 1},
    "iso_launch": {"iso_min_cells": 1, "tail_cells": 0, "graph": 0},
    "iso_no_post": {"iso_min_cells": 1, "post_ghost": 0},
    "iso_staged": {"iso_min_cells": 1, "relax_lds": 1},
    "iso_one_stream": {"iso_min_cells": 1, "iso_side": 0},
    "cgl_fine": {"cgl_fine_cells": 0},
    "cgl_fine_iso_launch": {"cgl_fine_cells": 0, "iso_min_cells": 1, "tail_cells": 0},
    "cgl_lin": {"cgl_fine_cells": -1},
}


def pair(layout, fields, names=("q", "f", "qc"), tune=None):
    eng = Engine(layout, device=0, tune=tune, **KW)
    ora = orc.Oracle(layout, **KW)
    for n, p in enumerate(ora.patches):
        for name in names:
            if name in fields[p.id]:
                eng.upload(n, name, fields[p.id][name])
                ora.field(n, name)[...] = fields[p.id][name]
    return eng, ora


def assert_same(eng, ora, names=("q",)):
    for n, p in enumerate(ora.patches):
        for name in names:
            got, want = eng.download(n, name), ora.field(n, name)
            if not np.array_equal(got, want):
                bad = np.argwhere(got != want)
                raise AssertionError(
                    f"patch {p.id} (level {p.level}) field {name}: {len(bad)} cells differ, "
                    f"first {bad[0].tolist()} got {got[tuple(bad[0])]!r} want {want[tuple(bad[0])]!r}")


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_set_ghost_all_levels(name):
    lay = LAYOUTS[name]()
    for level in range(lay.minlevel, lay.finestlevel + 1):
        for ipass in (0, 1, 2):
            eng, ora = pair(lay, random_fields(lay, seed=(level - lay.minlevel) * 3 + ipass))
            eng.set_ghost(level, ipass)
            ora.set_ghost(level, ipass)
            assert_same(eng, ora)
            eng.close()


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_level_operations(name):
    lay = LAYOUTS[name]()
    for level in range(lay.finestlevel, lay.minlevel, -1):
        # fields each operation writes (the oracle also leaves scratch in rf
        # and, after correct, the correction in qc; the engine does not)
        ops = [("relax", (level, 1), ("q",)), ("relax", (level, 2), ("q",)),
               ("restrict", (level,), ("q", "qc")), ("coarse_ghost", (level,), ("q", "qc")),
               ("coarse_load", (level,), ("q", "f")), ("correct", (level,), ("q",))]
        for op, args, names in ops:
            eng, ora = pair(lay, random_fields(lay, seed=(level - lay.minlevel) + 17))
            getattr(eng, op)(*args)
            getattr(ora, op)(*args)
            assert_same(eng, ora, names=names)
            eng.close()


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_fill_down_and_source(name):
    lay = LAYOUTS[name]()
    for field in (1, 2):
        eng, ora = pair(lay, random_fields(lay, seed=field))
        eng.fill_down(field)
        ora.fill_down(field)
        assert_same(eng, ora, names=("q", "f"))
        eng.close()
    eng, ora = pair(lay, random_fields(lay, seed=5))
    for n, p in enumerate(ora.patches):
        qo = random_fields(lay, seed=9)[p.id]["q"]
        eng.upload(n, "qold", qo)
        ora.field(n, "qold")[...] = qo
    eng.source()
    ora.source()
    assert_same(eng, ora, names=("f", "ftmp"))
    eng.close()


@pytest.mark.parametrize("sched", sorted(SCHEDULES))
@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_vcycle_bitwise(name, sched):
    lay = LAYOUTS[name]()
    eng, ora = pair(lay, random_fields(lay, seed=3), tune=SCHEDULES[sched])
    eng.vcycle(lay.finestlevel)
    ora.vcycle(lay.finestlevel)
    assert_same(eng, ora, names=("q", "f"))
    eng.close()


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_error_afas(name):
    lay = LAYOUTS[name]()
    eng, ora = pair(lay, random_fields(lay, seed=4))
    for level in range(lay.minlevel, lay.finestlevel + 1):
        r, want = eng.error_afas(level), ora.error_afas(level)
        assert abs(r - want) <= 1e-12 * want


@pytest.mark.parametrize("sched", sorted(SCHEDULES))
@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_time_steps(name, sched):
    lay = LAYOUTS[name]()
    q0 = noise_q(lay, seed=11)
    eng, ora = pair(lay, {k: {"q": v} for k, v in q0.items()}, names=("q",),
                    tune=SCHEDULES[sched])
    for step in range(3):
        hist, conv = eng.time_step()
        ref = ora.time_step()
        assert conv
        assert len(hist) == len(ref), (step, hist, ref)
        np.testing.assert_allclose(hist, ref, rtol=1e-12, atol=0)
        assert_same(eng, ora, names=("q", "qold"))
    eng.close()


def test_tail_schedule_places_levels_in_lds():
    """The default schedule runs the small single-patch levels in the tail
    kernel with their arrays in LDS; tail_lds = 0 keeps them in global memory."""
    lay = layout_amr2d()
    eng = Engine(lay, device=0, **KW)
    assert eng.tune_get("tail_top") >= lay.minlevel
    assert eng.tune_get("tail_lds_fp64") > 0
    eng.close()
    eng = Engine(lay, device=0, tune={"tail_lds": 0}, **KW)
    assert eng.tune_get("tail_lds_fp64") == 0
    eng.close()


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_truncation_error_and_flags(name):
    """EstimateLevelErrors' truncation error (bit-identical qrte, q, qold, qc)
    and the default error flags at three tolerances around the data's scale,
    on the root level and every refined level (AMR estimates from the root
    up)."""
    lay = LAYOUTS[name]()
    for level in range(max(lay.minlevel + 1, 0), lay.finestlevel + 1):
        eng, ora = pair(lay, random_fields(lay, seed=23 + level))
        eng.truncation_error(level)
        ora.truncation_error(level)
        assert_same(eng, ora, names=("q", "qold", "qc", "qrte"))
        pats = [n for n, p in enumerate(ora.patches) if p.level == level]
        h = ora.patches[pats[0]].dx
        scale = max(float(np.abs(ora.field(n, "qrte")).max()) for n in pats) * h ** lay.ndims
        for frac in (0.01, 0.3, 2.0):
            eng.error_flags(level, frac * scale)
            ora.error_flags(level, frac * scale)
            for n in pats:
                assert np.array_equal(eng.flags(n), ora.flags(n)), (level, frac, n)
        eng.close()


@pytest.mark.parametrize("name", ["amr2d", "amr3d", "periodic2d_xy"])
def test_recommit_same_hierarchy(name):
    """A regrid that hands over the hierarchy already committed keeps its
    plans and graphs (commits_reused counts it) and restarts from zeroed
    fields: two time steps after it equal a fresh context's bit for bit. A
    changed hierarchy is committed afresh."""
    lay = LAYOUTS[name]()
    q0 = noise_q(lay, seed=31)
    q1 = noise_q(lay, seed=32)
    eng = Engine(lay, device=0, **KW)
    for n, p in enumerate(eng.patches):
        eng.upload(n, "q", q0[p.id])
    eng.time_step()
    eng.describe(LAYOUTS[name]())
    assert eng.tune_get("commits_reused") == 1
    fresh = Engine(lay, device=0, **KW)
    for e in (eng, fresh):
        for n, p in enumerate(e.patches):
            e.upload(n, "q", q1[p.id])
    for _ in range(2):
        h1, _ = eng.time_step()
        h2, _ = fresh.time_step()
        assert h1 == h2
    for n in range(len(eng.patches)):
        for field in ("q", "qold", "f"):
            assert np.array_equal(eng.download(n, field), fresh.download(n, field)), (n, field)
    fresh.close()
    root = lay.patches[0]
    other = Layout(lay.ndims, root.mx, root.dx, xlower=root.xlower, mthbc=root.mthbc,
                   minlevel=lay.minlevel)
    eng.describe(other)
    assert eng.tune_get("commits_reused") == 1
    eng.close()


@pytest.mark.parametrize("name", ["amr2d", "amr3d", "periodic2d_x_user_y"])
def test_frame_export(tmp_path, name):
    """A frame written from the engine's q after two time steps is the frame
    the oracle's q gives, byte for byte, and reads back to those arrays."""
    from bsam_amd import frames
    lay = LAYOUTS[name]()
    q0 = noise_q(lay, seed=41)
    eng, ora = pair(lay, {k: {"q": v} for k, v in q0.items()}, names=("q",))
    for _ in range(2):
        eng.time_step()
        ora.time_step()
    root = lay.patches[0]
    xu = [root.xlower[d] + root.mx[d] * root.dx for d in range(lay.ndims)]
    frames.write_frame(str(tmp_path / "gpu.dat"), 0.002, lay.finestlevel,
                       frames.frame_from_engine(eng, xu))
    qo = {p.id: ora.field(n, "q") for n, p in enumerate(ora.patches) if p.level >= 0}
    frames.write_frame(str(tmp_path / "ora.dat"), 0.002, lay.finestlevel,
                       frames.frame_from_arrays(lay, qo, xu))
    assert (tmp_path / "gpu.dat").read_bytes() == (tmp_path / "ora.dat").read_bytes()
    _, _, ps = frames.read_frame(str(tmp_path / "gpu.dat"))
    for fp in ps:
        assert np.isfinite(fp.q).all()
    eng.close()


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_user_error_flags(name):
    """ErrFlag's user tagging (SetErrFlagsUser2D/3D of the Cahn-Hilliard
    module) after SetGhost(level, 0), as AMR runs it, on the root and every
    refined level at three tolerances: the flags equal the oracle's."""
    lay = LAYOUTS[name]()
    for level in range(max(lay.minlevel, 0), lay.finestlevel + 1):
        eng, ora = pair(lay, random_fields(lay, seed=51 + level))
        eng.set_ghost(level, 0)
        ora.set_ghost(level, 0)
        pats = [n for n, p in enumerate(ora.patches) if p.level == level]
        for tol in (0.4, 1.0, 1.8):
            eng.error_flags_user(level, tol)
            ora.error_flags_user(level, tol)
            on = size = 0
            for n in pats:
                got = eng.flags(n)
                assert np.array_equal(got, ora.flags(n)), (level, tol, n)
                on, size = on + int(got.sum()), size + got.size
            assert 0 < on < size or tol != 1.0, (level, on, size)
        eng.close()


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_after_step_stats(name):
    """Mass and (2D) energy of the composite grid after two time steps, as
    TakeTimeSteps computes them (AfterStepStatistics with AfterStep2D/3D).
    The device reduces each patch's sums in another order than the
    reference's sequential SUM, so they agree to 1e-12 of the sum of the
    magnitudes of the terms (the bound of a reordered sum)."""
    lay = LAYOUTS[name]()
    q0 = noise_q(lay, seed=61)
    eng, ora = pair(lay, {k: {"q": v} for k, v in q0.items()}, names=("q",))
    for _ in range(2):
        eng.time_step()
        ora.time_step()
        m, e = eng.after_step_stats()
        mo, eo = ora.after_step_stats()
        # magnitude of the terms: h^d sum |q1| per patch plus its parent's cells
        scale = 0.0
        for n, p in enumerate(ora.patches):
            if p.level < 0:
                continue
            q = ora.field(n, "q")
            inner = q[(slice(1, -1),) * lay.ndims + (0,) * (3 - lay.ndims) + (0,)]
            scale += p.dx ** lay.ndims * np.abs(inner).sum() * (2 if p.level > 0 else 1)
        assert abs(m - mo) <= 1e-12 * scale, (m, mo, scale)
        if lay.ndims == 2:
            assert e > 0 and abs(e - eo) <= 1e-12 * max(abs(eo), 1.0) * 8, (e, eo)
        else:
            assert e == 0.0 and eo == 0.0
    eng.close()


def test_regrid_entry_points_report_errors():
    """The regrid's field entry points refuse what does not fit: levels
    whose patches differ, a level without a parent, contexts of another
    shape."""
    from bsam_amd.engine import BsamError
    a = Engine(layout_amr2d(), device=0, **KW)
    b = Engine(layout_single(32), device=0, **KW)
    with pytest.raises(BsamError):
        a.copy_levels_from(b, 1)            # level 1 differs
    with pytest.raises(BsamError):
        a.init_fields(0)                    # the root has no parent grid
    c = Engine(layout_single(8, ndims=3), device=0, **KW)
    with pytest.raises(BsamError):
        a.transfer_values_from(c, 0)        # 2D from 3D
    a.copy_levels_from(Engine(layout_amr2d(), device=0, **KW), 0)
    for e in (a, b, c):
        e.close()


def test_solve_status():
    """A solve reports how it ended: converged, stopped at maxvcycles (a
    tolerance it cannot reach), or stopped at the first non-finite residual
    (a NaN in the initial state) instead of iterating on."""
    lay = layout_amr2d()
    q0 = noise_q(lay, seed=71)
    kw = dict(KW)
    eng = Engine(lay, device=0, **kw)
    for n, p in enumerate(eng.patches):
        eng.upload(n, "q", q0[p.id])
    hist, conv = eng.time_step()
    assert conv and eng.solve_status() == "converged"
    kw.update(maxvcycles=2, qerrortol=1e-30)
    eng2 = Engine(lay, device=0, **kw)
    for n, p in enumerate(eng2.patches):
        eng2.upload(n, "q", q0[p.id])
    hist, conv = eng2.time_step()
    assert not conv and len(hist) == 2 and eng2.solve_status() == "maxcycles"
    bad = q0[lay.patches[0].id].copy()
    bad[3, 3, 0, 0] = np.nan
    eng3 = Engine(lay, device=0, **KW)
    for n, p in enumerate(eng3.patches):
        eng3.upload(n, "q", bad if p.id == 0 else q0[p.id])
    hist, conv = eng3.time_step()
    assert not conv and len(hist) == 1 and not np.isfinite(hist[0])
    assert eng3.solve_status() == "nonfinite"
    for e in (eng, eng2, eng3):
        e.close()
