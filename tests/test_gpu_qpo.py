"""Quasi-periodic offsets on the GPU, through the C ABI.

qpo(1:nrvars) (nodeinfodef.f90:101, the griddata namelist's qpo) shifts every
value copied across the periodic x boundary: +qpo(n) into the high ghosts,
-qpo(n) into the low ones (PeriodicBC, boundary.f90:441-461;
PeriodicGhostOverlap with GetPeriodicOffsets' qoffset, :578-591,
:965-1044). The engine carries the offset in its ghost-update records; fields
must match the FP64 oracle bit for bit in every schedule that applies ghost
updates differently (level launches, the single-workgroup tail, updates
fused into the relaxation, the isolated-patch smoothing out of LDS, the
32-byte records). Mixed precision applies no offset to its FP32 correction
(the offset cancels in the difference of two states) and is held to the
result bar of the mixed-precision tests.
"""

import numpy as np
import pytest

import orc
from bsam_amd.engine import PREC_MIXED, Engine
from helpers import layout_periodic2d, layout_periodic3d, noise_q, random_fields

pytestmark = pytest.mark.gpu

KW = dict(dt=1e-3, eps=0.03, alpha=1.0, maxvcycles=30, qerrortol=1e-8)
QPO = (0.375, -1.25)
LAYOUTS = {
    "periodic2d_xy": lambda: layout_periodic2d((2, 2, 2, 2)),
    "periodic2d_x_user_y": lambda: layout_periodic2d((2, 2, 10, 10)),
    "periodic3d_xyz": layout_periodic3d,
    "periodic3d_xz": lambda: layout_periodic3d((2, 2, 1, 1, 2, 2)),
}
SCHEDULES = {
    "launch": {"tail_cells": 0, "graph": 0},
    "default": {},
    "all_tail": {"tail_cells": 1 << 30, "graph": 1},
    "unfused": {"fuse_ghost": 0, "graph": 0},
    "staged": {"relax_lds": 1},
    "long_records": {"short_ghost": 0},
    "iso_launch": {"iso_min_cells": 1, "tail_cells": 0, "graph": 0},
    "iso": {"iso_min_cells": 1},
}


def pair(layout, fields, qpo, names=("q", "f", "qc"), tune=None, **kw):
    args = dict(KW, **kw)
    eng = Engine(layout, device=0, tune=tune, qpo=qpo, **args)
    ora = orc.Oracle(layout, qpo=qpo, **args)
    for n, p in enumerate(ora.patches):
        for name in names:
            if name in fields[p.id]:
                eng.upload(n, name, fields[p.id][name])
                ora.field(n, name)[...] = fields[p.id][name]
    return eng, ora


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def assert_same(eng, ora, names=("q",)):
    for n, p in enumerate(ora.patches):
        for name in names:
            got, want = eng.download(n, name), ora.field(n, name)
            if not np.array_equal(bits(got), bits(want)):
                bad = np.argwhere(bits(got) != bits(want))
                raise AssertionError(
                    f"patch {p.id} (level {p.level}) field {name}: {len(bad)} cells differ, "
                    f"first {bad[0].tolist()} got {got[tuple(bad[0])]!r} want {want[tuple(bad[0])]!r}")


def signed_zero_fields(lay, seed):
    """Random fields with half the cells +-0.0: the +0.0 offset of images
    shifted in y or z only turns -0.0 into +0.0, as the reference's
    `+ qoffset` does."""
    rng = np.random.default_rng(seed)
    out = random_fields(lay, seed=seed)
    for d in out.values():
        q = d["q"]
        z = rng.uniform(size=q.shape) < 0.5
        q[z] = rng.choice([-0.0, 0.0], size=int(z.sum()))
    return out


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_set_ghost_with_qpo(name):
    lay = LAYOUTS[name]()
    for level in range(lay.minlevel, lay.finestlevel + 1):
        for ipass in (0, 1, 2):
            seed = (level - lay.minlevel) * 3 + ipass
            eng, ora = pair(lay, signed_zero_fields(lay, seed), QPO)
            eng.set_ghost(level, ipass)
            ora.set_ghost(level, ipass)
            assert_same(eng, ora)
            eng.close()


@pytest.mark.parametrize("sched", sorted(SCHEDULES))
@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_vcycle_with_qpo(name, sched):
    lay = LAYOUTS[name]()
    eng, ora = pair(lay, random_fields(lay, seed=3), QPO, tune=SCHEDULES[sched])
    eng.vcycle(lay.finestlevel)
    ora.vcycle(lay.finestlevel)
    assert_same(eng, ora, names=("q", "f"))
    eng.close()


@pytest.mark.parametrize("sched", ["default", "launch", "iso"])
@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_time_steps_with_qpo(name, sched):
    lay = LAYOUTS[name]()
    q0 = noise_q(lay, seed=11)
    qpo = (0.02, -0.001)
    eng, ora = pair(lay, {k: {"q": v} for k, v in q0.items()}, qpo, names=("q",),
                    tune=SCHEDULES[sched])
    for step in range(2):
        hist, conv = eng.time_step()
        ref = ora.time_step()
        assert conv
        assert len(hist) == len(ref), (step, hist, ref)
        np.testing.assert_allclose(hist, ref, rtol=1e-12, atol=0)
        assert_same(eng, ora, names=("q", "qold"))
    eng.close()


def test_set_qpo_after_commit():
    """A new qpo takes effect on the next solve: the captured graphs that
    hold the old problem constants are dropped."""
    lay = LAYOUTS["periodic2d_xy"]()
    q0 = noise_q(lay, seed=12)
    eng, ora = pair(lay, {k: {"q": v} for k, v in q0.items()}, None, names=("q",))
    for qpo in ((0.0, 0.0), (0.02, 0.0), (-0.01, 0.003)):
        eng.set_qpo(qpo)
        orc.lib().orc_set_qpo(ora.h, *qpo)
        hist, conv = eng.time_step()
        ref = ora.time_step()
        assert conv
        np.testing.assert_allclose(hist, ref, rtol=1e-12, atol=0)
        assert_same(eng, ora, names=("q",))
    eng.close()


@pytest.mark.parametrize("name", ["periodic2d_xy", "periodic3d_xyz"])
def test_mixed_with_qpo(name):
    """FP32 smoothing inside FP64 defect correction with qpo != 0: the
    FP64 residual meets the stop test and the fields stay within 2e-6 of the
    FP64 oracle's."""
    lay = LAYOUTS[name]()
    kw = dict(dt=1e-3, eps=0.05 if lay.ndims == 3 else 0.03, well="plusminus",
              maxvcycles=40, qerrortol=1e-8)
    qpo = (0.02, -0.001)
    q0 = noise_q(lay, seed=21, mean=0.1, amp=0.05)
    eng = Engine(lay, device=0, precision=PREC_MIXED, qpo=qpo, **kw)
    ora = orc.Oracle(lay, qpo=qpo, **kw)
    for n, p in enumerate(ora.patches):
        eng.upload(n, "q", q0[p.id])
        ora.field(n, "q")[...] = q0[p.id]
    for step in range(2):
        hist, conv = eng.time_step()
        ora.time_step()
        assert conv, hist
        assert hist[-1] < 1e-8
        worst = 0.0
        for n, p in enumerate(ora.patches):
            if p.level < 0:
                continue
            inner = (slice(1, -1),) * lay.ndims
            got, want = eng.download(n, "q"), ora.field(n, "q")
            worst = max(worst, float(np.abs(got[inner] - want[inner]).max()))
        assert worst < 2e-6, (step, worst)
    eng.close()
