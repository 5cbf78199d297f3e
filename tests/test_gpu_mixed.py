"""Mixed precision (FP32 smoothing inside FP64 defect correction) on the GPU.

The mixed solve converges to the same discrete solution as the FP64 path;
the bar stated here is on the result, not on bits: after each time step the
FP64 residual (ErrorAFAS on the finest level, the FP64 path's stop test) is
below qerrortol, phi and mu differ from the FP64 oracle's by less than
TOL_FIELD (a multiple of the solver tolerance), and the composite mass
h^d sum(phi) over the root (which FillDown makes the composite integral)
differs from the oracle's by less than TOL_MASS per step taken.

Mass bound: summing the phi equation phi - dt Lap(mu) = f over the unit
domain, two solutions whose residuals r1, r2 both meet the stop test differ in
mass by |mean(r1 - r2)| <= 2 max RMS(r) (the Laplacian sums to zero up to the
coarse-fine flux mismatch, second order in the field difference), and the
stop test bounds RMS(r) by qerrortol, so each step adds at most
2 * qerrortol.
"""

import numpy as np
import pytest

import orc
from bsam_amd.engine import PREC_MIXED, Engine
from helpers import (layout_amr2d, layout_amr2d_tiles, layout_amr3d, layout_periodic2d,
                     layout_single, noise_q)

pytestmark = pytest.mark.gpu

TOL_FIELD = 2e-6     # max |phi - phi_fp64|, |mu - mu_fp64| after a step
QERRORTOL = 1e-8
TOL_MASS = 2 * QERRORTOL   # |mass - mass_fp64| added per step (domain area 1)
LAYOUTS = {
    "single32": lambda: layout_single(32),
    "amr2d": layout_amr2d,
    "amr2d_tiles": layout_amr2d_tiles,
    "amr3d": layout_amr3d,
    "periodic2d": lambda: layout_periodic2d((2, 2, 2, 2)),
}


def root_mass(lay, q_root):
    p = lay.patches[0]
    inner = q_root[(slice(1, -1),) * lay.ndims + ((0,) if lay.ndims == 2 else ()) + (0,)]
    return float(inner.sum()) * p.dx ** lay.ndims


# inner FP32 cycles per defect-correction step, and the LDS-staged relaxation
@pytest.mark.parametrize("inner,staged", [(1, 0), (2, 0), (3, 0), (2, 1)])
@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_mixed_matches_fp64(name, inner, staged):
    lay = LAYOUTS[name]()
    kw = dict(dt=1e-3, eps=0.05 if lay.ndims == 3 else 0.03, well="plusminus",
              maxvcycles=40, qerrortol=QERRORTOL)
    q0 = noise_q(lay, seed=21, mean=0.1, amp=0.05)
    eng = Engine(lay, device=0, precision=PREC_MIXED,
                 tune={"inner_cycles": inner, "relax_lds": staged}, **kw)
    ora = orc.Oracle(lay, **kw)
    for n, p in enumerate(ora.patches):
        eng.upload(n, "q", q0[p.id])
        ora.field(n, "q")[...] = q0[p.id]
    root = lay.index_map()[0]
    for step in range(2):
        hist, conv = eng.time_step()
        ref = ora.time_step()
        assert conv, hist
        assert hist[-1] < 1e-8
        worst = 0.0
        for n, p in enumerate(ora.patches):
            if p.level < 0:
                continue
            got, want = eng.download(n, "q"), ora.field(n, "q")
            inner_sl = (slice(1, -1),) * lay.ndims
            worst = max(worst, float(np.abs(got[inner_sl] - want[inner_sl]).max()))
        assert worst < TOL_FIELD, (step, worst)
        m_got = root_mass(lay, eng.download(root, "q"))
        m_want = root_mass(lay, ora.field(root, "q"))
        assert abs(m_got - m_want) < TOL_MASS * (step + 1), (m_got, m_want)
    eng.close()


@pytest.mark.parametrize("inner", [1, 3])
@pytest.mark.parametrize("name", ["amr2d_tiles", "amr3d", "blocks3d"])
def test_mixed_iso_matches_mixed(name, inner):
    """The FP32 smoothing of isolated patches out of LDS, and the coarse
    loading over fine tiles, evaluate the per-colour launches' and the
    per-coarse-cell loading's arithmetic in the same order: the mixed solve
    with both (on every level, coarse single-patch levels run by level
    launches included) leaves q, cycle counts and residuals bit-identical to
    the mixed solve without them."""
    from bsam_amd.workloads import nested_block_layout_3d
    lay = (nested_block_layout_3d(32, 2, (0.25, 0.25), seed=5, sides=(8, 16))
           if name == "blocks3d" else LAYOUTS[name]())
    kw = dict(dt=1e-3, eps=0.05 if lay.ndims == 3 else 0.03, well="plusminus",
              maxvcycles=40, qerrortol=QERRORTOL)
    q0 = noise_q(lay, seed=23, mean=0.1, amp=0.05)
    runs = []
    for tune in ({"iso_min_cells": 0, "tail_cells": 0, "cgl_fine_cells": -1},
                 {"iso_min_cells": 1, "tail_cells": 0, "cgl_fine_cells": 0}):
        eng = Engine(lay, device=0, precision=PREC_MIXED,
                     tune=dict(tune, inner_cycles=inner), **kw)
        if tune["iso_min_cells"]:
            assert eng.tune_get("iso_patches") > 0
        for n, p in enumerate(lay.ordered()):
            eng.upload(n, "q", q0[p.id])
        hists = [eng.time_step()[0] for _ in range(2)]
        runs.append((hists, [eng.download(n, "q") for n in range(len(lay.ordered()))]))
        eng.close()
    (h0, q0s), (h1, q1s) = runs
    assert h0 == h1
    for n, (a, b) in enumerate(zip(q0s, q1s)):
        assert np.array_equal(a, b), n
