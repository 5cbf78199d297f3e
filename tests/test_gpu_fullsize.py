"""The GPU engine against the FP64 oracle at the benchmark sizes.

The parity tests use small hierarchies; these run the benchmark workloads
themselves (bsam_amd/workloads.py) once: a whole time step of C2 (root
256^2 + 2 refined levels, 466 patches), a V-cycle of C3 (root 512^2 + 3
levels, ~2000 uneven patches) and of C4 (3D root 128^3 + 2 levels, 14 M
cells). Fields must be bit-identical, residuals equal to 1e-12.
"""

import numpy as np
import pytest

import orc
from bsam_amd import workloads as W
from bsam_amd.engine import Engine

pytestmark = pytest.mark.gpu


def setup(layout, q0, params, maxv=50):
    kw = dict(dt=params["dt"], eps=params["eps"], well=params["well"],
              nsmoothingpasses=params["nsmoothingpasses"], maxvcycles=maxv,
              qerrortol=params["qerrortol"])
    eng = Engine(layout, device=0, **kw)
    ora = orc.Oracle(layout, **kw)
    for n, p in enumerate(eng.patches):
        eng.upload(n, "q", q0[p.id])
        ora.field(n, "q")[...] = q0[p.id]
    return eng, ora


def same(eng, ora, names=("q",)):
    for n, p in enumerate(ora.patches):
        for name in names:
            a, b = eng.download(n, name), ora.field(n, name)
            if not np.array_equal(a, b):
                bad = np.argwhere(a != b)
                raise AssertionError(f"patch {p.id} level {p.level} {name}: {len(bad)} cells "
                                     f"differ, first {bad[0].tolist()}")


def test_c2_time_step():
    lay, q0 = W.config_c2()
    eng, ora = setup(lay, q0, W.C2_PARAMS)
    hist, conv = eng.time_step()
    ref = ora.time_step()
    assert conv and len(hist) == len(ref) >= 3
    np.testing.assert_allclose(hist, ref, rtol=1e-12, atol=0)
    same(eng, ora, ("q", "qold"))
    eng.close()


@pytest.mark.parametrize("cfg", ["c3", "c4"])
def test_vcycle(cfg):
    if cfg == "c3":
        lay, q0 = W.config_c3()
        params = W.C3_PARAMS
    else:
        lay, q0 = W.config_c4()
        params = W.C4_PARAMS
    eng, ora = setup(lay, q0, params)
    for e in (eng, ora):
        e.copy_qold()
        e.fill_down(2)
        e.source()
    eng.vcycle(lay.finestlevel)
    ora.vcycle(lay.finestlevel)
    same(eng, ora, ("q", "f"))
    r, want = eng.error_afas(lay.finestlevel), ora.error_afas(lay.finestlevel)
    assert abs(r - want) <= 1e-12 * want
    eng.close()
