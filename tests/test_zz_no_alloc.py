"""The allocation-free after-step statistics (bsam_after_step_stats_noalloc)
against bsam_after_step_stats, through the C ABI: two calls in a row on
different hierarchies, the second smaller than the first, each bit-identical
to the allocating entry point on the same fields."""

import ctypes

import numpy as np
import pytest

from bsam_amd import engine as E
from bsam_amd import workloads as W
from helpers import layout_amr2d, noise_q

pytestmark = pytest.mark.gpu


def stats(eng, fn):
    m, e = ctypes.c_double(), ctypes.c_double()
    rc = fn(eng.ctx, ctypes.byref(m), ctypes.byref(e))
    assert rc == 0, E.lib().bsam_last_error(eng.ctx)
    return m.value, e.value


def run(lay, q0, ndims_eps):
    eng = E.Engine(lay, dt=1e-3, eps=ndims_eps, well="plusminus", device=0)
    for n, p in enumerate(eng.patches):
        eng.upload(n, "q", q0[p.id])
    eng.time_step()
    L = E.lib()
    L.bsam_after_step_stats_noalloc.argtypes = L.bsam_after_step_stats.argtypes
    want = stats(eng, L.bsam_after_step_stats)
    got = stats(eng, L.bsam_after_step_stats_noalloc)
    again = stats(eng, L.bsam_after_step_stats_noalloc)
    eng.close()
    return want, got, again


def test_noalloc_stats_match_allocating_entry_point():
    big, qbig = W.config_c2()                        # 466 patches, 2D
    small = layout_amr2d()                            # 7 patches
    qsmall = noise_q(small, seed=5, mean=0.1, amp=0.05)
    for lay, q0 in ((big, qbig), (small, qsmall)):
        want, got, again = run(lay, q0, 0.03)
        assert got == want and again == want, (want, got, again)
        assert want[0] != 0.0 and want[1] != 0.0


def test_noalloc_stats_3d():
    lay, q0 = W.config_c4()
    want, got, _ = run(lay, q0, 0.02)
    assert got == want and want[1] == 0.0
