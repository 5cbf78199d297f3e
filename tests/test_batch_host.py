"""Batched hierarchies (several requests in one engine context, one group
each) checked on the CPU: the ghost-fill programs of the batch, applied on
host arrays, reproduce every request's own SetGhost (the oracle on that
request alone) bit for bit, on every level and colour pass. So the groups
share no ghost cell and each keeps its own geometry."""

import numpy as np
import pytest

import orc
from bsam_amd import engine as E
from bsam_amd.layout import Layout


def requests(ndims, side, n):
    return [Layout(ndims, (side,) * ndims, 1.0 / side) for _ in range(n)]


@pytest.mark.parametrize("ndims,side", [(2, 32), (3, 8)])
def test_batch_ghost_programs_match_each_request(ndims, side):
    lays = requests(ndims, side, 3)
    eps = [0.02, 0.03, 0.04]
    eng = E.BatchEngine(lays, eps, dt=1e-3, well="plusminus", device=-1)
    assert E.lib().bsam_num_groups(eng.ctx) == 3
    rng = np.random.default_rng(1)
    lo = lays[0].minlevel
    for level in range(lo, lays[0].finestlevel + 1):
        for ipass in (0, 1, 2):
            oras = [orc.Oracle(l, dt=1e-3, eps=e, well="plusminus") for l, e in zip(lays, eps)]
            qs = [None] * len(eng.patches)
            for g, (lay, ora) in enumerate(zip(lays, oras)):
                for n, p in enumerate(lay.ordered()):
                    a = rng.uniform(-1, 1, lay.field_shape(p, "q"))
                    ora.field(n, "q")[...] = a
                    qs[eng.index[(g, n)]] = np.asfortranarray(a.copy())
            eng.ghost_apply_host(level, ipass, qs)
            for g, (lay, ora) in enumerate(zip(lays, oras)):
                ora.set_ghost(level, ipass)
                for n, p in enumerate(lay.ordered()):
                    assert np.array_equal(ora.field(n, "q"), qs[eng.index[(g, n)]]), \
                        (level, ipass, g, n)
    eng.close()


def test_groups_refuse_periodic_roots():
    lays = [Layout(2, (16, 16), 1.0 / 16, mthbc=(2, 2, 1, 1)) for _ in range(2)]
    with pytest.raises(E.BsamError, match="periodic"):
        E.BatchEngine(lays, [0.03, 0.03], dt=1e-3, well="plusminus", device=-1)


def test_batch_needs_one_shape_class():
    with pytest.raises(E.BsamError, match="minlevel"):
        E.BatchEngine([Layout(2, (16, 16), 1.0 / 16), Layout(2, (32, 32), 1.0 / 32)],
                      [0.03, 0.03], dt=1e-3, well="plusminus", device=-1)
