"""bsam_set_dt: the time step changed on a committed context (TakeTimeSteps'
syncelliptic pass runs with dt = 0 and restores dt after it,
lib/bsamroutines.f90:620-622, :748-750). A context whose dt is changed, after
it has run steps with captured graphs and the tail kernel, must step exactly
as a context created with the new dt, and dt = 0 must leave c unchanged."""

import numpy as np
import pytest

from bsam_amd.engine import BsamError, Engine
from helpers import layout_amr2d, layout_amr3d, noise_q

pytestmark = pytest.mark.gpu

KW = dict(eps=0.03, maxvcycles=40, qerrortol=1e-8, device=0)


@pytest.mark.parametrize("mk", [layout_amr2d, layout_amr3d])
def test_set_dt_matches_fresh_context(mk):
    lay = mk()
    q0 = noise_q(lay, seed=91)
    a = Engine(lay, dt=2e-3, **KW)
    b = Engine(lay, dt=1e-3, **KW)
    for e in (a, b):
        for n, p in enumerate(e.patches):
            e.upload(n, "q", q0[p.id])
    a.time_step()                        # graphs and tail state at dt 2e-3
    for n, p in enumerate(a.patches):    # back to the same start
        a.upload(n, "q", q0[p.id])
    a.set_dt(1e-3)
    for _ in range(2):
        ra, _ = a.time_step()
        rb, _ = b.time_step()
        assert ra == rb
    for n, _ in enumerate(a.patches):
        assert np.array_equal(a.download(n, "q"), b.download(n, "q")), n
    # dt = 0: phi stays what it was (only mu is solved for)
    before = [a.download(n, "q")[..., 0].copy() for n, _ in enumerate(a.patches)]
    a.set_dt(0.0)
    a.time_step()
    for n, p in enumerate(a.patches):
        if p.level >= 0:
            got = a.download(n, "q")[..., 0]
            inner = (slice(1, -1),) * lay.ndims
            assert np.allclose(got[inner], before[n][inner], rtol=0, atol=1e-9), n
    with pytest.raises(BsamError):
        a.set_dt(-1.0)
    a.close()
    b.close()
