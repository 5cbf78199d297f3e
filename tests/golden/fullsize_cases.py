"""Full-size cases of the benchmark configurations whose FP64 results are
recorded by make_fullsize.py (from the FP64 oracle) and checked on the GPU by
tests/test_gpu_configs.py.

At these sizes the oracle takes minutes (C4: about 30 s per time step on one
core), so the GPU tests do not rerun it: the fixture keeps, per time step,
the V-cycle count, the residual history and the after-step statistics (mass,
2D energy), and after the last step a SHA-256 of every level's q (all
patches of the level, ghost-extended, in patch order, the bytes of the
Fortran-order float64 arrays). Bit-identical fields give identical digests.
"""

from __future__ import annotations

import hashlib

from bsam_amd import serve
from bsam_amd import workloads as W

# C5 requests at the largest sizes of the serving configuration: a 1024^2
# 2D and a 128^3 3D request, 20 time steps each (bsam_amd.serve.solve_request
# parameters: dt 1e-3, maxvcycles 60, qerrortol 1e-8, +-1 well).
C5_BIG = {
    "c5_2d_1024": serve.Request(1001, 2, 1024, 0.02, 101, steps=20),
    "c5_3d_128": serve.Request(1002, 3, 128, 0.04, 102, steps=20),
}


def case(name):
    """(layout, initial q per patch id, solver parameters, steps) of a case."""
    if name == "c4":
        lay, q0 = W.config_c4()
        p = dict(W.C4_PARAMS, maxvcycles=50)
        return lay, q0, p, 10
    r = C5_BIG[name]
    lay = serve.request_layout(r)
    q0 = W.noise_q(lay, r.seed, 0.0, 0.05)
    p = dict(eps=r.eps, dt=r.dt, well="plusminus", nsmoothingpasses=2, qerrortol=1e-8,
             maxvcycles=60)
    return lay, q0, p, r.steps


CASES = ("c4", "c5_2d_1024", "c5_3d_128")


def solver_kw(p):
    return dict(dt=p["dt"], eps=p["eps"], well=p["well"], nsmoothingpasses=p["nsmoothingpasses"],
                maxvcycles=p["maxvcycles"], qerrortol=p["qerrortol"])


def level_digests(layout, get_q):
    """SHA-256 (hex) of each level's q: get_q(n) returns patch index n's
    ghost-extended q (Fortran order)."""
    out = {}
    hs = {}
    for n, p in enumerate(layout.ordered()):
        a = get_q(n)
        h = hs.setdefault(p.level, hashlib.sha256())
        h.update(a.tobytes(order="F"))
    for lev, h in sorted(hs.items()):
        out[str(lev)] = h.hexdigest()
    return out
