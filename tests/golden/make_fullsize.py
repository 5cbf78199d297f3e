"""Write tests/golden/fullsize.json: the FP64 oracle's results on the
full-size cases of fullsize_cases.py (C4: 10 time steps of the 3D AMR
workload, 14.2 M cells; the largest C5 requests, 1024^2 and 128^3, 20 time
steps each).

The oracle (oracle/liborc.so, the C restatement pinned bit for bit against
the reference's own modules by tests/test_oracle_golden.py) runs each case
from its seeded initial state on one core; this takes several minutes, so
the GPU tests read the recorded results instead of rerunning it. Run in the
build container after `make -C oracle orc`:

    python tests/golden/make_fullsize.py [case ...]
"""

from __future__ import annotations

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, HERE, os.path.join(ROOT, "oracle")):
    sys.path.insert(0, p)

import orc  # noqa: E402
from fullsize_cases import CASES, case, level_digests, solver_kw  # noqa: E402

OUT = os.path.join(HERE, "fullsize.json")


def run(name):
    lay, q0, params, steps = case(name)
    o = orc.Oracle(lay, **solver_kw(params))
    for n, p in enumerate(o.patches):
        o.field(n, "q")[...] = q0[p.id]
    rec = {"steps": []}
    t0 = time.time()
    for s in range(steps):
        hist = o.time_step()
        mass, energy = o.after_step_stats()
        rec["steps"].append({"cycles": len(hist), "hist": hist, "mass": mass,
                             "energy": energy})
        print(f"{name} step {s + 1}/{steps}: {len(hist)} V-cycles, residual {hist[-1]:.3e}, "
              f"{time.time() - t0:.0f} s", flush=True)
    # level "0" is the root patch alone
    rec["q_sha256"] = level_digests(lay, lambda n: o.field(n, "q"))
    return rec


def main(names):
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as fh:
            data = json.load(fh)
    for name in names:
        data[name] = run(name)
        with open(OUT, "w") as fh:
            json.dump(data, fh, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:] or list(CASES))
