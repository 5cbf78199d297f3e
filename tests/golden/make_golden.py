"""Write the golden fixtures of tests/golden/ from the reference itself.

Runs the reference BSAM modules (compiled unmodified by `make -C oracle ref`
into oracle/_ref/harness) on small seeded inputs and stores inputs and
outputs as .npz (numpy arrays only, no pickles). Run in the build container:

    python tests/golden/make_golden.py [case ...]

With case names, writes only those fixtures (the layout listing is always
rewritten).
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import refrun  # noqa: E402
from golden_cases import LAYOUTS, OPS_CASES, QPO, STEP_CASES  # noqa: E402


def pack_fields(prefix, layout, fields, names, out):
    for n, p in enumerate(layout.ordered()):
        for name in names:
            out[f"{prefix}/{n}/{name}"] = np.asarray(fields[p.id][name])


def main(only=()):
    if not refrun.available():
        raise SystemExit("build the reference harness first: make -C oracle ref")
    listing = {}
    for name, mk in LAYOUTS.items():
        lay = mk()
        zero = {p.id: {k: np.zeros(lay.field_shape(p, k), order="F") for k in ("q", "f", "qc")}
                for p in lay.ordered()}
        _, _, lst = refrun.run(lay, zero, [])
        listing[name] = lst
    with open(os.path.join(HERE, "layouts_reference.json"), "w") as fh:
        json.dump(listing, fh, indent=0)

    for case, (lname, seed, ops) in OPS_CASES.items():
        if only and case not in only:
            continue
        lay = LAYOUTS[lname]()
        rng = np.random.default_rng(seed)
        fields = {p.id: {k: np.asfortranarray(rng.uniform(-1, 1, lay.field_shape(p, k)))
                         for k in ("q", "f", "qc")} for p in lay.ordered()}
        out, res, _ = refrun.run(lay, fields, ops, qpo=QPO.get(case))
        arrs = {}
        pack_fields("in", lay, fields, ("q", "f", "qc"), arrs)
        names = ("q", "f", "qc")
        if any(op[0] == "truncerr" for op in ops):
            names += ("qold", "qrte")
        if any(op[0] == "userflags" for op in ops):
            names += ("errorflags",)
        stats = [r for s, _, r in res if s == -1]
        if stats:
            arrs["stats"] = np.array(stats)
        pack_fields("out", lay, out, names, arrs)
        np.savez_compressed(os.path.join(HERE, f"ops_{case}.npz"), **arrs)

    for case, spec in STEP_CASES.items():
        if only and case not in only:
            continue
        lay = LAYOUTS[spec["layout"]]()
        fields = spec["init"](lay)
        kw = dict(dt=spec["dt"], alpha=spec["alpha"], eps=spec["eps"], maxv=spec["maxv"],
                  tol=spec["tol"], qpo=QPO.get(case))
        out, res, _ = refrun.run(lay, fields, [("step", spec["steps"], 0)], **kw)
        arrs = {"residuals": np.array([[s, it, r] for s, it, r in res])}
        pack_fields("in", lay, fields, ("q",), arrs)
        keep = spec.get("keep_levels")
        for n, p in enumerate(lay.ordered()):
            if keep is None or p.level in keep:
                arrs[f"out/{n}/q"] = out[p.id]["q"]
        np.savez_compressed(os.path.join(HERE, f"steps_{case}.npz"), **arrs)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
