"""Schedule autotuner: times one workload under each combination of the
engine's schedule knobs and reports the fastest.

Knobs (bsam_tune): the level kernels' tile shape (tile_x x tile_y, in 3D
times the extent that fills 256 threads), tail_cells (which single-patch
coarse levels run inside the one-workgroup tail kernel), tail_threads (its
workgroup size) and tail_lds (whether the tail keeps its levels in LDS).
Every candidate solves the same time steps from the same initial state;
results are bit-identical across candidates (the schedule changes where work
runs, never the arithmetic), which the sweep checks on the final root field.

    python -m bsam_amd.autotune [--config c2|c3|c4] [--steps 5] [--knobs tiles|tail]
"""

from __future__ import annotations

import argparse
import itertools
import json
import time
from typing import Dict, List

import numpy as np

from . import engine as E

TAIL_GRID = {
    "tail_cells": [0, 256, 1024, 4096, 16384],
    "tail_threads": [256, 512, 1024],
    "tail_lds": [-1, 0],
}
TILES_2D = [(32, 8), (64, 4), (16, 16), (128, 2), (256, 1)]
TILES_3D = [(8, 8), (16, 4), (32, 4), (16, 8), (32, 2), (64, 2), (16, 16), (32, 8)]


def time_schedule(layout, q0, params, tune: Dict[str, float], steps: int, warmup: int = 1):
    eng = E.Engine(layout, dt=params["dt"], eps=params["eps"], well=params["well"],
                   nsmoothingpasses=params["nsmoothingpasses"], maxvcycles=50,
                   qerrortol=params["qerrortol"], device=0, tune=tune)
    for n, p in enumerate(eng.patches):
        eng.upload(n, "q", q0[p.id])
    for _ in range(warmup):
        eng.time_step()
    eng.sync()
    cycles = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        hist, _ = eng.time_step()
        cycles += len(hist)
    eng.sync()
    dt = time.perf_counter() - t0
    root = eng.download(layout.index_map()[0], "q")
    eng.close()
    return dt / cycles, cycles, root


def workload(name: str):
    from . import workloads as W
    if name == "c2":
        lay, q0 = W.config_c2()
        return lay, q0, dict(W.C2_PARAMS)
    if name == "c3":
        lay, q0 = W.config_c3()
        return lay, q0, dict(W.C3_PARAMS)
    lay, q0 = W.config_c4()
    return lay, q0, dict(W.C4_PARAMS)


def candidates(knobs: str, ndims: int):
    if knobs == "tiles":
        for tx, ty in (TILES_2D if ndims == 2 else TILES_3D):
            yield {"tile_x": tx, "tile_y": ty}
        return
    keys = list(TAIL_GRID)
    for combo in itertools.product(*(TAIL_GRID[k] for k in keys)):
        tune = dict(zip(keys, combo))
        if tune["tail_cells"] == 0 and (tune["tail_threads"] != 1024 or tune["tail_lds"] != -1):
            continue   # no tail: the other tail knobs do nothing
        yield tune


def sweep(config: str = "c2", steps: int = 5, knobs: str = "tail"):
    layout, q0, params = workload(config)
    out, ref = [], None
    for tune in candidates(knobs, layout.ndims):
        per_v, cycles, root = time_schedule(layout, q0, params, tune, steps)
        same = True
        if ref is None:
            ref = root
        else:
            same = bool(np.array_equal(root, ref))
        out.append({"tune": tune, "us_per_vcycle": per_v * 1e6, "vcycles": cycles,
                    "bitwise_same": same})
        print(json.dumps(out[-1]), flush=True)
    best = min(out, key=lambda r: r["us_per_vcycle"])
    return best, out


def main():
    ap = argparse.ArgumentParser(description="sweep the engine's schedule knobs")
    ap.add_argument("--config", choices=("c2", "c3", "c4"), default="c2")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--knobs", choices=("tiles", "tail"), default="tiles")
    args = ap.parse_args()
    best, _ = sweep(args.config, args.steps, args.knobs)
    print(json.dumps({"best": best}))


if __name__ == "__main__":
    main()
