#!/usr/bin/env python3
"""Benchmark of the FAS multigrid engine on one MI355X.

Workload (one "step" = one implicit time step, i.e. one MultigridIterations
solve to the 1e-8 residual on a fixed hierarchy): 2D AMR Cahn-Hilliard,
root 256^2 on [0,1]^2 with homogeneous Neumann boundaries plus two refined
levels tagged where |grad phi| > 0.5 max (blocks of 16..64 cells per side),
eps = 0.03, dt = 1e-3, +-1 double well, phi0 = tanh((0.25-r)/(sqrt(2) eps))
with r the distance to 5 seeded centres; plain FP64 V(2,2) cycles.

value = relaxation cell-updates per second over the timed steps (a
cell-update is one full red-black relaxation of one cell's (phi, mu); per
V-cycle sum over levels of 2*nu*N_L, nu*N at the coarsest level), inputs
resident in HBM. The roofline object is for the relaxation kernel: after the
timed steps, the V-cycle's mix of relaxation launches (both colours on every
level above the tail) is replayed back to back between two HIP events on the
engine's stream, giving its mean launch duration; achieved = algorithmic bytes
per launch / that duration, at 48 B per cell-update (q read 16, f read 16, q
written 16), i.e. 24 B per cell of the level for one colour's half-sweep.
traffic is the measured HBM bytes per relaxation launch recorded in
profiles/k_relax_c2_pmc.md (rocprofv3 FETCH_SIZE and WRITE_SIZE, calibrated on
a copy kernel), null when that file is absent.
cpu_baseline is the reference itself (its Fortran modules built from the
reference sources, oracle/_ref/harness) taking time steps of the same workload
on one host core; where that binary is absent it is the FP64 oracle (the
single-threaded C restatement of the reference).
"""

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "V-cycle cell-updates/s + HBM roofline %; solves/s & p99 latency at 1e-8 resid"
HBM_PEAK_GBS = 8000.0


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * p / 100.0
    lo, hi = int(k), min(int(k) + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def measured_traffic():
    """HBM bytes per relaxation launch from the committed counter summary."""
    path = os.path.join(ROOT, "profiles", "k_relax_c2_pmc.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def cpu_baseline_reference(layout, q0, params, budget_s):
    """The reference's own Fortran modules (oracle/_ref/harness, built from the
    reference sources with flang -O2) taking time steps of the same workload
    on one host core, or None where the harness is not there. The +-1 well is
    the reference's well under c = (phi+1)/2, alpha = 8, eps_ref = sqrt(2) eps,
    dt_ref = dt/2 (mu unchanged). A 1-step run is subtracted from an N-step run
    so that reading the input and building the tree do not count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    if not refrun.available() or params["well"] != "plusminus":
        return None
    arrays = {}
    for p in layout.ordered():
        q = np.array(q0[p.id], dtype=np.float64)
        q[..., 0] = (q[..., 0] + 1.0) / 2.0
        arrays[p.id] = {"q": q, "f": np.zeros(layout.field_shape(p, "f")),
                        "qc": np.zeros(layout.field_shape(p, "qc"))}
    kw = dict(dt=params["dt"] / 2.0, alpha=8.0, eps=math.sqrt(2.0) * params["eps"],
              nsmooth=params["nsmoothingpasses"], maxv=50, tol=params["qerrortol"])

    def timed(steps):
        t0 = time.perf_counter()
        _, res, _ = refrun.run(layout, arrays, [("step", steps, 0)], **kw)
        return time.perf_counter() - t0, len(res)

    try:
        t1, c1 = timed(1)
        t3, _ = timed(3)
        steps = max(3, int(budget_s / max((t3 - t1) / 2.0, 1e-3)) + 1)
        tn, cn = timed(steps)
    except (RuntimeError, OSError):
        return None
    if tn <= t1 or cn <= c1:
        return None
    per_v = layout.relax_cell_updates_per_vcycle(params["nsmoothingpasses"])
    return {"value": (cn - c1) * per_v / (tn - t1), "unit": "cell-updates/s", "cores": 1,
            "kind": "reference",
            "sample": f"time steps 2..{steps} ({cn - c1} V-cycles) of the same workload from "
                      f"t=0: the reference's Fortran modules (flang -O2, single-threaded) "
                      f"on the mapped well, {tn - t1:.1f} s"}


def cpu_baseline(layout, q0, params, budget_s):
    """Oracle time steps from the same initial state until budget_s elapses."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    o = orc.Oracle(layout, dt=params["dt"], eps=params["eps"], well=params["well"],
                   nsmoothingpasses=params["nsmoothingpasses"], maxvcycles=50,
                   qerrortol=params["qerrortol"])
    for n, p in enumerate(o.patches):
        o.field(n, "q")[...] = q0[p.id]
    per_v = layout.relax_cell_updates_per_vcycle(params["nsmoothingpasses"])
    t0 = time.perf_counter()
    steps = cycles = 0
    while True:
        cycles += len(o.time_step())
        steps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": cycles * per_v / dt, "unit": "cell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"{steps} time steps ({cycles} V-cycles) of the same workload from "
                      f"t=0, single-threaded FP64 C oracle, {dt:.1f} s"}


WORKLOADS = {
    "c2": ("2D AMR Cahn-Hilliard, root 256^2 + 2 refined levels (ratio 2), "
           "FAS V(2,2) to residual 1e-8, one time step per step"),
    "c3": ("2D AMR Cahn-Hilliard, root 512^2 + 3 refined levels, ~2000 uneven patches "
           "8^2..128^2, FAS V(2,2) to residual 1e-8, one time step per step"),
    "c4": ("3D AMR Cahn-Hilliard, root 128^3 + 2 refined levels of 16^3/32^3 blocks, "
           "FAS V(2,2) to residual 1e-8, one time step per step"),
}


def load_workload(name):
    from bsam_amd import workloads as W
    if name == "c2":
        lay, q0 = W.config_c2()
        return lay, q0, dict(W.C2_PARAMS)
    if name == "c3":
        lay, q0 = W.config_c3()
        return lay, q0, dict(W.C3_PARAMS)
    lay, q0 = W.config_c4()
    return lay, q0, dict(W.C4_PARAMS)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-budget", type=float, default=15.0,
                    help="seconds of CPU-baseline work (0 to skip)")
    ap.add_argument("--config", choices=sorted(WORKLOADS), default="c2",
                    help="workload (c2 is the headline configuration)")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--check", action="store_true",
                    help="mixed precision: also run the FP64 path and report the "
                         "difference (max |phi|, |mu| and root mass) after the steps")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the solver path here is single-GPU "
              "(run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    from bsam_amd import engine as E

    layout, q0, params = load_workload(args.config)
    prec = E.PREC_MIXED if args.precision == "mixed" else E.PREC_FP64

    def make_engine(precision):
        eng = E.Engine(layout, dt=params["dt"], eps=params["eps"], well=params["well"],
                       nsmoothingpasses=params["nsmoothingpasses"], maxvcycles=50,
                       qerrortol=params["qerrortol"], precision=precision, device=0)
        for n, p in enumerate(eng.patches):
            eng.upload(n, "q", q0[p.id])
        return eng

    eng = make_engine(prec)
    per_v = layout.relax_cell_updates_per_vcycle(params["nsmoothingpasses"])
    if prec == E.PREC_MIXED:
        per_v *= eng.tune_get("inner_cycles")

    for _ in range(args.warmup):
        eng.time_step()
    eng.sync()

    lat, cycles, conv_all = [], 0, True
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        hist, conv = eng.time_step()   # returns after the convergence test
        lat.append(time.perf_counter() - s0)
        cycles += len(hist)
        conv_all = conv_all and conv
    eng.sync()
    elapsed = time.perf_counter() - t0

    check = None
    if args.check and prec == E.PREC_MIXED:
        ref = make_engine(E.PREC_FP64)
        for _ in range(args.warmup + args.steps):
            ref.time_step()
        worst = 0.0
        for n, p in enumerate(layout.ordered()):
            if p.level < 0:
                continue
            a, b = eng.download(n, "q"), ref.download(n, "q")
            inner = (slice(1, -1),) * layout.ndims
            worst = max(worst, float(np.abs(a[inner] - b[inner]).max()))
        m, e = eng.after_step_stats()
        mr, er = ref.after_step_stats()
        check = {"vs": "FP64 GPU path (bit-identical to the FP64 oracle)",
                 "max_abs_field_diff": worst,
                 # composite mass of phi (and 2D energy) as TakeTimeSteps' statistics
                 "mass": m, "mass_fp64": mr, "mass_abs_diff": abs(m - mr),
                 "mass_rel_diff": abs(m - mr) / abs(mr),
                 "energy_rel_diff": abs(e - er) / abs(er) if er else None,
                 "steps": args.warmup + args.steps}
        ref.close()

    # Roofline of the relaxation kernel: the V-cycle's relaxation launch mix,
    # back to back between two HIP events on the engine's stream.
    mass, energy = eng.after_step_stats()
    reps = 50
    avg_us, bytes_per_launch = eng.profile_kernel("relax", reps)
    avg_s = avg_us * 1e-6
    achieved = bytes_per_launch / avg_s / 1e9
    traffic = measured_traffic() if args.config == "c2" else None

    value = cycles * per_v / elapsed
    levels = layout.finestlevel
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == E.PREC_FP64 else "f32 smoothing in f64 defect correction",
        "data": "synthetic",
        "config": {
            "workload": WORKLOADS[args.config] + (
                "" if prec == E.PREC_FP64 else " (mixed precision)"),
            "patches": len(layout.ordered()),
            "cells": sum(layout.cells_on_level(l) for l in range(0, levels + 1)),
            "refined_levels": levels,
            "cell_updates_per_vcycle": per_v,
            "vcycles": cycles, "all_converged": conv_all,
            "solves_per_s": args.steps / elapsed,
            "p50_ms": 1e3 * percentile(lat, 50), "p99_ms": 1e3 * percentile(lat, 99),
            # TakeTimeSteps' after-step statistics of the final state
            "mass": mass, "energy": energy if layout.ndims == 2 else None,
        },
        "roofline": {
            "bound": "hbm", "kernel": f"k_relax<{layout.ndims}> (red-black half-sweep, FP64)",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "avg_launch_us": avg_us, "bytes_per_launch": bytes_per_launch,
        },
        "cpu_baseline": None,
    }
    if check:
        out["check"] = check
    eng.close()
    if args.cpu_budget > 0:
        out["cpu_baseline"] = (cpu_baseline_reference(layout, q0, params, args.cpu_budget) or
                               cpu_baseline(layout, q0, params, args.cpu_budget))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
