#!/usr/bin/env python3
"""Benchmark of the FAS multigrid engine on one MI355X.

Workload (default c4, the largest single-GPU configuration; one "step" = one
implicit time step, i.e. one MultigridIterations solve to the 1e-8 residual,
lib/afasroutines.f90:38-106, on a fixed hierarchy): 3D AMR Cahn-Hilliard,
root 128^3 on [0,1]^3 with homogeneous Neumann boundaries plus two refined
levels of 16^3 / 32^3 blocks (1963 patches, 14.2 M cells at levels >= 0),
eps = 0.02, dt = 5e-4, +-1 double well, phi0 = U(-0.05, 0.05) (seed 4);
plain FP64 V(2,2) cycles. --config c2 / c3 select the 2D AMR workloads,
--precision mixed the FP32-smoothing defect-correction path.

value = relaxation cell-updates per second over the timed steps (a
cell-update is one full red-black relaxation of one cell's (phi, mu); per
V-cycle the sum over levels of 2*nu*N_L, nu*N at the coarsest level), inputs
resident in HBM, the solve replayed as captured hipGraphs.

roofline: after the timed steps the same steps are run again from the same
initial state with HIP events around every kernel launch on the engine's
stream (bsam_profile_*; graphs off). The kernel kind with the largest summed
duration is the dominant kernel; achieved = its algorithmic bytes per launch
/ its mean launch duration (relaxation: 48 B per cell-update, i.e. 24 B per
cell of the level for one colour's half-sweep, plus 32 B per ghost update the
fused relaxation applies; the other kinds' figures are in
include/bsam_gpu.h), peak = the MI355X's 8 TB/s. traffic = the measured
HBM bytes per launch of that kernel on this workload, from the committed
rocprofv3 counter summary profiles/<kernel>_<workload>_pmc.md (FETCH_SIZE and
WRITE_SIZE calibrated on a copy kernel), null when there is none. The events
are stream markers and also count each launch's dispatch latency (a few us);
kernel_trace_avg_us / frac_kernel_trace give the same kernel's mean duration
from the committed kernel-trace summary profiles/<kernel>_<workload>_kernel_trace.md
and the fraction it implies. The line
also carries the measured copy rate of the device, the whole V-cycle's
algorithmic bytes per second (296 B per cell per V(2,2) level in 2D, 284 in
3D, 96 at the coarsest level, plus 104 B per finest-level cell per iteration)
and each kernel kind's share of the profiled time.

cpu_baseline: the reference itself (its Fortran modules built from the
reference sources, oracle/_ref/harness) on one host core: V-cycles of the
first solve of the same workload, timed as the difference of a 1-cycle and a
(1+m)-cycle MultigridIterations run so that input and tree set-up do not
count; where that binary is absent, the FP64 oracle (the single-threaded C
restatement of the reference).
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

# profile kind -> the kernel (template) names rocprofv3 reports for it
KERNELS = {
    "relax": "k_relax", "relax_fused": "k_relax_fused", "ghost": "k_ghost",
    "restrict": "k_restrict", "coarse_ghost": "k_coarse_ghost",
    "coarse_load": "k_coarse_ghost_load", "correct": "k_correct", "filldown": "k_filldown",
    "source": "k_source", "resnorm": "k_resnorm", "tail": "k_tail_vcycle", "copy": "k_copy",
    "defect": "k_defect_setup+k_apply_correction", "smooth_iso": "k_smooth_iso",
}

WORKLOADS = {
    "c2": ("2D AMR Cahn-Hilliard, root 256^2 + 2 refined levels (ratio 2), "
           "FAS V(2,2) to residual 1e-8, one time step per step"),
    "c3": ("2D AMR Cahn-Hilliard, root 512^2 + 3 refined levels, 2000 uneven patches "
           "8..128 cells per side (log-uniform), FAS V(2,2) to residual 1e-8, "
           "one time step per step"),
    "c4": ("3D AMR Cahn-Hilliard, root 128^3 + 2 refined levels of 16^3/32^3 blocks, "
           "FAS V(2,2) to residual 1e-8, one time step per step"),
}


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * p / 100.0
    lo, hi = int(k), min(int(k) + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    counter summary, or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def kernel_trace_avg_us(kernel, workload):
    """Mean kernel duration (us) of `kernel` on `workload` in the committed
    rocprofv3 kernel-trace summary, or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_kernel_trace.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("avg_ns"):
                    return float(line.split("=")[1]) / 1e3
    except OSError:
        pass
    return None


def vcycle_bytes(layout):
    """Algorithmic FP64 bytes of one multigrid iteration (V(2,2) cycle plus
    FillDown, source restore and residual norm on the finest level)."""
    per_cell = 296.0 if layout.ndims == 2 else 284.0
    lo, hi = layout.minlevel, layout.finestlevel
    b = 96.0 * layout.cells_on_level(lo)
    b += per_cell * sum(layout.cells_on_level(l) for l in range(lo + 1, hi + 1))
    return b + 104.0 * layout.cells_on_level(hi)


def mapped_reference_arrays(layout, q0):
    """The +-1 well is the reference's well under c = (phi+1)/2 with
    alpha = 8, eps_ref = sqrt(2) eps, dt_ref = dt/2 (mu unchanged)."""
    arrays = {}
    for p in layout.ordered():
        q = np.array(q0[p.id], dtype=np.float64)
        q[..., 0] = (q[..., 0] + 1.0) / 2.0
        arrays[p.id] = {"q": q, "f": np.zeros(layout.field_shape(p, "f")),
                        "qc": np.zeros(layout.field_shape(p, "qc"))}
    return arrays


def cpu_baseline_reference(layout, q0, params, budget_s):
    """The reference's own Fortran modules (oracle/_ref/harness, flang -O2)
    on one host core: MultigridIterations of the first time step with
    maxvcycles 1 and 1 + m; the difference times m V-cycles."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import refrun
    if not refrun.available() or params["well"] != "plusminus":
        return None
    arrays = mapped_reference_arrays(layout, q0)

    def timed(maxv):
        kw = dict(dt=params["dt"] / 2.0, alpha=8.0, eps=math.sqrt(2.0) * params["eps"],
                  nsmooth=params["nsmoothingpasses"], maxv=maxv, tol=params["qerrortol"])
        t0 = time.perf_counter()
        _, res, _ = refrun.run(layout, arrays, [("copyqold", 0, 0), ("mgiter", 0, 0)], **kw)
        return time.perf_counter() - t0, len(res)

    try:
        t1, c1 = timed(1)
        t2, c2 = timed(2)
        per = max(t2 - t1, 1e-3)
        m = max(2, int(budget_s / per))
        tm, cm = timed(1 + m)
    except (RuntimeError, OSError):
        return None
    if tm <= t1 or cm <= c1:
        return None
    per_v = layout.relax_cell_updates_per_vcycle(params["nsmoothingpasses"])
    return {"value": (cm - c1) * per_v / (tm - t1), "unit": "cell-updates/s", "cores": 1,
            "kind": "reference",
            "sample": f"V-cycles 2..{cm} of the first time step's solve of the same workload "
                      f"(MultigridIterations with maxvcycles {cm} minus maxvcycles 1): the "
                      f"reference's Fortran modules (flang -O2, single-threaded) on the mapped "
                      f"well, {tm - t1:.1f} s"}


def cpu_baseline_oracle(layout, q0, params, budget_s):
    """The FP64 C oracle on one core: time steps of the same workload from
    the same initial state until budget_s elapses."""
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
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="seconds of CPU-baseline V-cycles (0 to skip)")
    ap.add_argument("--config", choices=sorted(WORKLOADS), default="c4",
                    help="workload (c4, the largest single-GPU configuration, is the headline)")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--check", action="store_true",
                    help="mixed precision: also run the FP64 path and report the "
                         "difference (max |phi|, |mu| and mass) after the steps")
    ap.add_argument("--inner-cycles", type=int, default=3,
                    help="mixed precision: FP32 V-cycles per FP64 defect-correction step")
    ap.add_argument("--serving", type=int, default=256,
                    help="requests of the serving configuration to serve after the solve "
                         "measurement (0 to skip)")
    ap.add_argument("--serving-rate", type=float, default=50.0,
                    help="Poisson arrival rate of the served requests (per second)")
    ap.add_argument("--no-profile", action="store_true",
                    help="skip the event-timed pass (roofline null)")
    ap.add_argument("--tile", default=None,
                    help="level-kernel tile as X,Y (powers of two; the z extent fills 256 "
                         "cells in 3D) instead of the shape timed at commit, e.g. for "
                         "counter runs whose serialised kernels would distort that timing")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="engine tuning key set before the hierarchy (repeatable), "
                         "e.g. short_ghost=0")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the solver path here is single-GPU "
              "(run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    from bsam_amd import engine as E

    layout, q0, params = load_workload(args.config)
    prec = E.PREC_MIXED if args.precision == "mixed" else E.PREC_FP64
    finest = layout.finestlevel

    def make_engine(precision):
        tune = {"inner_cycles": args.inner_cycles} if precision == E.PREC_MIXED else {}
        if args.tile:
            tx, ty = (int(v) for v in args.tile.split(","))
            tune.update(tile_x=tx, tile_y=ty)
        for kv in args.tune:
            key, val = kv.split("=", 1)
            tune[key] = float(val)
        eng = E.Engine(layout, dt=params["dt"], eps=params["eps"], well=params["well"],
                       nsmoothingpasses=params["nsmoothingpasses"], maxvcycles=50,
                       qerrortol=params["qerrortol"], precision=precision, device=0,
                       tune=tune)
        for n, p in enumerate(eng.patches):
            eng.upload(n, "q", q0[p.id])
        return eng

    eng = make_engine(prec)
    per_v = layout.relax_cell_updates_per_vcycle(params["nsmoothingpasses"])
    if prec == E.PREC_MIXED:
        per_v *= eng.tune_get("inner_cycles")

    def run_steps(n):
        lat, cycles, conv_all, last = [], 0, True, None
        for _ in range(n):
            s0 = time.perf_counter()
            hist, conv = eng.time_step()   # returns after the convergence test
            lat.append(time.perf_counter() - s0)
            cycles += len(hist)
            conv_all = conv_all and conv
            last = hist[-1]
        return lat, cycles, conv_all, last

    run_steps(args.warmup)
    eng.sync()
    t0 = time.perf_counter()
    lat, cycles, conv_all, last_res = run_steps(args.steps)
    eng.sync()
    elapsed = time.perf_counter() - t0
    rms_f = eng.source_norm(finest)
    mass, energy = eng.after_step_stats()

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
        mr, er = ref.after_step_stats()
        check = {"vs": "FP64 GPU path (bit-identical to the FP64 oracle)",
                 "max_abs_field_diff": worst,
                 # composite mass of phi (and 2D energy) as TakeTimeSteps' statistics
                 "mass": mass, "mass_fp64": mr, "mass_abs_diff": abs(mass - mr),
                 "energy_rel_diff": abs(energy - er) / abs(er) if er else None,
                 "steps": args.warmup + args.steps}
        ref.close()

    # Event-timed pass: the same steps from the same initial state, every
    # launch bracketed by HIP events on the engine's stream.
    roof, kinds = None, None
    if not args.no_profile:
        for n, p in enumerate(eng.patches):
            eng.upload(n, "q", q0[p.id])
        eng.profile_begin(E.Engine.PROF_KINDS)
        _, pcycles, _, _ = run_steps(args.warmup + args.steps)
        stats = eng.profile_end()
        total_ms = sum(v["ms"] for v in stats.values())
        kinds = {k: {"kernel": KERNELS[k], "launches": v["launches"],
                     "ms_per_step": v["ms"] / (args.warmup + args.steps),
                     "share": v["ms"] / total_ms,
                     "GBs": v["bytes"] / (v["ms"] * 1e-3) / 1e9 if v["bytes"] else None}
                 for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"])}
        dom = max((k for k, v in stats.items() if v["bytes"] > 0), key=lambda k: stats[k]["ms"])
        v = stats[dom]
        avg_s = v["ms"] * 1e-3 / v["launches"]
        bpl = v["bytes"] / v["launches"]
        achieved = bpl / avg_s / 1e9
        kname = KERNELS[dom]
        wname = args.config + ("_mixed" if prec == E.PREC_MIXED else "")
        traffic = measured_traffic(kname, wname)
        trace_us = kernel_trace_avg_us(kname, wname)
        copy_gbs = E.copy_rate(0, 1 << 30, 20)
        vb = vcycle_bytes(layout)
        roof = {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": kname + ("" if prec == E.PREC_FP64 else " (FP32 defect sweeps included)"),
            "profile_kind": dom, "avg_launch_us": avg_s * 1e6, "bytes_per_launch": bpl,
            "launches": v["launches"], "share_of_gpu_time": v["ms"] / total_ms,
            "copy_peak_GBs": copy_gbs, "frac_of_copy_peak": achieved / copy_gbs,
            "traffic_ratio": traffic / bpl if traffic else None,
            # the committed kernel trace's mean duration of the same kernel:
            # the events above also count each launch's dispatch latency
            "kernel_trace_avg_us": trace_us,
            "frac_kernel_trace": bpl / (trace_us * 1e-6) / 1e9 / HBM_PEAK_GBS if trace_us
            else None,
            "vcycle_bytes": vb,
            "vcycle_GBs": vb * cycles / elapsed / 1e9,
            "vcycle_frac": vb * cycles / elapsed / 1e9 / HBM_PEAK_GBS,
            "profiled_pass_vcycles": pcycles,
        }

    value = cycles * per_v / elapsed
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
            "name": args.config,
            "patches": len(layout.ordered()),
            "cells": sum(layout.cells_on_level(l) for l in range(0, finest + 1)),
            "refined_cells": sum(layout.cells_on_level(l) for l in range(1, finest + 1)),
            "refined_levels": finest,
            "cell_updates_per_vcycle": per_v,
            "vcycles": cycles, "all_converged": conv_all,
            "solves_per_s": args.steps / elapsed,
            "p50_ms": 1e3 * percentile(lat, 50), "p99_ms": 1e3 * percentile(lat, 99),
            # the reference's stop test (absolute RMS of f - A(q), finest level)
            # and the relative residual RMS(f - A(q)) / RMS(f)
            "final_residual": last_res,
            "relative_residual": last_res / rms_f if rms_f else None,
            # TakeTimeSteps' after-step statistics of the final state
            "mass": mass, "energy": energy if layout.ndims == 2 else None,
            # level-kernel tile shape, chosen at commit by timing the relaxation
            # under each candidate shape on this hierarchy (tile_auto)
            "tile": [eng.tune_get("tile_x"), eng.tune_get("tile_y")] +
                    ([eng.tune_get("tile_z")] if layout.ndims == 3 else []),
            "tile_auto": {"on": bool(eng.tune_get("tile_auto")),
                          "candidates_timed": int(eng.tune_get("tile_auto_candidates")),
                          "relax_rounds_us": eng.tune_get("tile_auto_us")},
            # patches whose whole smoothing runs in one workgroup out of LDS
            "iso": {"patches": int(eng.tune_get("iso_patches")),
                    "cells": int(eng.tune_get("iso_cells"))},
            "tune": dict(kv.split("=", 1) for kv in args.tune) or None,
        },
        "roofline": roof,
        "kernels": kinds,
        "cpu_baseline": None,
    }
    if check:
        out["check"] = check
    eng.close()
    if args.serving > 0:
        out["serving"] = serving(args.serving, args.serving_rate)
    if args.cpu_budget > 0:
        out["cpu_baseline"] = (cpu_baseline_reference(layout, q0, params, args.cpu_budget) or
                               cpu_baseline_oracle(layout, q0, params, args.cpu_budget))
    print(json.dumps(out))


def serving(n, rate):
    """The serving configuration: n Cahn-Hilliard solve requests with Poisson
    arrivals at `rate` per second (2D side in {64..1024} or 3D side in
    {32, 64, 128}, eps ~ U(0.01, 0.05), seeded noise, 20 time steps each to
    the 1e-8 residual), served with continuous batching (batches of up to 8
    requests of one shape class per context, 8 contexts at a time, admission
    by HBM footprint) in mixed precision; beside it the same requests solved
    one after another in plain FP64 (the serial baseline). At a rate the
    server keeps up with, solves/s is set by the arrivals (n * 20 / the
    arrival span); the same requests are therefore also served as a burst,
    all waiting at t = 0, which measures the server's capacity."""
    from bsam_amd import engine as E
    from bsam_amd import serve
    reqs = serve.make_requests(n, rate, seed=5, steps=20)
    srv = serve.Server(budget_bytes=64 << 30, max_active=8, precision=E.PREC_MIXED, batch=8)
    got = srv.run(reqs, realtime=True)
    cap_srv = serve.Server(budget_bytes=64 << 30, max_active=8, precision=E.PREC_MIXED, batch=8)
    cap = cap_srv.run(reqs, realtime=False)
    base = serve.serial(reqs, E.PREC_FP64)
    return {
        "requests": n, "arrival_rate_per_s": rate, "steps_per_request": 20,
        "mode": "mixed precision, batches of <= 8 same-shape requests, <= 8 active contexts",
        "solves_per_s": got["solves_per_s"], "requests_per_s": got["requests_per_s"],
        "p50_latency_s": got["p50_latency_s"], "p99_latency_s": got["p99_latency_s"],
        "makespan_s": got["makespan_s"], "all_converged": got["all_converged"],
        "peak_active_contexts": srv.peak_active, "peak_bytes": srv.peak_used,
        "serial_fp64_solves_per_s": base["solves_per_s"],
        "speedup_vs_serial_fp64": got["solves_per_s"] / base["solves_per_s"],
        # all requests waiting at t = 0: the served rate is the capacity
        "burst_solves_per_s": cap["solves_per_s"],
        "burst_p50_latency_s": cap["p50_latency_s"], "burst_p99_latency_s": cap["p99_latency_s"],
        "burst_all_converged": cap["all_converged"],
        "burst_speedup_vs_serial_fp64": cap["solves_per_s"] / base["solves_per_s"],
    }


if __name__ == "__main__":
    main()
