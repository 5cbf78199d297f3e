"""Continuous-batching server for Cahn-Hilliard time-step requests.

A request is a whole solve sequence: one hierarchy (2D or 3D), its own eps,
a seeded noise initial condition and `steps` implicit time steps, each solved
to the residual tolerance. The server

* admits waiting requests (FIFO) as soon as both an execution slot and
  enough device memory are free: admission is by the hierarchy's HBM
  footprint as the engine reports it (bsam_footprint_bytes), against a
  budget;
* keeps every admitted hierarchy resident on the device until its last step;
* runs admitted requests concurrently, each on its own engine context and
  HIP stream, so their kernels overlap on the GPU while new requests keep
  arriving (continuous batching, not a queue);
* reports per-request latency (completion - arrival) p50/p99, throughput in
  requests/s and solves (time steps)/s.

Results do not depend on what else runs: contexts share no device state and
every kernel is deterministic, so a request's final fields are bit-identical
alone and in the mix (tests/test_serve.py).
"""

from __future__ import annotations

import argparse
import json
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import engine as E
from .layout import Layout
from .workloads import noise_q


@dataclass
class Request:
    rid: int
    ndims: int
    side: int
    eps: float
    seed: int
    arrival: float = 0.0          # seconds after the server starts
    steps: int = 20
    dt: float = 1e-3


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    cycles: List[int] = field(default_factory=list)
    converged: bool = True
    footprint: int = 0
    phi: Optional[np.ndarray] = None   # final root q (kept when asked)
    mass: List[float] = field(default_factory=list)   # per step (kept when asked)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def make_requests(n: int = 256, rate: float = 50.0, seed: int = 5, steps: int = 20,
                  sides2d=(64, 128, 256, 512, 1024), sides3d=(32, 64, 128),
                  frac3d: float = 0.3) -> List[Request]:
    """n requests with Poisson arrivals at `rate` per second; 2D with side
    uniform in sides2d or (with probability frac3d) 3D with side in sides3d;
    eps ~ U(0.01, 0.05)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate)
        if rng.uniform() < frac3d:
            nd, side = 3, int(rng.choice(sides3d))
        else:
            nd, side = 2, int(rng.choice(sides2d))
        out.append(Request(rid, nd, side, float(rng.uniform(0.01, 0.05)),
                           int(rng.integers(1, 2**31)), t, steps))
    return out


def request_layout(r: Request) -> Layout:
    return Layout(r.ndims, (r.side,) * r.ndims, 1.0 / r.side)


def solve_request(r: Request, precision: int, device: int = 0, keep_phi: bool = False,
                  tune: Optional[Dict[str, float]] = None, on_commit=None) -> Result:
    """Runs one request start to finish on its own engine context. With
    keep_phi the result also keeps the root patch's final q and the composite
    mass after every step (TakeTimeSteps' statistic, one stream
    synchronisation per step)."""
    lay = request_layout(r)
    eng = E.Engine(lay, dt=r.dt, eps=r.eps, well="plusminus", maxvcycles=60,
                   qerrortol=1e-8, precision=precision, device=device, tune=tune)
    res = Result(r.rid, r.arrival, 0.0, 0.0, footprint=eng.footprint_bytes())
    if on_commit:
        on_commit(res)
    q0 = noise_q(lay, r.seed, 0.0, 0.05)
    for n, p in enumerate(eng.patches):
        eng.upload(n, "q", q0[p.id])
    for _ in range(r.steps):
        hist, conv = eng.time_step()
        res.cycles.append(len(hist))
        res.converged = res.converged and conv
        if keep_phi:
            res.mass.append(eng.after_step_stats()[0])
    if keep_phi:
        res.phi = eng.download(lay.index_map()[0], "q")
    eng.close()
    return res


def footprint_estimate(r: Request, precision: int) -> int:
    """Device bytes of a request's hierarchy (host-only commit, no GPU)."""
    eng = E.Engine(request_layout(r), dt=r.dt, eps=r.eps, well="plusminus",
                   precision=precision, device=-1)
    b = eng.footprint_bytes()
    eng.close()
    return b


def solve_batch(reqs: List[Request], precision: int, device: int = 0, keep_phi: bool = False,
                tune: Optional[Dict[str, float]] = None) -> List[Result]:
    """Runs requests of one shape class (same ndims, side, steps and dt)
    start to finish in one engine context, one group each
    (engine.BatchEngine): every launch of a time step spans all of them, each
    request keeps its own eps and residual test and is left untouched once
    converged, so its fields end as they would alone (tests/test_serve.py)."""
    r0 = reqs[0]
    for r in reqs:
        if (r.ndims, r.side, r.steps, r.dt) != (r0.ndims, r0.side, r0.steps, r0.dt):
            raise ValueError("a batch holds requests of one shape class")
    lays = [request_layout(r) for r in reqs]
    eng = E.BatchEngine(lays, [r.eps for r in reqs], dt=r0.dt, well="plusminus",
                        maxvcycles=60, qerrortol=1e-8, precision=precision, device=device,
                        tune=tune)
    foot = eng.footprint_bytes()
    out = [Result(r.rid, r.arrival, 0.0, 0.0, footprint=foot // len(reqs)) for r in reqs]
    for g, (r, lay) in enumerate(zip(reqs, lays)):
        q0 = noise_q(lay, r.seed, 0.0, 0.05)
        for n, p in enumerate(lay.ordered()):
            eng.upload(n, "q", q0[p.id], group=g)
    for _ in range(r0.steps):
        for g, (hist, conv) in enumerate(eng.time_steps()):
            out[g].cycles.append(len(hist))
            out[g].converged = out[g].converged and conv
            if keep_phi:
                out[g].mass.append(eng.after_step_stats(g)[0])
    if keep_phi:
        for g, lay in enumerate(lays):
            out[g].phi = eng.download(lay.index_map()[0], "q", group=g)
    eng.close()
    return out


def shape_class(r: Request):
    return (r.ndims, r.side, r.steps, r.dt)


class Server:
    """Continuous-batching server. With batch = 1 every admitted request runs
    on its own context; with batch > 1 the requests waiting when a slot frees
    are admitted in batches of up to `batch` requests of the head request's
    shape class (solve_batch), so small requests share their launches."""

    def __init__(self, budget_bytes: int = 64 << 30, max_active: int = 8,
                 precision: int = E.PREC_MIXED, device: int = 0,
                 tune: Optional[Dict[str, float]] = None, keep_phi: bool = False,
                 solve_fn=None, footprint_fn=None, batch: int = 1, batch_fn=None):
        self.solve_fn = solve_fn or solve_request
        self.batch_fn = batch_fn or solve_batch
        self.footprint_fn = footprint_fn or footprint_estimate
        self.budget = budget_bytes
        self.max_active = max_active
        self.precision = precision
        self.device = device
        self.tune = tune
        self.keep_phi = keep_phi
        self.batch = max(1, int(batch))

    def run(self, requests: List[Request], realtime: bool = True) -> Dict:
        """Serve `requests`; with realtime, each is released at its arrival
        time, otherwise all are waiting at t = 0."""
        pending = sorted(requests, key=lambda r: r.arrival)
        foot = {r.rid: self.footprint_fn(r, self.precision) for r in pending}
        for r in pending:
            if foot[r.rid] > self.budget:
                raise ValueError(f"request {r.rid} needs {foot[r.rid]} B > budget")
        cv = threading.Condition()
        used = [0]
        active = [0]
        peak = [0, 0]
        results: List[Result] = []
        errors: List[BaseException] = []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0

        def worker(unit: List[Request], admitted: float):
            try:
                if len(unit) == 1 and self.batch == 1:
                    res = [self.solve_fn(unit[0], self.precision, self.device, self.keep_phi,
                                         self.tune)]
                else:
                    res = self.batch_fn(unit, self.precision, self.device, self.keep_phi,
                                        self.tune)
                fin = now()
                for r, x in zip(unit, res):
                    x.arrival = r.arrival if realtime else 0.0
                    x.admitted = admitted
                    x.finished = fin
                with cv:
                    results.extend(res)
            except BaseException as e:  # reported after the run
                with cv:
                    errors.append(e)
            finally:
                with cv:
                    used[0] -= sum(foot[r.rid] for r in unit)
                    active[0] -= 1
                    cv.notify_all()

        def next_unit(queue: List[Request]) -> List[Request]:
            """The queue head and the waiting requests of its shape class,
            up to `batch`, as many as the free memory takes (FIFO)."""
            head = queue[0]
            unit = [head] + [r for r in queue[1:] if shape_class(r) == shape_class(head)]
            unit = unit[:self.batch]
            while unit and used[0] + sum(foot[r.rid] for r in unit) > self.budget:
                unit.pop()
            return unit

        threads = []
        queue: List[Request] = []
        idx = 0
        with cv:
            while idx < len(pending) or queue or active[0]:
                t = now()
                while idx < len(pending) and (not realtime or pending[idx].arrival <= t):
                    queue.append(pending[idx])
                    idx += 1
                while queue and active[0] < self.max_active:
                    unit = next_unit(queue)
                    if not unit:
                        break
                    for r in unit:
                        queue.remove(r)
                    used[0] += sum(foot[r.rid] for r in unit)
                    active[0] += 1
                    peak[0] = max(peak[0], used[0])
                    peak[1] = max(peak[1], active[0])
                    th = threading.Thread(target=worker, args=(unit, now()), daemon=True)
                    th.start()
                    threads.append(th)
                if idx >= len(pending) and not queue and not active[0]:
                    break
                wait = 0.05
                if realtime and idx < len(pending):
                    wait = max(0.0, min(wait, pending[idx].arrival - now()))
                cv.wait(timeout=wait)
        for th in threads:
            th.join()
        if errors:
            raise errors[0]
        makespan = now()
        lat = sorted(r.latency for r in results)
        steps = sum(len(r.cycles) for r in results)
        self.peak_used = peak[0]
        self.peak_active = peak[1]
        return {
            "requests": len(results),
            "makespan_s": makespan,
            "requests_per_s": len(results) / makespan,
            "solves_per_s": steps / makespan,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "all_converged": all(r.converged for r in results),
            "vcycles": int(sum(sum(r.cycles) for r in results)),
            "results": results,
        }


def serial(requests: List[Request], precision: int = E.PREC_FP64, device: int = 0,
           keep_phi: bool = False) -> Dict:
    """Baseline: the same requests one after another, no overlap."""
    t0 = time.perf_counter()
    results = []
    for r in requests:
        s = time.perf_counter() - t0
        res = solve_request(r, precision, device, keep_phi)
        res.arrival, res.admitted, res.finished = 0.0, s, time.perf_counter() - t0
        results.append(res)
    makespan = time.perf_counter() - t0
    steps = sum(len(r.cycles) for r in results)
    return {"requests": len(results), "makespan_s": makespan,
            "requests_per_s": len(results) / makespan, "solves_per_s": steps / makespan,
            "all_converged": all(r.converged for r in results),
            "vcycles": int(sum(sum(r.cycles) for r in results)), "results": results}


def write_log(path: str, reqs: List[Request], results: List[Result]):
    """One JSON line per served request, in completion order."""
    by_id = {r.rid: r for r in reqs}
    with open(path, "w") as fh:
        for res in sorted(results, key=lambda x: x.finished):
            q = by_id[res.rid]
            fh.write(json.dumps({
                "rid": res.rid, "ndims": q.ndims, "side": q.side, "eps": q.eps,
                "steps": len(res.cycles), "arrival_s": res.arrival, "admitted_s": res.admitted,
                "finished_s": res.finished, "latency_s": res.latency,
                "vcycles_per_step": res.cycles, "converged": res.converged,
                "footprint_bytes": res.footprint}) + "\n")


def main():
    ap = argparse.ArgumentParser(description="serve a stream of CH solve requests")
    ap.add_argument("--requests", type=int, default=256)
    ap.add_argument("--rate", type=float, default=50.0)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--max-active", type=int, default=8)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="mixed")
    ap.add_argument("--baseline", action="store_true", help="also run the FP64 serial baseline")
    ap.add_argument("--burst", action="store_true",
                    help="all requests waiting at t = 0 (capacity rather than arrival-bound)")
    ap.add_argument("--batch", type=int, default=1,
                    help="requests of one shape class admitted together into one context")
    ap.add_argument("--log", default=None,
                    help="write one JSON line per served request (arrival, admission, "
                         "completion, latency, V-cycles per step, converged)")
    args = ap.parse_args()
    reqs = make_requests(args.requests, args.rate, steps=args.steps)
    prec = E.PREC_MIXED if args.precision == "mixed" else E.PREC_FP64
    srv = Server(int(args.budget_gb * (1 << 30)), args.max_active, prec, batch=args.batch)
    out = srv.run(reqs, realtime=not args.burst)
    out["peak_active"] = srv.peak_active
    out["peak_bytes"] = srv.peak_used
    if args.log:
        write_log(args.log, reqs, out["results"])
    out.pop("results")
    line = {"served": out}
    if args.baseline:
        b = serial(reqs, E.PREC_FP64)
        b.pop("results")
        line["serial_fp64"] = b
    print(json.dumps(line))


if __name__ == "__main__":
    main()
