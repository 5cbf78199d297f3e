"""The request server: admission, continuous batching, and results that do
not depend on the mix."""

import threading
import time

import numpy as np
import pytest

from bsam_amd import engine as E
from bsam_amd import serve


def test_requests_follow_the_configuration():
    reqs = serve.make_requests(400, rate=20.0, seed=3)
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert abs(gaps.mean() - 1 / 20.0) < 0.01          # Poisson arrivals at rate 20/s
    assert all(r.side in (64, 128, 256, 512, 1024) for r in reqs if r.ndims == 2)
    assert all(r.side in (32, 64, 128) for r in reqs if r.ndims == 3)
    assert all(0.01 <= r.eps <= 0.05 for r in reqs)
    assert 0.2 < np.mean([r.ndims == 3 for r in reqs]) < 0.4


def test_request_log_lines(tmp_path):
    """The server's request log: one JSON line per served request, in
    completion order, with its latency and V-cycles per step."""
    import json
    reqs = serve.make_requests(3, 50.0, steps=2)
    results = [serve.Result(rid=r.rid, arrival=r.arrival, admitted=r.arrival + 0.1,
                            finished=r.arrival + 1.0 - 0.1 * n, cycles=[5, 4])
               for n, r in enumerate(reqs)]
    path = str(tmp_path / "served.jsonl")
    serve.write_log(path, reqs, results)
    with open(path) as fh:
        lines = [json.loads(l) for l in fh]
    assert [l["finished_s"] for l in lines] == sorted(r.finished for r in results)
    for l in lines:
        q = next(r for r in reqs if r.rid == l["rid"])
        assert (l["ndims"], l["side"], l["steps"], l["vcycles_per_step"]) == (q.ndims, q.side, 2, [5, 4])
        assert abs(l["latency_s"] - (l["finished_s"] - l["arrival_s"])) < 1e-12


def test_footprint_grows_with_size_and_precision():
    small = serve.Request(0, 2, 64, 0.03, 1)
    big = serve.Request(1, 2, 256, 0.03, 1)
    f64 = serve.footprint_estimate(small, E.PREC_FP64)
    assert serve.footprint_estimate(big, E.PREC_FP64) > 15 * f64
    assert serve.footprint_estimate(small, E.PREC_MIXED) > f64


def test_admission_respects_budget_and_slots():
    """A fake solve (no GPU) that sleeps: the server never exceeds the memory
    budget or the slot count, admits in arrival order, and admits new work
    while earlier requests are still running."""
    reqs = [serve.Request(i, 2, 64, 0.03, i, arrival=0.01 * i, steps=1) for i in range(12)]
    foot = {r.rid: (3 if r.rid % 3 == 0 else 1) * 100 for r in reqs}
    live, lock, overlaps, order = [0], threading.Lock(), [0], []

    def fake_solve(r, prec, dev, keep, tune):
        with lock:
            live[0] += foot[r.rid]
            assert live[0] <= 500
            order.append(r.rid)
            if live[0] > foot[r.rid]:
                overlaps[0] += 1
        time.sleep(0.05)
        with lock:
            live[0] -= foot[r.rid]
        return serve.Result(r.rid, r.arrival, 0.0, 0.0, cycles=[5])

    srv = serve.Server(budget_bytes=500, max_active=3, solve_fn=fake_solve,
                       footprint_fn=lambda r, p: foot[r.rid])
    out = srv.run(reqs, realtime=True)
    assert out["requests"] == 12
    assert order == sorted(order)
    assert srv.peak_used <= 500 and srv.peak_active <= 3
    assert overlaps[0] > 0
    with pytest.raises(ValueError):
        serve.Server(budget_bytes=50, footprint_fn=lambda r, p: 100,
                     solve_fn=fake_solve).run(reqs[:1])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [E.PREC_FP64, E.PREC_MIXED])
def test_results_same_alone_and_in_the_mix(precision):
    reqs = [serve.Request(0, 2, 64, 0.02, 11, steps=3), serve.Request(1, 2, 128, 0.04, 12, steps=3),
            serve.Request(2, 3, 16, 0.05, 13, steps=3), serve.Request(3, 2, 32, 0.03, 14, steps=3),
            serve.Request(4, 3, 32, 0.04, 15, steps=2), serve.Request(5, 2, 64, 0.01, 16, steps=3)]
    alone = {r.rid: serve.solve_request(r, precision, keep_phi=True) for r in reqs}
    srv = serve.Server(budget_bytes=1 << 30, max_active=4, precision=precision, keep_phi=True)
    out = srv.run(reqs, realtime=False)
    assert srv.peak_active > 1
    for res in out["results"]:
        a = alone[res.rid]
        assert res.converged and a.converged
        assert res.cycles == a.cycles
        assert np.array_equal(res.phi, a.phi), res.rid


def test_batches_form_by_shape_class_within_budget():
    """A fake batch solve (no GPU): admitted units hold one shape class, at
    most `batch` requests, and never exceed the memory budget."""
    reqs = [serve.Request(i, 2 if i % 3 else 3, 64 if i % 2 else 32, 0.03, i,
                          arrival=0.0, steps=1) for i in range(14)]
    units, lock = [], threading.Lock()

    def fake_batch(unit, prec, dev, keep, tune):
        with lock:
            units.append([r.rid for r in unit])
        time.sleep(0.02)
        return [serve.Result(r.rid, r.arrival, 0.0, 0.0, cycles=[4]) for r in unit]

    srv = serve.Server(budget_bytes=300, max_active=2, batch=3, batch_fn=fake_batch,
                       footprint_fn=lambda r, p: 100)
    out = srv.run(reqs, realtime=False)
    assert out["requests"] == 14 and srv.peak_used <= 300
    by = {r.rid: r for r in reqs}
    seen = []
    for u in units:
        assert 1 <= len(u) <= 3
        assert len({serve.shape_class(by[i]) for i in u}) == 1
        seen += u
    assert sorted(seen) == list(range(14))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [E.PREC_FP64, E.PREC_MIXED])
@pytest.mark.parametrize("ndims,side", [(2, 64), (3, 16)])
def test_batched_solves_match_solo(precision, ndims, side):
    """Requests solved together in one context (one group each) end with the
    fields, V-cycle counts and per-step mass of each request solved alone,
    bit for bit, although their eps (hence their convergence) differ."""
    reqs = [serve.Request(i, ndims, side, 0.01 + 0.012 * i, 40 + i, steps=3) for i in range(4)]
    batch = serve.solve_batch(reqs, precision, keep_phi=True)
    for r, b in zip(reqs, batch):
        a = serve.solve_request(r, precision, keep_phi=True)
        assert b.rid == r.rid and a.converged and b.converged
        assert b.cycles == a.cycles, (r.rid, a.cycles, b.cycles)
        assert b.mass == a.mass, r.rid
        assert np.array_equal(b.phi, a.phi), r.rid
    assert len({tuple(b.cycles) for b in batch}) > 1   # the groups stopped apart


@pytest.mark.gpu
def test_batched_server_matches_unbatched():
    reqs = [serve.Request(i, 2, 64 if i % 2 else 128, 0.02 + 0.004 * i, 60 + i, steps=2)
            for i in range(8)]
    one = {r.rid: r for r in serve.Server(budget_bytes=1 << 30, max_active=4,
                                          precision=E.PREC_FP64, keep_phi=True)
           .run(reqs, realtime=False)["results"]}
    srv = serve.Server(budget_bytes=1 << 30, max_active=4, precision=E.PREC_FP64,
                       keep_phi=True, batch=4)
    out = srv.run(reqs, realtime=False)
    assert out["requests"] == 8
    for res in out["results"]:
        a = one[res.rid]
        assert res.cycles == a.cycles and np.array_equal(res.phi, a.phi), res.rid
