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
