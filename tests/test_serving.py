"""Continuous-batching server: admission, routing, refusal (CPU, fake engine)
and results identical alone and in the mix (GPU)."""
import random
import threading
import time

import pytest

import oracle
import rabung_amd
from rabung_amd import serve, workloads


class FakeEngine:
    """Deterministic stand-in for a library context (CPU tests only)."""

    def __init__(self, delay=0.01, fail_on=None):
        self.delay = delay
        self.fail_on = fail_on
        self.calls = []
        self.lock = threading.Lock()

    def estimate_bytes(self, primes, method=0):
        return 1000 * len(primes)

    def profile_batch(self, primes, method=0):
        with self.lock:
            self.calls.append(list(primes))
        time.sleep(self.delay)
        if self.fail_on is not None and self.fail_on in primes:
            raise rabung_amd.RabungError(1, "refused")
        return [{"p": p, "mask": rabung_amd.codec.admissible_mask(p), "M": [p % 97] * 8,
                 "L": [1] * 8, "W": [p % 13] * 8, "K": [0] * 8, "A4": p % 5} for p in primes]

    def close(self):
        pass


def test_requests_batched_in_order_and_routed():
    engs = []

    def factory():
        engs.append(FakeEngine(delay=0.05))
        return engs[-1]

    with serve.Server(slots=1, hbm_budget=10**9, max_primes=50, small=0,
                      engine_factory=factory) as srv:
        first = srv.submit([101])
        time.sleep(0.01)  # the slot is now busy with `first`
        reqs = [srv.submit([103 + 2 * i, 107 + 2 * i]) for i in range(40)]
        for r in [first] + reqs:
            r.wait(5)
    calls = engs[0].calls
    assert calls[0] == [101]
    # later requests were admitted together, in arrival order, <= 50 primes
    flat = [p for c in calls[1:] for p in c]
    assert flat == [p for r in reqs for p in r.primes]
    assert all(len(c) <= 50 for c in calls) and len(calls) >= 3
    for r in reqs:
        assert [rec["p"] for rec in r.records] == r.primes
        assert r.done >= r.admitted >= r.arrival


def test_hbm_budget_limits_batches():
    with serve.Server(slots=2, hbm_budget=2 * 5000, max_primes=10**6, small=0,
                      engine_factory=lambda: FakeEngine(delay=0.02)) as srv:
        reqs = [srv.submit([3, 5, 7]) for _ in range(10)]  # 3000 bytes each
        for r in reqs:
            r.wait(5)
        assert all(b[2] <= 3 for b in srv.batches)  # one request (3 primes) per batch


def test_small_requests_first_and_chunked_intervals():
    engs = []

    def factory():
        engs.append(FakeEngine(delay=0.05))
        return engs[-1]

    with serve.Server(slots=1, hbm_budget=10**9, max_primes=8, chunk=4, small=2,
                      engine_factory=factory) as srv:
        first = srv.submit([3])
        time.sleep(0.01)
        big = srv.submit(list(range(1000, 1020)))  # 20 primes -> 5 pieces of 4
        single = srv.submit([7])
        assert [r["p"] for r in big.wait(5)] == list(range(1000, 1020))
        single.wait(5)
        first.wait(5)
    calls = engs[0].calls
    assert calls[0] == [3]
    assert calls[1] == [7]  # the later single gets an express batch ahead of the pieces
    assert all(len(c) <= 8 for c in calls)
    assert single.done < big.done


def test_refusal_is_per_batch_and_visible():
    with serve.Server(slots=1, hbm_budget=10**9, engine_factory=lambda: FakeEngine(
            delay=0.05, fail_on=13)) as srv:
        ok = srv.submit([11])
        time.sleep(0.01)
        bad = [srv.submit([13]), srv.submit([17])]
        assert ok.wait(5)[0]["p"] == 11
        for r in bad:
            with pytest.raises(rabung_amd.RabungError):
                r.wait(5)


def test_percentile():
    assert serve.percentile([1, 2, 3, 4], 0.5) == 2.5
    assert serve.percentile(list(range(101)), 0.99) == 99


@pytest.mark.gpu
def test_mixed_stream_matches_alone():
    stream = workloads.serving_stream(seed=11, qps=200.0, seconds=0.4)
    assert len(stream) > 20
    with serve.Server(slots=2) as srv:
        reqs, summary = serve.run_stream(srv, stream)
    assert summary["max_batch_requests"] > 1  # requests really were mixed
    with rabung_amd.Engine() as eng:
        for r in reqs:
            assert r.records == eng.profile_batch(r.primes), r.primes[:3]
    small = [p for r in reqs for p in r.primes if p < 3_000_000][:20]
    if small:
        alone = {rec["p"]: rec for rec in oracle.profile_many(small)}
        for r in reqs:
            for rec in r.records:
                if rec["p"] in alone:
                    assert rec == alone[rec["p"]]
