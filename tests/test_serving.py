"""Continuous-batching server: admission, routing, refusal (CPU, fake engine)
and results identical alone and in the mix (GPU)."""
import bisect
import random
import threading
import time

import pytest

import oracle
import rabung_amd
from rabung_amd import serve, workloads


def _host_primes(limit=200000):
    sieve = bytearray([1]) * limit
    sieve[0:2] = b"\x00\x00"
    for q in range(2, int(limit ** 0.5) + 1):
        if sieve[q]:
            sieve[q * q::q] = bytearray(len(sieve[q * q::q]))
    return [i for i in range(limit) if sieve[i]]


HOST_PRIMES = _host_primes()


class FakeEngine:
    """Deterministic stand-in for a library context (CPU tests only); its
    prime table holds the primes below 200,000."""

    def __init__(self, delay=0.01, fail_on=None):
        self.delay = delay
        self.fail_on = fail_on
        self.calls = []
        self.selects = []
        self.lock = threading.Lock()

    def estimate_bytes(self, primes, method=0):
        return 1000 * len(primes)

    def prime_count(self):
        return len(HOST_PRIMES)

    def prime_rank(self, x):
        return bisect.bisect_left(HOST_PRIMES, x)

    def primes_by_rank(self, first, n):
        with self.lock:
            self.selects.append((first, n))
        return HOST_PRIMES[first:first + n]

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


def test_interval_resolved_by_rank():
    engs = []

    def factory():
        engs.append(FakeEngine(delay=0.005))
        return engs[-1]

    with serve.Server(slots=2, hbm_budget=10**9, max_primes=32, chunk=16, small=4,
                      engine_factory=factory) as srv:
        big = srv.submit(serve.Interval(1000, 50))
        tiny = srv.submit(serve.Interval(10**6, 5))   # beyond the fake table: empty
        edge = srv.submit_interval(HOST_PRIMES[-3], 10)  # clipped to the last 3 primes
        assert [r["p"] for r in big.wait(5)] == rabung_amd.primes_from(1000, 50)
        assert big.primes == rabung_amd.primes_from(1000, 50)
        assert tiny.wait(5) == [] and tiny.primes == []
        assert [r["p"] for r in edge.wait(5)] == HOST_PRIMES[-3:]
    first = bisect.bisect_left(HOST_PRIMES, 1000)
    selects = sorted(s for e in engs[:2] for s in e.selects)
    assert selects[:4] == [(first, 16), (first + 16, 16), (first + 32, 16), (first + 48, 2)]


def test_bulk_work_not_starved_by_singles():
    """Continuous single-prime arrivals (one every 2 ms, each batch 10 ms)
    would keep the express queue busy forever; bulk pieces still get one
    batch in every `bulk_every`."""
    eng = FakeEngine(delay=0.01)
    with serve.Server(slots=1, hbm_budget=10**9, max_primes=8, chunk=4, small=2, bulk_every=3,
                      bulk_wait=60.0, engine_factory=lambda: eng) as srv:
        stop = threading.Event()
        singles = []

        def arrivals():
            k = 0
            while not stop.is_set():
                singles.append(srv.submit([HOST_PRIMES[k % 1000]]))
                k += 1
                time.sleep(0.002)

        th = threading.Thread(target=arrivals)
        th.start()
        time.sleep(0.05)
        bulk = srv.submit(HOST_PRIMES[5000:5040])  # 10 pieces of 4
        try:
            bulk.wait(2.0)  # with no bound it would wait until the arrivals stop
            still_arriving = not stop.is_set()
        finally:
            stop.set()
            th.join()
        for r in singles:
            r.wait(5)
    assert still_arriving
    assert [r["p"] for r in bulk.records] == HOST_PRIMES[5000:5040]
    assert bulk.latency < 2.0
    # an express batch (single primes) ran between any two bulk batches
    kinds = ["bulk" if any(p >= HOST_PRIMES[5000] for p in c) else "express" for c in eng.calls]
    first_bulk = kinds.index("bulk")
    assert "express" in kinds[first_bulk:]


def test_bulk_admitted_when_overdue():
    eng = FakeEngine(delay=0.01)
    with serve.Server(slots=1, hbm_budget=10**9, max_primes=8, chunk=4, small=2,
                      bulk_every=10**6, bulk_wait=0.1, engine_factory=lambda: eng) as srv:
        stop = threading.Event()

        def arrivals():
            while not stop.is_set():
                srv.submit([3])
                time.sleep(0.002)

        th = threading.Thread(target=arrivals)
        th.start()
        time.sleep(0.03)
        bulk = srv.submit(HOST_PRIMES[5000:5012])
        try:
            bulk.wait(2.0)
        finally:
            stop.set()
            th.join()
    assert bulk.latency < 2.0


def test_request_ids_unique_across_threads():
    with serve.Server(slots=2, hbm_budget=10**9, engine_factory=lambda: FakeEngine(
            delay=0.0)) as srv:
        reqs = []
        lock = threading.Lock()

        def client(k):
            mine = [srv.submit([HOST_PRIMES[k * 50 + i]]) for i in range(50)]
            with lock:
                reqs.extend(mine)

        ths = [threading.Thread(target=client, args=(k,)) for k in range(8)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        for r in reqs:
            r.wait(5)
    assert sorted(r.rid for r in reqs) == list(range(400))


def test_percentile():
    assert serve.percentile([1, 2, 3, 4], 0.5) == 2.5
    assert serve.percentile(list(range(101)), 0.99) == 99


@pytest.mark.gpu
def test_mixed_stream_matches_alone():
    stream = workloads.serving_stream(seed=11, qps=200.0, seconds=0.4, intervals=True)
    assert len(stream) > 20 and any(isinstance(w, serve.Interval) for _, w in stream)
    with serve.Server(slots=2) as srv:
        reqs, summary = serve.run_stream(srv, stream)
    assert summary["max_batch_requests"] > 1  # requests really were mixed
    with rabung_amd.Engine() as eng:
        for (_, what), r in zip(stream, reqs):
            if isinstance(what, serve.Interval):  # device ranks = host enumeration
                assert r.primes == rabung_amd.primes_from(what.start, what.count)
            assert r.records == eng.profile_batch(r.primes), r.primes[:3]
    small = [p for r in reqs for p in r.primes if p < 3_000_000][:20]
    if small:
        alone = {rec["p"]: rec for rec in oracle.profile_many(small)}
        for r in reqs:
            for rec in r.records:
                if rec["p"] in alone:
                    assert rec == alone[rec["p"]]
