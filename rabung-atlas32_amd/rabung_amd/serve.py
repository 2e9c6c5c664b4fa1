"""Continuous-batching request loop over the census library.

Requests (a single prime or an interval of consecutive primes) arrive at any
time.  Each of `slots` worker threads owns one library context: a HIP stream
and its resident device workspaces (root tables, label tables, colour planes,
inverses and run keys stay allocated between batches).  A request is cut
into pieces of at most `chunk` primes.  Whenever a slot is free it admits
either every waiting small request (at most `small` primes each: an express
batch) or, when none waits, bulk pieces in arrival order, while the batch stays within
`max_primes` primes and within the slot's share of the HBM budget (estimated
with the library's own workspace model, `rbg_estimate_bytes`); pieces that
arrive while batches run join the next free slot.  Short batches and the
small-first rule bound how long a single-prime query waits behind an
interval query.
Batches of different slots run concurrently on their own streams, and the
host-side work of one (root tables, elliptic A4 counts) overlaps the device
work of another.  Results are returned per request, identical to profiling
the request alone (every record depends on its prime only).

The reference has no serving path; its workers read 128-prime batches from
stdin (src/census/worker32.cu:10-45).
"""
import collections
import threading
import time

from . import METHOD_BATCHED, Engine


class Request:
    __slots__ = ("rid", "primes", "arrival", "admitted", "done", "records", "error", "pending",
                 "_event")

    def __init__(self, rid, primes, arrival):
        self.rid = rid
        self.primes = list(primes)
        self.arrival = arrival
        self.admitted = None
        self.done = None
        self.records = [None] * len(self.primes)
        self.error = None
        self.pending = 0
        self._event = threading.Event()

    def wait(self, timeout=None):
        self._event.wait(timeout)
        if self.error:
            raise self.error
        return self.records

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


class _Piece:
    __slots__ = ("req", "offset", "primes", "bytes")

    def __init__(self, req, offset, primes, nbytes):
        self.req = req
        self.offset = offset
        self.primes = primes
        self.bytes = nbytes


class Server:
    def __init__(self, slots=2, hbm_budget=64 << 30, max_primes=256, chunk=64, small=16,
                 method=METHOD_BATCHED, engine_factory=None, clock=time.perf_counter):
        self.clock = clock
        self.method = method
        self.max_primes = max_primes
        self.chunk = max(1, chunk)
        self.small = small
        self.slot_budget = hbm_budget // max(1, slots)
        factory = engine_factory or (lambda: Engine(0, hbm_budget=self.slot_budget))
        self.engines = [factory() for _ in range(slots)]
        self.small_queue = collections.deque()
        self.queue = collections.deque()
        self.cond = threading.Condition()
        self.stopping = False
        self.batches = []  # (slot, n_requests, n_primes, t_start, t_end)
        self._next_id = 0
        self.threads = [threading.Thread(target=self._worker, args=(i,), daemon=True)
                        for i in range(slots)]
        for t in self.threads:
            t.start()

    def submit(self, primes, arrival=None):
        """Queue a request; returns a Request whose wait() gives its records."""
        req = Request(self._next_id, primes, self.clock() if arrival is None else arrival)
        self._next_id += 1
        pieces = []
        for off in range(0, len(req.primes), self.chunk):
            part = req.primes[off:off + self.chunk]
            pieces.append(_Piece(req, off, part, self.engines[0].estimate_bytes(part, self.method)))
        req.pending = len(pieces)
        if not pieces:  # nothing to compute
            req.done = req.arrival
            req._event.set()
            return req
        with self.cond:
            (self.small_queue if len(req.primes) <= self.small else self.queue).extend(pieces)
            self.cond.notify()
        return req

    def _take(self):
        """Admit, in arrival order, either the waiting small requests (an
        express batch) or bulk pieces, while the batch fits (caller holds
        cond)."""
        batch, primes, nbytes = [], 0, 0
        q = self.small_queue if self.small_queue else self.queue
        while q:
            piece = q[0]
            if batch and (primes + len(piece.primes) > self.max_primes
                          or nbytes + piece.bytes > self.slot_budget):
                break
            batch.append(q.popleft())
            primes += len(piece.primes)
            nbytes += piece.bytes
        return batch

    def _worker(self, slot):
        eng = self.engines[slot]
        while True:
            with self.cond:
                while not self.queue and not self.small_queue and not self.stopping:
                    self.cond.wait()
                if self.stopping and not self.queue and not self.small_queue:
                    return
                batch = self._take()
            t0 = self.clock()
            for piece in batch:
                if piece.req.admitted is None:
                    piece.req.admitted = t0
            flat = [p for piece in batch for p in piece.primes]
            try:
                records = eng.profile_batch(flat, method=self.method)
                err = None
            except Exception as exc:  # refuse the whole batch, visibly
                records, err = None, exc
            t1 = self.clock()
            self.batches.append((slot, len({id(pc.req) for pc in batch}), len(flat), t0, t1))
            pos = 0
            finished = []
            with self.cond:
                for piece in batch:
                    req = piece.req
                    if err is None:
                        req.records[piece.offset:piece.offset + len(piece.primes)] = \
                            records[pos:pos + len(piece.primes)]
                    else:
                        req.error = err
                    pos += len(piece.primes)
                    req.pending -= 1
                    if req.pending == 0:
                        req.done = t1
                        finished.append(req)
            for req in finished:
                req._event.set()

    def close(self):
        with self.cond:
            self.stopping = True
            self.cond.notify_all()
        for t in self.threads:
            t.join()
        for eng in self.engines:
            if hasattr(eng, "close"):
                eng.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def percentile(values, q):
    vals = sorted(values)
    if not vals:
        return None
    k = (len(vals) - 1) * q
    lo = int(k)
    hi = min(lo + 1, len(vals) - 1)
    return vals[lo] + (vals[hi] - vals[lo]) * (k - lo)


def run_stream(server, stream, clock=time.perf_counter, sleep=time.sleep):
    """Replay [(arrival_s, primes)] against the server in real time.

    Returns (requests, summary) with p50/p99 latency in milliseconds and the
    profile throughput over the whole run."""
    t0 = clock()
    reqs = []
    for arrival, primes in stream:
        now = clock() - t0
        if arrival > now:
            sleep(arrival - now)
        reqs.append(server.submit(primes, arrival=t0 + arrival))
    for r in reqs:
        r.wait()
    t_end = max(r.done for r in reqs) if reqs else t0
    lat = [r.latency * 1000.0 for r in reqs]
    single = [r.latency * 1000.0 for r in reqs if len(r.primes) == 1]
    profiles = sum(bin(rec["mask"]).count("1") for r in reqs for rec in r.records)
    summary = {
        "requests": len(reqs),
        "primes": sum(len(r.primes) for r in reqs),
        "profiles": profiles,
        "p50_ms": percentile(lat, 0.50),
        "p99_ms": percentile(lat, 0.99),
        "p50_single_ms": percentile(single, 0.50),
        "p99_single_ms": percentile(single, 0.99),
        "wall_s": t_end - t0,
        "profiles_per_s": profiles / (t_end - t0) if t_end > t0 else None,
        "batches": len(server.batches),
        "max_batch_requests": max((b[1] for b in server.batches), default=0),
    }
    return reqs, summary
