"""Continuous-batching request loop over the census library.

Requests arrive at any time: a list of primes, or an interval query (`count`
consecutive primes from the first prime >= `start`).  An interval is
resolved on the device when it is submitted: the library keeps every prime
below 2^32 as a bitmap in HBM (`rbg_prime_table`), the request becomes the
rank range [rank(start), rank(start) + count), and its pieces are rank
sub-ranges whose primes the device selects when a slot admits them
(`rbg_primes_by_rank`).  No host primality scan runs for an interval, and
the interval is never re-enumerated: each piece carries its own ranks.

Each of `slots` worker threads owns one library context: a HIP stream and
its resident device workspaces (root tables, label tables, colour planes,
inverses and run keys stay allocated between batches).  A request is cut
into pieces of at most `chunk` primes.  Whenever a slot is free it admits
every waiting small request (at most `small` primes: an express batch)
and, when no small request waits, or when bulk work has been passed over
`bulk_every` batches in a row or has waited `bulk_wait` seconds, bulk
pieces in arrival order, while the batch stays within `max_primes` primes
and the slot's share of the HBM budget (the library's own workspace model,
`rbg_estimate_bytes`; interval pieces are priced at the worst prime below
2^32).  So single-prime queries overtake bulk work, and bulk work still
advances by at least one piece every `bulk_every` batches.
Batches of different slots run concurrently on their own streams, and the
host-side work of one (root tables, elliptic A4 counts) overlaps the device
work of another.  Results are returned per request, identical to profiling
the request alone (every record depends on its prime only).

The reference has no serving path; its workers read 128-prime batches from
stdin (src/census/worker32.cu:10-45), and its engine keeps one persistent
device workspace per slot (src/census/engine32/src/frozen_entry.cu:66-77).
"""
import collections
import threading
import time

from . import METHOD_BATCHED, Engine

# The largest prime below 2^32 with p = 1 (mod 2520): every order 2..9 is
# admissible and its cover is the largest, so its workspace bounds any
# other prime's.
WORST_PRIME = 4294964521


class Interval:
    """`count` consecutive primes from the first prime >= start (< 2^32)."""
    __slots__ = ("start", "count")

    def __init__(self, start, count):
        self.start = int(start)
        self.count = int(count)

    def __repr__(self):
        return "Interval(%d, %d)" % (self.start, self.count)


class Request:
    __slots__ = ("rid", "primes", "first_rank", "arrival", "admitted", "done", "records",
                 "error", "pending", "_event")

    def __init__(self, rid, primes, arrival, first_rank=None, count=None):
        self.rid = rid
        self.first_rank = first_rank
        # an interval's primes are known once its pieces have been selected
        self.primes = list(primes) if first_rank is None else [None] * count
        self.arrival = arrival
        self.admitted = None
        self.done = None
        self.records = [None] * len(self.primes)
        self.error = None
        self.pending = 0
        self._event = threading.Event()

    def wait(self, timeout=None):
        if not self._event.wait(timeout):
            raise TimeoutError("request %d still running" % self.rid)
        if self.error:
            raise self.error
        return self.records

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


class _Piece:
    __slots__ = ("req", "offset", "n", "primes", "rank", "bytes", "queued")

    def __init__(self, req, offset, n, primes, rank, nbytes, queued):
        self.req = req
        self.offset = offset
        self.n = n
        self.primes = primes  # explicit primes, or None for a rank range
        self.rank = rank
        self.bytes = nbytes
        self.queued = queued


class Server:
    def __init__(self, slots=2, hbm_budget=64 << 30, max_primes=256, chunk=64, small=16,
                 bulk_every=4, bulk_wait=0.5, method=METHOD_BATCHED, engine_factory=None,
                 clock=time.perf_counter):
        self.clock = clock
        self.method = method
        self.max_primes = max_primes
        self.chunk = max(1, chunk)
        self.small = small
        self.bulk_every = max(1, bulk_every)
        self.bulk_wait = bulk_wait
        self.slot_budget = hbm_budget // max(1, slots)
        factory = engine_factory or (lambda: Engine(0, hbm_budget=self.slot_budget))
        self.engines = [factory() for _ in range(slots)]
        # a context of its own for submission-time queries (ranks, estimates)
        self.resolver = factory()
        self._resolve_lock = threading.Lock()
        with self._resolve_lock:
            self.prime_bytes = self.resolver.estimate_bytes([WORST_PRIME], method)
        self.small_queue = collections.deque()
        self.queue = collections.deque()
        self.cond = threading.Condition()
        self.stopping = False
        self.batches = []  # (slot, n_requests, n_primes, t_start, t_end)
        self._passed_over = 0  # batches in a row that left bulk work waiting
        self._next_id = 0
        self.threads = [threading.Thread(target=self._worker, args=(i,), daemon=True)
                        for i in range(slots)]
        for t in self.threads:
            t.start()

    def _new_id(self):
        with self.cond:
            rid = self._next_id
            self._next_id += 1
        return rid

    def submit(self, primes, arrival=None):
        """Queue a request (a list of primes or an Interval); returns a
        Request whose wait() gives its records."""
        arrival = self.clock() if arrival is None else arrival
        if isinstance(primes, Interval):
            return self.submit_interval(primes.start, primes.count, arrival)
        req = Request(self._new_id(), primes, arrival)
        pieces = []
        for off in range(0, len(req.primes), self.chunk):
            part = req.primes[off:off + self.chunk]
            with self._resolve_lock:
                nbytes = self.resolver.estimate_bytes(part, self.method)
            pieces.append(_Piece(req, off, len(part), part, None, nbytes, arrival))
        return self._enqueue(req, pieces, small=len(req.primes) <= self.small)

    def submit_interval(self, start, count, arrival=None):
        """Queue an interval query, resolved by rank on the device."""
        arrival = self.clock() if arrival is None else arrival
        with self._resolve_lock:
            first = self.resolver.prime_rank(min(int(start), 1 << 32))
            total = self.resolver.prime_count()
        count = max(0, min(int(count), total - first))
        req = Request(self._new_id(), (), arrival, first_rank=first, count=count)
        pieces = [_Piece(req, off, min(self.chunk, count - off), None, first + off,
                         min(self.chunk, count - off) * self.prime_bytes, arrival)
                  for off in range(0, count, self.chunk)]
        return self._enqueue(req, pieces, small=count <= self.small)

    def _enqueue(self, req, pieces, small):
        req.pending = len(pieces)
        if not pieces:  # nothing to compute
            req.done = req.arrival
            req._event.set()
            return req
        with self.cond:
            (self.small_queue if small else self.queue).extend(pieces)
            self.cond.notify()
        return req

    def _fill(self, batch, q, primes, nbytes):
        """Move pieces from q while the batch fits; the first piece of q
        always goes (a piece alone may exceed max_primes; the library splits
        a batch over its HBM budget by itself)."""
        added = 0
        while q:
            piece = q[0]
            if added and (primes + piece.n > self.max_primes
                          or nbytes + piece.bytes > self.slot_budget):
                break
            added += 1
            batch.append(q.popleft())
            primes += piece.n
            nbytes += piece.bytes
        return primes, nbytes

    def _take(self):
        """Admit the waiting small requests, then bulk pieces when no small
        request waits or bulk work is overdue (caller holds cond)."""
        batch = []
        primes, nbytes = self._fill(batch, self.small_queue, 0, 0)
        if self.queue:
            overdue = (self._passed_over >= self.bulk_every
                       or self.clock() - self.queue[0].queued >= self.bulk_wait)
            if not batch or overdue:
                before = len(batch)
                self._fill(batch, self.queue, primes, nbytes)
                self._passed_over = 0 if len(batch) > before else self._passed_over + 1
            else:
                self._passed_over += 1
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
            try:
                flat = []
                for piece in batch:
                    if piece.primes is None:  # a rank range: select on the device
                        piece.primes = eng.primes_by_rank(piece.rank, piece.n)
                        if len(piece.primes) != piece.n:
                            raise RuntimeError("rank range %d+%d beyond the table"
                                               % (piece.rank, piece.n))
                    flat.extend(piece.primes)
                records = eng.profile_batch(flat, method=self.method)
                err = None
            except Exception as exc:  # refuse the whole batch, visibly
                records, err = None, exc
            t1 = self.clock()
            self.batches.append((slot, len({id(pc.req) for pc in batch}),
                                 sum(pc.n for pc in batch), t0, t1))
            pos = 0
            finished = []
            with self.cond:
                for piece in batch:
                    req = piece.req
                    if err is None:
                        end = piece.offset + piece.n
                        req.records[piece.offset:end] = records[pos:pos + piece.n]
                        req.primes[piece.offset:end] = piece.primes
                    else:
                        req.error = err
                    pos += piece.n
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
        for eng in self.engines + [self.resolver]:
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
    """Replay [(arrival_s, primes or Interval)] against the server in real
    time.

    Returns (requests, summary) with p50/p99 latency in milliseconds and the
    profile throughput over the whole run."""
    t0 = clock()
    reqs = []
    for arrival, what in stream:
        now = clock() - t0
        if arrival > now:
            sleep(arrival - now)
        reqs.append(server.submit(what, arrival=t0 + arrival))
    for r in reqs:
        r.wait()
    t_end = max(r.done for r in reqs) if reqs else t0
    lat = [r.latency * 1000.0 for r in reqs]
    single = [r.latency * 1000.0 for r in reqs if len(r.primes) == 1]
    bulk = [r.latency * 1000.0 for r in reqs if len(r.primes) > server.small]
    profiles = sum(bin(rec["mask"]).count("1") for r in reqs for rec in r.records)
    offered = stream[-1][0] if stream else 0.0
    summary = {
        "requests": len(reqs),
        "primes": sum(len(r.primes) for r in reqs),
        "profiles": profiles,
        "p50_ms": percentile(lat, 0.50),
        "p99_ms": percentile(lat, 0.99),
        "p50_single_ms": percentile(single, 0.50),
        "p99_single_ms": percentile(single, 0.99),
        "p50_bulk_ms": percentile(bulk, 0.50),
        "p99_bulk_ms": percentile(bulk, 0.99),
        "wall_s": t_end - t0,
        "drain_s": max(0.0, t_end - t0 - offered),
        "profiles_per_s": profiles / (t_end - t0) if t_end > t0 else None,
        "batches": len(server.batches),
        "max_batch_requests": max((b[1] for b in server.batches), default=0),
    }
    return reqs, summary
