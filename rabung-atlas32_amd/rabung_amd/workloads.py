"""Seeded synthetic workloads of the census path (deterministic prime lists).

* top(n): the n largest primes below 2^32 (top-of-range throughput);
* block(start, n): n consecutive primes from the first prime >= start;
* batch_blocks(seed, nblocks): blocks of U{1..4096} consecutive primes from
  starts drawn log-uniform in [2^20, 2^32) (the cross-block batching case);
* serving_stream(seed, qps, seconds): Poisson arrivals; 80 % single-prime
  queries (the next prime >= a uniform 32-bit integer), 20 % interval queries
  of U{16..4096} consecutive primes from a log-uniform start in [2^16, 2^32).

Random numbers come from Python's `random.Random(seed)` (Mersenne Twister),
so every list is reproducible from its seed.
"""
import random

from . import primes_from

TOP = 1 << 32


def top(n):
    out = []
    x = TOP - 1
    from . import is_prime
    while len(out) < n:
        if is_prime(x):
            out.append(x)
        x -= 1
    return sorted(out)


def top_wide(n):
    """The n largest primes below 2^33 (the wider domain's hardest case)."""
    out = []
    x = (1 << 33) - 1
    from . import is_prime
    while len(out) < n:
        if is_prime(x):
            out.append(x)
        x -= 2 if x % 2 else 1
    return sorted(out)


def block(start, n):
    return primes_from(start, n)


def _log_uniform(rng, lo_bits, hi_bits):
    return int(2.0 ** (lo_bits + (hi_bits - lo_bits) * rng.random()))


def _prime_near(x):
    """Next prime >= x, or the largest prime below 2^32 when none is left."""
    got = primes_from(x, 1)
    return got[0] if got else 4294967291


def batch_blocks(seed=2026, nblocks=64):
    rng = random.Random(seed)
    blocks = []
    for _ in range(nblocks):
        n = rng.randint(1, 4096)
        start = _log_uniform(rng, 20, 32)
        got = block(start, n)
        if not got:
            got = [4294967291]
        blocks.append(got)
    return blocks


def serving_stream(seed=4, qps=20.0, seconds=10.0, intervals=False):
    """[(arrival_seconds, [primes...]), ...] sorted by arrival.  With
    `intervals`, an interval query is given as serve.Interval(start, n) (the
    server resolves it on the device) instead of its list of primes; the
    random draws, and so the queries, are the same."""
    from .serve import Interval
    rng = random.Random(seed)
    t = 0.0
    out = []
    while True:
        t += rng.expovariate(qps)
        if t >= seconds:
            break
        if rng.random() < 0.8:
            primes = [_prime_near(rng.randrange(2, TOP))]
        else:
            n = rng.randint(16, 4096)
            start = _log_uniform(rng, 16, 32)
            if intervals:
                primes = Interval(start, n)
            else:
                primes = block(start, n) or [4294967291]
        out.append((t, primes))
    return out
