"""GPU: the device-resident prime table (segmented sieve in HBM) and the
profiles of primes selected on the device by rank.

Sources of truth: the known values of the prime-counting function
pi(10^k) and pi(2^32) = 203,280,221; the host's deterministic Miller-Rabin
enumeration (rbg_primes_from); the reference's sample blocks 0 and 49900,
whose 4096 and 2313 primes are exactly the primes of ranks 0..4095 and
203,277,908..203,280,220.
"""
import random

import pytest

import golden_data
import rabung_amd
from rabung_amd import codec

pytestmark = pytest.mark.gpu

PI_2_32 = 203280221


@pytest.fixture(scope="module")
def engine():
    eng = rabung_amd.Engine()
    yield eng
    eng.close()


def test_prime_count(engine):
    assert engine.prime_count() == PI_2_32


@pytest.mark.parametrize("x,pi", [(0, 0), (2, 0), (3, 1), (4, 2), (100, 25), (10**6, 78498),
                                  (10**7, 664579), (10**8, 5761455), (10**9, 50847534),
                                  (4294967291, PI_2_32 - 1), (4294967292, PI_2_32),
                                  (1 << 32, PI_2_32)])
def test_rank_known(engine, x, pi):
    assert engine.prime_rank(x) == pi


def test_select_matches_host(engine):
    assert engine.primes_by_rank(0, 3000) == rabung_amd.primes_from(2, 3000)
    r = engine.prime_rank(10**9)
    assert engine.primes_by_rank(r, 2000) == rabung_amd.primes_from(10**9, 2000)
    top = engine.primes_by_rank(PI_2_32 - 500, 600)
    assert top == rabung_amd.primes_from(top[0], 600) and len(top) == 500
    assert top[-1] == 4294967291
    assert engine.primes_by_rank(PI_2_32, 10) == []


def test_rank_select_consistency(engine):
    rng = random.Random(7)
    for _ in range(64):
        x = rng.randrange(2, 4294967291)
        (p,) = engine.primes_by_rank(engine.prime_rank(x), 1)
        assert p == rabung_amd.primes_from(x, 1)[0]
        assert engine.prime_rank(p + 1) == engine.prime_rank(x) + 1


@pytest.mark.parametrize("block,first", [(0, 0), (49900, PI_2_32 - 2313)])
def test_profile_ranked_sample_blocks(engine, block, first):
    rows, a4, entry = golden_data.sample_block(block)
    assert engine.prime_rank(entry["lo"]) == first
    recs = engine.profile_ranked(first, 4096)
    assert len(recs) == len(rows)
    got = codec.records_to_rows(recs)
    bad = [i for i in range(len(rows)) if got[i] != rows[i]]
    assert not bad, ("mismatch", rows[bad[0]]["p"], got[bad[0]], rows[bad[0]])
    assert [r["A4"] for r in recs] == a4


def test_profile_ranked_matches_block(engine):
    first = engine.prime_rank(3 * 10**9)
    ranked = engine.profile_ranked(first, 777, method=rabung_amd.METHOD_PLAIN)
    primes = engine.primes_by_rank(first, 777)
    assert [r["p"] for r in ranked] == primes
    assert ranked == engine.profile_block(primes)
    assert engine.profile_ranked(PI_2_32, 16) == []
