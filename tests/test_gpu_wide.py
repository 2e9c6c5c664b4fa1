"""GPU parity of the wider domain (primes between 2^32 and 2^33).

Sources of truth: the reference's 33-bit throughput windows (1024 consecutive
primes from 5.389e9, 6.023e9, 8.041e9 and 8.317e9) and its 33-bit night
measurements, committed under tests/golden/; the 33-bit Montgomery product
and power against Python integers; and the dense scan of the same primes,
which counts A4 directly and so pins the host's elliptic A4 of the wide
primes independently.
"""
import json
import random
import subprocess

import pytest

import golden_data
import rabung_amd
from rabung_amd import codec, workloads
from test_worker import WORKER

pytestmark = pytest.mark.gpu

R32 = 1 << 32


@pytest.fixture(scope="module")
def engine():
    eng = rabung_amd.Engine()
    yield eng
    eng.close()


def test_wide_product_and_power():
    rng = random.Random(33)
    for p in [4294967311, 5389000007, 8589934583]:
        assert rabung_amd.is_prime(p)
        a = [rng.randrange(p) for _ in range(2045)] + [0, 1, p - 1]
        b = [rng.randrange(p) for _ in range(2045)] + [p - 1, p - 1, p - 1]
        rinv = pow(R32, -1, p)
        assert rabung_amd.check_wide(a, b, p, 0) == [x * y * rinv % p for x, y in zip(a, b)]
        # powers of Montgomery forms: (x R)^e -> x^e R
        x = [rng.randrange(1, p) for _ in range(512)]
        e = [rng.randrange(1 << 34) for _ in range(510)] + [0, 1]
        got = rabung_amd.check_wide([v * R32 % p for v in x], e, p, 1)
        assert got == [pow(v, k, p) * R32 % p for v, k in zip(x, e)]


@pytest.mark.parametrize("lo", [5389000001, 6023000001, 8041000003, 8317000001])
def test_sparse_wide_window(engine, lo):
    w = [w for w in golden_data.holdout33_windows() if w["lo"] == lo][0]
    primes = [r["p"] for r in w["rows"]]
    print("wide window", lo, len(primes), flush=True)
    recs = engine.profile_block(primes, certify=True)
    got = codec.records_to_rows(recs)
    bad = [i for i in range(len(primes)) if got[i] != w["rows"][i]]
    assert not bad, ("mismatch", primes[bad[0]], got[bad[0]], w["rows"][bad[0]])
    assert all(r["p"] == p for r, p in zip(recs, primes))


def test_sparse_wide_night(engine):
    rows = golden_data.night33_rows()
    recs = engine.profile_block([r["p"] for r in rows], certify=True)
    assert codec.records_to_rows(recs) == rows


def test_dense_wide_against_sparse(engine):
    """Dense scan of the whole half interval against the sparse engine, A4
    included (dense: run summaries; sparse: host elliptic count)."""
    rows = golden_data.night33_rows()
    primes = [rows[0]["p"], golden_data.holdout33_windows()[0]["rows"][0]["p"]]
    print("dense wide", primes, flush=True)
    dense = engine.profile_block(primes, method=rabung_amd.METHOD_DENSE)
    sparse = engine.profile_block(primes)
    assert dense == sparse
    plain = engine.profile_block(primes[:1], method=rabung_amd.METHOD_PLAIN)
    assert plain == sparse[:1]


def test_dense_wide_top_of_domain(engine):
    """The largest prime below 2^33: its half interval (4294967291 positions)
    fills the 32-bit tile space to the last tile."""
    p = workloads.top_wide(1)
    assert p == [8589934583]
    dense = engine.profile_block(p, method=rabung_amd.METHOD_DENSE)
    assert dense == engine.profile_block(p, certify=True)


def test_random_wide_primes_sparse_vs_dense(engine):
    """Seeded primes drawn uniformly from (2^32, 2^33): the sparse engine and
    the dense scan (independent algorithms, both on 33-bit residues) agree bit
    for bit, A4 included, and the sparse records pass direct certification."""
    rng = random.Random(20261204)
    primes = sorted({rabung_amd.primes_from_wide(rng.randrange(R32, 2 * R32 - 2000), 1)[0]
                     for _ in range(6)})
    print("random wide", primes, flush=True)
    sparse = engine.profile_block(primes, certify=True)
    assert engine.last_timing()["sparse_primes"] == len(primes)
    dense = engine.profile_block(primes, method=rabung_amd.METHOD_DENSE)
    for s_, d_ in zip(sparse, dense):
        assert s_ == d_, (s_["p"], s_, d_)


def test_block_straddling_2_32(engine):
    """One call with primes on both sides of 2^32: each domain's kernels,
    records in input order, equal to single-domain calls."""
    below = rabung_amd.primes_from(R32 - 400, 12)
    above = rabung_amd.primes_from_wide(R32, 12)
    mixed = [x for pair in zip(below, above) for x in pair]
    recs = engine.profile_block(mixed, certify=True)
    assert [r["p"] for r in recs] == mixed
    assert recs[0::2] == engine.profile_block(below)
    assert recs[1::2] == engine.profile_block(above)
    night = {r["p"]: r for r in golden_data.night_rows()}
    got = codec.records_to_rows(recs)
    for row in got:
        if row["p"] in night:
            assert row == night[row["p"]]


def test_worker_domain33_protocol():
    w = [w for w in golden_data.holdout33_windows() if w["lo"] == 6023000001][0]
    b0, b1 = w["rows"][:128], golden_data.night33_rows()
    text = "0 %s\n1 %s\nQUIT\n" % (" ".join(str(r["p"]) for r in b0),
                                   " ".join(str(r["p"]) for r in b1))
    res = subprocess.run([str(WORKER), "profile", "--domain-bits", "33"], input=text.encode(),
                         capture_output=True, timeout=300)
    assert res.returncode == 0, res.stderr
    lines = [json.loads(x) for x in res.stdout.decode().splitlines()]
    assert lines[0]["type"] == "ready" and lines[0]["domain_bits"] == 33
    meas = [x for x in lines if x["type"] == "measurement"]
    want = b0 + b1
    assert [m["p"] for m in meas] == [r["p"] for r in want]
    for m, row in zip(meas, want):
        assert [(v["r"], v["max"], v["prefix"], v["K"], v["witness"]) for v in m["profiles"]] == \
            [(v["r"], v["max"], v["prefix"], v["K"], v["witness"]) for v in row["profiles"]]
    assert lines[-1]["type"] == "bye"


@pytest.mark.parametrize("text", ["0 997\nQUIT\n", "0 8589934609\nQUIT\n",
                                  "0 5389000001\nQUIT\n"])
def test_worker_domain33_refuses(text):
    res = subprocess.run([str(WORKER), "profile", "--domain-bits", "33"], input=text.encode(),
                         capture_output=True, timeout=120)
    assert res.returncode == 2 and b"VDW_CAMPAIGN_FAILURE" in res.stderr
    assert b"measurement" not in res.stdout
