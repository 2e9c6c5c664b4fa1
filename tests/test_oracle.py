"""Pinning the CPU oracle (oracle/rbg_oracle.c) before trusting it.

Checks, chosen so that a plausible mistake (tie rule, reflection, prefix,
K branch, A4 window bounds, colour relabelling) fails at least one:
  * every profile and A4 value of the reference sample block 0 (4096 primes);
  * a seeded prime of sample block 49900 (just below 2^32, full 2^32 walk);
  * the literal definition (pow(x,(p-1)/r,p) colours) for every prime < 700;
  * the worked example p=101, r=2 -> (M, L, K, w) = (7, 1, 8, 19);
  * reference night/holdout measurement profiles (seeded subsets);
  * the reference's own CPU programs built from its sources into oracle/_ref
    (independent_profile for (M, L, K, w); moment_worker's elliptic A4).
"""
import json
import os
import random
import subprocess
from pathlib import Path

import pytest

import brute
import golden_data
import oracle

REF_BIN = Path(oracle.__file__).resolve().parent / "_ref"


def test_sample_block0_complete():
    rows, a4, _ = golden_data.sample_block(0)
    recs = oracle.profile_many([r["p"] for r in rows])
    assert oracle.rows(recs) == rows
    assert [r["A4"] for r in recs] == a4


def test_brute_force_small_primes():
    primes = [p for p in range(2, 700) if oracle.is_prime(p)]
    recs = oracle.profile_many(primes)
    for p, rec in zip(primes, recs):
        assert rec == brute.profile(p), p


def test_worked_example_101():
    rec = oracle.profile(101, 1)
    assert (rec["M"][0], rec["L"][0], rec["K"][0], rec["W"][0]) == (7, 1, 8, 19)


@pytest.mark.slow
def test_sample_block49900_seeded_prime():
    rows, a4, _ = golden_data.sample_block(49900)
    i = random.Random(2313).randrange(len(rows))
    rec = oracle.profile(rows[i]["p"])
    assert oracle.rows([rec])[0] == rows[i]
    assert rec["A4"] == a4[i]


def test_night_smallest_primes():
    rows = golden_data.night_rows()[:3]  # 1.16e8 .. : full walks of ~1e8
    recs = [oracle.profile(r["p"]) for r in rows]
    assert oracle.rows(recs) == rows


def test_holdout_lowest_window_seeded():
    w = golden_data.holdout_windows()[0]  # primes near 2.57e8
    pick = random.Random(7).sample(range(len(w["rows"])), 2)
    for i in pick:
        rec = oracle.profile(w["rows"][i]["p"])
        assert oracle.rows([rec])[0] == w["rows"][i]


def test_planted_mistakes_are_caught():
    """Variants of the contract that golden vectors must reject."""
    rows, _, _ = golden_data.sample_block(0)
    good = oracle.rows(oracle.profile_many([r["p"] for r in rows[:600]]))
    assert good == rows[:600]
    # the reflected start matters: the plain start of the last longest run
    # (the mirror of the first) must differ somewhere
    latest = 0
    for row in rows[:600]:
        p = row["p"]
        for v in row["profiles"]:
            if v["witness"] != brute_witness_latest(p, v["r"]):
                latest += 1
    assert latest > 0
    # the K branch on r | (p-1)/2 matters
    assert any(v["K"] != max(v["max"] + 1, v["prefix"] + 2)
               for row in rows for v in row["profiles"])


def brute_witness_latest(p, r):
    """Plain (unreflected) start of the last longest run."""
    col = [None] + [pow(x, (p - 1) // r, p) for x in range(1, p)]
    best, wit, start = 0, p, 1
    for x in range(2, p + 1):
        if x == p or col[x] != col[x - 1]:
            length = x - start
            if length >= best:
                best, wit = length, start
            start = x
    return wit


def _ref(name):
    path = REF_BIN / name
    if not path.exists():
        pytest.skip("reference CPU programs not built (make -C oracle ref)")
    return path


def test_reference_independent_profile():
    exe = _ref("independent_profile")
    rng = random.Random(5)
    primes = oracle.next_primes(rng.randrange(2_000_000, 20_000_000), 3)
    recs = oracle.profile_many(primes)
    for rec in recs:
        p = rec["p"]
        for r in range(2, 10):
            if rec["mask"] & (1 << (r - 2)):
                k = max(3, rec["K"][r - 2])
                out = subprocess.run([str(exe), str(p), str(r), str(k)], capture_output=True,
                                     text=True, timeout=120)
                assert out.returncode in (0, 1), out.stderr
                ref = json.loads(out.stdout)
                assert (ref["M"], ref["L"], ref["K"]) == (rec["M"][r - 2], rec["L"][r - 2],
                                                          rec["K"][r - 2])
                if r == 2:  # least-start witness coincides with the reflected one
                    assert min(ref["w"], p - (ref["w"] + ref["M"] - 1)) >= rec["W"][0]


def moment_worker_a4(primes):
    exe = _ref("moment_worker")
    req = ("0 -1 %d %s\nQUIT\n" % (len(primes), " ".join(map(str, primes)))).encode()
    out = subprocess.run([str(exe)], input=req, capture_output=True, timeout=300).stdout
    ready, payload = out.split(b"\n", 1)
    import struct
    header = struct.Struct("<8sQIIQQIIQQ")
    fields = header.unpack(payload[:header.size])
    assert fields[0] == b"A4BATCH1" and fields[2] == len(primes)
    return list(struct.unpack("<%dI" % len(primes), payload[header.size:header.size + 4 * len(primes)]))


def test_reference_moment_worker_a4_near_2_24():
    primes = oracle.next_primes(1 << 24, 4096)
    pick = sorted(random.Random(24).sample(primes, 32))
    recs = oracle.profile_many(pick)
    assert [r["A4"] for r in recs] == moment_worker_a4(pick)


def test_block2_24_fixture_consistent():
    """The committed 2^24 block fixture: the prime list (count, ends, hash),
    the masks, the output contract's bounds and K, and the oracle itself on
    a seeded subset (the fixture was written by the oracle and cross-checked
    against the reference's CPU programs when it was made)."""
    import hashlib
    import struct
    from rabung_amd import codec
    rows, a4, doc = golden_data.block2_24()
    primes = [r["p"] for r in rows]
    assert primes == oracle.next_primes(1 << 24, 4096)
    assert hashlib.sha256(b"".join(struct.pack("<I", p) for p in primes)).hexdigest() == \
        doc["primes_sha256"]
    assert sum(len(r["profiles"]) for r in rows) == doc["profiles"] == 13741
    for row, mask_row, value in zip(rows, doc["rows"], a4):
        p = row["p"]
        assert mask_row[1] == codec.admissible_mask(p)
        assert 0 <= value <= (p - 3) // 2
        for v in row["profiles"]:
            assert 1 <= v["prefix"] <= v["max"] <= (p - 1) // v["r"]
            assert 1 <= v["witness"] <= (p - 1) // 2
            assert v["K"] == codec.threshold_k(p, v["r"], v["max"], v["prefix"])
    pick = sorted(random.Random(4096).sample(range(4096), 12))
    recs = oracle.profile_many([primes[i] for i in pick])
    assert oracle.rows(recs) == [rows[i] for i in pick]
    assert [r["A4"] for r in recs] == [a4[i] for i in pick]
