"""GPU parity: the HIP library against the golden vectors and the CPU oracle.

Bit-exact comparison of (mask, M, L, witness, K) and A4 for every admissible
order.  Sources of truth:
  * reference sample block 0 (4096 primes 2..38873) and block 49900 (primes
    just below 2^32), frames re-encoded byte-identically in tests/golden/;
  * the reference's holdout and night measurement profiles near 2^32;
  * the CPU oracle (discrete-log walk, oracle/rbg_oracle.c) on seeded inputs.
"""
import random

import pytest

import golden_data
import oracle
import rabung_amd
from rabung_amd import codec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    eng = rabung_amd.Engine()
    yield eng
    eng.close()


def as_rows(records):
    return codec.records_to_rows(records)


@pytest.mark.parametrize("method", [rabung_amd.METHOD_BATCHED, rabung_amd.METHOD_PLAIN,
                                    rabung_amd.METHOD_DENSE])
def test_sample_block0_all(engine, method):
    rows, a4, _ = golden_data.sample_block(0)
    primes = [r["p"] for r in rows]
    recs = engine.profile_block(primes, method=method)
    got = as_rows(recs)
    bad = [i for i in range(len(rows)) if got[i] != rows[i]]
    assert not bad, ("mismatch", primes[bad[0]], got[bad[0]], rows[bad[0]])
    assert [r["A4"] for r in recs] == a4


def test_sample_block0_hashes(engine):
    rows, a4, entry = golden_data.sample_block(0)
    recs = engine.profile_block([r["p"] for r in rows])
    got = as_rows(recs)
    assert codec.science_hash(got).hex() == entry["science_hash"]
    frame = codec.encode_scan_frame(got, entry["lo"], entry["hi"], True, entry["zlib_level_scan"])
    import hashlib
    assert hashlib.sha256(frame).hexdigest() == entry["scan_sha256"]


def test_sample_block49900_subset(engine):
    rows, a4, _ = golden_data.sample_block(49900)
    rng = random.Random(49900)
    idx = sorted(rng.sample(range(len(rows)), 24)) + [len(rows) - 1]
    primes = [rows[i]["p"] for i in idx]
    recs = engine.profile_block(primes)
    got = as_rows(recs)
    for j, i in enumerate(idx):
        assert got[j] == rows[i], (primes[j], got[j], rows[i])
        assert recs[j]["A4"] == a4[i], primes[j]


def test_night_profiles(engine):
    rows = golden_data.night_rows()
    rng = random.Random(151)
    pick = sorted(rng.sample(range(len(rows)), 24))
    primes = [rows[i]["p"] for i in pick]
    got = as_rows(engine.profile_block(primes))
    for j, i in enumerate(pick):
        assert got[j] == rows[i], (primes[j], got[j], rows[i])


def test_holdout_profiles(engine):
    for w in golden_data.holdout_windows():
        rng = random.Random(w["lo"])
        pick = sorted(rng.sample(range(len(w["rows"])), 8))
        primes = [w["rows"][i]["p"] for i in pick]
        got = as_rows(engine.profile_block(primes))
        for j, i in enumerate(pick):
            assert got[j] == w["rows"][i], (primes[j], got[j], w["rows"][i])


def test_oracle_near_2_24(engine):
    """Seeded primes of the 4096-prime block starting at 2^24 against the oracle."""
    primes = oracle.next_primes(1 << 24, 4096)
    rng = random.Random(24)
    pick = sorted(rng.sample(primes, 48))
    exp = oracle.profile_many(pick)
    got = engine.profile_block(pick)
    for g, e in zip(got, exp):
        assert g == e, (g["p"], g, e)


def test_methods_agree_and_are_deterministic(engine):
    primes = oracle.next_primes(3_000_000_000, 16) + [4294967291, 4273768801]
    a = engine.profile_block(primes, method=rabung_amd.METHOD_BATCHED)
    t = engine.last_timing()
    assert t["sparse_primes"] == len(primes) and t["fallbacks"] == 0
    b = engine.profile_block(primes, method=rabung_amd.METHOD_PLAIN)
    c = engine.profile_block(primes, method=rabung_amd.METHOD_DENSE)
    d = engine.profile_block(primes, method=rabung_amd.METHOD_BATCHED)
    assert a == b == c == d


def test_tile_geometry_invariance():
    """Results must not depend on segment length or persistent grid size."""
    primes = [3, 5, 7, 11, 13, 101, 65537, 1000003] + oracle.next_primes(1 << 24, 8)
    ref = None
    for chunks, bpc in [(1, 1), (3, 2), (64, 8), (200, 4)]:
        with rabung_amd.Engine(tile_chunks=chunks, blocks_per_cu=bpc) as eng:
            got = eng.profile_block(primes, method=rabung_amd.METHOD_DENSE)
            assert eng.profile_block(primes) == got  # sparse grid size too
        if ref is None:
            ref = got
        assert got == ref, (chunks, bpc)
    assert ref == oracle.profile_many(primes)


def test_order_of_block_and_duplicates(engine):
    primes = [4294967291, 2, 3, 101, 4294967291, 16777259]
    got = engine.profile_block(primes)
    assert [g["p"] for g in got] == primes
    assert got[0] == got[4]
    assert got[1]["mask"] == 0 and got[1]["A4"] == 0
    assert got[3]["M"][0] == 7 and got[3]["L"][0] == 1 and got[3]["K"][0] == 8 and got[3]["W"][0] == 19


def test_certify_flag(engine):
    primes = oracle.next_primes(2_000_000_000, 8)
    a = engine.profile_block(primes, certify=True)
    assert a == engine.profile_block(primes)


def test_refusals(engine):
    with pytest.raises(rabung_amd.RabungError) as e:
        engine.profile_block([101, 100])
    assert e.value.status == 1
    with pytest.raises(rabung_amd.RabungError):
        engine.profile_block([1])
    with pytest.raises(rabung_amd.RabungError):
        engine.profile_block([7] * 4097)
    assert engine.profile_block([]) == []


def test_baseline_bridge():
    rc, out, pre = rabung_amd.baseline_profile(4294967291, 0, True)
    assert rc == 0
    exp = oracle.profile_many([1000003])[0]
    rc, out, pre = rabung_amd.baseline_profile(1000003, 0, True)
    assert rc == 0
    for r in range(2, 10):
        if exp["mask"] & (1 << (r - 2)):
            assert out[r] == exp["M"][r - 2] and pre[r] == exp["L"][r - 2]
    rc, out, pre = rabung_amd.baseline_profile(1000003, 1 << (7 - 2), False)
    assert rc == 2  # 7 does not divide 1000002


def test_sparse_block49900_large_subset(engine):
    """Sparse engine on 400 seeded primes just below 2^32 (profiles and A4)."""
    rows, a4, _ = golden_data.sample_block(49900)
    rng = random.Random(400)
    idx = sorted(rng.sample(range(len(rows)), 400))
    recs = engine.profile_block([rows[i]["p"] for i in idx])
    t = engine.last_timing()
    assert t["sparse_primes"] == 400
    got = as_rows(recs)
    for j, i in enumerate(idx):
        assert got[j] == rows[i], (rows[i]["p"], got[j], rows[i])
        assert recs[j]["A4"] == a4[i]


def test_top256_headline_block(engine):
    """The benchmark's workload, the 256 largest primes below 2^32 (871
    profiles), in one batched call: every profile and A4 equal the committed
    sample block's, and the autotuned launch shapes give the same records."""
    from rabung_amd import workloads
    rows, a4, _ = golden_data.sample_block(49900)
    index = {r["p"]: (r, v) for r, v in zip(rows, a4)}
    primes = workloads.top(256)
    assert all(p in index for p in primes)
    recs = engine.profile_block(primes)
    t = engine.last_timing()
    assert t["sparse_primes"] == 256 and t["sparse_profiles"] == 871 and t["fallbacks"] == 0
    for rec, row in zip(recs, as_rows(recs)):
        want, want_a4 = index[row["p"]]
        assert row == want, row["p"]
        assert rec["A4"] == want_a4
    with rabung_amd.Engine() as eng:
        eng.autotune(primes[:32])
        assert eng.profile_block(primes) == recs


def test_sparse_all_orders_and_small_domain(engine):
    """Primes = 1 mod 2520 (every order admissible) across the sparse domain."""
    primes = [p for p in range(2520 * 27 + 1, 2520 * 3000, 2520 * 7) if oracle.is_prime(p)][:24]
    top = (4294967295 // 2520) * 2520 + 1 - 2520 * 4000  # = 1 mod 2520, near 2^32
    primes += [p for p in range(top, 4294967295, 2520) if oracle.is_prime(p)][:4]
    sparse = engine.profile_block(primes)
    dense = engine.profile_block(primes, method=rabung_amd.METHOD_DENSE)
    assert sparse == dense
    small = [p for p in primes if p < 10_000_000]
    assert engine.profile_block(small) == oracle.profile_many(small)


def test_autotune_keeps_results():
    primes = oracle.next_primes(4_200_000_000, 12) + oracle.next_primes(1 << 24, 12)
    with rabung_amd.Engine() as eng:
        before = eng.profile_block(primes)
        cfg = eng.autotune(primes[:16])
        assert cfg["blocks_per_cu"] in (6, 8, 12) and cfg["xcd_group"] in (1, 2, 4)
        assert eng.profile_block(primes) == before
        dcfg = eng.autotune(primes[-8:], rabung_amd.METHOD_DENSE)
        assert dcfg["tile_chunks"] in (16, 32, 64, 128, 256)
        assert eng.profile_block(primes, method=rabung_amd.METHOD_DENSE) == before


def test_batch_across_blocks(engine):
    blocks = [oracle.next_primes(s, n) for s, n in [(1 << 20, 300), (3 << 30, 40), (70001, 900)]]
    flat = [p for b in blocks for p in b]
    got = engine.profile_batch(flat)
    t = engine.last_timing()
    assert t["sparse_primes"] == len(flat)
    assert got == [r for b in blocks for r in engine.profile_block(b)]
    raw = engine.profile_batch_raw(flat)
    assert [rabung_amd.record_to_dict(raw[i]) for i in range(len(flat))] == got


@pytest.mark.parametrize("method", [rabung_amd.METHOD_BATCHED, rabung_amd.METHOD_PLAIN,
                                    rabung_amd.METHOD_DENSE])
def test_graph_capture_replays(method):
    """Launch chains recorded into a hipGraph give the eager results on every
    replay, with eager work of other shapes and syncs between the replays
    (the cached executable graph is updated or rebuilt each time)."""
    a = oracle.next_primes(4_000_000_000, 6) + oracle.next_primes(1 << 20, 10)
    b = oracle.next_primes(3_000_000_000, 9) + [101, 65537]
    with rabung_amd.Engine() as eng:
        ea, eb = eng.profile_block(a, method=method), eng.profile_block(b, method=method)
        for rep in range(4):
            assert eng.profile_block(a, method=method, graph=True) == ea, rep
            assert eng.profile_block(b, method=method) == eb
            assert eng.profile_block(b, method=method, graph=True) == eb, rep
            assert eng.profile_block(a, method=method, graph=True) == ea, rep


def test_random_primes_sparse_vs_dense(engine):
    """Seeded primes drawn log-uniformly over the sparse domain [2^16, 2^32)
    (and near its lower edge): the sparse engine and the dense scan are
    independent algorithms and must agree bit for bit; the oracle pins the
    smaller ones."""
    rng = random.Random(20261203)
    primes = set()
    while len(primes) < 40:
        x = int(2.0 ** rng.uniform(16, 32))
        got = rabung_amd.primes_from(x, 1)
        if got:
            primes.add(got[0])
    primes |= set(rabung_amd.primes_from(65520, 6))  # both sides of the sparse threshold
    primes = sorted(primes)
    sparse = engine.profile_block(primes)
    t = engine.last_timing()
    assert t["sparse_primes"] >= sum(p >= 1 << 16 for p in primes) - t["fallbacks"]
    dense = engine.profile_block(primes, method=rabung_amd.METHOD_DENSE)
    for s, d in zip(sparse, dense):
        assert s == d, (s["p"], s, d)
    small = [p for p in primes if p < 1 << 24]
    assert engine.profile_block(small) == oracle.profile_many(small)
