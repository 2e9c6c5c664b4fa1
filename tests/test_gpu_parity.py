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
    """Every profile of the reference's night measurement primes."""
    rows = golden_data.night_rows()
    got = as_rows(engine.profile_block([r["p"] for r in rows]))
    bad = [i for i in range(len(rows)) if got[i] != rows[i]]
    assert not bad, ("mismatch", got[bad[0]], rows[bad[0]])


def test_holdout_profiles(engine):
    """Every prime of the reference's four 32-bit throughput windows (1024
    consecutive primes from 0.257e9, 1.337e9, 2.971e9, 4.237e9), one call
    each."""
    for w in golden_data.holdout_windows():
        primes = [r["p"] for r in w["rows"]]
        got = as_rows(engine.profile_block(primes))
        bad = [i for i in range(len(primes)) if got[i] != w["rows"][i]]
        assert not bad, ("mismatch", primes[bad[0]], got[bad[0]], w["rows"][bad[0]])


@pytest.mark.parametrize("method", [rabung_amd.METHOD_BATCHED, rabung_amd.METHOD_PLAIN,
                                    rabung_amd.METHOD_DENSE])
def test_block2_24_complete(engine, method):
    """The plain-method workload, all 4096 consecutive primes from the first
    prime >= 2^24 (13,741 profiles), in one call per method: every (M, L, W,
    K) and A4 equals the committed fixture (oracle values, cross-checked
    against the reference's moment_worker and independent_profile)."""
    rows, a4, doc = golden_data.block2_24()
    primes = [r["p"] for r in rows]
    recs = engine.profile_block(primes, method=method)
    got = as_rows(recs)
    bad = [i for i in range(len(rows)) if got[i] != rows[i] or recs[i]["A4"] != a4[i]]
    assert not bad, ("mismatch", primes[bad[0]], got[bad[0]], rows[bad[0]], recs[bad[0]]["A4"],
                     a4[bad[0]])
    assert codec.science_hash(got) == codec.science_hash(rows)
    t = engine.last_timing()
    if method == rabung_amd.METHOD_BATCHED:
        assert t["sparse_primes"] == 4096 and t["sparse_profiles"] == doc["profiles"]
        assert t["fallbacks"] == 0


def test_oracle_near_2_24(engine):
    """Seeded primes of the 4096-prime block starting at 2^24 against the oracle."""
    primes = oracle.next_primes(1 << 24, 4096)
    rng = random.Random(24)
    pick = sorted(rng.sample(primes, 48))
    exp = oracle.profile_many(pick)
    got = engine.profile_block(pick)
    for g, e in zip(got, exp):
        assert g == e, (g["p"], g, e)


BATCH64_SHA256 = "f4ebfa462d8f7a50593e5a31f4265609788fdb4492f929ad2323afb7c5b1b8e1"


def test_batch64_workload(engine):
    """The batching workload: 64 seeded blocks of U{1..4096} consecutive
    primes from log-uniform starts in [2^20, 2^32) (152,696 primes, 508,348
    profiles) in ONE rbg_profile_batch call.  Checked four ways: equal to
    one profile_block call per block; equal to every committed golden value
    the blocks happen to contain; equal to the dense scan on a seeded subset
    of 48 primes; equal to the CPU oracle on a seeded subset of its primes
    below 2^25."""
    import hashlib
    import struct
    from rabung_amd import workloads
    blocks = workloads.batch_blocks(seed=2026, nblocks=64)
    flat = [p for b in blocks for p in b]
    assert hashlib.sha256(b"".join(struct.pack("<I", p) for p in flat)).hexdigest() == \
        BATCH64_SHA256
    got = engine.profile_batch(flat)
    t = engine.last_timing()
    assert t["sparse_profiles"] == 508348 and t["fallbacks"] == 0
    # per block
    k = 0
    for b in blocks:
        assert engine.profile_block(b) == got[k:k + len(b)], b[0]
        k += len(b)
    # golden values inside the blocks
    golden = {}
    for block in (0, 49900):
        rows, a4, _ = golden_data.sample_block(block)
        golden.update({r["p"]: (r, v) for r, v in zip(rows, a4)})
    rows, a4, _ = golden_data.block2_24()
    golden.update({r["p"]: (r, v) for r, v in zip(rows, a4)})
    for w in golden_data.holdout_windows():
        golden.update({r["p"]: (r, None) for r in w["rows"]})
    golden.update({r["p"]: (r, None) for r in golden_data.night_rows()})
    index = {p: i for i, p in enumerate(flat)}
    for p in set(golden) & set(index):
        row, value = golden[p]
        rec = got[index[p]]
        assert as_rows([rec])[0] == row, p
        assert value is None or rec["A4"] == value, p
    # dense scan on a seeded subset
    rng = random.Random(64)
    pick = sorted(rng.sample(range(len(flat)), 48))
    dense = engine.profile_block([flat[i] for i in pick], method=rabung_amd.METHOD_DENSE)
    for i, d in zip(pick, dense):
        assert got[i] == d, flat[i]
    # the oracle on small primes of the workload
    small = [i for i in range(len(flat)) if flat[i] < 1 << 25]
    pick = sorted(rng.sample(small, min(16, len(small))))
    exp = oracle.profile_many([flat[i] for i in pick])
    for i, e in zip(pick, exp):
        assert got[i] == e, flat[i]


def test_forced_refusal_falls_back_to_dense():
    """With an extension limit of one step every passing anchor whose run
    needs more steps refuses its prime (the refusal the reference takes on
    queue overflow, gpu_pipeline.cuh:335,342,363); the prime is recomputed by
    the dense scan and the records equal the default engine's and the dense
    method's, golden values included."""
    rows, a4, _ = golden_data.sample_block(49900)
    pick = sorted(random.Random(12).sample(range(len(rows)), 12))
    primes = [rows[i]["p"] for i in pick] + oracle.next_primes(1 << 24, 8) + \
        oracle.next_primes(70001, 4)
    with rabung_amd.Engine() as eng:
        normal = eng.profile_block(primes)
        assert eng.last_timing()["fallbacks"] == 0
        dense = eng.profile_block(primes, method=rabung_amd.METHOD_DENSE)
    with rabung_amd.Engine(extend_limit=1) as eng:
        forced = eng.profile_block(primes)
        t = eng.last_timing()
    assert t["fallbacks"] > 0 and t["dense_primes"] >= t["fallbacks"]
    assert forced == normal == dense
    got = as_rows(forced[:12])
    for j, i in enumerate(pick):
        assert got[j] == rows[i] and forced[j]["A4"] == a4[i]


def test_speculative_threshold_misses_fall_back():
    """A speculative threshold above every profile's longest run (spec_boost
    40 lifts it to the cap Q) makes every profile miss: each prime is
    recomputed by the dense scan, and the records equal the default
    engine's, the golden values and the records without speculation."""
    rows, a4, _ = golden_data.sample_block(49900)
    pick = sorted(random.Random(14).sample(range(len(rows)), 6))
    primes = [rows[i]["p"] for i in pick] + oracle.next_primes(1 << 24, 6) + \
        oracle.next_primes(70001, 4)
    with rabung_amd.Engine() as eng:
        ref = eng.profile_block(primes)
        t = eng.last_timing()
        assert t["spec_misses"] == 0 and t["fallbacks"] == 0
    for j, i in enumerate(pick):
        assert as_rows([ref[j]])[0] == rows[i] and ref[j]["A4"] == a4[i]
    with rabung_amd.Engine(spec_boost=40) as eng:
        got = eng.profile_block(primes)
        t = eng.last_timing()
    assert got == ref
    assert t["spec_misses"] == t["sparse_profiles"] > 0
    assert t["fallbacks"] == t["sparse_primes"] == len(primes)
    with rabung_amd.Engine(spec_off=1) as eng:
        assert eng.profile_block(primes) == ref
        assert eng.last_timing()["spec_misses"] == 0


@pytest.mark.parametrize("q", [12, 24, 63])
def test_cover_threshold_does_not_change_records(q):
    """Any certified cover gives the same exact profiles: threshold caps Q of
    12, 24 and 63 against the default 48 (the seeded primes span the sparse
    domain)."""
    rng = random.Random(q)
    primes = sorted({rabung_amd.primes_from(int(2.0 ** rng.uniform(17, 32)), 1)[0]
                     for _ in range(24)})
    with rabung_amd.Engine() as eng:
        ref = eng.profile_block(primes)
    with rabung_amd.Engine(cover_q=q) as eng:
        got = eng.profile_block(primes)
        t = eng.last_timing()
    assert got == ref
    assert t["sparse_primes"] > 0


def test_cover_threshold_outside_range_refused():
    with rabung_amd.Engine(cover_q=64) as eng:
        with pytest.raises(rabung_amd.RabungError) as e:
            eng.profile_block([101])
    assert e.value.status == 1


@pytest.mark.parametrize("div,chunk", [(3, 1), (4, 8), (8, 256)])
def test_cover_shape_and_chunk_do_not_change_records(div, chunk):
    """Cover shapes n = 3, 4, 8 (fewer denominators, wider tables; near 2^32
    n = 8 hits the 2^20 table limit and D grows back) and task chunks of 1,
    8 and 256 give the golden records of the top of block 49900 and the
    default engine's records on seeded primes across the sparse domain."""
    rows, a4, _ = golden_data.sample_block(49900)
    tail = rows[-40:]
    rng = random.Random(div * 1000 + chunk)
    primes = sorted({rabung_amd.primes_from(int(2.0 ** rng.uniform(17, 32)), 1)[0]
                     for _ in range(24)})
    with rabung_amd.Engine() as eng:
        ref = eng.profile_block(primes)
    with rabung_amd.Engine(cover_div=div, filter_chunk=chunk) as eng:
        got = eng.profile_block(primes)
        top = eng.profile_block([r["p"] for r in tail])
        t = eng.last_timing()
    assert got == ref
    assert codec.records_to_rows(top) == tail
    assert [g["A4"] for g in top] == a4[-40:]
    assert t["sparse_primes"] == 40 and t["fallbacks"] == 0


@pytest.mark.parametrize("kw", [{"cover_div": 1}, {"cover_div": 9}, {"filter_chunk": 257}])
def test_cover_shape_and_chunk_outside_range_refused(kw):
    with rabung_amd.Engine(**kw) as eng:
        with pytest.raises(rabung_amd.RabungError) as e:
            eng.profile_block([101])
    assert e.value.status == 1

def test_methods_agree_and_are_deterministic(engine):
    primes = oracle.next_primes(3_000_000_000, 16) + [4294967291, 4273768801]
    a = engine.profile_block(primes, method=rabung_amd.METHOD_BATCHED)
    t = engine.last_timing()
    assert t["sparse_primes"] == len(primes) and t["fallbacks"] == 0
    b = engine.profile_block(primes, method=rabung_amd.METHOD_PLAIN)
    c = engine.profile_block(primes, method=rabung_amd.METHOD_DENSE)
    d = engine.profile_block(primes, method=rabung_amd.METHOD_BATCHED)
    assert a == b == c == d


def test_run_floor_rescan_keeps_results(engine):
    """Run floors (runs too short to matter are not tracked inside segments)
    never change results: with every floor raised past the longest run the
    library rescans the dense primes exactly and returns the same records."""
    primes = oracle.next_primes(3_000_000_000, 3) + [1000003, 65537, 4294967291]
    for method in (rabung_amd.METHOD_PLAIN, rabung_amd.METHOD_DENSE):
        want = engine.profile_block(primes, method=method)
        assert engine.last_timing()["rescans"] == 0
        with rabung_amd.Engine(floor_boost=40) as eng:
            assert eng.profile_block(primes, method=method) == want
            assert eng.last_timing()["rescans"] == 1


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
    """baseline_profile keeps the reference bridge's contract
    (baseline_bridge.cu:21-27): only the caller's orders are written, 2 for
    a mask it cannot serve, and a failure code other than 2."""
    exp = oracle.profile_many([1000003])[0]  # 1000002 = 2 * 3 * 166667: r = 2, 3, 6
    adm = exp["mask"]
    unset = 0xFFFFFFFF
    rc, out, pre = rabung_amd.baseline_profile(1000003, 0xFF, True, fill=unset)
    assert rc == 0
    for r in range(2, 10):
        if adm & (1 << (r - 2)):
            assert out[r] == exp["M"][r - 2] and pre[r] == exp["L"][r - 2]
        else:
            assert out[r] == 0 and pre[r] == 0  # asked for, not admissible
    assert out[0] == out[1] == unset
    rc, out, pre = rabung_amd.baseline_profile(1000003, 1 << (3 - 2), True, fill=unset)
    assert rc == 0 and out[3] == exp["M"][1] and pre[3] == exp["L"][1]
    assert all(out[r] == unset for r in range(10) if r != 3)
    rc, out, pre = rabung_amd.baseline_profile(1000003, (1 << 0) | (1 << 4), False, fill=unset)
    assert rc == 0 and (out[2], out[6]) == (exp["M"][0], exp["M"][4]) and out[3] == unset
    assert rabung_amd.baseline_profile(1000003, 1 << (7 - 2), False)[0] == 2  # 7 does not divide
    assert rabung_amd.baseline_profile(1000003, 0, False)[0] == 2
    assert rabung_amd.baseline_profile(1000001, 1, True)[0] == -1  # 1000001 = 101 * 9901
    rc, out, pre = rabung_amd.baseline_profile(4294967291, 0xFF, True)
    rows, _, _ = golden_data.sample_block(49900)
    assert rc == 0 and out[2] == rows[-1]["profiles"][0]["max"]

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


def test_batch_of_top_primes_splits_by_table_offsets():
    """About 8000 primes just below 2^32 hold more than 2^32 label entries
    (B ~ 562k each), so a batch with room in its HBM budget must still cut
    them into calls whose table offsets fit 32 bits.  The top 4096 are sample
    block 49900; the primes below them must equal one block call."""
    rows, a4, _ = golden_data.sample_block(49900)
    lo = rows[0]["p"]
    below = [p for p in oracle.next_primes(lo - 88000, 4400) if p < lo]
    assert 3800 < len(below) <= 4096
    flat = below + [r["p"] for r in rows]
    with rabung_amd.Engine(hbm_budget=64 << 30) as eng:
        got = eng.profile_batch(flat)
        assert eng.last_timing()["sparse_primes"] == len(flat)
        assert as_rows(got[len(below):]) == rows
        assert [g["A4"] for g in got[len(below):]] == a4
        assert got[:len(below)] == eng.profile_block(below)


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
