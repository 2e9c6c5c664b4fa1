"""Wider arithmetic domain (primes between 2^32 and 2^33): host-side checks.

The device path is in test_gpu_wide.py.  Here: the committed 33-bit fixtures
(the reference's 33-bit throughput windows and night measurements) are
consistent with this library's prime, mask and threshold rules, every sampled
witness run and prefix is confirmed by direct powers in Python integers (the
reference's certify_witness rule, VBB/src/advance.cpp:7-13), and the widened
host arithmetic (Miller-Rabin, A4 point count) agrees with independent
computations.
"""
import ctypes
import random

import golden_data
import rabung_amd
from rabung_amd import codec


def small_primes(limit):
    sieve = bytearray([1]) * (limit + 1)
    sieve[0:2] = b"\x00\x00"
    for i in range(2, int(limit ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = bytearray(len(sieve[i * i::i]))
    return [i for i in range(limit + 1) if sieve[i]]


SMALL = small_primes(92_700)  # covers sqrt(2^33)


def trial_prime(n):
    if n < 2:
        return False
    for q in SMALL:
        if q * q > n:
            return True
        if n % q == 0:
            return n == q
    return True


def sample_rows():
    rows = list(golden_data.night33_rows())
    for w in golden_data.holdout33_windows():
        rows += w["rows"][:24] + w["rows"][-8:]
    return rows


def test_fixture_windows_are_consecutive_primes():
    windows = golden_data.holdout33_windows()
    assert [w["lo"] for w in windows] == [5389000001, 6023000001, 8041000003, 8317000001]
    for w in windows:
        primes = [r["p"] for r in w["rows"]]
        assert len(primes) == 1024
        assert primes == rabung_amd.primes_from_wide(w["lo"], 1024)
        assert primes[-1] <= w["hi"] < (1 << 33)


def test_primality_against_trial_division():
    rng = random.Random(20261204)
    for _ in range(300):
        n = rng.randrange(1 << 32, 1 << 33) | 1
        assert rabung_amd.is_prime(n) == trial_prime(n), n
    # products of two primes near 2^16.5 and a strong pseudoprime base set
    for n in [92_681 * 92_683, 65_537 * 131_071, 4_759_123_141, 8_589_934_583]:
        assert rabung_amd.is_prime(n) == trial_prime(n), n
    assert not rabung_amd.is_prime((1 << 33) + 17)  # outside the domain


def test_masks_and_thresholds_match_profiles():
    for row in sample_rows():
        p = row["p"]
        m = rabung_amd.mask(p)
        assert [v["r"] for v in row["profiles"]] == [r for r in range(2, 10) if m >> (r - 2) & 1]
        for v in row["profiles"]:
            assert v["K"] == codec.threshold_k(p, v["r"], v["max"], v["prefix"])
            assert 1 <= v["prefix"] <= v["max"] <= (p - 1) // v["r"]
            assert 1 <= v["witness"] <= (p - 1) // 2


def test_golden_witness_runs_by_direct_powers():
    """The run [W, W+M-1] is monochromatic and maximal, and 1..L is the
    identity run (by reflection the witness segment is a run of length M
    whichever end the run was found at)."""
    for row in sample_rows():
        p = row["p"]
        for v in row["profiles"]:
            e = (p - 1) // v["r"]
            W, M, L = v["witness"], v["max"], v["prefix"]
            colour = pow(W, e, p)
            assert all(pow(W + j, e, p) == colour for j in range(1, M)), (p, v)
            assert W == 1 or pow(W - 1, e, p) != colour, (p, v)
            assert pow(W + M, e, p) != colour, (p, v)
            assert all(pow(x, e, p) == 1 for x in range(1, L + 1)), (p, v)
            assert pow(L + 1, e, p) != 1, (p, v)


def test_wide_a4_count_matches_reference_frames():
    """The widened A4 point count (128-bit field products) gives the
    reference's A4 values on a sample block."""
    rows, values, _ = golden_data.sample_block(49900)
    primes = [r["p"] for r in rows[:40]]
    lib = rabung_amd.load()
    arr = (ctypes.c_uint64 * len(primes))(*primes)
    out = (ctypes.c_uint32 * len(primes))()
    assert lib.rbg_a4_host_wide(arr, len(primes), out, 0) == 0
    assert list(out) == values[:40]


def test_wide_a4_within_character_sum_bound():
    """A4 = (p - 10 - Bd + S) / 8 with |S| = |N - p - 2| <= 2 sqrt(p) + 1 and
    |Bd| <= 12: 8 A4 stays within 2 sqrt(p) + 23 of p - 10."""
    primes = [row["p"] for row in golden_data.night33_rows()[:6]]
    a4 = rabung_amd.a4_host(primes)
    for p, a in zip(primes, a4):
        assert 0 < a <= (p - 3) // 2
        assert abs(8 * a - (p - 10)) <= 2 * int(p ** 0.5) + 23, (p, a)
