"""CPU check of the sparse anchor search design (tests/sparse_model.py).

The model uses the kernels' cover, fold, band schedule, wheel masks,
thresholds and completion with plain integers; it must reproduce the oracle's
(M, witness, L) for every admissible order.
"""
import random

import oracle
import sparse_model


def _check(primes):
    for p in primes:
        rec = oracle.profile_many([p])[0]
        for r in range(2, 10):
            if rec["mask"] & (1 << (r - 2)):
                got = sparse_model.sparse_profile(p, r)
                assert got == (rec["M"][r - 2], rec["W"][r - 2], rec["L"][r - 2]), (p, r, got)


def test_model_random_primes():
    rng = random.Random(1)
    _check(oracle.next_primes(65537, 4) +
           [oracle.next_primes(rng.randrange(70000, 400000), 1)[0] for _ in range(8)])


def test_model_all_orders():
    _check([p for p in range(2520 * 30 + 1, 2520 * 400, 2520) if oracle.is_prime(p)][:6])


def test_cover_certificate():
    for p in oracle.next_primes(65537, 3) + [16777259, 4294967291]:
        B, D = sparse_model.cover(p)
        H = 47
        assert 4 * B * D - H * D * D > 4 * p and H * D <= 2 * B and B < p
        # the least B: one less fails the area inequality
        assert 4 * (B - 1) * D - H * D * D <= 4 * p


def test_screen_prunes_without_changing_results():
    for p in oracle.next_primes(200003, 3):
        rec = oracle.profile_many([p])[0]
        a, b = {}, {}
        assert sparse_model.sparse_profile(p, 2, screen=True, stats=a) == \
            sparse_model.sparse_profile(p, 2, screen=False, stats=b) == \
            (rec["M"][0], rec["W"][0], rec["L"][0])
        assert a["completions"] < b["completions"]


def test_spans_hold_exactly_the_scheduled_anchors():
    """The kernel's fused spans and the reference's two-sided band schedule
    select the same anchors, for every denominator of several covers and
    thresholds of both parities (odd and even t, d <= 2, s = 1)."""
    import random
    rng = random.Random(11)
    for p in (65537 + 2, 1000003, 257000003, 4294967291):
        B, D = sparse_model.cover(p)
        ds = sorted(set(range(1, min(D, 60) + 1)) | {D, D - 1} | {rng.randint(1, D) for _ in range(20)})
        for q in (3, 4, 5, 6, 7, 10, 17, 24, 25, 33, 47, 48):
            t = min(q, max(3, (q + 1) // 2))
            sp = q - t + 1
            phi = ((t - 1) // 2) % sp
            for d in ds:
                bmin = ((t - 1) * d + 1) // 2
                bands = sparse_model.anchors(sparse_model.band_ranges(d, t, sp, phi, bmin, B))
                spans = sparse_model.anchors(sparse_model.span_ranges(d, t, sp, phi, bmin, B))
                assert bands == spans, (p, q, d)


def test_word_masks_keep_the_scheduled_anchors_coprime_to_the_wheel():
    """The filter's words (96 anchors from each span's start) with the task's
    wheel mask built from s_lo alone keep exactly the scheduled anchors that
    are not excluded by gcd(b, d) in {2, 3}: the word starts are congruent to
    s_lo mod 2 and mod 3 whenever the wheel applies."""
    import random
    rng = random.Random(5)
    for p in (65537 + 2, 1000003, 4294967291):
        B, D = sparse_model.cover(p)
        ds = sorted(set(range(1, 25)) | {D, D - 1, D - 2} | {rng.randint(1, D) for _ in range(6)})
        for q in (3, 4, 17, 24, 25, 48):
            t = min(q, max(3, (q + 1) // 2))
            sp = q - t + 1
            phi = ((t - 1) // 2) % sp
            for d in ds:
                bmin = ((t - 1) * d + 1) // 2
                want = {b for b in sparse_model.anchors(
                            sparse_model.span_ranges(d, t, sp, phi, bmin, B))
                        if not (d % 2 == 0 and b % 2 == 0) and not (d % 3 == 0 and b % 3 == 0)}
                got = sparse_model.span_words(d, t, sp, phi, bmin, B)
                assert got == want, (p, q, d)


def test_class_row_words_hold_the_scheduled_anchors():
    """The filter's words over the wheel's class rows (one row per class c
    mod g coprime to g = gcd(d, 6)) keep exactly the scheduled anchors that
    gcd(b, d) in {2, 3} does not exclude; every anchor's row bit maps back to
    it, and its stage windows b - j d (j < t) are row bits k - j d/g inside
    its class's colours (numerators >= -B), as are the screen's reads up to
    q - t positions beyond."""
    import random
    rng = random.Random(6)
    for p in (65537 + 2, 1000003, 16777259, 4294967291):
        B, D = sparse_model.cover(p)
        # d = 1, 2 (one span over [bmin, B]) only where B is small
        ds = sorted(set(range(1 if B < 40000 else 3, 30)) | {D, D - 1, D - 2, D - 3, D - 4, D - 5} |
                    {rng.randint(3, D) for _ in range(5)})
        for q in ((3, 4, 17, 24, 25, 48) if p < 1 << 30 else (3, 24, 25, 48)):
            t = min(q, max(3, (q + 1) // 2))
            sp = q - t + 1
            phi = ((t - 1) // 2) % sp
            for d in ds:
                bmin = ((t - 1) * d + 1) // 2
                want = {b for b in sparse_model.anchors(
                            sparse_model.span_ranges(d, t, sp, phi, bmin, B))
                        if not (d % 2 == 0 and b % 2 == 0) and not (d % 3 == 0 and b % 3 == 0)}
                for np_ in ((1, 3) if d % 6 == 0 else (1,)):
                    got, info = sparse_model.class_row_words(d, t, sp, phi, bmin, B, np_)
                    assert set(got) == want, (p, q, d, np_)
                    g, dg, lay = info["g"], info["dg"], info["layout"]
                    for b, (ci, k) in got.items():
                        row = lay[info["row0"] + ci]
                        start = ci * info["cstride"]
                        assert (b + B - row["rho"]) % g == 0 and (b + B - row["rho"]) // g == k - start
                        # stage t-1 and the screen's right reads stay in the row
                        low = k - (t - 1) * dg
                        assert low >= start, (p, q, d)
                        assert k - start < row["bits"]
                        assert row["words"] * 32 >= row["bits"] + 256
