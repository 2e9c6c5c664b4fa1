"""Host-side verification: the A4 audit certificate and the direct check of
records (no GPU needed).

* Certificates of seeded primes below 2^32 are re-verified point by point
  (tests/a4_certificate.py); a planted change to any certificate field is
  rejected.
* rbg_certify_records (the check behind RBG_CERTIFY: witness run, its
  neighbours, prefix, K, mask by direct powers) accepts the golden records and
  rejects each planted mistake (a witness moved by one, M one too long or
  too short, L off by one, K off by one, a wrong mask).
"""
import copy
import random

import pytest

import a4_certificate
import golden_data
import rabung_amd
from rabung_amd import codec


def _primes(seed, n, lo, hi):
    rng = random.Random(seed)
    out = set()
    while len(out) < n:
        got = rabung_amd.primes_from(rng.randrange(lo, hi), 1)
        if got:
            out.add(got[0])
    return sorted(out)


def test_a4_certificates_verify():
    rows, a4, _ = golden_data.sample_block(49900)
    primes = _primes(11, 60, 1000, 1 << 32) + [rows[i]["p"] for i in (0, 100, 2312)]
    for p in primes:
        value, cert = rabung_amd.a4_certificate(p)
        assert a4_certificate.verify(cert), p
        assert value == cert["quadruples"] == rabung_amd.a4_host([p])[0]
    index = {r["p"]: v for r, v in zip(rows, a4)}
    for p in primes[-3:]:
        assert rabung_amd.a4_certificate(p)[0] == index[p]


def test_a4_certificate_small_prime():
    value, cert = rabung_amd.a4_certificate(997)
    assert cert["certificate"] == "small_dense" and cert["quadruples"] == value


@pytest.mark.parametrize("field,change", [
    ("quadruples", lambda c: c["quadruples"] + 1),
    ("elliptic_points", lambda c: c["elliptic_points"] + 2),
    ("order_lcm", lambda c: c["order_lcm"] * 2),
    ("twist_parameter", lambda c: 1 if c["twist_parameter"] != 1 else 2),
    ("hasse_upper", lambda c: c["hasse_upper"] + 1),
])
def test_a4_certificate_planted_mistakes(field, change):
    _, cert = rabung_amd.a4_certificate(4294967291)
    bad = copy.deepcopy(cert)
    bad[field] = change(cert)
    with pytest.raises(AssertionError):
        a4_certificate.verify(bad)


def test_a4_certificate_planted_point():
    _, cert = rabung_amd.a4_certificate(4294967279)
    for mutate in (lambda w: w.update(y=(w["y"] + 1) % cert["p"]),
                   lambda w: w.update(order=w["order"] * 2,
                                      prime_factors=sorted(set(w["prime_factors"] + [2])))):
        bad = copy.deepcopy(cert)
        mutate(bad["certificate_points"][0])
        with pytest.raises(AssertionError):
            a4_certificate.verify(bad)


def _golden_records(block, count, seed):
    rows, a4, _ = golden_data.sample_block(block)
    pick = sorted(random.Random(seed).sample(range(len(rows)), count))
    recs = []
    for i in pick:
        row = rows[i]
        rec = {"p": row["p"], "mask": codec.admissible_mask(row["p"]), "M": [0] * 8,
               "L": [0] * 8, "W": [0] * 8, "K": [0] * 8, "A4": a4[i]}
        for v in row["profiles"]:
            k = v["r"] - 2
            rec["M"][k], rec["L"][k], rec["W"][k], rec["K"][k] = (v["max"], v["prefix"],
                                                                  v["witness"], v["K"])
        recs.append(rec)
    return recs


def test_certify_records_accepts_golden():
    recs = _golden_records(0, 300, 1) + _golden_records(49900, 40, 2)
    assert rabung_amd.certify_records(recs) is None


def _plant(rec, k, how):
    rec = copy.deepcopy(rec)
    if how == "witness+1":
        rec["W"][k] += 1
    elif how == "witness-1":
        rec["W"][k] -= 1
    elif how == "M+1":
        rec["M"][k] += 1
        rec["K"][k] = max(rec["K"][k], rec["M"][k] + 1)
    elif how == "M-1":
        rec["M"][k] -= 1
    elif how == "L+1":
        rec["L"][k] += 1
        rec["M"][k] = max(rec["M"][k], rec["L"][k])
    elif how == "L-1":
        rec["L"][k] -= 1
    elif how == "K+1":
        rec["K"][k] += 1
    elif how == "mask":
        rec["mask"] ^= 1 << 7
    return rec


@pytest.mark.parametrize("how", ["witness+1", "witness-1", "M+1", "M-1", "L+1", "L-1", "K+1",
                                 "mask"])
def test_certify_records_rejects_planted(how):
    recs = _golden_records(49900, 12, 3)
    # an order whose run and prefix are long enough for every change to
    # stay in range (M >= 3, L >= 2 for "L-1", W >= 2 for "witness-1")
    target = None
    for i, rec in enumerate(recs):
        for k in range(8):
            if rec["mask"] >> k & 1 and rec["M"][k] >= 3 and rec["W"][k] >= 2 and \
                    (how != "L-1" or rec["L"][k] >= 2):
                target = (i, k)
                break
        if target:
            break
    if target is None:
        pytest.skip("no order in the sample allows this change")
    i, k = target
    bad = list(recs)
    bad[i] = _plant(recs[i], k, how)
    assert rabung_amd.certify_records(bad) == i
    assert rabung_amd.certify_records(recs) is None
