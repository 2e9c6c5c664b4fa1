#!/usr/bin/env python3
"""Write tests/golden/block2_24.json.gz: every profile and A4 value of the
4096 consecutive primes from the first prime >= 2^24 (16,777,259 ..
16,844,819; 13,741 profiles), the plain-method workload.

The reference commits no vector for this block, so the values come from the
CPU oracle (oracle/rbg_oracle.c), cross-checked in the same run against the
reference's own CPU programs built from its sources into oracle/_ref
(`make -C oracle ref`, development container only):

  * moment_worker (src/census/moment_worker.cpp:19-52, elliptic point count
    plus the closed form of src/census/theorie/elliptic4_mestre.cpp:42-67)
    for all 4096 A4 values, in one request of the A4BATCH1 protocol;
  * independent_profile (reproducibility/bounds_update/independent_profile.cpp
    :43-117: its own walk, colour populations, reflection and boundary
    checks) for (M, L, K) of every admissible order of a seeded subset of
    primes, and the least-start witness it prints, which bounds the
    reflected witness from above (the reflected witness is the least
    min(s, p - e) over runs of length M).

    python tests/golden/make_block2_24.py [--subset 64]

The fixture holds data only: the prime list's SHA-256, then one row per
prime [p, mask, [[r, M, L, K, W], ...], A4].
"""
import argparse
import gzip
import hashlib
import json
import random
import struct
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import oracle  # noqa: E402

REF_BIN = ROOT / "oracle" / "_ref"
OUT = HERE / "block2_24.json.gz"
START, COUNT = 1 << 24, 4096


def prime_list_sha256(primes):
    return hashlib.sha256(b"".join(struct.pack("<I", p) for p in primes)).hexdigest()


def a4_from_moment_worker(primes):
    """A4 values of the reference moment worker (audit disabled) for one
    strictly increasing request of at most 4096 primes."""
    request = "0 -1 %d %s\nQUIT\n" % (len(primes), " ".join(map(str, primes)))
    done = subprocess.run([str(REF_BIN / "moment_worker")], input=request.encode(),
                          capture_output=True, timeout=1800, check=True)
    _ready, payload = done.stdout.split(b"\n", 1)
    head = struct.Struct("<8sQIIQQIIQQ")
    magic, _rid, n = head.unpack_from(payload)[:3]
    if magic != b"A4BATCH1" or n != len(primes):
        raise SystemExit("moment_worker answered an unexpected header")
    return list(struct.unpack_from("<%dI" % n, payload, head.size))


def independent(p, r, k):
    done = subprocess.run([str(REF_BIN / "independent_profile"), str(p), str(r), str(k)],
                          capture_output=True, text=True, timeout=600)
    if done.returncode not in (0, 1):
        raise SystemExit("independent_profile refused p=%d r=%d: %s" % (p, r, done.stderr))
    return json.loads(done.stdout)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--subset", type=int, default=64,
                    help="primes checked with independent_profile (seeded)")
    args = ap.parse_args()
    for exe in ("moment_worker", "independent_profile"):
        if not (REF_BIN / exe).exists():
            raise SystemExit("oracle/_ref/%s missing: run make -C oracle ref" % exe)

    primes = oracle.next_primes(START, COUNT)
    assert len(primes) == COUNT and primes[0] == 16777259 and primes[-1] == 16844819
    print("profiling %d primes with the oracle ..." % COUNT, flush=True)
    recs = oracle.profile_many(primes)
    nprof = sum(bin(r["mask"]).count("1") for r in recs)
    assert nprof == 13741, nprof

    print("A4 of every prime with the reference moment worker ...", flush=True)
    ref_a4 = a4_from_moment_worker(primes)
    bad = [p for p, r, a in zip(primes, recs, ref_a4) if r["A4"] != a]
    if bad:
        raise SystemExit("A4 differs from moment_worker at p=%d" % bad[0])

    pick = sorted(random.Random(2024).sample(range(COUNT), args.subset))
    print("independent_profile on %d seeded primes ..." % len(pick), flush=True)
    checked = 0
    for i in pick:
        rec = recs[i]
        p = rec["p"]
        for r in range(2, 10):
            if not rec["mask"] & (1 << (r - 2)):
                continue
            k = r - 2
            ref = independent(p, r, max(3, rec["K"][k]))
            if (ref["M"], ref["L"], ref["K"]) != (rec["M"][k], rec["L"][k], rec["K"][k]):
                raise SystemExit("independent_profile differs at p=%d r=%d" % (p, r))
            if min(ref["w"], p - (ref["w"] + ref["M"] - 1)) < rec["W"][k]:
                raise SystemExit("a run of length M starts nearer an end than W at p=%d r=%d"
                                 % (p, r))
            checked += 1

    rows = []
    for rec in recs:
        prof = [[r, rec["M"][r - 2], rec["L"][r - 2], rec["K"][r - 2], rec["W"][r - 2]]
                for r in range(2, 10) if rec["mask"] & (1 << (r - 2))]
        rows.append([rec["p"], rec["mask"], prof, rec["A4"]])
    doc = {"workload": "4096 consecutive primes from the first prime >= 2^24",
           "first": primes[0], "last": primes[-1], "profiles": nprof,
           "primes_sha256": prime_list_sha256(primes),
           "checked": {"moment_worker_a4": COUNT, "independent_profile_orders": checked},
           "rows": rows}
    with gzip.GzipFile(OUT, "wb", mtime=0) as f:
        f.write(json.dumps(doc, separators=(",", ":")).encode())
    print("wrote %s: %d primes, %d profiles, %d orders cross-checked"
          % (OUT.name, COUNT, nprof, checked))


if __name__ == "__main__":
    main()
