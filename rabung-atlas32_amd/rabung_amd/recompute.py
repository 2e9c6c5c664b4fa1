"""Re-derive primes of a committed sample block and diff them against its frames.

    python -m rabung_amd.recompute [--block 0|49900] [--first N | --sample N]
                                   [--method batched|plain|dense]
                                   [--route library|worker] [--frames DIR]

The block's scan frame and A4 frame are read with this package's decoder; the
selected primes (the first N, or N drawn with a fixed seed) are profiled by
librabung, either in-process through the ctypes shim (route `library`, one
call per 4096 primes) or through the worker executable's two line protocols
(route `worker`: `rbg_worker profile` in batches of 128 primes,
`rbg_worker a4` in one batch).  Every (M, L, K, witness) of every admissible
order and every A4 value is compared; when the whole block is recomputed on
the library route, the library's frame writer must also reproduce the
committed frame's SHA-256.

The reference's harness for the same purpose is scripts/recompute_small.py
(worker processes, first N primes of block 0, A4 header checks); this one
drives the library directly by default and lists every difference instead of
stopping at the first.  Output: one JSON summary line; exit status 0 when
nothing differs, 1 otherwise.
"""
import argparse
import hashlib
import json
import random
import struct
import subprocess
import sys
from pathlib import Path

from . import codec

PACKAGE_ROOT = Path(__file__).resolve().parent.parent
DEFAULT_FRAMES = PACKAGE_ROOT.parent / "tests" / "golden"
WORKER = PACKAGE_ROOT / "rbg_worker"
FIELDS = (("max", "M"), ("prefix", "L"), ("K", "K"), ("witness", "W"))
PROFILE_BATCH = 128  # primes per request of the profile protocol
A4_REPLY = struct.Struct("<8sQIIQQIIQQ")  # A4BATCH1 reply header (moment worker protocol)


def load_block(frames, block):
    """(rows, a4 values, manifest entry) of a committed sample block."""
    manifest = json.loads((frames / "sample_manifest.json").read_text())
    entry = next((b for b in manifest["blocks"] if b["block"] == block), None)
    if entry is None:
        raise SystemExit("no sample block %d under %s" % (block, frames))
    rows, _ = codec.decode_scan_frame((frames / entry["scan_file"]).read_bytes())
    values, _ = codec.decode_a4_frame((frames / entry["a4_file"]).read_bytes())
    return rows, values, entry


def by_library(primes, method):
    """Records of `primes` from librabung in calls of at most 4096 primes."""
    import rabung_amd
    methods = {"batched": rabung_amd.METHOD_BATCHED, "plain": rabung_amd.METHOD_PLAIN,
               "dense": rabung_amd.METHOD_DENSE}
    out = []
    with rabung_amd.Engine() as engine:
        for lo in range(0, len(primes), rabung_amd.MAX_BLOCK):
            out += engine.profile_block(primes[lo:lo + rabung_amd.MAX_BLOCK], method=methods[method])
    return codec.records_to_rows(out), [rec["A4"] for rec in out], out


def _run_worker(mode, request, *options):
    done = subprocess.run([str(WORKER), mode, *options], input=request, capture_output=True,
                          timeout=600)
    if done.returncode:
        raise SystemExit("rbg_worker %s failed (%d): %s"
                         % (mode, done.returncode, done.stderr.decode(errors="replace")[-400:]))
    return done.stdout


def by_worker(primes):
    """Rows and A4 values of `primes` through the worker's two protocols."""
    request = "".join("%d %s\n" % (k, " ".join(map(str, primes[lo:lo + PROFILE_BATCH])))
                      for k, lo in enumerate(range(0, len(primes), PROFILE_BATCH)))
    # floor 2: the sample blocks start at p = 2, below the GPU worker's 1000
    lines = [json.loads(x) for x in _run_worker("profile", (request + "QUIT\n").encode(),
                                                "--floor", "2").decode().splitlines()]
    rows = [{"p": x["p"], "profiles": x["profiles"]} for x in lines if x["type"] == "measurement"]
    a4_request = "0 -1 %d %s\nQUIT\n" % (len(primes), " ".join(map(str, primes)))
    reply = _run_worker("a4", a4_request.encode())
    body = reply[reply.index(b"\n") + 1:]
    head = A4_REPLY.unpack_from(body)
    if head[0] != b"A4BATCH1" or head[2] != len(primes):
        raise SystemExit("rbg_worker a4 answered an unexpected header")
    values = list(struct.unpack_from("<%dI" % len(primes), body, A4_REPLY.size))
    if (head[8], head[9], head[5]) != (primes[0], primes[-1], sum(values)):
        raise SystemExit("rbg_worker a4 header disagrees with its values")
    return rows, values


def differences(want_rows, want_a4, got_rows, got_a4):
    """Every disagreement as (p, r or 'A4', field, frame value, recomputed)."""
    out = []
    if [r["p"] for r in got_rows] != [r["p"] for r in want_rows]:
        return [(None, None, "prime list", len(want_rows), len(got_rows))]
    for want, got, a, b in zip(want_rows, got_rows, want_a4, got_a4):
        mine = {v["r"]: v for v in got["profiles"]}
        for v in want["profiles"]:
            other = mine.pop(v["r"], None)
            if other is None:
                out.append((want["p"], v["r"], "order", "present", "missing"))
                continue
            for key, label in FIELDS:
                if key in v and v[key] != other.get(key):
                    out.append((want["p"], v["r"], label, v[key], other.get(key)))
        out += [(want["p"], r, "order", "absent", "present") for r in mine]
        if a != b:
            out.append((want["p"], "A4", "A4", a, b))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--block", type=int, default=0, choices=(0, 49900))
    pick = ap.add_mutually_exclusive_group()
    pick.add_argument("--first", type=int, help="recompute the block's first N primes")
    pick.add_argument("--sample", type=int, help="recompute N primes drawn with --seed")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--method", default="batched", choices=("batched", "plain", "dense"))
    ap.add_argument("--route", default="library", choices=("library", "worker"))
    ap.add_argument("--frames", type=Path, default=DEFAULT_FRAMES)
    args = ap.parse_args(argv)

    rows, values, entry = load_block(args.frames, args.block)
    if args.sample is not None:
        if not 1 <= args.sample <= len(rows):
            ap.error("--sample must lie in 1..%d" % len(rows))
        index = sorted(random.Random(args.seed).sample(range(len(rows)), args.sample))
    else:
        count = len(rows) if args.first is None else args.first
        if not 1 <= count <= len(rows):
            ap.error("--first must lie in 1..%d" % len(rows))
        index = list(range(count))
    want_rows = [rows[i] for i in index]
    want_a4 = [values[i] for i in index]
    primes = [r["p"] for r in want_rows]

    frame_match = None
    if args.route == "library":
        got_rows, got_a4, records = by_library(primes, args.method)
        if len(index) == len(rows):
            import rabung_amd
            frame = rabung_amd.encode_scan_frame(records, entry["lo"], entry["hi"], True,
                                                 entry["zlib_level_scan"])
            frame_match = hashlib.sha256(frame).hexdigest() == entry["scan_sha256"]
    else:
        got_rows, got_a4 = by_worker(primes)

    diff = differences(want_rows, want_a4, got_rows, got_a4)
    for item in diff[:20]:
        print("differs: p=%s r=%s %s frame=%s recomputed=%s" % item, file=sys.stderr)
    ok = not diff and frame_match is not False
    print(json.dumps({"block": args.block, "route": args.route, "method": args.method,
                      "primes": len(primes), "largest": primes[-1],
                      "profiles": sum(len(r["profiles"]) for r in want_rows),
                      "a4_values": len(want_a4), "differences": len(diff),
                      "frame_sha256_match": frame_match,
                      "result": "identical" if ok else "different"}))
    return 0 if ok else 1


if __name__ == "__main__":
    raise SystemExit(main())
