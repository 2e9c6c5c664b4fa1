"""Recompute the first primes of a committed sample block with the library's
worker and compare them with the block's frames.

    python -m rabung_amd.recompute [--count 64] [--block 0|49900] [--frames DIR]

The counterpart of the reference's scripts/recompute_small.py:15-48: the
profile worker (line protocol, batches of 128 primes) must reproduce the scan
frame's rows (p, r, M, L, K, witness) and the A4 worker (binary protocol,
one batch) the A4 frame's values; the A4 header's count, first and last
prime and sum are checked too.  Here both workers are rbg_worker (profiles on
the GPU, A4 by point counts on the host) and the block may also be the one
just below 2^32.  Prints one PASS JSON line; any disagreement raises.
"""
import argparse
import json
import struct
import subprocess
from pathlib import Path

from . import codec

PKG = Path(__file__).resolve().parent.parent
REPO = PKG.parent
A4_HEADER = struct.Struct("<8sQIIQQIIQQ")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--count", type=int, default=64, help="number of initial primes; 1..4096")
    ap.add_argument("--block", type=int, default=0, choices=[0, 49900])
    ap.add_argument("--frames", type=Path, default=REPO / "tests" / "golden")
    ap.add_argument("--worker", type=Path, default=PKG / "rbg_worker")
    args = ap.parse_args(argv)
    if not 1 <= args.count <= 4096:
        ap.error("the bounded reproduction count is 1..4096")
    tag = "%05d" % args.block
    rows, _ = codec.decode_scan_frame((args.frames / ("scan-%s.vdw" % tag)).read_bytes())
    golden_a4, _ = codec.decode_a4_frame((args.frames / ("quadruplets-%s.a4" % tag)).read_bytes())
    if args.count > len(rows):
        ap.error("block %d holds %d primes" % (args.block, len(rows)))
    rows = rows[:args.count]
    primes = [v["p"] for v in rows]
    requests = ["%d %s" % (i // 128, " ".join(map(str, primes[i:i + 128])))
                for i in range(0, len(primes), 128)]
    worker = subprocess.run([str(args.worker), "profile"], input="\n".join(requests) + "\nQUIT\n",
                            text=True, capture_output=True, check=True, timeout=120)
    generated = []
    a4_inline = []
    for line in worker.stdout.splitlines():
        item = json.loads(line)
        if item["type"] == "measurement":
            generated.append({"p": item["p"], "profiles": item["profiles"]})
            a4_inline.append(item["A4"])
    assert generated == rows, "profile calculation disagrees with the sample block"
    request = ("0 -1 %d %s\nQUIT\n" % (len(primes), " ".join(map(str, primes)))).encode()
    a4 = subprocess.run([str(args.worker), "a4"], input=request, capture_output=True, check=True,
                        timeout=120).stdout
    ready, payload = a4.split(b"\n", 1)
    assert json.loads(ready)["type"] == "ready"
    magic, ident, n, cert_len, _ns, total, _att, _small, first, last = A4_HEADER.unpack(
        payload[:A4_HEADER.size])
    assert (magic, ident, n, cert_len, first, last) == (b"A4BATCH1", 0, len(primes), 0, primes[0],
                                                        primes[-1])
    assert len(payload) == A4_HEADER.size + 4 * n
    values = list(struct.unpack("<%dI" % n, payload[A4_HEADER.size:]))
    assert values == golden_a4[:n] == a4_inline and sum(values) == total
    print(json.dumps({"status": "PASS", "block": args.block, "primes": n, "last_prime": last,
                      "profiles": sum(len(v["profiles"]) for v in generated), "sum_A4": total}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
