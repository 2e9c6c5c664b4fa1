"""Worker front-end (rbg_worker): the reference's two process protocols.

The A4 protocol (src/census/moment_worker.cpp:19-52) runs on the host and is
tested here on the CPU; the profile protocol (src/census/worker32.cu:10-45,
small_cpu.cpp:34-40) needs the GPU.  Expected values come from the sample
frames (tests/golden).
"""
import json
import struct
import subprocess
from pathlib import Path

import pytest

import a4_certificate
import golden_data

WORKER = Path(__file__).resolve().parent.parent / "rabung-atlas32_amd" / "rbg_worker"
HEADER = struct.Struct("<8sQIIQQIIQQ")


def run(mode, text, *args, timeout=120):
    if not WORKER.exists():
        pytest.fail("rbg_worker is not built (run rabung-atlas32_amd/build.py)")
    return subprocess.run([str(WORKER), mode, *args], input=text.encode(), capture_output=True,
                          timeout=timeout)


def parse_a4(data):
    nl = data.index(b"\n")
    ready = json.loads(data[:nl])
    out, i = [], nl + 1
    while i < len(data):
        h = HEADER.unpack(data[i:i + HEADER.size])
        i += HEADER.size
        n, clen = h[2], h[3]
        values = list(struct.unpack("<%dI" % n, data[i:i + 4 * n]))
        i += 4 * n
        cert = data[i:i + clen].decode()
        i += clen
        out.append((h, values, cert))
    return ready, out


def test_a4_protocol_matches_sample_frames():
    rows0, a40, _ = golden_data.sample_block(0)
    rows1, a41, _ = golden_data.sample_block(49900)
    b0 = [r["p"] for r in rows0[:300]]
    b1 = [r["p"] for r in rows1[-64:]]
    text = "0 -1 %d %s\n1 5 %d %s\nQUIT\n" % (len(b0), " ".join(map(str, b0)), len(b1),
                                            " ".join(map(str, b1)))
    res = run("a4", text)
    assert res.returncode == 0, res.stderr
    ready, batches = parse_a4(res.stdout)
    assert ready["type"] == "ready" and ready["maximum_batch"] == 4096
    (h0, v0, c0), (h1, v1, c1) = batches
    assert v0 == a40[:300] and v1 == a41[-64:]
    assert h0[0] == b"A4BATCH1" and h0[1] == 0 and h0[2] == 300 and h0[3] == 0 and c0 == ""
    assert h0[5] == sum(v0) and h0[7] == sum(p < 1000 for p in b0)
    assert (h0[8], h0[9]) == (b0[0], b0[-1]) and (h1[8], h1[9]) == (b1[0], b1[-1])
    cert = json.loads(c1)
    assert h1[3] == len(c1) and cert["p"] == b1[5] and cert["quadruples"] == v1[5]
    assert a4_certificate.verify(cert)


@pytest.mark.parametrize("text", ["0 -1 2 7 5\nQUIT\n", "0 -1 2 7 9\nQUIT\n",
                                  "0 3 2 5 7\nQUIT\n", "0 -1 1 5 7\nQUIT\n", "x\n"])
def test_a4_protocol_refuses(text):
    res = run("a4", text)
    assert res.returncode == 2 and b"VDW_A4_FAILURE" in res.stderr


def _profile_rows(stdout):
    lines = [json.loads(x) for x in stdout.decode().splitlines()]
    assert lines[0]["type"] == "ready" and lines[-1]["type"] == "bye"
    return lines[1:-1]


@pytest.mark.gpu
def test_profile_protocol_matches_sample_frames():
    rows0, a40, _ = golden_data.sample_block(0)
    rows1, a41, _ = golden_data.sample_block(49900)
    b0, b1 = rows0[1000:1128], rows1[-40:]
    text = "0 %s\n1 %s\nQUIT\n" % (" ".join(str(r["p"]) for r in b0),
                                   " ".join(str(r["p"]) for r in b1))
    res = run("profile", text, "--with-a4")
    assert res.returncode == 0, res.stderr
    lines = _profile_rows(res.stdout)
    meas = [x for x in lines if x["type"] == "measurement"]
    done = [x for x in lines if x["type"] == "batch_done"]
    assert [d["id"] for d in done] == [0, 1] and [d["count"] for d in done] == [128, 40]
    want = b0 + b1
    assert [m["p"] for m in meas] == [r["p"] for r in want]
    for m, row, a4 in zip(meas, want, a40[1000:1128] + a41[-40:]):
        assert m["status"] == "PASS" and m["A4"] == a4
        assert [(v["r"], v["max"], v["prefix"], v["K"], v["witness"]) for v in m["profiles"]] == \
            [(v["r"], v["max"], v["prefix"], v["K"], v["witness"]) for v in row["profiles"]]


@pytest.mark.gpu
@pytest.mark.parametrize("text", ["1 7\nQUIT\n", "0 9\nQUIT\n", "0 4294967311\nQUIT\n",
                                  "0 " + " ".join(["7"] * 129) + "\nQUIT\n"])
def test_profile_protocol_refuses(text):
    res = run("profile", text)
    assert res.returncode == 2 and b"VDW_CAMPAIGN_FAILURE" in res.stderr
    assert b"measurement" not in res.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("args,primes,frame", [
    (["--block", "0", "--first", "256", "--route", "worker"], 256, None),
    (["--block", "49900", "--sample", "130", "--seed", "7"], 130, None),
    (["--block", "0", "--method", "plain"], 4096, True),
])
def test_recompute_sample_block(args, primes, frame):
    """The recompute harness: the first 256 primes (2..1619) through the
    worker protocols, a seeded sample of the block below 2^32 through the
    library, and all of block 0 with its frame re-encoded byte-identically."""
    import sys
    root = Path(__file__).resolve().parent.parent
    res = subprocess.run([sys.executable, "-m", "rabung_amd.recompute", *args],
                         cwd=root / "rabung-atlas32_amd", capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stderr[-2000:]
    out = json.loads(res.stdout.strip().splitlines()[-1])
    assert out["result"] == "identical" and out["differences"] == 0
    assert out["primes"] == primes and out["frame_sha256_match"] == frame


@pytest.mark.gpu
def test_profile_capacity_refusal_replays_in_halves():
    """A workspace budget (12 MB) below what a 64-prime batch near 2^32 needs
    makes the library refuse with RBG_ECAPACITY; the worker replays the
    batch in halves (the reference's replay rule, GPU_CONFIG.json:164-170)
    and its output equals an unconstrained run's."""
    rows, a4, _ = golden_data.sample_block(49900)
    batch = rows[-64:]
    text = "0 %s\nQUIT\n" % " ".join(str(r["p"]) for r in batch)
    free = run("profile", text)
    tight = run("profile", text, "--hbm-budget", "12000000")
    assert free.returncode == 0 and tight.returncode == 0, tight.stderr
    assert b"capacity refusal of 64 primes" in tight.stderr
    assert b"capacity refusal" not in free.stderr

    def profiles(res):
        return [(m["p"], m["profiles"]) for m in _profile_rows(res.stdout)
                if m["type"] == "measurement"]
    assert profiles(tight) == profiles(free) == [
        (r["p"], [{k: v[k] for k in ("r", "max", "prefix", "K", "witness")}
                  for v in r["profiles"]]) for r in batch]


@pytest.mark.gpu
def test_profile_protocol_floor_and_a4_field():
    """The default floor is the GPU worker's 1000 (worker32.cu:24); --floor 2
    takes the CPU worker's small primes; A4 is reported only on request."""
    res = run("profile", "0 101\nQUIT\n")
    assert res.returncode == 2 and b"worker prime range" in res.stderr
    res = run("profile", "0 101 1009\nQUIT\n", "--floor", "2", "--with-a4")
    assert res.returncode == 0, res.stderr
    meas = [m for m in _profile_rows(res.stdout) if m["type"] == "measurement"]
    assert [m["p"] for m in meas] == [101, 1009] and all("A4" in m for m in meas)
    rows0, a40, _ = golden_data.sample_block(0)
    want = {r["p"]: v for r, v in zip(rows0, a40)}
    assert [m["A4"] for m in meas] == [want[101], want[1009]]
    res = run("profile", "0 1009\nQUIT\n")
    assert all("A4" not in m for m in _profile_rows(res.stdout))
