#!/usr/bin/env python3
"""Benchmark of the census-worker hot path on one MI355X.

A step is one pass of the hot path over one block of primes already resident
in HBM: every admissible order r in 2..9 of every prime gets (M, L, witness,
K) and every prime its A4 count.  Default workload: the 256 largest primes
below 2^32 (871 profiles), batched persistent launch; the plain method (one
launch per prime) is timed beside it on the same GPU.  Output: one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload top256|block2_24]
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "rabung-atlas32_amd"))

METRIC = "exact (p,r) profiles/sec near 2^32, plain vs batched; p50/p99 query latency"
HBM_PEAK_GBS = 8000.0
# sp_filter's unit of work: one 32-anchor dword of a plane window compared
# with the base window (one funnel shift, one xor, one or, and a share of the
# address arithmetic), counted only for the stages a word needed (through the
# stage that killed its last anchor), not for the rest of its last step
OPS_PER_WINDOW_DWORD = 4
# scan_tiles' unit of work: one Montgomery product (a 32x32->64 multiply, a
# low and a high 32-bit multiply, a subtract with borrow, a select, an add)
OPS_PER_PRODUCT = 6


def top_primes(count):
    import rabung_amd
    out = []
    x = (1 << 32) - 1
    while len(out) < count:
        if rabung_amd.is_prime(x):
            out.append(x)
        x -= 1
    return sorted(out)


def block_from(start, count):
    import rabung_amd
    out = []
    x = start
    while len(out) < count:
        if rabung_amd.is_prime(x):
            out.append(x)
        x += 1
    return out


WORKLOADS = {
    "top256": ("top-of-range: the 256 largest primes below 2^32, all admissible r, "
               "full profile (M, L, witness, K) + A4", lambda: top_primes(256)),
    "block2_24": ("plain block: 4096 consecutive primes from the first prime >= 2^24, all "
                  "admissible r, full profile + A4", lambda: block_from(1 << 24, 4096)),
}
BATCH_DESC = ("batching: 64 blocks of U{1..4096} consecutive primes from log-uniform starts in "
              "[2^20, 2^32) (seed 2026), all admissible r, one profile_batch call")


BATCH64_SHA256 = "f4ebfa462d8f7a50593e5a31f4265609788fdb4492f929ad2323afb7c5b1b8e1"


def measure_batch64(steps, warmup):
    """The cross-block batching workload: one rbg_profile_batch call per step
    over the 64 seeded blocks (records returned as the C record array in host
    memory).  Returns the measurement as a dict."""
    import hashlib
    import struct
    import torch
    import rabung_amd
    from rabung_amd import workloads
    blocks = workloads.batch_blocks(seed=2026, nblocks=64)
    flat = [p for b in blocks for p in b]
    digest = hashlib.sha256(b"".join(struct.pack("<I", p) for p in flat)).hexdigest()
    if digest != BATCH64_SHA256:
        raise SystemExit("bench: the batch64 prime list differs from the committed hash")
    profiles = sum(bin(rabung_amd.mask(p)).count("1") for p in flat)
    with rabung_amd.Engine(0, hbm_budget=64 << 30) as eng:
        for _ in range(warmup):
            eng.profile_batch_raw(flat)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            raw = eng.profile_batch_raw(flat)  # the C records, host buffers
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        t = eng.last_timing()
        first = [rabung_amd.record_to_dict(raw[i]) for i in range(len(blocks[0]))]
        if eng.profile_block(blocks[0]) != first:
            raise SystemExit("bench: batch64 records differ from the per-block call")
    recs = [rabung_amd.record_to_dict(raw[i]) for i in range(len(flat))]
    checked = golden_check(flat, recs)
    return {"workload": BATCH_DESC, "profiles_per_s": round(profiles / dt, 3),
            "ms_per_step": round(dt * 1000.0, 3), "steps": steps, "warmup": warmup,
            "blocks": len(blocks), "primes": len(flat), "profiles": profiles,
            "primes_sha256": digest,
            "records": "rbg_record array in host memory (the C-ABI output, no per-prime "
                       "Python objects in the timed region)",
            "sparse_primes": t["sparse_primes"], "dense_primes": t["dense_primes"],
            "fallbacks": t["fallbacks"], "spec_misses": t["spec_misses"],
            "golden_checked": checked,
            "stage_ms": {"prep": round(t["sparse_prep_ms"], 3),
                         "seed": round(t["sparse_seed_ms"], 3),
                         "filter": round(t["sparse_filter_ms"], 3),
                         "dense": round(t["dense_scan_ms"], 3),
                         "a4_host": round(t["a4_host_ms"], 3),
                         "a4_wait": round(t["a4_wait_ms"], 3),
                         "device_total": round(t["total_ms"], 3)}}


def bench_batch64(args):
    """The batching workload as the bench line's own workload."""
    m = measure_batch64(args.steps, args.warmup)
    line = {"metric": METRIC, "value": m["profiles_per_s"], "unit": "profiles/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": m["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
            "data": "synthetic: seeded prime blocks",
            "config": {k: v for k, v in m.items() if k not in ("profiles_per_s", "ms_per_step",
                                                               "stage_ms", "steps", "warmup")},
            "stage_ms": m["stage_ms"]}
    print(json.dumps(line), flush=True)
    return 0


# The reference's published batch throughput (RTX 5090): 1024 consecutive
# primes per window, all admissible r, eight 128-prime batches through its
# worker protocol, sum of batch wall times (paper/GPU_METHOD_EN.tex:839-857:
# four windows on the 32-bit path, two on the 33-bit path).
REFERENCE_WINDOWS = {257000003: 0.1461, 1337000003: 0.4461, 2971000001: 0.8447,
                     4237000001: 1.1118}
REFERENCE_WINDOWS_33 = {6023000001: 1.1628, 8041000003: 1.4750}


def bench_holdout(args):
    """The reference's throughput windows through this build's worker: the
    same 8 x 128-prime batches of the same line protocol, timed by the driver
    from sending a batch to its batch_done line (process start excluded);
    results checked against the windows' committed profiles."""
    import json as _json
    import subprocess
    sys.path.insert(0, str(ROOT / "tests"))
    import golden_data
    worker = ROOT / "rabung-atlas32_amd" / "rbg_worker"
    windows = []  # (window, reference seconds, worker domain bits)
    for w in golden_data.holdout_windows():
        if w["lo"] in REFERENCE_WINDOWS:
            windows.append((w, REFERENCE_WINDOWS[w["lo"]], 32))
    for w in golden_data.holdout33_windows():
        if w["lo"] in REFERENCE_WINDOWS_33:
            windows.append((w, REFERENCE_WINDOWS_33[w["lo"]], 33))
    procs, batch_ids = {}, {32: 0, 33: 0}
    for bits in (32, 33):
        procs[bits] = subprocess.Popen([str(worker), "profile", "--domain-bits", str(bits)],
                                       stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True,
                                       bufsize=1)
        ready = _json.loads(procs[bits].stdout.readline())
        assert ready["type"] == "ready" and ready["domain_bits"] == bits

    def run_window(rows, bits):
        proc = procs[bits]
        primes = [r["p"] for r in rows]
        got, total = [], 0.0
        for i in range(0, len(primes), 128):
            t0 = time.perf_counter()
            proc.stdin.write("%d %s\n" % (batch_ids[bits], " ".join(map(str, primes[i:i + 128]))))
            proc.stdin.flush()
            while True:
                item = _json.loads(proc.stdout.readline())
                if item["type"] == "measurement":
                    got.append({"p": item["p"], "profiles": item["profiles"]})
                elif item["type"] == "batch_done":
                    break
            total += time.perf_counter() - t0
            batch_ids[bits] += 1
        if got != rows:
            raise SystemExit("bench: worker results differ from the window's profiles")
        return total

    out, ref_total, our_total, profiles = [], 0.0, 0.0, 0
    for w, ref, bits in windows:
        run_window(w["rows"], bits)  # warm-up pass
        reps = [run_window(w["rows"], bits) for _ in range(max(1, args.steps))]
        mean = sum(reps) / len(reps)
        nprof = sum(len(r["profiles"]) for r in w["rows"])
        out.append({"window_start": w["lo"], "path_bits": bits, "primes": len(w["rows"]),
                    "profiles": nprof, "seconds": round(mean, 4),
                    "profiles_per_s": round(nprof / mean, 1), "reference_seconds": ref,
                    "speedup_vs_reference": round(ref / mean, 2)})
        print("bench: window %d (%d-bit path) %.4f s, reference %.4f s"
              % (w["lo"], bits, mean, ref), file=sys.stderr, flush=True)
        ref_total += ref
        our_total += mean
        profiles += nprof
    for proc in procs.values():
        proc.stdin.write("QUIT\n")
        proc.stdin.flush()
        proc.wait(timeout=60)
    line = {"metric": METRIC, "value": round(profiles / our_total, 3), "unit": "profiles/s",
            "n_gpus": 1, "steps": args.steps, "warmup": 1,
            "ms_per_step": round(our_total * 1000.0, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": round(ref_total / our_total, 3), "dtype": "u32",
            "data": "reference throughput windows (committed profiles)",
            "config": {"workload": "reference batch-throughput windows: 1024 consecutive primes "
                                   "from 0.257e9, 1.337e9, 2.971e9, 4.237e9 (32-bit path) and "
                                   "6.023e9, 8.041e9 (33-bit path), all admissible r, "
                                   "8 x 128-prime batches through rbg_worker's line protocol "
                                   "(certified), sum of batch wall times per window",
                       "windows": out,
                       "baseline": "paper/GPU_METHOD_EN.tex:839-857, RTX 5090, selected engine"}}
    print(json.dumps(line), flush=True)
    return 0


def bench_wide(args):
    """The wider domain (2^32 < p < 2^33): the 256 largest primes below 2^33
    through rbg_profile_block_wide (host buffers, batched method); records
    certified by direct powers once, outside the timed steps."""
    import rabung_amd
    from rabung_amd import workloads
    primes = workloads.top_wide(256)
    profiles = sum(bin(rabung_amd.mask(p)).count("1") for p in primes)
    eng = rabung_amd.Engine(0)
    info = eng.device_info()
    ref = eng.profile_block(primes, certify=True)
    for _ in range(args.warmup):
        eng.profile_block(primes)
    timings = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if eng.profile_block(primes) != ref:
            raise SystemExit("bench: wide records differ between steps")
        timings.append(eng.last_timing())
    dt = time.perf_counter() - t0
    peak_tops = info["cus"] * 128 * info["clock_khz"] * 1e3 / 1e12
    kern_ms = sum(t["sparse_filter_ms"] for t in timings) / len(timings)
    ops = timings[-1]["filter_windows"] * OPS_PER_WINDOW_DWORD
    t = timings[-1]
    line = {"metric": METRIC, "value": round(profiles / (dt / args.steps), 3),
            "unit": "profiles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(dt * 1000.0 / args.steps, 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u32/u64",
            "data": "synthetic (deterministic prime list), certified",
            "config": {"workload": "top of the wider domain: the 256 largest primes below 2^33 "
                                   "(%d profiles), all admissible r, full profile + A4, "
                                   "rbg_profile_block_wide" % profiles,
                       "primes": len(primes)},
            "roofline": {"bound": "valu", "kernel": "sp_filter<Dom33>",
                         "achieved": round(ops / (kern_ms / 1000.0) / 1e12, 3),
                         "peak": round(peak_tops, 3), "unit": "TOP/s",
                         "frac": round(ops / (kern_ms / 1000.0) / 1e12 / peak_tops, 4),
                         "traffic": None, "avg_launch_ms": round(kern_ms, 3),
                         "algorithmic_ops_per_launch": ops},
            "stage_ms": {"prep": round(t["sparse_prep_ms"], 3),
                         "seed": round(t["sparse_seed_ms"], 3),
                         "filter": round(t["sparse_filter_ms"], 3),
                         "a4_host": round(t["a4_host_ms"], 3),
                         "a4_wait": round(t["a4_wait_ms"], 3),
                         "device_total": round(t["total_ms"], 3)}}
    eng.close()
    print(json.dumps(line), flush=True)
    return 0


def mm_products(e):
    if e < 2:
        return 0
    return e.bit_length() + bin(e).count("1") - 2


def lcm_mask(mask):
    from math import gcd
    m = 1
    for r in range(2, 10):
        if mask & (1 << (r - 2)):
            m = m * r // gcd(m, r)
    return m


def int32_ops(primes):
    """Algorithmic INT32 work of the dense scan: nh * mm(e) * OPS_PER_PRODUCT
    per prime (mm(e) Montgomery products per position)."""
    import rabung_amd
    total = 0
    for p in primes:
        if p < 3:
            continue
        m = lcm_mask(rabung_amd.mask(p))
        total += ((p - 1) // 2) * mm_products((p - 1) // m) * OPS_PER_PRODUCT
    return total


def golden_check(primes, records):
    """Bit-exact check against the committed sample block 49900 when the
    workload lies inside it; returns the number of primes checked."""
    sys.path.insert(0, str(ROOT / "tests"))
    import golden_data
    from rabung_amd import codec
    rows, a4, entry = golden_data.sample_block(49900)
    index = {r["p"]: (r, v) for r, v in zip(rows, a4)}
    checked = 0
    got = codec.records_to_rows(records)
    for g, rec in zip(got, records):
        if g["p"] in index:
            row, v = index[g["p"]]
            if g != row or rec["A4"] != v:
                raise SystemExit("bench: result differs from the golden vector at p=%d" % g["p"])
            checked += 1
    return checked


def memory_path(timing, kern_ms, info):
    """The filter against the unit that bounds it: every wave-level 4-dword
    load moves 64 lanes x 16 bytes through a CU's vector-memory path (TA/TD,
    whatever the active lanes), which delivers 64 bytes per clock per CU
    (profiles/valu_and_load_rates.md: 16-17 cycles per such load)."""
    loads = timing.get("filter_loads", 0)
    if not loads:
        return None
    bytes_moved = loads * 64 * 16
    peak = info["cus"] * 64 * info["clock_khz"] * 1e3 / 1e12  # TB/s
    achieved = bytes_moved / (kern_ms / 1000.0) / 1e12
    return {"bound": "vector-memory path (TA/TD)", "wave_loads": loads,
            "bytes_per_launch": bytes_moved, "achieved": round(achieved, 3),
            "peak": round(peak, 3), "unit": "TB/s", "frac": round(achieved / peak, 4),
            "peak_basis": "%d CUs x 64 B x %.0f MHz" % (info["cus"], info["clock_khz"] / 1e3)}


def host_threads():
    """Host cores this process may run on (the box's CPU share)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _small_cpu_seconds(primes):
    """Wall time of the reference's own CPU worker (src/census/small_cpu.cpp,
    built from its sources into oracle/_ref) over `primes` through its line
    protocol (batches of 128), with its output rows; None if not built."""
    import subprocess
    exe = ROOT / "oracle" / "_ref" / "small_cpu"
    if not exe.exists():
        return None, None
    text = "".join("%d %s\n" % (k, " ".join(map(str, primes[i:i + 128])))
                   for k, i in enumerate(range(0, len(primes), 128))) + "QUIT\n"
    t0 = time.perf_counter()
    done = subprocess.run([str(exe)], input=text, capture_output=True, text=True, timeout=300)
    dt = time.perf_counter() - t0
    if done.returncode:
        return None, None
    rows = [json.loads(x) for x in done.stdout.splitlines()]
    return dt, [{"p": r["p"], "profiles": r["profiles"]} for r in rows if r["type"] == "measurement"]


def cpu_baseline(eng, top):
    """The CPU oracle (a port of the reference's dense CPU worker: discrete-log
    walk + full-interval scan, threaded) timed on the host cores of this run,
    with its records compared with the GPU's:
      * headline: a seeded 8-prime subset of the top-256 workload;
      * the first 256 primes (2..1619), also with the reference's own CPU
        worker when it has been built from its sources;
      * a seeded 256-prime sample of the 4096-prime block from 2^24.
    """
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    import random
    from rabung_amd import codec
    threads = host_threads()

    def timed(primes, many):
        t0 = time.perf_counter()
        recs = oracle.profile_many(primes, threads) if many else \
            [oracle.profile(p, threads) for p in primes]
        dt = time.perf_counter() - t0
        if eng.profile_block(primes) != recs:
            raise SystemExit("bench: CPU oracle and GPU records differ")
        return recs, dt

    def nprof(recs):
        return sum(bin(r["mask"]).count("1") for r in recs)

    print("bench: CPU baseline, 8 seeded top-of-range primes on %d threads" % threads,
          file=sys.stderr, flush=True)
    pick = sorted(random.Random(5).sample(top, 8))
    recs, dt = timed(pick, many=False)
    out = {"value": round(nprof(recs) / dt, 4), "unit": "profiles/s", "cores": threads,
           "kind": "port",
           "sample": "seeded 8 of the 256 largest primes below 2^32 (%s), %d profiles in %.1f s "
                     "on %d threads (oracle: dense discrete-log walk + full-interval scan, "
                     "threaded over each prime's walk); records equal to the GPU's"
                     % (",".join(map(str, pick)), nprof(recs), dt, threads),
           "os_cpu_count": os.cpu_count(), "threads_used": threads, "matches_gpu": True}
    print("bench: CPU baseline, first 256 primes and a 2^24 sample", file=sys.stderr, flush=True)
    c1 = oracle.next_primes(2, 256)
    recs1, dt1 = timed(c1, many=True)
    out["first_256_primes"] = {"profiles": nprof(recs1), "seconds": round(dt1, 4),
                               "profiles_per_s": round(nprof(recs1) / dt1, 1), "kind": "port",
                               "matches_gpu": True}
    ref_dt, ref_rows = _small_cpu_seconds(c1)
    if ref_dt is not None:
        if ref_rows != codec.records_to_rows(recs1):
            raise SystemExit("bench: the reference CPU worker disagrees on the first 256 primes")
        out["first_256_primes_reference_worker"] = {
            "seconds": round(ref_dt, 4), "profiles_per_s": round(nprof(recs1) / ref_dt, 1),
            "kind": "reference", "cores": 1, "matches_gpu": True,
            "program": "src/census/small_cpu.cpp built from its sources (oracle/_ref/small_cpu), "
                       "process start included"}
    block = oracle.next_primes(1 << 24, 4096)
    sample = sorted(random.Random(24).sample(block, 256))
    recs2, dt2 = timed(sample, many=True)
    out["block_2_24_sample"] = {"primes": len(sample), "profiles": nprof(recs2),
                                "seconds": round(dt2, 3),
                                "profiles_per_s": round(nprof(recs2) / dt2, 2),
                                "full_block_seconds_estimate": round(dt2 * len(block) / len(sample), 1),
                                "kind": "port", "cores": threads, "matches_gpu": True}
    return out


def calibrate_copy(nbytes=256 << 20, reps=3):
    """A plain device copy of known size: the reading of rocprof's FETCH_SIZE
    and WRITE_SIZE for it converts the counters of the filter into bytes."""
    import torch
    src = torch.ones(nbytes, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read": nbytes, "bytes_written": nbytes,
                      "reps": reps}), flush=True)
    return 0


def microbench():
    """VALU issue rates (TOP/s) and CU cycles per wave-wide load (one JSON line)."""
    import rabung_amd
    names = ["add_or_u32", "mul_lo_or_u32", "mul_hi_or_u32", "alignbit_b32", "montgomery_product",
             "add_u32", "mul_lo_u32"]
    rates = {n: round(rabung_amd.microbench(k) / 1e12, 3) for k, n in enumerate(names)}
    layout = {4: "lanes_4B_apart", 8: "lanes_8B_apart", 12: "lanes_12B_apart",
              4160: "line_per_lane"}
    loads = {"%d_dword%s_%s" % (w, "s" if w > 1 else "", layout[st]):
             round(rabung_amd.microbench(10 + w, st), 2)
             for st in (4, 8, 12, 4160) for w in (1, 2, 3, 4)}
    print(json.dumps({"valu_tops": rates, "cu_cycles_per_wave_load": loads}), flush=True)
    return 0


def profiled_traffic(path=ROOT / "profiles" / "sp_filter_top256.md"):
    """HBM bytes per filter launch on the top-256 workload, as measured by the
    two rocprof counter runs recorded in the profile summary (FETCH_SIZE and
    WRITE_SIZE, scaled by the same counters' reading for a device copy)."""
    try:
        for line in path.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="top256",
                    choices=sorted(WORKLOADS) + ["batch64", "holdout", "wide33"])
    ap.add_argument("--no-plain", action="store_true", help="skip timing the plain method")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch64", action="store_true",
                    help="skip the batching workload's sub-measurement")
    ap.add_argument("--serve-seconds", type=float, default=8.0,
                    help="length of each Poisson serving run (0 skips serving)")
    ap.add_argument("--qps", default="20,60",
                    help="serving arrival rates, comma separated (one run each)")
    ap.add_argument("--serve-chunk", type=int, default=64, help="primes per request piece")
    ap.add_argument("--serve-max", type=int, default=256, help="primes per serving batch")
    ap.add_argument("--microbench", action="store_true",
                    help="only measure VALU instruction rates and vector-load cost")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB device copy (rocprof FETCH/WRITE_SIZE calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy()
    if args.microbench:
        return microbench()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the path would shard across primes, "
              "each GPU profiling its own blocks with no exchange); run with --gpus 1",
              file=sys.stderr)
        return 2
    if args.workload == "batch64":
        return bench_batch64(args)
    if args.workload == "holdout":
        return bench_holdout(args)
    if args.workload == "wide33":
        return bench_wide(args)
    import torch
    import rabung_amd
    from rabung_amd import Record, RECORD_SIZE

    desc, make = WORKLOADS[args.workload]
    primes = make()
    n = len(primes)
    profiles = sum(bin(rabung_amd.mask(p)).count("1") for p in primes)
    dev = torch.device("cuda:0")
    d_primes = torch.tensor([p - (1 << 32) if p >= (1 << 31) else p for p in primes],
                            dtype=torch.int32, device=dev)
    d_out = torch.zeros(n * RECORD_SIZE, dtype=torch.uint8, device=dev)
    eng = rabung_amd.Engine(0)
    info = eng.device_info()
    # autotune the launch shape on a seeded sample of the workload (untimed)
    import random
    sample = sorted(random.Random(7).sample(primes, min(32, n)))
    tuned = eng.autotune(sample, rabung_amd.METHOD_BATCHED)

    def run(method, steps, warmup):
        for _ in range(warmup):
            eng.profile_block_device(d_primes.data_ptr(), n, d_out.data_ptr(), method)
        torch.cuda.synchronize()
        timings = []
        t0 = time.perf_counter()
        for _ in range(steps):
            eng.profile_block_device(d_primes.data_ptr(), n, d_out.data_ptr(), method)
            timings.append(eng.last_timing())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        return dt, timings

    print("bench: batched method, %d primes, %d steps" % (n, args.steps), file=sys.stderr,
          flush=True)
    dt, timings = run(rabung_amd.METHOD_BATCHED, args.steps, args.warmup)
    timing = timings[-1]
    raw = d_out.cpu().numpy().tobytes()
    records = [rabung_amd.record_to_dict(Record.from_buffer_copy(raw[i * RECORD_SIZE:(i + 1) * RECORD_SIZE]))
               for i in range(n)]
    checked = golden_check(primes, records)
    ms_step = dt * 1000.0 / args.steps
    value = profiles / (dt / args.steps)

    # INT32 VALU peak: each CU retires 128 32-bit lane operations per clock
    # (the same rate as vector FP32, 157.3 TFLOP/s as FMA); the live probe
    # below measures what dependent chains of adds, multiplies and whole
    # Montgomery products reach on this device
    peak_tops = info["cus"] * 128 * info["clock_khz"] * 1e3 / 1e12
    peak_basis = "%d CUs x 128 lanes x %.0f MHz, INT32 VALU ops" % (info["cus"],
                                                                    info["clock_khz"] / 1e3)
    probe = {"add_u32_tops": round(rabung_amd.microbench(5) / 1e12, 3),
             "mul_lo_u32_tops": round(rabung_amd.microbench(6) / 1e12, 3),
             "montgomery_products_per_s": round(rabung_amd.microbench(4), 0)}
    plain = None
    if not args.no_plain:
        psteps = max(1, args.steps // 10)
        print("bench: plain method, %d step(s)" % psteps, file=sys.stderr, flush=True)
        pdt, ptimings = run(rabung_amd.METHOD_PLAIN, psteps, 0)
        raw2 = d_out.cpu().numpy().tobytes()
        if raw2 != raw:
            raise SystemExit("bench: plain and batched methods disagree")
        # the library's own count of the powers it computed: with label
        # reuse only the seed positions and the positions coprime to 6 take
        # one (the others add a label gathered from HBM)
        ops = ptimings[-1]["products"] * OPS_PER_PRODUCT
        scan_s = sum(t["dense_scan_ms"] for t in ptimings) / len(ptimings) / 1000.0
        plain = {"method": "dense Montgomery scan with label reuse (powers for the positions "
                           "coprime to 6, labels of the others from x/2 and x/3), one "
                           "launch per phase over groups of primes",
                 "positions_per_step": ptimings[-1]["positions"],
                 "products_per_step": ptimings[-1]["products"],
                 "profiles_per_s": round(profiles / (pdt / psteps), 3),
                 "ms_per_step": round(pdt * 1000.0 / psteps, 3),
                 "scan_launches_per_step": ptimings[-1]["launches"],
                 "roofline": {"bound": "valu",
                              # label reuse (scan_reuse) takes fewer products
                              # than one power per position (scan_tiles)
                              "kernel": ("scan_reuse" if ops < int32_ops(primes)
                                         else "scan_tiles"),
                              "achieved": round(ops / scan_s / 1e12, 3),
                              "peak": round(peak_tops, 3), "unit": "TOP/s",
                              "frac": round(ops / scan_s / 1e12 / peak_tops, 4),
                              "algorithmic_ops_per_step": ops,
                              # priced at the 6 instructions this build's
                              # REDC issues; at the 8 of a textbook REDC
                              # the same time gives frac x 8/6
                              "ops_per_product": OPS_PER_PRODUCT,
                              "frac_at_8_ops_per_product": round(
                                  ops * 8 / OPS_PER_PRODUCT / scan_s / 1e12 / peak_tops, 4),
                              "scan_ms_per_step": round(scan_s * 1000.0, 3),
                              "peak_basis": peak_basis,
                              # products/s against the measured product rate
                              "montgomery_products_per_s": round(ops / OPS_PER_PRODUCT / scan_s, 0),
                              "frac_of_measured_product_rate": round(
                                  ops / OPS_PER_PRODUCT / scan_s / probe["montgomery_products_per_s"], 4),
                              "positions_per_s": round(ptimings[-1]["positions"] / scan_s, 0),
                              # the same time priced at one power per position
                              # (the scan without label reuse): the work the
                              # labels replaced counted as done
                              "frac_one_power_per_position": round(
                                  int32_ops(primes) / scan_s / 1e12 / peak_tops, 4),
                              "probe": probe}}

    # dominant kernel of the batched path: the sparse anchor filter (or the
    # dense scan when the whole workload falls below the sparse domain)
    if timing["sparse_primes"]:
        kern_ms = sum(t["sparse_filter_ms"] for t in timings) / len(timings)
        windows = timing["filter_windows"]
        ops = windows * OPS_PER_WINDOW_DWORD
        roofline = {"bound": "valu", "kernel": "sp_filter",
                    "achieved": round(ops / (kern_ms / 1000.0) / 1e12, 3),
                    "peak": round(peak_tops, 3), "unit": "TOP/s",
                    "frac": round(ops / (kern_ms / 1000.0) / 1e12 / peak_tops, 4),
                    "traffic": profiled_traffic() if args.workload == "top256" else None,
                    "avg_launch_ms": round(kern_ms, 3),
                    "algorithmic_ops_per_launch": ops,
                    "work": {"window_dword_compares": windows, "words": timing["filter_words"],
                             "completions": timing["filter_completions"],
                             "ops_per_window_dword_compare": OPS_PER_WINDOW_DWORD},
                    "stage_ms": {"prep": round(timing["sparse_prep_ms"], 3),
                                 "seed": round(timing["sparse_seed_ms"], 3),
                                 "filter": round(timing["sparse_filter_ms"], 3),
                                 "dense_fallback": round(timing["dense_scan_ms"], 3),
                                 # A4 point counts run on host threads beside
                                 # the device; a4_wait is the part the call
                                 # waited for after the device was done
                                 "a4_host": round(timing["a4_host_ms"], 3),
                                 "a4_wait": round(timing["a4_wait_ms"], 3),
                                 "call_total": round(timing["total_ms"], 3)},
                    "peak_basis": peak_basis, "probe": probe,
                    "memory_path": memory_path(timing, kern_ms, info)}
    else:
        ops = int32_ops(primes)
        kern_ms = sum(t["dense_scan_ms"] for t in timings) / len(timings)
        roofline = {"bound": "valu", "kernel": "scan_tiles",
                    "achieved": round(ops / (kern_ms / 1000.0) / 1e12, 3),
                    "peak": round(peak_tops, 3), "unit": "TOP/s",
                    "frac": round(ops / (kern_ms / 1000.0) / 1e12 / peak_tops, 4),
                    "traffic": None, "avg_launch_ms": round(kern_ms, 3),
                    "algorithmic_ops_per_launch": ops, "peak_basis": peak_basis}
    serving = None
    if args.serve_seconds > 0:
        from rabung_amd import serve, workloads
        serving = []
        for qps in [float(q) for q in args.qps.split(",") if q]:
            print("bench: serving at %.1f QPS" % qps, file=sys.stderr, flush=True)
            stream = workloads.serving_stream(seed=4, qps=qps, seconds=args.serve_seconds,
                                              intervals=True)
            with serve.Server(slots=2, chunk=args.serve_chunk, max_primes=args.serve_max) as srv:
                serve.run_stream(srv, workloads.serving_stream(seed=5, qps=qps, seconds=1.0,
                                                               intervals=True))
                srv.batches.clear()
                _, summ = serve.run_stream(srv, stream)
            serving.append({
                "workload": "Poisson arrivals at %.1f QPS for %.0f s (seed 4): 80%% single "
                            "primes (next prime >= uniform u32), 20%% interval queries of "
                            "U{16..4096} consecutive primes from log-uniform starts in "
                            "[2^16, 2^32), resolved by rank on the device; 2 slots, continuous "
                            "batching, pieces of %d primes, batches of at most %d primes"
                            % (qps, args.serve_seconds, args.serve_chunk, args.serve_max),
                "qps": qps,
                # work offered per second of arrivals; a drain well above zero
                # means the arrivals outran the server (saturation)
                "offered_profiles_per_s": round(summ["profiles"] / args.serve_seconds, 1),
                **{k: (round(v, 3) if isinstance(v, float) else v) for k, v in summ.items()}})
    batch64 = None
    if not args.no_batch64:
        print("bench: batch64 workload", file=sys.stderr, flush=True)
        batch64 = measure_batch64(max(1, args.steps // 10), 1)
    cpu = None if args.no_cpu else cpu_baseline(eng, primes)
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "profiles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic: deterministic prime block (%d primes %d..%d)" % (n, primes[0], primes[-1]),
        "config": {"workload": desc, "primes": n, "profiles": profiles, "autotuned": tuned,
                   "method": "batched: sparse anchor filter, one persistent launch, dense fallback",
                   "sparse_primes": timing["sparse_primes"], "fallbacks": timing["fallbacks"],
                   "spec_misses": timing["spec_misses"], "golden_checked": checked},
        "plain": plain,
        "batched_vs_plain": round(value / plain["profiles_per_s"], 2) if plain else None,
        "roofline": roofline,
        "batch64": batch64,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
