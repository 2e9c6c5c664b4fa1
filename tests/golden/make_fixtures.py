#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ from the reference's data files.

Run in the development container only (the reference tree is not present on
GPU machines).  Nothing here imports or runs reference code: the sample frames
are decoded with this repository's own codec, which must reproduce every
manifest commitment (frame SHA256, prime_hash, science_hash, values_hash), and
the frames are then re-encoded by this repository's encoder, which must give
the identical bytes.  The JSON measurement files are read as plain JSON.

Outputs:
  sample-<block>.vdw / quadruplets-<block>.a4   frames re-encoded here (bytes
                                                equal to the reference's)
  sample_manifest.json                          the commitments they must meet
  holdout32.json.gz                             profiles of 4096 primes in four
                                                windows up to 4.237e9
  night32.json.gz                               151 primes 1.16e8 .. 4294967291
  holdout33.json.gz                             profiles of 4096 primes in four
                                                windows 5.389e9 .. 8.317e9 (wide
                                                arithmetic path, p < 2^33)
  night33.json.gz                               16 primes 4.4e9 .. 8.6e9
"""
import gzip
import hashlib
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1] / "rabung-atlas32_amd"))
from rabung_amd import codec  # noqa: E402

REF = Path("/root/reference")


def compact_rows(rows):
    """rows -> {'p': [...], 'prof': [[r, M, L, K, W], ...] per prime}"""
    return [[row["p"], [[v["r"], v["max"], v["prefix"], v["K"], v["witness"]]
                        for v in row["profiles"]]] for row in rows]


def main():
    sample = REF / "data/sample"
    manifest = json.loads((sample / "manifest.json").read_text())
    out_manifest = {"blocks": []}
    for block in manifest["blocks"]:
        frames = {}
        for item in block["frames"]:
            data = (sample / item["file"]).read_bytes()
            assert len(data) == item["bytes"] and hashlib.sha256(data).hexdigest() == item["sha256"]
            frames[item["file"]] = data
        scan_name = [f for f in frames if f.endswith(".vdw")][0]
        a4_name = [f for f in frames if f.endswith(".a4")][0]
        rows, meta = codec.decode_scan_frame(frames[scan_name])
        values, ameta = codec.decode_a4_frame(frames[a4_name])
        primes = [r["p"] for r in rows]
        assert len(rows) == block["primes"] == len(values)
        assert codec.prime_hash(primes).hex() == block["prime_hash"] == ameta["prime_hash"]
        assert codec.science_hash(rows).hex() == block["science_hash"] == ameta["profile_science_hash"]
        assert ameta["values_sha256"] == block["values_hash"]
        # re-encode with this repository's encoder and find the zlib level that
        # reproduces the reference bytes
        level_scan = level_a4 = None
        for level in range(10):
            if codec.encode_scan_frame(rows, block["lo"], block["hi"], True, level) == frames[scan_name]:
                level_scan = level
                break
        for level in range(10):
            enc = codec.encode_a4_frame(primes, values, block["lo"], block["hi"],
                                        codec.science_hash(rows), level)
            if enc == frames[a4_name]:
                level_a4 = level
                break
        assert level_scan is not None and level_a4 is not None, "re-encoding differs"
        scan_bytes = codec.encode_scan_frame(rows, block["lo"], block["hi"], True, level_scan)
        a4_bytes = codec.encode_a4_frame(primes, values, block["lo"], block["hi"],
                                         codec.science_hash(rows), level_a4)
        (HERE / scan_name).write_bytes(scan_bytes)
        (HERE / a4_name).write_bytes(a4_bytes)
        out_manifest["blocks"].append({
            "block": block["block"], "lo": block["lo"], "hi": block["hi"], "primes": block["primes"],
            "prime_hash": block["prime_hash"], "science_hash": block["science_hash"],
            "values_hash": block["values_hash"], "scan_file": scan_name, "a4_file": a4_name,
            "scan_sha256": hashlib.sha256(scan_bytes).hexdigest(),
            "a4_sha256": hashlib.sha256(a4_bytes).hexdigest(),
            "zlib_level_scan": level_scan, "zlib_level_a4": level_a4})
        print(block["block"], "ok", len(rows), "levels", level_scan, level_a4)
    (HERE / "sample_manifest.json").write_text(json.dumps(out_manifest, indent=1) + "\n")

    gm = REF / "reproducibility/gpu_method"
    hold = json.loads((gm / "throughput/holdout-32-mr_selected.json").read_text())
    windows = []
    for w in hold["windows"]:
        rows = w["profiles_full"]
        for row in rows:  # sanity: K consistent with the contract
            for v in row["profiles"]:
                assert v["K"] == codec.threshold_k(row["p"], v["r"], v["max"], v["prefix"])
        windows.append({"lo": w["lo"], "hi": w["hi"], "rows": compact_rows(rows)})
    with gzip.open(HERE / "holdout32.json.gz", "wt") as f:
        json.dump({"source": "reproducibility/gpu_method/throughput/holdout-32-mr_selected.json",
                   "windows": windows}, f, separators=(",", ":"))
    night = json.loads((gm / "night/MESURES_FINALES_32_BITS.json").read_text())
    rows = []
    seen = set()
    for cohort in night["cohorts"].values():
        for row in cohort["per_prime"]:
            if row["p"] not in seen:
                seen.add(row["p"])
                rows.append(row)
    rows.sort(key=lambda r: r["p"])
    with gzip.open(HERE / "night32.json.gz", "wt") as f:
        json.dump({"source": "reproducibility/gpu_method/night/MESURES_FINALES_32_BITS.json",
                   "rows": compact_rows(rows)}, f, separators=(",", ":"))
    print("holdout", sum(len(w["rows"]) for w in windows), "night", len(rows))

    # wide path: the final 33-bit cohort (the paper's throughput windows
    # 6.023e9 and 8.041e9) and the earlier 33-bit windows 5.389e9, 8.317e9
    windows33 = []
    for name in ("holdout-33-selected_selected_binary48.json", "holdout-33-mr_selected.json"):
        hold = json.loads((gm / "throughput" / name).read_text())
        assert hold["bits"] == 33
        for w in hold["windows"]:
            rows = w["profiles_full"]
            for row in rows:
                assert (1 << 32) < row["p"] < (1 << 33)
                for v in row["profiles"]:
                    assert v["K"] == codec.threshold_k(row["p"], v["r"], v["max"], v["prefix"])
            windows33.append({"lo": w["lo"], "hi": w["hi"], "source": name,
                              "rows": compact_rows(rows)})
    windows33.sort(key=lambda w: w["lo"])
    with gzip.open(HERE / "holdout33.json.gz", "wt") as f:
        json.dump({"source": "reproducibility/gpu_method/throughput/holdout-33-*.json",
                   "windows": windows33}, f, separators=(",", ":"))
    night = json.loads((gm / "night/MESURES_FINALES_33_BITS.json").read_text())
    rows = sorted(night["per_prime"], key=lambda r: r["p"])
    for row in rows:
        for v in row["profiles"]:
            assert v["K"] == codec.threshold_k(row["p"], v["r"], v["max"], v["prefix"])
    with gzip.open(HERE / "night33.json.gz", "wt") as f:
        json.dump({"source": "reproducibility/gpu_method/night/MESURES_FINALES_33_BITS.json",
                   "rows": compact_rows(rows)}, f, separators=(",", ":"))
    print("holdout33", sum(len(w["rows"]) for w in windows33), "night33", len(rows))


if __name__ == "__main__":
    main()
