"""Loaders for the committed golden fixtures (tests/golden/)."""
import gzip
import json
from pathlib import Path

from rabung_amd import codec

GOLDEN = Path(__file__).resolve().parent / "golden"


def sample_block(block):
    """(rows, a4_values, manifest entry) of a reference sample block (0 or 49900)."""
    man = json.loads((GOLDEN / "sample_manifest.json").read_text())
    entry = [b for b in man["blocks"] if b["block"] == block][0]
    rows, meta = codec.decode_scan_frame((GOLDEN / entry["scan_file"]).read_bytes())
    values, ameta = codec.decode_a4_frame((GOLDEN / entry["a4_file"]).read_bytes())
    return rows, values, entry


def _expand(compact):
    out = []
    for p, prof in compact:
        out.append({"p": p, "profiles": [{"r": r, "max": m, "prefix": l, "K": k, "witness": w}
                                         for r, m, l, k, w in prof]})
    return out


def holdout_windows():
    with gzip.open(GOLDEN / "holdout32.json.gz", "rt") as f:
        data = json.load(f)
    return [{"lo": w["lo"], "hi": w["hi"], "rows": _expand(w["rows"])} for w in data["windows"]]


def night_rows():
    with gzip.open(GOLDEN / "night32.json.gz", "rt") as f:
        data = json.load(f)
    return _expand(data["rows"])


def holdout33_windows():
    """Four windows of 1024 consecutive primes between 2^32 and 2^33."""
    with gzip.open(GOLDEN / "holdout33.json.gz", "rt") as f:
        data = json.load(f)
    return [{"lo": w["lo"], "hi": w["hi"], "source": w["source"], "rows": _expand(w["rows"])}
            for w in data["windows"]]


def night33_rows():
    with gzip.open(GOLDEN / "night33.json.gz", "rt") as f:
        data = json.load(f)
    return _expand(data["rows"])


def block2_24():
    """(rows, a4, document) of the 4096 consecutive primes from the first
    prime >= 2^24, written by tests/golden/make_block2_24.py (oracle values,
    cross-checked there against the reference's moment_worker for every A4
    and independent_profile on a seeded subset)."""
    with gzip.open(GOLDEN / "block2_24.json.gz", "rt") as f:
        doc = json.load(f)
    rows, a4 = [], []
    for p, _mask, prof, value in doc["rows"]:
        rows.append({"p": p, "profiles": [{"r": r, "max": m, "prefix": l, "K": k, "witness": w}
                                          for r, m, l, k, w in prof]})
        a4.append(value)
    return rows, a4, doc
