"""Pack-frame codec: commitments, byte-identical re-encoding, corruption.

Mirrors the reference's tests/test_integrity.py:12-23 (sample pairs decode;
truncated, extended or bit-flipped frames raise) and the sample checks of
scripts/verify_atlas.py:18-45 (prime/science/values hashes, A4 range).
"""
import hashlib
import struct
import re

import pytest

import golden_data
from rabung_amd import codec


@pytest.mark.parametrize("block", [0, 49900])
def test_sample_commitments(block):
    rows, values, entry = golden_data.sample_block(block)
    assert len(rows) == entry["primes"] == len(values)
    primes = [r["p"] for r in rows]
    assert codec.prime_hash(primes).hex() == entry["prime_hash"]
    assert codec.science_hash(rows).hex() == entry["science_hash"]
    assert hashlib.sha256(codec.encode_a4_frame(
        primes, values, entry["lo"], entry["hi"], codec.science_hash(rows),
        entry["zlib_level_a4"])).hexdigest() == entry["a4_sha256"]
    for row, a4 in zip(rows, values):
        p = row["p"]
        assert a4 == 0 if p == 2 else 0 <= a4 <= (p - 3) // 2
        assert all(lo <= p < hi for lo, hi in [(entry["lo"], entry["hi"])])


@pytest.mark.parametrize("block", [0, 49900])
def test_reencode_byte_identical(block):
    rows, values, entry = golden_data.sample_block(block)
    frame = codec.encode_scan_frame(rows, entry["lo"], entry["hi"], True, entry["zlib_level_scan"])
    assert hashlib.sha256(frame).hexdigest() == entry["scan_sha256"]
    assert frame == (golden_data.GOLDEN / entry["scan_file"]).read_bytes()


@pytest.mark.parametrize("name,decoder", [("scan-00000.vdw", codec.decode_scan_frame),
                                          ("quadruplets-00000.a4", codec.decode_a4_frame)])
def test_damaged_frames_rejected(name, decoder):
    original = (golden_data.GOLDEN / name).read_bytes()
    for damaged in [original[:-1], original + b"0", bytes([original[0] ^ 1]) + original[1:],
                    original[:-10] + bytes([original[-10] ^ 0x40]) + original[-9:]]:
        with pytest.raises(ValueError):
            decoder(damaged)


def test_varint_canonical():
    for v in [0, 1, 127, 128, 300, 2 ** 32 - 1, 2 ** 63, 2 ** 64 - 1]:
        out = bytearray()
        codec.put_varint(out, v)
        assert codec.read_varint(bytes(out), 0) == (v, len(out))
    with pytest.raises(ValueError):
        codec.read_varint(b"\x80\x00", 0)  # non-canonical trailing zero
    with pytest.raises(ValueError):
        codec.read_varint(b"\x80", 0)  # truncated


def test_k_formula_worked_example():
    # p = 101, r = 2: (M, L, K, w) = (7, 1, 8, 19) (GPU_METHOD_EN.tex:300-314)
    assert codec.threshold_k(101, 2, 7, 1) == 8
    assert codec.admissible_mask(101) == 0b1101  # r = 2, 4, 5 (100 = 2^2 * 5^2)
    assert codec.admissible_mask(2) == 0


def test_chains_are_order_sensitive():
    rows0, v0, e0 = golden_data.sample_block(0)
    rows1, v1, e1 = golden_data.sample_block(49900)
    z = bytes(32)
    ph0 = codec.prime_hash([r["p"] for r in rows0])
    sh0 = codec.science_hash(rows0)
    ph1 = codec.prime_hash([r["p"] for r in rows1])
    sh1 = codec.science_hash(rows1)
    a = codec.chain_scan(codec.chain_scan(z, ph0, sh0, e0["lo"], e0["hi"]), ph1, sh1, e1["lo"], e1["hi"])
    b = codec.chain_scan(codec.chain_scan(z, ph1, sh1, e1["lo"], e1["hi"]), ph0, sh0, e0["lo"], e0["hi"])
    assert a != b and len(a) == 32


def _records(rows, values):
    recs = []
    for row, a4 in zip(rows, values):
        rec = {"p": row["p"], "mask": codec.admissible_mask(row["p"]), "M": [0] * 8,
               "L": [0] * 8, "W": [0] * 8, "K": [0] * 8, "A4": a4}
        for v in row["profiles"]:
            k = v["r"] - 2
            rec["M"][k], rec["L"][k], rec["W"][k], rec["K"][k] = (v["max"], v["prefix"],
                                                                  v["witness"], v["K"])
        recs.append(rec)
    return recs


@pytest.mark.parametrize("block", [0, 49900])
def test_library_encoder_byte_identical(block):
    """The C-ABI writer (rbg_encode_*_frame) reproduces the sample frames."""
    import rabung_amd
    rows, values, entry = golden_data.sample_block(block)
    recs = _records(rows, values)
    scan = rabung_amd.encode_scan_frame(recs, entry["lo"], entry["hi"], True,
                                        entry["zlib_level_scan"])
    assert scan == (golden_data.GOLDEN / entry["scan_file"]).read_bytes()
    a4 = rabung_amd.encode_a4_frame(recs, entry["lo"], entry["hi"], entry["zlib_level_a4"])
    assert hashlib.sha256(a4).hexdigest() == entry["a4_sha256"]
    assert rabung_amd.science_hash(recs).hex() == entry["science_hash"]
    nowit = rabung_amd.encode_scan_frame(recs, entry["lo"], entry["hi"], False, 6)
    assert nowit == codec.encode_scan_frame(rows, entry["lo"], entry["hi"], False, 6)


def test_library_chains_and_values_hash():
    """The C-ABI commitments of both sample blocks equal the manifest's prime,
    science and values hashes, and the C-ABI chain steps (rbg_chain_scan,
    rbg_chain_a4) equal the Python codec's over block 0 then block 49900
    and in the other order (which gives different chains)."""
    import rabung_amd
    z = bytes(32)
    steps = []
    for block in (0, 49900):
        rows, values, entry = golden_data.sample_block(block)
        recs = _records(rows, values)
        ph, sh = rabung_amd.prime_hash(recs), rabung_amd.science_hash(recs)
        vh = rabung_amd.values_hash(recs)
        assert (ph.hex(), sh.hex(), vh.hex()) == (entry["prime_hash"], entry["science_hash"],
                                                  entry["values_hash"])
        assert vh == hashlib.sha256(struct.pack("<%dI" % len(values), *values)).digest()
        steps.append((ph, sh, vh, entry["lo"], entry["hi"]))
    for order in (steps, steps[::-1]):
        sc_c = sc_py = ac_c = ac_py = z
        for ph, sh, vh, lo, hi in order:
            sc_c = rabung_amd.chain_scan(sc_c, ph, sh, lo, hi)
            sc_py = codec.chain_scan(sc_py, ph, sh, lo, hi)
            ac_c = rabung_amd.chain_a4(ac_c, ph, sh, vh, lo, hi)
            ac_py = codec.chain_a4(ac_py, ph, sh, vh, lo, hi)
        assert sc_c == sc_py and ac_c == ac_py and sc_c != ac_c
    forward = rabung_amd.chain_scan(rabung_amd.chain_scan(z, *steps[0][:2], *steps[0][3:]),
                                    *steps[1][:2], *steps[1][3:])
    backward = rabung_amd.chain_scan(rabung_amd.chain_scan(z, *steps[1][:2], *steps[1][3:]),
                                     *steps[0][:2], *steps[0][3:])
    assert forward != backward


def test_library_encoder_refuses_bad_records():
    import rabung_amd
    rows, values, entry = golden_data.sample_block(0)
    recs = _records(rows[:40], values[:40])
    lo, hi = entry["lo"], entry["hi"]
    bad = [dict(r, M=list(r["M"]), L=list(r["L"])) for r in recs]
    bad[5]["L"][0] = bad[5]["M"][0] + 1  # L > M
    for args in [(bad, lo, hi), (recs[::-1], lo, hi), (recs, recs[3]["p"], hi),
                 (recs, lo, recs[-1]["p"])]:
        with pytest.raises(rabung_amd.RabungError) as e:
            rabung_amd.encode_scan_frame(*args)
        assert e.value.status == 1
    worse = [dict(r) for r in recs]
    worse[7] = dict(worse[7], A4=worse[7]["p"])  # beyond (p-3)/2
    with pytest.raises(rabung_amd.RabungError):
        rabung_amd.encode_a4_frame(worse, lo, hi)


def _scan_frame_of(raw, lo, hi, count, body=None, raw_len=None):
    """A scan frame around an arbitrary payload (hash and lengths made to
    match unless overridden), for the decoder's record checks."""
    import zlib
    body = zlib.compress(raw, 6) if body is None else body
    header = codec.SCAN_HEADER.pack(codec.SCAN_MAGIC, len(body),
                                    len(raw) if raw_len is None else raw_len, lo, hi, count, 1,
                                    hashlib.sha256(raw).digest())
    return header + body


def _payload(*records):
    out = bytearray()
    for delta, mask, triples in records:
        codec.put_varint(out, delta)
        out.append(mask)
        for value in triples:
            codec.put_varint(out, value)
    return bytes(out)


def test_decoder_record_checks():
    """p = 101 (mask 0b1101: r = 2, 4, 5) then p = 103 (r = 2, 3, 6): each
    broken field is reported, the intact payload decodes."""
    good = [(101 - 99, 0b1101, [7, 1, 19, 3, 1, 10, 3, 1, 7]), (2, 0x13, [5, 1, 9, 4, 2, 3, 2, 1, 7])]
    rows, meta = codec.decode_scan_frame(_scan_frame_of(_payload(*good), 100, 110, 2))
    assert [r["p"] for r in rows] == [101, 103] and meta["count"] == 2
    assert rows[0]["profiles"][0] == {"r": 2, "max": 7, "prefix": 1, "K": 8, "witness": 19}
    cases = {
        "repeats the previous prime": [good[0], (0, 0x13, [5, 1, 9, 4, 2, 3, 2, 1, 7])],
        "outside [100, 110)": [good[0], (10, 0x13, [5, 1, 9, 4, 2, 3, 2, 1, 7])],
        "stores mask": [(2, 0b1100, [7, 1, 19, 3, 1, 10, 3, 1, 7])],
        "outside 1 <= L <= M": [(2, 0b1101, [7, 8, 19, 3, 1, 10, 3, 1, 7])],
        "witness 0 outside": [(2, 0b1101, [7, 1, 0, 3, 1, 10, 3, 1, 7])],
        "witness 51 outside": [(2, 0b1101, [7, 1, 51, 3, 1, 10, 3, 1, 7])],
    }
    for text, recs in cases.items():
        with pytest.raises(codec.FrameError, match=re.escape(text)):
            codec.decode_scan_frame(_scan_frame_of(_payload(*recs), 100, 110, len(recs)))
    raw = _payload(*good)
    with pytest.raises(codec.FrameError, match="payload bytes after record"):
        codec.decode_scan_frame(_scan_frame_of(raw + b"\x01", 100, 110, 2))
    with pytest.raises(codec.FrameError, match="ends inside record 1"):
        codec.decode_scan_frame(_scan_frame_of(raw[:-2], 100, 110, 2))


def test_decoder_stream_checks():
    import zlib
    raw = _payload((2, 0b1101, [7, 1, 19, 3, 1, 10, 3, 1, 7]))
    body = zlib.compress(raw, 6)
    with pytest.raises(codec.FrameError, match="follow the compressed stream"):
        codec.decode_scan_frame(_scan_frame_of(raw, 100, 110, 1, body=body + b"\x00\x00"))
    with pytest.raises(codec.FrameError, match="cut short|corrupt"):
        codec.decode_scan_frame(_scan_frame_of(raw, 100, 110, 1, body=body[:-5]))
    with pytest.raises(codec.FrameError, match="length other than the declared"):
        codec.decode_scan_frame(_scan_frame_of(raw, 100, 110, 1, raw_len=len(raw) - 1))
    with pytest.raises(codec.FrameError, match="length other than the declared"):
        codec.decode_scan_frame(_scan_frame_of(raw, 100, 110, 1, raw_len=len(raw) + 1))
    with pytest.raises(codec.FrameError, match="exceeds"):
        codec.decode_scan_frame(_scan_frame_of(raw, 100, 110, 1, raw_len=codec.MAX_FRAME + 1))
    frame = bytearray(_scan_frame_of(raw, 100, 110, 1))
    frame[-1] ^= 0x01  # adler-32 trailer of the stream
    with pytest.raises(codec.FrameError):
        codec.decode_scan_frame(bytes(frame))
    values = [0, 3, 5]
    a4 = codec.encode_a4_frame([2, 3, 5], values, 2, 7, bytes(32), 6)
    assert codec.decode_a4_frame(a4)[0] == values
    with pytest.raises(codec.FrameError, match="flags"):
        codec.decode_a4_frame(a4[:36] + b"\x01" + a4[37:])
