"""Pack-record codec for census profiles (scan frames and A4 frames).

Written from the layout specification in the reference's docs/DATA_FORMAT.md
(lines 9-47 for frames, 63-77 for chains); the reference's own decoder is
scripts/atlas_codec.py:20-60.  The byte layout must be identical to the
reference's so that frames written here decode there and vice versa, which is
pinned by re-encoding the reference sample blocks byte for byte.

Scan frame: header struct '<8sQQQQII32s' =
    (b'VDWSM1\\0\\0', compressed_len, raw_len, lo, hi, count, witness_flag,
     sha256(raw)), then zlib(raw).
raw: per prime, varint(delta) (first delta relative to lo-1), one mask byte,
     then for each admissible r = 2..9 ascending varints M, L[, witness].
A4 frame: header struct '<8sQQIIII32s32s32s' =
    (b'VDWA4V1\\0', lo, hi, n, compressed_len, raw_len, flags=0, prime_hash,
     profile_hash, values_hash), then zlib(n x uint32 LE).
"""
import hashlib
import struct
import zlib

SCAN_MAGIC = b"VDWSM1\0\0"
SCAN_HEADER = struct.Struct("<8sQQQQII32s")
A4_MAGIC = b"VDWA4V1\0"
A4_HEADER = struct.Struct("<8sQQIIII32s32s32s")
MAX_FRAME = 128 * 1024 ** 2
MAX_A4_PRIMES = 4096


class FrameError(ValueError):
    """A frame failed a structural, hash or range check."""


def admissible_mask(p):
    """Bit r-2 set iff r | p-1, for r in 2..9 (p=2 has mask 0)."""
    if p < 3:
        return 0
    return sum(1 << (r - 2) for r in range(2, 10) if (p - 1) % r == 0)


def threshold_k(p, r, m, l):
    """Rabung threshold: max(M+1, 2L+2) if r | (p-1)/2 else max(M+1, L+2)."""
    return max(m + 1, 2 * l + 2 if ((p - 1) // 2) % r == 0 else l + 2)


def put_varint(out, value):
    if value < 0 or value >= 1 << 64:
        raise FrameError("varint out of range")
    while True:
        byte = value & 0x7F
        value >>= 7
        if value:
            out.append(byte | 0x80)
        else:
            out.append(byte)
            return


def get_varint(raw, index):
    value = 0
    for shift_index in range(10):
        if index >= len(raw):
            raise FrameError("truncated integer")
        byte = raw[index]
        index += 1
        value |= (byte & 0x7F) << (7 * shift_index)
        if byte < 0x80:
            if value >= 1 << 64 or (shift_index and byte == 0):
                raise FrameError("noncanonical integer")
            return value, index
    raise FrameError("oversized integer")


def encode_scan_raw(records, lo, witnesses=True):
    """records: iterable of dicts {'p', 'profiles': [{'r','max','prefix','witness'}...]}"""
    out = bytearray()
    previous = lo - 1
    for rec in records:
        p = rec["p"]
        delta = p - previous
        if delta <= 0:
            raise FrameError("record ordering")
        previous = p
        mask = admissible_mask(p)
        put_varint(out, delta)
        out.append(mask)
        by_r = {v["r"]: v for v in rec["profiles"]}
        if sorted(by_r) != [r for r in range(2, 10) if mask & (1 << (r - 2))]:
            raise FrameError("profiles do not match the admissible mask")
        for r in sorted(by_r):
            v = by_r[r]
            put_varint(out, v["max"])
            put_varint(out, v["prefix"])
            if witnesses:
                put_varint(out, v["witness"])
    return bytes(out)


def encode_scan_frame(records, lo, hi, witnesses=True, level=9):
    records = list(records)
    raw = encode_scan_raw(records, lo, witnesses)
    comp = zlib.compress(raw, level)
    header = SCAN_HEADER.pack(SCAN_MAGIC, len(comp), len(raw), lo, hi, len(records),
                              1 if witnesses else 0, hashlib.sha256(raw).digest())
    return header + comp


def _inflate(payload, size):
    d = zlib.decompressobj()
    raw = d.decompress(payload, size + 1)
    if len(raw) != size or not d.eof or d.unused_data:
        raise FrameError("frame hash/length")
    return raw


def decode_scan_frame(frame):
    """Returns (rows, meta); rows are {'p', 'profiles': [{'r','max','prefix','K'[, 'witness']}]}."""
    if len(frame) < SCAN_HEADER.size:
        raise FrameError("truncated header")
    magic, n, size, lo, hi, count, flags, digest = SCAN_HEADER.unpack(frame[:SCAN_HEADER.size])
    if magic != SCAN_MAGIC or flags not in (0, 1) or n != len(frame) - SCAN_HEADER.size:
        raise FrameError("frame header")
    if size > MAX_FRAME or n > MAX_FRAME:
        raise FrameError("oversized frame")
    try:
        raw = _inflate(frame[SCAN_HEADER.size:], size)
    except zlib.error as exc:
        raise FrameError("frame payload: %s" % exc) from None
    if hashlib.sha256(raw).digest() != digest:
        raise FrameError("frame hash/length")
    rows = []
    i = 0
    p = lo - 1
    for _ in range(count):
        delta, i = get_varint(raw, i)
        p += delta
        if delta == 0 or not lo <= p < hi or i >= len(raw):
            raise FrameError("record ordering")
        mask = raw[i]
        i += 1
        if mask != admissible_mask(p):
            raise FrameError("decoded mask")
        profiles = []
        for r in range(2, 10):
            if mask & (1 << (r - 2)):
                m, i = get_varint(raw, i)
                l, i = get_varint(raw, i)
                v = {"r": r, "max": m, "prefix": l, "K": threshold_k(p, r, m, l)}
                if not 1 <= l <= m <= (p - 1) // r:
                    raise FrameError("decoded profile bound")
                if flags:
                    v["witness"], i = get_varint(raw, i)
                    if not 1 <= v["witness"] <= (p - 1) // 2:
                        raise FrameError("decoded witness")
                profiles.append(v)
        rows.append({"p": p, "profiles": profiles})
    if i != len(raw):
        raise FrameError("trailing record bytes")
    meta = {"lo": lo, "hi": hi, "count": count, "witnesses": bool(flags),
            "raw_sha256": digest.hex(), "raw_bytes": size, "compressed_bytes": n}
    return rows, meta


def prime_hash(primes):
    h = hashlib.sha256(b"vdw-primes-u64le")
    for p in primes:
        h.update(struct.pack("<Q", p))
    return h.digest()


def science_hash(rows):
    """Profile commitment over (p, mask, M, L, K); witnesses are not part of it."""
    h = hashlib.sha256(b"vdw-p-profile-v1-u64le")
    for row in rows:
        p = row["p"]
        h.update(struct.pack("<QB", p, admissible_mask(p)))
        for v in row["profiles"]:
            h.update(struct.pack("<QQQ", v["max"], v["prefix"],
                                 v.get("K", threshold_k(p, v["r"], v["max"], v["prefix"]))))
    return h.digest()


def encode_a4_frame(primes, values, lo, hi, profile_hash, level=9):
    values = list(values)
    if len(values) > MAX_A4_PRIMES or len(values) != len(primes):
        raise FrameError("A4 count")
    raw = struct.pack("<%dI" % len(values), *values)
    comp = zlib.compress(raw, level)
    header = A4_HEADER.pack(A4_MAGIC, lo, hi, len(values), len(comp), len(raw), 0,
                            prime_hash(primes), profile_hash, hashlib.sha256(raw).digest())
    return header + comp


def decode_a4_frame(frame):
    if len(frame) < A4_HEADER.size:
        raise FrameError("A4 short frame")
    magic, lo, hi, n, csize, rsize, flags, ph, sh, digest = A4_HEADER.unpack(frame[:A4_HEADER.size])
    if (magic != A4_MAGIC or n > MAX_A4_PRIMES or rsize != 4 * n or flags
            or csize != len(frame) - A4_HEADER.size):
        raise FrameError("A4 header")
    try:
        raw = _inflate(frame[A4_HEADER.size:], rsize)
    except zlib.error as exc:
        raise FrameError("A4 payload: %s" % exc) from None
    if hashlib.sha256(raw).digest() != digest:
        raise FrameError("A4 frame hash")
    values = list(struct.unpack("<%dI" % n, raw))
    meta = {"lo": lo, "hi": hi, "count": n, "prime_hash": ph.hex(),
            "profile_science_hash": sh.hex(), "values_sha256": digest.hex()}
    return values, meta


def chain_scan(previous, prime_digest, profile_digest, lo, hi):
    return hashlib.sha256(previous + prime_digest + profile_digest
                          + struct.pack("<QQ", lo, hi)).digest()


def chain_a4(previous, prime_digest, profile_digest, values_digest, lo, hi):
    return hashlib.sha256(previous + prime_digest + profile_digest + values_digest
                          + struct.pack("<QQ", lo, hi)).digest()


def records_to_rows(records):
    """Convert library records (dicts from the ctypes shim) to frame rows."""
    rows = []
    for rec in records:
        p = rec["p"]
        profiles = []
        for r in range(2, 10):
            if rec["mask"] & (1 << (r - 2)):
                k = r - 2
                profiles.append({"r": r, "max": rec["M"][k], "prefix": rec["L"][k],
                                 "K": rec["K"][k], "witness": rec["W"][k]})
        rows.append({"p": p, "profiles": profiles})
    return rows
