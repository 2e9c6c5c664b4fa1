"""Pack-record codec for census profiles (scan frames and A4 frames).

Written from the layout specification in the reference's docs/DATA_FORMAT.md
(lines 9-47 for frames, 63-77 for chains); frames written here must be the
reference's bytes (scripts/atlas_codec.py:20-60 reads the same layout), which
is pinned by re-encoding the reference sample blocks byte for byte.

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


# ---- decoding -------------------------------------------------------------
#
# A decoder trusts nothing in a frame: the header fields are checked against
# the bytes that follow, the payload is inflated to exactly its declared
# length and must hash to the header's digest, and every record must satisfy
# the output contract (ascending primes inside [lo, hi), the admissibility
# mask of its prime, 1 <= L <= M <= (p-1)/r, 1 <= W <= (p-1)/2).  Every
# failure raises FrameError naming the frame kind and what was wrong.


class PayloadReader:
    """Forward-only cursor over an inflated scan payload."""

    __slots__ = ("data", "offset")

    def __init__(self, data):
        self.data = data
        self.offset = 0

    def at_end(self):
        return self.offset == len(self.data)

    def byte(self, context):
        if self.offset >= len(self.data):
            raise FrameError("scan payload ends inside %s" % context)
        value = self.data[self.offset]
        self.offset += 1
        return value

    def uint(self, context):
        """One unsigned base-128 integer, least significant group first, in
        its shortest form and below 2^64."""
        first = self.offset
        value = 0
        for group in range(10):
            byte = self.byte(context)
            value |= (byte & 0x7F) << (7 * group)
            if not byte & 0x80:
                if value >> 64:
                    raise FrameError("%s at payload byte %d exceeds 64 bits" % (context, first))
                if group and not byte:
                    raise FrameError("%s at payload byte %d has a redundant zero group"
                                     % (context, first))
                return value
        raise FrameError("%s at payload byte %d runs past ten bytes" % (context, first))


def read_varint(data, offset=0):
    """(value, next offset) of the varint at `offset` of `data`."""
    reader = PayloadReader(data)
    reader.offset = offset
    value = reader.uint("integer")
    return value, reader.offset


def _split(frame, layout, magic, kind):
    """Header fields and compressed body of a frame whose first field is its
    magic."""
    if len(frame) < layout.size:
        raise FrameError("%s frame: %d bytes, shorter than its %d-byte header"
                         % (kind, len(frame), layout.size))
    fields = layout.unpack_from(frame)
    if fields[0] != magic:
        raise FrameError("%s frame: magic %r instead of %r" % (kind, fields[0], magic))
    return fields, bytes(frame[layout.size:])


def _inflate_exact(body, raw_len, digest, kind):
    """The payload of `body`, which must inflate to exactly raw_len bytes
    with the SHA-256 `digest`, using the whole compressed stream."""
    if raw_len > MAX_FRAME:
        raise FrameError("%s frame: declared payload of %d bytes exceeds %d"
                         % (kind, raw_len, MAX_FRAME))
    inflater = zlib.decompressobj()
    try:
        # output is capped at the declared length; input left over at the cap
        # is fed once more with room for one byte, which either ends the
        # stream or proves it longer than declared
        data = inflater.decompress(body, max(raw_len, 1))
        if not inflater.eof and inflater.unconsumed_tail:
            data += inflater.decompress(inflater.unconsumed_tail, 1)
    except zlib.error as exc:
        raise FrameError("%s frame: compressed payload is corrupt (%s)" % (kind, exc)) from None
    if len(data) != raw_len or inflater.unconsumed_tail:
        raise FrameError("%s frame: payload inflates to a length other than the declared %d"
                         % (kind, raw_len))
    if not inflater.eof:
        raise FrameError("%s frame: compressed stream is cut short" % kind)
    if inflater.unused_data:
        raise FrameError("%s frame: %d bytes follow the compressed stream"
                         % (kind, len(inflater.unused_data)))
    if hashlib.sha256(data).digest() != digest:
        raise FrameError("%s frame: payload SHA-256 differs from the header" % kind)
    return data


def decode_scan_frame(frame):
    """Rows {'p', 'profiles': [{'r', 'max', 'prefix', 'K'[, 'witness']}]}
    and the frame's metadata, from one scan frame (DATA_FORMAT.md:9-31)."""
    (_magic, comp_len, raw_len, lo, hi, count, witness_flag,
     digest), body = _split(frame, SCAN_HEADER, SCAN_MAGIC, "scan")
    if witness_flag > 1:
        raise FrameError("scan frame: witness flag %d is neither 0 nor 1" % witness_flag)
    if comp_len != len(body):
        raise FrameError("scan frame: header declares %d compressed bytes, %d follow"
                         % (comp_len, len(body)))
    reader = PayloadReader(_inflate_exact(body, raw_len, digest, "scan"))
    rows = []
    prime = lo - 1
    for index in range(count):
        where = "record %d" % index
        step = reader.uint(where + " (prime delta)")
        if not step:
            raise FrameError("scan frame: %s repeats the previous prime" % where)
        prime += step
        if prime >= hi:
            raise FrameError("scan frame: %s has p=%d outside [%d, %d)" % (where, prime, lo, hi))
        mask = reader.byte(where + " (mask)")
        if mask != admissible_mask(prime):
            raise FrameError("scan frame: %s stores mask %#04x, p=%d has %#04x"
                             % (where, mask, prime, admissible_mask(prime)))
        profiles = []
        for r in (r for r in range(2, 10) if mask >> (r - 2) & 1):
            longest = reader.uint("%s r=%d (M)" % (where, r))
            prefix = reader.uint("%s r=%d (L)" % (where, r))
            if not 1 <= prefix <= longest <= (prime - 1) // r:
                raise FrameError("scan frame: %s p=%d r=%d has M=%d, L=%d outside "
                                 "1 <= L <= M <= (p-1)/r" % (where, prime, r, longest, prefix))
            entry = {"r": r, "max": longest, "prefix": prefix,
                     "K": threshold_k(prime, r, longest, prefix)}
            if witness_flag:
                start = reader.uint("%s r=%d (witness)" % (where, r))
                if not 1 <= start <= (prime - 1) // 2:
                    raise FrameError("scan frame: %s p=%d r=%d has witness %d outside "
                                     "1..(p-1)/2" % (where, prime, r, start))
                entry["witness"] = start
            profiles.append(entry)
        rows.append({"p": prime, "profiles": profiles})
    if not reader.at_end():
        raise FrameError("scan frame: %d payload bytes after record %d"
                         % (len(reader.data) - reader.offset, count - 1))
    meta = {"lo": lo, "hi": hi, "count": count, "witnesses": witness_flag == 1,
            "raw_sha256": digest.hex(), "raw_bytes": raw_len, "compressed_bytes": comp_len}
    return rows, meta


def decode_a4_frame(frame):
    """A4 values (uint32 per prime) and the frame's metadata, from one A4
    frame (DATA_FORMAT.md:35-47)."""
    (_magic, lo, hi, count, comp_len, raw_len, flags, primes_digest, profile_digest,
     digest), body = _split(frame, A4_HEADER, A4_MAGIC, "A4")
    if flags:
        raise FrameError("A4 frame: flags %#x, the format defines only 0" % flags)
    if count > MAX_A4_PRIMES:
        raise FrameError("A4 frame: %d primes, a block holds at most %d" % (count, MAX_A4_PRIMES))
    if raw_len != 4 * count:
        raise FrameError("A4 frame: %d payload bytes declared for %d primes" % (raw_len, count))
    if comp_len != len(body):
        raise FrameError("A4 frame: header declares %d compressed bytes, %d follow"
                         % (comp_len, len(body)))
    data = _inflate_exact(body, raw_len, digest, "A4")
    values = [v for (v,) in struct.iter_unpack("<I", data)]
    meta = {"lo": lo, "hi": hi, "count": count, "prime_hash": primes_digest.hex(),
            "profile_science_hash": profile_digest.hex(), "values_sha256": digest.hex()}
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
