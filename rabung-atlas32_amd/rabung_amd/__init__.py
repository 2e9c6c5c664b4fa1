"""ctypes shim over librabung.so, the MI355X census-worker library.

The library is the product path; this module only marshals arguments.  It
refuses to run without the compiled HIP library (no CPU fallback).

Python-side mirror of the reference's callers:
  * Engine.profile_block(primes) -> records with M, L, W, K, A4 per order
    (the worker's per-batch loop, src/census/worker32.cu:20-41, and the moment
    worker's A4 values, src/census/moment_worker.cpp:35-41);
  * baseline_profile(p, mask, full) -> (M by r, prefix by r)
    (src/census/engine32/src/baseline_bridge.cu:21-27);
  * codec: pack frames (scripts/atlas_codec.py layout).
"""
import ctypes
import os
from pathlib import Path

from . import codec  # noqa: F401

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = Path(os.environ.get("RBG_LIBRARY", PKG_DIR / "librabung.so"))

METHOD_BATCHED = 0x0  # sparse anchor filter, one persistent launch, dense fallback
METHOD_PLAIN = 0x1    # dense scan, one launch per prime
METHOD_DENSE = 0x2    # dense scan, one persistent launch
CERTIFY = 0x10
GRAPH = 0x20
MAX_BLOCK = 4096

STATUS = {0: "ok", 1: "invalid argument", 2: "device error", 3: "verification failed",
          4: "capacity"}


class Record(ctypes.Structure):
    _fields_ = [("p", ctypes.c_uint32), ("mask", ctypes.c_uint32),
                ("M", ctypes.c_uint32 * 8), ("L", ctypes.c_uint32 * 8),
                ("W", ctypes.c_uint32 * 8), ("K", ctypes.c_uint64 * 8),
                ("A4", ctypes.c_uint32), ("status", ctypes.c_uint32)]


class RecordWide(ctypes.Structure):
    """rbg_record_wide: the record of the wider domain (p < 2^33)."""
    _fields_ = [("p", ctypes.c_uint64), ("mask", ctypes.c_uint32),
                ("M", ctypes.c_uint32 * 8), ("L", ctypes.c_uint32 * 8),
                ("W", ctypes.c_uint32 * 8), ("K", ctypes.c_uint64 * 8),
                ("A4", ctypes.c_uint32), ("status", ctypes.c_uint32)]


WIDE_LIMIT = 1 << 33  # primes up to here take the wide entry points


class Timing(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("dense_scan_ms", ctypes.c_double),
                ("sparse_prep_ms", ctypes.c_double), ("sparse_seed_ms", ctypes.c_double),
                ("sparse_filter_ms", ctypes.c_double), ("dense_primes", ctypes.c_uint64),
                ("sparse_primes", ctypes.c_uint64), ("sparse_profiles", ctypes.c_uint64),
                ("fallbacks", ctypes.c_uint64), ("positions", ctypes.c_uint64),
                ("products", ctypes.c_uint64), ("launches", ctypes.c_uint64),
                ("filter_windows", ctypes.c_uint64), ("filter_words", ctypes.c_uint64),
                ("filter_completions", ctypes.c_uint64), ("a4_host_ms", ctypes.c_double),
                ("a4_wait_ms", ctypes.c_double), ("rescans", ctypes.c_uint64),
                ("filter_loads", ctypes.c_uint64), ("spec_misses", ctypes.c_uint64)]


class Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("tile_chunks", ctypes.c_uint32),
                ("blocks_per_cu", ctypes.c_uint32), ("hbm_budget", ctypes.c_uint64),
                ("xcd_group", ctypes.c_uint32), ("cover_q", ctypes.c_uint32),
                ("extend_limit", ctypes.c_uint32), ("floor_boost", ctypes.c_uint32),
                ("cover_div", ctypes.c_uint32), ("filter_chunk", ctypes.c_uint32),
                ("spec_off", ctypes.c_uint32), ("spec_boost", ctypes.c_uint32)]


EXPORTS = ["rbg_open", "rbg_profile_block", "rbg_profile_block_device",
           "rbg_profile_block_noalloc", "rbg_profile_batch",
           "rbg_estimate_bytes", "rbg_autotune", "rbg_last_error",
           "rbg_last_timing", "rbg_device_info", "rbg_close", "rbg_mask", "rbg_is_prime",
           "rbg_version", "rbg_a4_host", "rbg_primes_from", "rbg_encode_scan_frame",
           "rbg_encode_a4_frame", "rbg_science_hash", "rbg_prime_hash", "rbg_values_hash",
           "rbg_chain_scan", "rbg_chain_a4", "rbg_a4_certificate", "rbg_microbench",
           "rbg_profile_block_wide", "rbg_profile_batch_wide", "rbg_estimate_bytes_wide",
           "rbg_a4_host_wide", "rbg_is_prime_wide", "rbg_mask_wide", "rbg_primes_from_wide",
           "rbg_check_wide", "rbg_certify_records", "rbg_prime_table", "rbg_prime_rank",
           "rbg_primes_by_rank", "rbg_profile_ranked", "rbg_check_mont32",
           "baseline_init",
           "baseline_profile"]


class RabungError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("%s (status %d: %s)" % (message, status, STATUS.get(status, "?")))
        self.status = status


_lib = None


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError("librabung.so is not built (run rabung-atlas32_amd/build.py); "
                           "there is no CPU fallback")
    lib = ctypes.CDLL(str(LIB_PATH))
    P = ctypes.POINTER
    lib.rbg_open.argtypes = [P(Config)]
    lib.rbg_open.restype = ctypes.c_void_p
    lib.rbg_profile_block.argtypes = [ctypes.c_void_p, P(ctypes.c_uint32), ctypes.c_uint32,
                                      ctypes.c_uint32, P(Record)]
    lib.rbg_profile_block.restype = ctypes.c_int
    lib.rbg_profile_block_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                             ctypes.c_uint32, ctypes.c_void_p]
    lib.rbg_profile_block_device.restype = ctypes.c_int
    lib.rbg_profile_block_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                              ctypes.c_uint32, ctypes.c_void_p]
    lib.rbg_profile_block_noalloc.restype = ctypes.c_int
    lib.rbg_profile_batch.argtypes = [ctypes.c_void_p, P(ctypes.c_uint32), ctypes.c_uint32,
                                      ctypes.c_uint32, P(Record)]
    lib.rbg_profile_batch.restype = ctypes.c_int
    lib.rbg_estimate_bytes.argtypes = [ctypes.c_void_p, P(ctypes.c_uint32), ctypes.c_uint32,
                                       ctypes.c_uint32, P(ctypes.c_uint64)]
    lib.rbg_estimate_bytes.restype = ctypes.c_int
    lib.rbg_autotune.argtypes = [ctypes.c_void_p, P(ctypes.c_uint32), ctypes.c_uint32,
                                 ctypes.c_uint32, P(Config)]
    lib.rbg_autotune.restype = ctypes.c_int
    lib.rbg_last_error.argtypes = [ctypes.c_void_p]
    lib.rbg_last_error.restype = ctypes.c_char_p
    lib.rbg_last_timing.argtypes = [ctypes.c_void_p, P(Timing)]
    lib.rbg_last_timing.restype = ctypes.c_int
    lib.rbg_device_info.argtypes = [ctypes.c_void_p, P(ctypes.c_uint32), P(ctypes.c_uint32),
                                    P(ctypes.c_uint64)]
    lib.rbg_device_info.restype = ctypes.c_int
    lib.rbg_close.argtypes = [ctypes.c_void_p]
    lib.rbg_close.restype = None
    lib.rbg_mask.argtypes = [ctypes.c_uint32]
    lib.rbg_mask.restype = ctypes.c_uint32
    lib.rbg_is_prime.argtypes = [ctypes.c_uint32]
    lib.rbg_is_prime.restype = ctypes.c_int
    lib.rbg_version.argtypes = []
    lib.rbg_version.restype = ctypes.c_char_p
    lib.rbg_primes_from.argtypes = [ctypes.c_uint64, ctypes.c_uint32, P(ctypes.c_uint32)]
    lib.rbg_primes_from.restype = ctypes.c_uint32
    lib.rbg_a4_host.argtypes = [P(ctypes.c_uint32), ctypes.c_uint32, P(ctypes.c_uint32),
                                ctypes.c_uint32]
    lib.rbg_a4_host.restype = ctypes.c_int
    lib.baseline_init.argtypes = [ctypes.c_uint]
    lib.baseline_init.restype = None
    lib.baseline_profile.argtypes = [ctypes.c_uint, ctypes.c_uint, P(ctypes.c_uint),
                                     P(ctypes.c_uint), ctypes.c_int]
    lib.baseline_profile.restype = ctypes.c_int
    for name in ("rbg_encode_scan_frame", "rbg_encode_a4_frame"):
        fn = getattr(lib, name)
        fn.argtypes = [P(Record), ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64]
        fn.argtypes += [ctypes.c_int] if name == "rbg_encode_scan_frame" else []
        fn.argtypes += [ctypes.c_int, ctypes.c_void_p, P(ctypes.c_size_t)]
        fn.restype = ctypes.c_int
    for name in ("rbg_science_hash", "rbg_prime_hash", "rbg_values_hash"):
        getattr(lib, name).argtypes = [P(Record), ctypes.c_uint32, ctypes.c_char_p]
        getattr(lib, name).restype = ctypes.c_int
    lib.rbg_chain_scan.argtypes = [ctypes.c_char_p] * 3 + [ctypes.c_uint64] * 2 + \
        [ctypes.c_char_p]
    lib.rbg_chain_scan.restype = ctypes.c_int
    lib.rbg_chain_a4.argtypes = [ctypes.c_char_p] * 4 + [ctypes.c_uint64] * 2 + \
        [ctypes.c_char_p]
    lib.rbg_chain_a4.restype = ctypes.c_int
    lib.rbg_microbench.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32,
                                   P(ctypes.c_double)]
    lib.rbg_microbench.restype = ctypes.c_int
    for name in ("rbg_profile_block_wide", "rbg_profile_batch_wide"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, P(ctypes.c_uint64), ctypes.c_uint32,
                                       ctypes.c_uint32, P(RecordWide)]
        getattr(lib, name).restype = ctypes.c_int
    lib.rbg_estimate_bytes_wide.argtypes = [ctypes.c_void_p, P(ctypes.c_uint64), ctypes.c_uint32,
                                            ctypes.c_uint32, P(ctypes.c_uint64)]
    lib.rbg_estimate_bytes_wide.restype = ctypes.c_int
    lib.rbg_a4_host_wide.argtypes = [P(ctypes.c_uint64), ctypes.c_uint32, P(ctypes.c_uint32),
                                     ctypes.c_uint32]
    lib.rbg_a4_host_wide.restype = ctypes.c_int
    lib.rbg_is_prime_wide.argtypes = [ctypes.c_uint64]
    lib.rbg_is_prime_wide.restype = ctypes.c_int
    lib.rbg_mask_wide.argtypes = [ctypes.c_uint64]
    lib.rbg_mask_wide.restype = ctypes.c_uint32
    lib.rbg_primes_from_wide.argtypes = [ctypes.c_uint64, ctypes.c_uint32, P(ctypes.c_uint64)]
    lib.rbg_primes_from_wide.restype = ctypes.c_uint32
    lib.rbg_check_wide.argtypes = [ctypes.c_int, P(ctypes.c_uint64), P(ctypes.c_uint64),
                                   ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                   P(ctypes.c_uint64)]
    lib.rbg_check_wide.restype = ctypes.c_int
    lib.rbg_check_mont32.argtypes = [ctypes.c_int, P(ctypes.c_uint32), P(ctypes.c_uint32),
                                     ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                     P(ctypes.c_uint32)]
    lib.rbg_check_mont32.restype = ctypes.c_int
    lib.rbg_a4_certificate.argtypes = [ctypes.c_uint32, P(ctypes.c_uint32), P(ctypes.c_uint64),
                                       ctypes.c_void_p, P(ctypes.c_size_t)]
    lib.rbg_a4_certificate.restype = ctypes.c_int
    lib.rbg_certify_records.argtypes = [P(Record), ctypes.c_uint32, ctypes.c_uint32,
                                        P(ctypes.c_uint32)]
    lib.rbg_certify_records.restype = ctypes.c_int
    lib.rbg_prime_table.argtypes = [ctypes.c_void_p, P(ctypes.c_uint64)]
    lib.rbg_prime_rank.argtypes = [ctypes.c_void_p, ctypes.c_uint64, P(ctypes.c_uint64)]
    lib.rbg_primes_by_rank.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                       P(ctypes.c_uint32), P(ctypes.c_uint32)]
    lib.rbg_profile_ranked.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                       ctypes.c_uint32, P(Record), P(ctypes.c_uint32)]
    for name in ("rbg_prime_table", "rbg_prime_rank", "rbg_primes_by_rank",
                 "rbg_profile_ranked"):
        getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def record_to_dict(rec):
    return {"p": rec.p, "mask": rec.mask, "M": list(rec.M), "L": list(rec.L),
            "W": list(rec.W), "K": list(rec.K), "A4": rec.A4}


def dict_to_record(d):
    rec = Record()
    rec.p, rec.mask, rec.A4 = d["p"], d["mask"], d.get("A4", 0)
    for k in range(8):
        rec.M[k], rec.L[k], rec.W[k], rec.K[k] = d["M"][k], d["L"][k], d["W"][k], d["K"][k]
    return rec


def _records_array(records):
    arr = (Record * max(1, len(records)))()
    for i, d in enumerate(records):
        arr[i] = d if isinstance(d, Record) else dict_to_record(d)
    return arr


def _frame(call, *args):
    size = ctypes.c_size_t(0)
    rc = call(*args, None, ctypes.byref(size))
    if rc not in (0, 4):
        raise RabungError(rc, "frame refused")
    buf = ctypes.create_string_buffer(size.value)
    rc = call(*args, buf, ctypes.byref(size))
    if rc:
        raise RabungError(rc, "frame refused")
    return buf.raw[:size.value]


def encode_scan_frame(records, lo, hi, witnesses=True, level=3):
    """Scan frame bytes of library records (dicts or Record), written by the
    library's own encoder (rbg_encode_scan_frame)."""
    lib = load()
    return _frame(lib.rbg_encode_scan_frame, _records_array(records), len(records), lo, hi,
                  1 if witnesses else 0, level)


def encode_a4_frame(records, lo, hi, level=3):
    lib = load()
    return _frame(lib.rbg_encode_a4_frame, _records_array(records), len(records), lo, hi, level)


def science_hash(records):
    out = ctypes.create_string_buffer(32)
    rc = load().rbg_science_hash(_records_array(records), len(records), out)
    if rc:
        raise RabungError(rc, "science hash")
    return out.raw


def prime_hash(records):
    out = ctypes.create_string_buffer(32)
    rc = load().rbg_prime_hash(_records_array(records), len(records), out)
    if rc:
        raise RabungError(rc, "prime hash")
    return out.raw


def values_hash(records):
    out = ctypes.create_string_buffer(32)
    rc = load().rbg_values_hash(_records_array(records), len(records), out)
    if rc:
        raise RabungError(rc, "values hash")
    return out.raw


def chain_scan(previous, prime_digest, profile_digest, lo, hi):
    """One step of the scan chain through the C ABI (rbg_chain_scan)."""
    out = ctypes.create_string_buffer(32)
    rc = load().rbg_chain_scan(bytes(previous), bytes(prime_digest), bytes(profile_digest),
                               lo, hi, out)
    if rc:
        raise RabungError(rc, "scan chain")
    return out.raw


def chain_a4(previous, prime_digest, profile_digest, values_digest, lo, hi):
    """One step of the A4 chain through the C ABI (rbg_chain_a4)."""
    out = ctypes.create_string_buffer(32)
    rc = load().rbg_chain_a4(bytes(previous), bytes(prime_digest), bytes(profile_digest),
                             bytes(values_digest), lo, hi, out)
    if rc:
        raise RabungError(rc, "A4 chain")
    return out.raw


class Engine:
    """One library context (device, stream, workspaces)."""

    def __init__(self, device=0, tile_chunks=0, blocks_per_cu=0, hbm_budget=0, xcd_group=0,
                 cover_q=0, extend_limit=0, floor_boost=0, cover_div=0, filter_chunk=0,
                 spec_off=0, spec_boost=0):
        lib = load()
        cfg = Config(device, tile_chunks, blocks_per_cu, hbm_budget, xcd_group, cover_q,
                     extend_limit, floor_boost, cover_div, filter_chunk, spec_off, spec_boost)
        self._ctx = lib.rbg_open(ctypes.byref(cfg))
        if not self._ctx:
            raise RabungError(2, "rbg_open failed (no HIP device %d?)" % device)
        self._lib = lib

    def close(self):
        if self._ctx:
            self._lib.rbg_close(self._ctx)
            self._ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _profile(self, name, primes, method, certify, graph):
        out = self._profile_raw(name, primes, method, certify, graph)
        return [record_to_dict(out[i]) for i in range(len(primes))]

    def _profile_raw(self, name, primes, method, certify, graph):
        primes = [int(p) for p in primes]
        n = len(primes)
        flags = method | (CERTIFY if certify else 0) | (GRAPH if graph else 0)
        # primes above 2^32 take the wide entry points (rbg_record_wide)
        wide = any(p >= (1 << 32) for p in primes)
        if wide:
            arr = (ctypes.c_uint64 * max(1, n))(*primes)
            out = (RecordWide * max(1, n))()
            rc = getattr(self._lib, name + "_wide")(self._ctx, arr, n, flags, out)
        else:
            arr = (ctypes.c_uint32 * max(1, n))(*primes)
            out = (Record * max(1, n))()
            rc = getattr(self._lib, name)(self._ctx, arr, n, flags, out)
        if rc:
            raise RabungError(rc, self._lib.rbg_last_error(self._ctx).decode())
        return out

    def profile_block(self, primes, method=METHOD_BATCHED, certify=False, graph=False):
        """Profiles of up to 4096 primes below 2^33; returns a list of dicts."""
        return self._profile("rbg_profile_block", primes, method, certify, graph)

    def profile_batch(self, primes, method=METHOD_BATCHED, certify=False, graph=False):
        """Profiles of up to 2^20 primes (many blocks), split by HBM budget."""
        return self._profile("rbg_profile_batch", primes, method, certify, graph)

    def profile_batch_raw(self, primes, method=METHOD_BATCHED, certify=False, graph=False):
        """profile_batch returning the C record array itself (ctypes Record,
        or RecordWide when a prime exceeds 2^32), without building one dict
        per prime: the form for callers that hand the records on to C."""
        return self._profile_raw("rbg_profile_batch", primes, method, certify, graph)

    def autotune(self, primes, method=METHOD_BATCHED):
        """Pick the fastest launch shape on these sample primes; returns it."""
        primes = [int(p) for p in primes]
        arr = (ctypes.c_uint32 * max(1, len(primes)))(*primes)
        cfg = Config()
        rc = self._lib.rbg_autotune(self._ctx, arr, len(primes), method, ctypes.byref(cfg))
        if rc:
            raise RabungError(rc, self._lib.rbg_last_error(self._ctx).decode())
        return {"tile_chunks": cfg.tile_chunks, "blocks_per_cu": cfg.blocks_per_cu,
                "xcd_group": cfg.xcd_group}

    def estimate_bytes(self, primes, method=METHOD_BATCHED):
        primes = [int(p) for p in primes]
        out = ctypes.c_uint64()
        if any(p >= (1 << 32) for p in primes):
            arr = (ctypes.c_uint64 * max(1, len(primes)))(*primes)
            rc = self._lib.rbg_estimate_bytes_wide(self._ctx, arr, len(primes), method,
                                                   ctypes.byref(out))
        else:
            arr = (ctypes.c_uint32 * max(1, len(primes)))(*primes)
            rc = self._lib.rbg_estimate_bytes(self._ctx, arr, len(primes), method,
                                              ctypes.byref(out))
        if rc:
            raise RabungError(rc, "estimate failed")
        return out.value

    def profile_block_device(self, d_primes_ptr, n, d_out_ptr, method=METHOD_BATCHED):
        """Device-resident variant: pointers to n uint32 primes and n records."""
        rc = self._lib.rbg_profile_block_device(self._ctx, ctypes.c_void_p(d_primes_ptr), n,
                                                method, ctypes.c_void_p(d_out_ptr))
        if rc:
            raise RabungError(rc, self._lib.rbg_last_error(self._ctx).decode())

    def profile_block_noalloc(self, d_primes_ptr, n, d_out_ptr, method=METHOD_BATCHED,
                              graph=False):
        """rbg_profile_block_noalloc: the device-resident call with every
        workspace carved from the library's static scratch array."""
        flags = method | (GRAPH if graph else 0)
        rc = self._lib.rbg_profile_block_noalloc(self._ctx, ctypes.c_void_p(d_primes_ptr), n,
                                                 flags, ctypes.c_void_p(d_out_ptr))
        if rc:
            raise RabungError(rc, self._lib.rbg_last_error(self._ctx).decode())

    def _check(self, rc):
        if rc:
            raise RabungError(rc, self._lib.rbg_last_error(self._ctx).decode())

    def prime_count(self):
        """Number of primes below 2^32, from the device prime table (built on
        first use: about 258 MiB of HBM kept until the process ends)."""
        out = ctypes.c_uint64()
        self._check(self._lib.rbg_prime_table(self._ctx, ctypes.byref(out)))
        return out.value

    def prime_rank(self, x):
        """Number of primes below x (x <= 2^32), resolved on the device."""
        out = ctypes.c_uint64()
        self._check(self._lib.rbg_prime_rank(self._ctx, int(x), ctypes.byref(out)))
        return out.value

    def primes_by_rank(self, first, n):
        """Primes of ranks first .. first+n-1 (rank 0 is 2) below 2^32."""
        out = (ctypes.c_uint32 * max(1, n))()
        found = ctypes.c_uint32()
        self._check(self._lib.rbg_primes_by_rank(self._ctx, int(first), n, out,
                                                 ctypes.byref(found)))
        return list(out[:found.value])

    def profile_ranked_raw(self, first, n, method=METHOD_BATCHED, certify=False, graph=False):
        """Records of the primes of ranks first .. first+n-1 (n <= 4096),
        selected on the device; returns (Record array, count)."""
        flags = method | (CERTIFY if certify else 0) | (GRAPH if graph else 0)
        out = (Record * max(1, n))()
        found = ctypes.c_uint32()
        self._check(self._lib.rbg_profile_ranked(self._ctx, int(first), n, flags, out,
                                                 ctypes.byref(found)))
        return out, found.value

    def profile_ranked(self, first, n, method=METHOD_BATCHED, certify=False, graph=False):
        out, k = self.profile_ranked_raw(first, n, method, certify, graph)
        return [record_to_dict(out[i]) for i in range(k)]

    def device_info(self):
        cus = ctypes.c_uint32()
        clk = ctypes.c_uint32()
        hbm = ctypes.c_uint64()
        self._lib.rbg_device_info(self._ctx, ctypes.byref(cus), ctypes.byref(clk), ctypes.byref(hbm))
        return {"cus": cus.value, "clock_khz": clk.value, "hbm_bytes": hbm.value}

    def last_timing(self):
        t = Timing()
        self._lib.rbg_last_timing(self._ctx, ctypes.byref(t))
        return {name: getattr(t, name) for name, _ in Timing._fields_}


RECORD_SIZE = ctypes.sizeof(Record)


def mask(p):
    p = int(p)
    return load().rbg_mask(p) if p < (1 << 32) else load().rbg_mask_wide(p)


def is_prime(p):
    """Deterministic primality below 2^33 (False from 2^33 on)."""
    p = int(p)
    if p < 0:
        return False
    return bool(load().rbg_is_prime(p) if p < (1 << 32) else load().rbg_is_prime_wide(p))


def primes_from_wide(start, count):
    """Up to `count` consecutive primes >= start below 2^33."""
    out = (ctypes.c_uint64 * max(1, count))()
    k = load().rbg_primes_from_wide(start, count, out)
    return list(out[:k])


def check_wide(a, b, p, kind, device=0):
    """Device results of the wide residue arithmetic (kind 0: Montgomery
    product a b 2^-32 mod p; kind 1: Montgomery power of a to b)."""
    n = len(a)
    A = (ctypes.c_uint64 * n)(*a)
    Bv = (ctypes.c_uint64 * n)(*b)
    out = (ctypes.c_uint64 * n)()
    rc = load().rbg_check_wide(device, A, Bv, n, p, kind, out)
    if rc:
        raise RabungError(rc, "rbg_check_wide")
    return list(out)


def check_mont32(a, b, p, kind, device=0):
    """Device results of the dense scan's 32-bit residue arithmetic (kind 0:
    Montgomery product a b 2^-32 mod p; kind 1: a^b mod p through the
    Montgomery domain; kind 2: Montgomery form of a + b by the position
    recurrence)."""
    n = len(a)
    A = (ctypes.c_uint32 * n)(*a)
    Bv = (ctypes.c_uint32 * n)(*b)
    out = (ctypes.c_uint32 * n)()
    rc = load().rbg_check_mont32(device, A, Bv, n, p, kind, out)
    if rc:
        raise RabungError(rc, "rbg_check_mont32")
    return list(out)


def primes_from(start, count):
    """Up to `count` consecutive primes >= start below 2^32."""
    out = (ctypes.c_uint32 * max(1, count))()
    k = load().rbg_primes_from(start, count, out)
    return list(out[:k])


def a4_host(primes, threads=0):
    """A4 values through the host elliptic point count (no GPU needed)."""
    lib = load()
    primes = [int(p) for p in primes]
    n = len(primes)
    out = (ctypes.c_uint32 * max(1, n))()
    if any(p >= (1 << 32) for p in primes):
        arr = (ctypes.c_uint64 * max(1, n))(*primes)
        rc = lib.rbg_a4_host_wide(arr, n, out, threads)
    else:
        arr = (ctypes.c_uint32 * max(1, n))(*primes)
        rc = lib.rbg_a4_host(arr, n, out, threads)
    if rc:
        raise RabungError(rc, "A4 point count failed")
    return list(out[:n])


def a4_certificate(p):
    """(A4, audit certificate dict) of one prime below 2^32 by the host point
    count (rbg_a4_certificate)."""
    import json
    lib = load()
    a4 = ctypes.c_uint32()
    attempts = ctypes.c_uint64()
    size = ctypes.c_size_t(0)
    rc = lib.rbg_a4_certificate(p, ctypes.byref(a4), ctypes.byref(attempts), None,
                                ctypes.byref(size))
    if rc not in (0, 4):
        raise RabungError(rc, "A4 certificate")
    buf = ctypes.create_string_buffer(size.value)
    rc = lib.rbg_a4_certificate(p, ctypes.byref(a4), ctypes.byref(attempts), buf,
                                ctypes.byref(size))
    if rc:
        raise RabungError(rc, "A4 certificate")
    return a4.value, json.loads(buf.raw[:size.value].decode())


def certify_records(records, threads=0):
    """Host-only direct verification of library records (rbg_certify_records):
    returns None when all pass, else the index of the first failing record."""
    arr = _records_array(records)
    bad = ctypes.c_uint32(0)
    rc = load().rbg_certify_records(arr, len(records), threads, ctypes.byref(bad))
    if rc == 0:
        return None
    if rc == 3:
        return bad.value
    raise RabungError(rc, "certify_records")


def baseline_profile(p, mask_bits=0, full=True, fill=0):
    """Reference bridge: returns (status, M by r[10], prefix by r[10]); the
    arrays start as `fill` (entries the call does not write keep it)."""
    lib = load()
    out = (ctypes.c_uint * 10)(*([fill] * 10))
    pre = (ctypes.c_uint * 10)(*([fill] * 10))
    rc = lib.baseline_profile(p, mask_bits, out, pre, 1 if full else 0)
    return rc, list(out), list(pre)


def rows(records):
    """Library records -> codec rows."""
    return codec.records_to_rows(records)


def microbench(kind, param=0, device=0):
    """One measurement probe of rbg_microbench (see include/rabung.h)."""
    out = ctypes.c_double(0)
    rc = load().rbg_microbench(device, kind, param, ctypes.byref(out))
    if rc:
        raise RabungError(rc, "microbench")
    return out.value
