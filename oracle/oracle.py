"""ctypes loader for the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the bench's cpu_baseline leg may use this
module, and only as the checker.  The product library never imports it.
"""
import ctypes
import os
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


class OrcRecord(ctypes.Structure):
    _fields_ = [("p", ctypes.c_uint32), ("mask", ctypes.c_uint32),
                ("M", ctypes.c_uint32 * 8), ("L", ctypes.c_uint32 * 8),
                ("W", ctypes.c_uint32 * 8), ("K", ctypes.c_uint64 * 8),
                ("A4", ctypes.c_uint32), ("status", ctypes.c_uint32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError("oracle not built: run make -C oracle")
        _lib = ctypes.CDLL(str(LIB_PATH))
        _lib.orc_profile.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(OrcRecord)]
        _lib.orc_profile.restype = ctypes.c_int
        _lib.orc_profile_many.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_int,
                                          ctypes.c_int, ctypes.POINTER(OrcRecord)]
        _lib.orc_profile_many.restype = ctypes.c_int
        _lib.orc_is_prime.argtypes = [ctypes.c_uint64]
        _lib.orc_is_prime.restype = ctypes.c_int
        _lib.orc_mask.argtypes = [ctypes.c_uint64]
        _lib.orc_mask.restype = ctypes.c_uint32
        _lib.orc_next_primes.argtypes = [ctypes.c_uint64, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_uint64)]
        _lib.orc_next_primes.restype = ctypes.c_int
        _lib.orc_powmod.argtypes = [ctypes.c_uint64] * 3
        _lib.orc_powmod.restype = ctypes.c_uint64
    return _lib


def to_dict(rec):
    return {"p": rec.p, "mask": rec.mask, "M": list(rec.M), "L": list(rec.L),
            "W": list(rec.W), "K": list(rec.K), "A4": rec.A4}


def profile(p, threads=None):
    """Dense profile of one prime (multithreaded walk)."""
    if threads is None:
        threads = os.cpu_count() or 1
    rec = OrcRecord()
    rc = lib().orc_profile(p, threads, ctypes.byref(rec))
    if rc:
        raise ValueError("oracle refused p=%d (code %d)" % (p, rc))
    return to_dict(rec)


def profile_many(primes, threads=None):
    """Profiles of many (small) primes, one prime per thread at a time."""
    if threads is None:
        threads = os.cpu_count() or 1
    n = len(primes)
    arr = (ctypes.c_uint64 * n)(*primes)
    out = (OrcRecord * n)()
    rc = lib().orc_profile_many(arr, n, threads, out)
    if rc:
        raise ValueError("oracle refused a prime (code %d)" % rc)
    return [to_dict(out[i]) for i in range(n)]


def next_primes(start, n):
    out = (ctypes.c_uint64 * n)()
    k = lib().orc_next_primes(start, n, out)
    return list(out[:k])


def is_prime(n):
    return bool(lib().orc_is_prime(n))


def powmod(a, e, p):
    return lib().orc_powmod(a, e, p)


def rows(records):
    """Oracle records -> codec rows ({'p', 'profiles': [...]})."""
    out = []
    for rec in records:
        prof = []
        for r in range(2, 10):
            if rec["mask"] & (1 << (r - 2)):
                k = r - 2
                prof.append({"r": r, "max": rec["M"][k], "prefix": rec["L"][k],
                             "K": rec["K"][k], "witness": rec["W"][k]})
        out.append({"p": rec["p"], "profiles": prof})
    return out
