"""The allocation-free entry point (rbg_profile_block_noalloc) against the
device entry point it varies (rbg_profile_block_device), through the C ABI.

Both take device-resident primes and write device-resident records; the
records must be byte-identical.  Two calls in a row (the second with a
smaller block) check that the shared scratch array, which keeps the last
call's contents, is cleared where it must be; a block beyond the scratch is
refused with RBG_ECAPACITY before any work; graph mode is refused.
"""
import ctypes

import numpy as np
import pytest

import golden_data
import oracle
import rabung_amd

pytestmark = pytest.mark.gpu

RECORD = ctypes.sizeof(rabung_amd.Record)


def _hip():
    """The HIP runtime the library itself runs on (the libamdhip64 mapped
    into this process once librabung is loaded): the test allocates its
    input and output buffers with it, as a caller of the entry point would."""
    rabung_amd.load()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    assert paths, "libamdhip64 not mapped"
    hip = ctypes.CDLL(sorted(paths)[0])
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


class DeviceBuffer:
    """Caller-owned device memory (hipMalloc / hipMemcpy / hipFree)."""

    H2D, D2H = 1, 2

    def __init__(self, hip, nbytes):
        self.hip, self.nbytes = hip, max(nbytes, 1)
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0

    def write(self, data):
        assert len(data) <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, data, len(data), self.H2D) == 0

    def read(self, nbytes):
        out = ctypes.create_string_buffer(nbytes)
        assert self.hip.hipMemcpy(out, self.ptr, nbytes, self.D2H) == 0
        return out.raw

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _records(eng, primes, method, noalloc, graph=False):
    hip = _hip()
    n = len(primes)
    d_primes, d_out = DeviceBuffer(hip, 4 * n), DeviceBuffer(hip, RECORD * n)
    try:
        d_primes.write(np.asarray(primes, dtype=np.uint32).tobytes())
        d_out.write(b"\xa5" * (RECORD * n))  # stale bytes the call must overwrite
        if noalloc:
            eng.profile_block_noalloc(d_primes.ptr.value, n, d_out.ptr.value, method=method,
                                      graph=graph)
        else:
            eng.profile_block_device(d_primes.ptr.value, n, d_out.ptr.value, method=method)
        return d_out.read(RECORD * n)
    finally:
        d_primes.free()
        d_out.free()


def _top(n):
    rows, _, _ = golden_data.sample_block(49900)
    return [r["p"] for r in rows[-n:]]


@pytest.mark.parametrize("method", [rabung_amd.METHOD_BATCHED, rabung_amd.METHOD_DENSE,
                                    rabung_amd.METHOD_PLAIN])
def test_noalloc_matches_device_entry_point(method):
    # the dense workspace holds one summary per 16384 positions and order, so
    # its blocks are smaller primes than the sparse engine's
    if method == rabung_amd.METHOD_BATCHED:
        first, second = rabung_amd.primes_from((1 << 30) + 12345, 48) + [101, 65537, 2], _top(12)
    else:
        first = rabung_amd.primes_from((1 << 27) + 777, 48) + [101, 65537, 2]
        second = oracle.next_primes(1 << 24, 20)
    with rabung_amd.Engine() as eng:
        for primes in (first, second):  # the second block is smaller
            want = _records(eng, primes, method, noalloc=False)
            got = _records(eng, primes, method, noalloc=True)
            assert got == want, (method, len(primes))
            host = eng.profile_block(primes, method=method)
            arr = (rabung_amd.Record * len(primes)).from_buffer_copy(got)
            assert [rabung_amd.record_to_dict(arr[i]) for i in range(len(primes))] == host


def test_noalloc_refuses_beyond_scratch_and_graph():
    with rabung_amd.Engine() as eng:
        with pytest.raises(rabung_amd.RabungError) as e:
            _records(eng, _top(256), rabung_amd.METHOD_BATCHED, noalloc=True)
        assert e.value.status == 4 and "rbg_profile_block_noalloc" in str(e.value)
        with pytest.raises(rabung_amd.RabungError) as e:
            _records(eng, [101], rabung_amd.METHOD_BATCHED, noalloc=True, graph=True)
        assert e.value.status == 1
        # the context still works after the refusals
        small = [101, 1000003]
        assert _records(eng, small, rabung_amd.METHOD_BATCHED, noalloc=True) == \
            _records(eng, small, rabung_amd.METHOD_BATCHED, noalloc=False)
