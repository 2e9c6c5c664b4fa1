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


def _device_block(primes):
    import torch
    host = np.asarray(primes, dtype=np.uint32).view(np.int32)
    d_primes = torch.from_numpy(host.copy()).to("cuda:0")
    return d_primes


def _records(eng, primes, method, noalloc):
    import torch
    d_primes = _device_block(primes)
    d_out = torch.full((len(primes) * RECORD,), 0xA5, dtype=torch.uint8, device="cuda:0")
    torch.cuda.synchronize()
    call = eng.profile_block_noalloc if noalloc else eng.profile_block_device
    call(d_primes.data_ptr(), len(primes), d_out.data_ptr(), method=method)
    return d_out.cpu().numpy().tobytes()


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
        import torch
        d_primes = _device_block([101])
        d_out = torch.zeros(RECORD, dtype=torch.uint8, device="cuda:0")
        with pytest.raises(rabung_amd.RabungError) as e:
            eng.profile_block_noalloc(d_primes.data_ptr(), 1, d_out.data_ptr(), graph=True)
        assert e.value.status == 1
        # the context still works after the refusals
        small = [101, 1000003]
        assert _records(eng, small, rabung_amd.METHOD_BATCHED, noalloc=True) == \
            _records(eng, small, rabung_amd.METHOD_BATCHED, noalloc=False)
