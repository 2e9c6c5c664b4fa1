"""The C-ABI library loads and exports every symbol its header declares.

No compute call is made here (CPU-only); GPU behaviour is in the gpu tests.
"""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

import rabung_amd

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "rabung.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "defined")))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for n in ["rbg_open", "rbg_profile_block", "rbg_close", "rbg_last_error",
              "baseline_init", "baseline_profile"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    if not rabung_amd.LIB_PATH.exists():
        pytest.fail("librabung.so not built")
    lib = ctypes.CDLL(str(rabung_amd.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(rabung_amd.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(declared_functions()) <= exported


def test_shim_matches_exports():
    assert set(rabung_amd.EXPORTS) == set(declared_functions())


def test_pure_host_entry_points():
    lib = rabung_amd.load()
    assert rabung_amd.mask(101) == 0b1101
    assert rabung_amd.mask(2) == 0
    assert rabung_amd.is_prime(4294967291) and not rabung_amd.is_prime(4294967295)
    assert not rabung_amd.is_prime(1) and rabung_amd.is_prime(2)
    # strong pseudoprimes to small bases must be rejected
    for n in [2047, 3215031751, 25326001, 4759123141 % (1 << 32)]:
        assert rabung_amd.is_prime(n) == all(n % q for q in range(2, int(n ** 0.5) + 1))
    assert b"gfx950" in lib.rbg_version()


def test_record_layout():
    # 2 + 8*3 u32, 8 u64, 2 u32 with natural alignment
    assert rabung_amd.RECORD_SIZE == 4 * 26 + 8 * 8 + 4 * 2
