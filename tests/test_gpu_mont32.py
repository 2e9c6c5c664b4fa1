"""GPU known-answer test of the dense scan's 32-bit residue arithmetic at
the reference's self-check moduli (src/census/engine32/build/
original_scan_gpu.cu:550-574): the six moduli, 20,000 samples each drawn
with the same 64-bit LCG, exponents cycling through (p-1)/2, (p-1)/m and a
random 32-bit value, position offsets of up to 2047.

Truth: Python's exact integers (pow, modular inverse) and, for the prime
moduli, the Jacobi symbol by quadratic reciprocity (Euler's criterion ties
b^((p-1)/2) to it)."""
import math

import pytest

import rabung_amd

pytestmark = pytest.mark.gpu

MODULI = [3459826103, 3476732783, 4294967197, 4294967231, 4294967279, 4294967291]
SAMPLES = 20000


def jacobi(a, n):
    a %= n
    t = 1
    while a:
        while a % 2 == 0:
            a //= 2
            if n % 8 in (3, 5):
                t = -t
        a, n = n, a
        if a % 4 == 3 and n % 4 == 3:
            t = -t
        a %= n
    return t if n == 1 else 0


def samples(p):
    """(bases, exponents, offsets) of the reference's self-check."""
    m = 1
    for r in range(2, 10):
        if (p - 1) % r == 0:
            m = m * r // math.gcd(m, r)
    ve = (p - 1) // m
    state = 0x6a09e667f3bcc909
    mul, inc, mask = 6364136223846793005, 1442695040888963407, (1 << 64) - 1
    b, e, st = [], [], []
    for i in range(SAMPLES):
        state = (state * mul + inc) & mask
        b.append(1 + state % (p - 1))
        state = (state * mul + inc) & mask
        e.append((p - 1) // 2 if i % 4 == 0 else ve if i % 4 == 1 else state & 0xFFFFFFFF)
        st.append((state >> 48) & 2047)
    return b, e, st


@pytest.mark.parametrize("p", MODULI)
def test_mont32_power(p):
    b, e, _ = samples(p)
    got = rabung_amd.check_mont32(b, e, p, 1)
    assert got == [pow(x, y, p) for x, y in zip(b, e)]
    if rabung_amd.is_prime(p):  # Euler's criterion against the Jacobi symbol
        for i in range(0, SAMPLES, 4):
            j = jacobi(b[i], p)
            assert got[i] == (p - 1 if j < 0 else j)


@pytest.mark.parametrize("p", MODULI)
def test_mont32_product_and_recurrence(p):
    b, e, st = samples(p)
    c = [x % p for x in e]
    rinv = pow(1 << 32, -1, p)
    assert rabung_amd.check_mont32(b, c, p, 0) == [x * y * rinv % p for x, y in zip(b, c)]
    assert rabung_amd.check_mont32(b, st, p, 2) == [(x + s) * (1 << 32) % p
                                                   for x, s in zip(b, st)]
