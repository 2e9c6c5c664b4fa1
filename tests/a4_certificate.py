"""Independent check of an A4 audit certificate (test helper).

A certificate claims N = #E(F_p) for E: y^2 = x (x - 1)(x - 9) and derives
A4 from it.  It lists points of one curve E_s: y^2 = x (x - s)(x - 9 s)
(s = 1, or a quadratic non-residue for the twist, whose count is
2p + 2 - N) with their exact orders.  The lcm L of those orders divides
#E_s(F_p), which lies in the Hasse window [p + 1 - floor(2 sqrt p),
p + 1 + floor(2 sqrt p)]; when L has a single multiple there, that multiple
is #E_s(F_p).  A4 then follows from the character-sum identity
A4 = (p - 10 - Bd + N - p - 2) / 8 with
Bd = 4 chi(2) + 2 chi(6) + 2 chi(3) + 2 chi(-1) + 2 chi(-2)
(the reference's closed form, src/census/theorie/elliptic4_mestre.cpp:64-67).

Every step is recomputed here with Python integers: points on the curve,
n P = O and (n/q) P != O for each listed prime q | n (so n is the exact
order), the factor lists, the lcm, the window, N and A4.
"""
from math import gcd, isqrt


def _factors(n):
    out, q = [], 2
    while q * q <= n:
        if n % q == 0:
            out.append(q)
            while n % q == 0:
                n //= q
        q += 1 if q == 2 else 2
    if n > 1:
        out.append(n)
    return out


class _Curve:
    """y^2 = x^3 - 10 s x^2 + 9 s^2 x over F_p; None is the point at infinity."""

    def __init__(self, p, s):
        self.p, self.a2, self.a4 = p, (-10 * s) % p, (9 * s * s) % p

    def on(self, pt):
        x, y = pt
        return (y * y - (x * x * x + self.a2 * x * x + self.a4 * x)) % self.p == 0

    def add(self, u, v):
        p = self.p
        if u is None or v is None:
            return v if u is None else u
        (x1, y1), (x2, y2) = u, v
        if x1 == x2:
            if (y1 + y2) % p == 0:
                return None
            lam = (3 * x1 * x1 + 2 * self.a2 * x1 + self.a4) * pow(2 * y1, -1, p) % p
        else:
            lam = (y2 - y1) * pow(x2 - x1, -1, p) % p
        x3 = (lam * lam - self.a2 - x1 - x2) % p
        return x3, (lam * (x1 - x3) - y1) % p

    def times(self, pt, n):
        acc = None
        while n:
            if n & 1:
                acc = self.add(acc, pt)
            pt = self.add(pt, pt)
            n >>= 1
        return acc


def _chi(x, p):
    x %= p
    return 0 if x == 0 else (1 if pow(x, (p - 1) // 2, p) == 1 else -1)


def verify(cert):
    """Raise AssertionError unless the certificate proves its A4 value."""
    p, s = cert["p"], cert["twist_parameter"]
    assert p >= 1000 and 1 <= s < p
    assert s == 1 or _chi(s, p) == -1, "twist parameter is a residue"
    curve = _Curve(p, s)
    lcm = 1
    assert cert["certificate_points"], "no points"
    for w in cert["certificate_points"]:
        pt, n = (w["x"], w["y"]), w["order"]
        assert 0 <= pt[0] < p and 0 <= pt[1] < p and curve.on(pt), "point off the curve"
        assert w["prime_factors"] == _factors(n), "factor list"
        assert curve.times(pt, n) is None, "n P != O"
        assert all(curve.times(pt, n // q) is not None for q in w["prime_factors"]), "order not exact"
        lcm = lcm * n // gcd(lcm, n)
    assert lcm == cert["order_lcm"]
    r = isqrt(4 * p)
    lo, hi = p + 1 - r, p + 1 + r
    assert (cert["hasse_lower"], cert["hasse_upper"]) == (lo, hi)
    first = -(-lo // lcm) * lcm
    assert first <= hi < first + lcm, "lcm does not fix the count"
    n_points = first if s == 1 else 2 * p + 2 - first
    assert cert["elliptic_points"] == n_points
    bd = 4 * _chi(2, p) + 2 * _chi(6, p) + 2 * _chi(3, p) + 2 * _chi(-1, p) + 2 * _chi(-2, p)
    num = p - 10 - bd + n_points - p - 2
    assert num >= 0 and num % 8 == 0
    assert cert["quadruples"] == num // 8
    return True
