"""Plain-integer model of the sparse anchor search (test helper, small p).

Follows the same cover, fold, band schedule, wheel masks, thresholds and
completion as the HIP kernels in rabung-atlas32_amd/csrc/rbg_sparse.hip, with
colours read one by one instead of through 64-bit planes.  Used to check the
algorithmic design against the oracle on the CPU.
"""
from math import gcd, isqrt


def cover(p, Q=48):
    H = Q - 1
    d = isqrt(p // H)
    while d > 1 and H * d * d > 4 * p:
        d -= 1
    if d < 1:
        return None
    b = (4 * p + H * d * d) // (4 * d) + 1
    if b >= p or d >= p or b > (1 << 20) or H * d > 2 * b:
        return None
    if 4 * b * d - H * d * d <= 4 * p:
        return None
    if 2 * b + 2 >= p - 1:
        return None
    return b, d


def labels_of(p, m):
    """label j(y) = ind(y) mod m relative to some element of order m."""
    e = (p - 1) // m
    # element of exact order m
    for a in range(2, p):
        z = pow(a, e, p)
        if all(pow(z, m // q, p) != 1 for q in (2, 3, 5, 7) if m % q == 0):
            break
    root = {}
    cur = 1
    for j in range(m):
        root[cur] = j
        cur = cur * z % p
    return lambda y: root[pow(y, e, p)]


def band_ranges(d, t, sp, phi, bmin, B):
    """Scheduled anchors as the reference states them: for each side of the
    band (kind 0: c in [1, (d-1)/2], n = phi mod s; kind 1: c in
    [d/2 + 1, d-1], n = t-2-phi mod s) the ranges [n d + c_lo, n d + c_hi]
    clipped to [bmin, B] (GPU_METHOD_EN.tex:380-455; d <= 2: one range)."""
    if d <= 2:
        return [(max(bmin, t if d == 1 else t - 1), B)]
    out = []
    ca, cb = (d - 1) // 2, d // 2 + 1
    for kind in (0, 1):
        c_lo, c_hi = (cb, d - 1) if kind else (1, ca)
        if c_lo > c_hi:
            continue
        res = (t - 2 - phi) % sp if kind else phi
        n_lo, n_hi = bmin // d, B // d
        n = n_lo + (res + sp - n_lo % sp) % sp
        while n <= n_hi:
            out.append((max(n * d + c_lo, bmin), min(n * d + c_hi, B)))
            n += sp
    return out


def span_ranges(d, t, sp, phi, bmin, B):
    """The same anchors as the kernel enumerates them (spans_of): the two
    sides of a band are adjacent, because t-2-phi = phi (t even) or phi - 1
    (t odd) mod s, so they form spans [m d + s_lo, m d + s_lo + width - 1]
    for m = rm mod s with one unscheduled point inside (c = d/2 for even t
    and even d, c = 0 for odd t).  Returns (lo, hi, excluded or None)."""
    if d <= 2:
        return [(max(bmin, t if d == 1 else t - 1), B, None)]
    odd_t = t % 2 == 1
    rm = (phi + sp - 1) % sp if odd_t else phi
    s_lo = d // 2 + 1 if odd_t else 1
    width = d + (d - 1) // 2 - d // 2 if odd_t else d - 1
    excl = d - s_lo if odd_t else (d // 2 - 1 if d % 2 == 0 else None)
    x = bmin + 1 - s_lo - width
    m = 0 if x <= 0 else (x + d - 1) // d
    m += (rm + sp - m % sp) % sp
    out = []
    while m * d + s_lo <= B:
        lo = m * d + s_lo
        out.append((max(lo, bmin), min(lo + width - 1, B),
                    None if excl is None else lo + excl))
        m += sp
    return out


def span_words(d, t, sp, phi, bmin, B, word=96):
    """The anchors the kernel marks valid when it cuts the spans into words
    of `word` consecutive anchors from each span's unclipped start (filter
    word hand-out in rbg_sparse.hip): the word's range clipped to the span
    and to [bmin, B], the task's wheel mask built from s_lo alone (every word
    start is = s_lo mod 2 when d is even and mod 3 when 3 | d), and the
    excluded point cleared in word excl // word of every span."""
    if d <= 2:
        lo_b = max(bmin, t if d == 1 else t - 1)
        s_lo, width, excl, starts = lo_b, (B - lo_b + 1 if lo_b <= B else 0), None, [lo_b]
        if lo_b > B:
            starts = []
    else:
        odd_t = t % 2 == 1
        rm = (phi + sp - 1) % sp if odd_t else phi
        s_lo = d // 2 + 1 if odd_t else 1
        width = d + (d - 1) // 2 - d // 2 if odd_t else d - 1
        excl = d - s_lo if odd_t else (d // 2 - 1 if d % 2 == 0 else None)
        x = bmin + 1 - s_lo - width
        m = 0 if x <= 0 else (x + d - 1) // d
        m += (rm + sp - m % sp) % sp
        starts = []
        while m * d + s_lo <= B:
            starts.append(m * d + s_lo)
            m += sp
    # task wheel: bit i of a word stands for b0 + i
    wheel = (1 << word) - 1
    if d % 2 == 0:
        wheel &= int(("01" if s_lo % 2 else "10") * (word // 2), 2)  # keep odd b
    if d % 3 == 0:
        wheel &= sum(1 << i for i in range(word) if (s_lo + i) % 3 != 0)
    got = set()
    nwords = (width + word - 1) // word
    for lo in starts:
        for w in range(nwords):
            b0 = lo + word * w
            first, last = max(bmin, b0), min(lo + width - 1, B, b0 + word - 1)
            if first > last:
                continue
            valid = sum(1 << (b - b0) for b in range(first, last + 1)) & wheel
            if excl is not None and w == excl // word:
                valid &= ~(1 << (excl % word))
            got.update(b0 + i for i in range(word) if valid >> i & 1)
    return got


ROWS = [(1, 0), (2, 1), (3, 1), (3, 2), (6, 1), (6, 5)]  # (g, class) of rows 0..5


def row_layout(B, np_):
    """Row sizes of one profile as the kernels lay them out (rbg_sparse.hpp):
    per row (g, class, rho, colour bits, words per plane, first word)."""
    out, off = [], 0
    for g, c in ROWS:
        rho = (c + B) % g
        words = ((2 * B // g + 1 + 256 + 31) // 32 + 15) // 16 * 16
        out.append({"g": g, "c": c, "rho": rho, "bits": (2 * B - rho) // g + 1,
                    "words": words, "off": off})
        off += np_ * words
    return out


def class_row_words(d, t, sp, phi, bmin, B, np_=1, word=96):
    """The filter's words over class rows (class_words and the word hand-out
    of rbg_sparse.hip, integer for integer): for g = gcd(d, 6), each span's
    anchors of a class c coprime to g are cut into words of `word`
    consecutive row bits; word w of span `span` starts at row bit
    klo + span * sp * (d / g) + word * w, bits below kmin / above kmax and
    the unscheduled point are cleared.  Returns {anchor b: (class index,
    row bit k)} for every valid bit, k counted from the first class's row
    (the second class lies cstride bits further)."""
    g = (1 if d % 2 else 2) * (1 if d % 3 else 3)
    row0 = {1: 0, 2: 1, 3: 2, 6: 4}[g]
    nc = 2 if g >= 3 else 1
    lay = row_layout(B, np_)
    cstride = np_ * lay[row0]["words"] * 32
    dg = d // g
    if d <= 2:
        lo_b = max(bmin, t if d == 1 else t - 1)
        s_lo, width, excl, m_first, nsb = lo_b, (B - lo_b + 1 if lo_b <= B else 0), None, 0, \
            (1 if lo_b <= B else 0)
    else:
        odd_t = t % 2 == 1
        rm = (phi + sp - 1) % sp if odd_t else phi
        s_lo = d // 2 + 1 if odd_t else 1
        width = d + (d - 1) // 2 - d // 2 if odd_t else d - 1
        excl = d - s_lo if odd_t else (d // 2 - 1 if d % 2 == 0 else None)
        x = bmin + 1 - s_lo - width
        m_lo = 0 if x <= 0 else (x + d - 1) // d
        m_first = m_lo + (rm + sp - m_lo % sp) % sp
        m_hi = (B - s_lo) // d
        nsb = 0 if m_first > m_hi else (m_hi - m_first) // sp + 1
    classes = []
    for ci in range(nc):
        cls = 0 if g == 1 else (1 if g == 2 else (g - 1 if ci else 1))
        rho = lay[row0 + ci]["rho"]
        assert rho == (cls + B) % g
        delta = (cls + g - s_lo % g) % g
        n = (width - 1 - delta) // g + 1 if width > delta else 0
        start = ci * cstride
        ex = None
        if excl is not None and excl >= delta and (excl - delta) % g == 0:
            ex = divmod((excl - delta) // g, word)
        classes.append({"n": n, "W": (n + word - 1) // word, "start": start, "rho": rho,
                        "klo": start + (m_first * d + s_lo + delta + B - rho) // g,
                        "kmin": start + (bmin + B - rho + g - 1) // g,
                        "kmax": start + (2 * B - rho) // g, "ex": ex})
    wsum = sum(c["W"] for c in classes)
    got = {}
    for my in range(nsb * wsum):
        span, rem = divmod(my, wsum)
        ci = 1 if rem >= classes[0]["W"] else 0
        C = classes[ci]
        w = rem - (classes[0]["W"] if ci else 0)
        k0 = C["klo"] + span * sp * dg + word * w
        nvalid = min(word, C["n"] - word * w)
        first = C["kmin"] - k0 if C["kmin"] > k0 else 0
        if C["kmax"] < k0:
            continue
        last = min(nvalid - 1, C["kmax"] - k0)
        for i in range(first, last + 1):
            if C["ex"] == (w, i):
                continue
            k = k0 + i
            b = g * (k - C["start"]) + C["rho"] - B
            assert b not in got
            got[b] = (ci, k)
    return got, {"g": g, "dg": dg, "cstride": cstride, "row0": row0, "layout": lay}


def anchors(ranges):
    got = set()
    for lo, hi, *ex in ranges:
        got.update(range(lo, hi + 1))
        if ex and ex[0] is not None:
            got.discard(ex[0])
    return got


def sparse_profile(p, r, Q=48, screen=True, stats=None):
    """(M, witness, L) of order r by the anchor search; None if refused."""
    if stats is None:
        stats = {}
    stats["completions"] = 0
    c = cover(p, Q)
    if c is None:
        return None
    B, D = c
    mask = sum(1 << (k - 2) for k in range(2, 10) if (p - 1) % k == 0)
    m = 1
    for k in range(2, 10):
        if mask & (1 << (k - 2)):
            m = m * k // gcd(m, k)
    lab = labels_of(p, m)
    e = (p - 1) // m
    jneg = m // 2 if e & 1 else 0
    tab = [0] * (B + 1)
    for y in range(1, B + 1):
        tab[y] = lab(y)

    def col_num(a):  # colour of a signed numerator, |a| <= B, a != 0
        return tab[a] % r if a > 0 else (tab[-a] + jneg) % m % r

    def col_abs(z):
        return lab(z) % r

    # seed: runs of 1..B in order
    best, bw, start, prefix = 0, 0, 1, None
    for y in range(2, B + 2):
        if y == B + 1 or tab[y] % r != tab[y - 1] % r:
            length = y - start
            if prefix is None:
                prefix = length
            if length > best:
                best, bw = length, start
            start = y
    if prefix == B or best < 3:
        return None
    key = (best, -bw)

    def complete(b, d, t):
        dinv = pow(d, p - 2, p)
        x = p - b * dinv % p
        c0 = col_num(-b)
        lo = x
        a = -b - d
        table = True
        abs0 = None
        while lo > 1:
            if table and a >= -B:
                same = col_num(a) == c0
            else:
                table = False
                if abs0 is None:
                    abs0 = col_abs(x)
                same = col_abs(lo - 1) == abs0
            if not same:
                break
            lo -= 1
            a -= d
        hi = x + t - 1
        a = -b + t * d
        table = True
        while hi < p - 1:
            if table and a <= B:
                same = col_num(a) == c0
            else:
                table = False
                if abs0 is None:
                    abs0 = col_abs(x)
                same = col_abs(hi + 1) == abs0
            if not same:
                break
            hi += 1
            a += d
        return hi - lo + 1, min(lo, p - hi)

    def passes_screen(b, d, t, q, c0):
        """L + R >= q - t, numerators outside [-B, B] count as matches."""
        h = q - t
        L = 0
        while L < h:
            a = b + (L + 1) * d
            if a <= B and col_num(a) != c0:
                break
            L += 1
        R = 0
        while R < h:
            a = b - (t - 1 + R + 1) * d
            if a >= -B and col_num(a) != c0:
                break
            R += 1
        return L + R >= h

    for d in range(1, D + 1):
        q = min(key[0], Q)
        t = min(q, max(3, (q + 1) // 2))
        sp = q - t + 1
        phi = ((t - 1) // 2) % sp
        bmin = ((t - 1) * d + 1) // 2
        for lo_b, hi_b, excl in span_ranges(d, t, sp, phi, bmin, B):
            for b in range(lo_b, hi_b + 1):
                if b == excl:
                    continue
                if d % 2 == 0 and b % 2 == 0:
                    continue
                if d % 3 == 0 and b % 3 == 0:
                    continue
                c0 = col_num(b)
                if all(col_num(b - j * d) == c0 for j in range(1, t)):
                    if screen and not passes_screen(b, d, t, q, c0):
                        continue
                    stats["completions"] += 1
                    length, w = complete(b, d, t)
                    key = max(key, (length, -w))
    return key[0], -key[1], prefix
