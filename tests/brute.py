"""Literal definition of the profile for small p (test helper).

Colours are compared through the values pow(x, (p-1)/r, p) themselves (equal
values <=> equal colours; the identity colour is the value 1), runs are read
in plain integer order over 1..p-1 -- the check_small.py scheme of the
reference (reproducibility/bounds_update/check_small.py:8-25).  Pure Python,
for p below a few thousand only.
"""


def profile(p):
    mask = 0 if p < 3 else sum(1 << (r - 2) for r in range(2, 10) if (p - 1) % r == 0)
    rec = {"p": p, "mask": mask, "M": [0] * 8, "L": [0] * 8, "W": [0] * 8, "K": [0] * 8, "A4": 0}
    for r in range(2, 10):
        if not mask & (1 << (r - 2)):
            continue
        col = [None] + [pow(x, (p - 1) // r, p) for x in range(1, p)]
        best, wit = 0, p
        start = 1
        for x in range(2, p + 1):
            if x == p or col[x] != col[x - 1]:
                end = x - 1
                length = end - start + 1
                w = min(start, p - end)
                if length > best or (length == best and w < wit):
                    best, wit = length, w
                start = x
        pre = 0
        while pre + 1 < p and col[pre + 1] == 1:
            pre += 1
        k = r - 2
        rec["M"][k], rec["L"][k], rec["W"][k] = best, pre, wit
        rec["K"][k] = max(best + 1, 2 * pre + 2 if ((p - 1) // 2) % r == 0 else pre + 2)
    if p > 2:
        chi = [None] + [pow(x, (p - 1) // 2, p) for x in range(1, p)]
        rec["A4"] = sum(1 for x in range(1, p - 3)
                        if chi[x] == chi[x + 1] == chi[x + 2] == chi[x + 3])
    return rec
