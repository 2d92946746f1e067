"""CPU model of the exact sums by chunk programs and an in-order walk
(hera_amd/csrc/dev/walk_sum.hip), for tests.

It restates the record rules the device programs follow -- a run of values
on one unit (a binade's ulp, or a row's grid for plain sums) adds a fixed
integer once the parity before its first round-half tie is known; the
first tie is added by itself; each segment carries the interval of starts
it is valid for and, where it matters, the unit its start must be a
multiple of -- and the walk that adds the records in order from +0,
checks every segment's true start and adds a failing segment's values one
by one.  The value it returns is exact whatever the guesses were, which
is what the tests check (bitwise against the sequential sum, also with
deliberately wrong guesses).  Integer arithmetic in Python, one value at a
time: for rows of a few thousand values.  It models the method, not the
kernel's thread layout (rows of 64 values here, one DPP row of threads on
the device).
"""
import math
import struct

import numpy as np

B52, B53 = 1 << 52, 1 << 53
TAU = 2.0 ** -30          # classification margin, relative (as WS_TAU)


def bits(x):
    return struct.unpack('<q', struct.pack('<d', float(x)))[0]


def key12(x):
    """Sign and exponent field."""
    return (bits(x) >> 52) & 0xFFF


def low_exp(x):
    """Exponent of the lowest set bit of a finite nonzero x (else None)."""
    b = bits(x) & 0x7FFFFFFFFFFFFFFF
    e, m = b >> 52, b & (B52 - 1)
    if e == 0x7FF:
        return None
    if e:
        m |= B52
    if m == 0:
        return None
    return max(e, 1) - 1075 + ((m & -m).bit_length() - 1)


def on_grid(s, G):
    """s is a signed zero or a multiple of 2^G."""
    if s == 0:
        return True
    le = low_exp(s)
    return le is not None and le >= G


def classify(y, g, M, row=64, use_grid=True):
    """Per value: (label, unit code, q, tie, kind) or None, from the
    approximate trajectory g + prefix sums (any rounding)."""
    n = len(y)
    a = g + np.cumsum(y)
    prev = np.concatenate(([g], a[:-1]))
    tau = (M + abs(g) + float(np.sum(np.abs(y)))) * TAU
    code = [((bits(max(abs(prev[k]), abs(a[k])) + tau) >> 52) & 0x7FF)
            for k in range(n)]
    out = [None] * n
    grid = [None] * n          # (label, anchor) of row-grid candidates
    for k in range(n):
        yk, ak, pk = float(y[k]), float(a[k]), float(prev[k])
        ka = key12(ak)
        be = ka & 0x7FF
        binade = None
        if ka == key12(pk) and 2 <= be <= 2044:
            blo = math.ldexp(1.0, be - 1023)
            bhi = 2 * blo
            if all(d > tau for d in (abs(ak) - blo, bhi - abs(ak),
                                     abs(pk) - blo, bhi - abs(pk))):
                xs = math.ldexp(yk, 1075 - be)
                if abs(xs) < 2.0 ** 52:
                    kf = math.floor(xs)
                    f = xs - kf
                    binade = (be, int(kf) + (f > 0.5), f == 0.5,
                              '-' if ka >> 11 else '+')
        if binade:
            be, q, tie, sg = binade
            out[k] = ((be, sg == '-'), be, q, tie, sg)
        if not use_grid:
            continue
        gc = max(code[(k // row) * row:(k // row + 1) * row])
        same = (ak > 0 and pk > 0) or (ak < 0 and pk < 0)
        if not (2 <= gc <= 2044 and same):
            continue
        exact = bits(yk) == 0 or (low_exp(yk) is not None and
                                  low_exp(yk) >= gc - 1075)
        top = binade is not None and binade[0] == gc
        if exact or top:
            anchor = abs(pk) - tau >= math.ldexp(1.0, gc - 1023)
            lab = (gc, ak < 0)
            if exact:
                val = (lab, gc, int(math.ldexp(yk, 1075 - gc)), False, 'X')
            else:
                val = (lab, gc, binade[1], binade[2], binade[3])
            grid[k] = (val, anchor)
    # a row-grid value counts from the first anchor of its stretch on
    seen, plab = False, None
    for k in range(n):
        if grid[k] is None:
            seen, plab = False, None
            continue
        val, anchor = grid[k]
        if val[0] != plab:
            seen = False
        seen = seen or anchor
        plab = val[0]
        if seen:
            out[k] = val
    return out


def program(y, g, M, row=64, use_grid=True):
    """Records of one chunk: ('v', beg, value) or ('s', beg, end, add, lo,
    hi, grid exponent the start must be a multiple of, or None)."""
    cls = classify(y, g, M, row, use_grid)
    n = len(y)
    # the first tie of each run of equal labels is added by itself
    single = [c is None for c in cls]
    plab, seen = None, False
    for k in range(n):
        if cls[k] is None:
            plab = None
            continue
        if cls[k][0] != plab:
            seen = False
        plab = cls[k][0]
        if cls[k][3]:
            if not seen:
                single[k] = True
            seen = True
    recs = []
    k = 0
    while k < n:
        if single[k]:
            recs.append(('v', k, float(y[k])))
            k += 1
            continue
        even = k > 0 and single[k - 1] and cls[k - 1] is not None and \
            cls[k - 1][3]
        lab, c = cls[k][0], cls[k][1]
        j, P, lo, hi, hasx = k, 0, -(1 << 70), 1 << 70, False
        while j < n and not single[j] and cls[j][0] == lab:
            _, _, q, tie, kind = cls[j]
            if tie:                       # running sum S + P, S even
                q += (P + q) & 1
            if kind == 'X':
                L, U, hasx = -(B53 - 1), B53 - 1, True
            elif kind == '+':
                L, U = B52 + 1, B53 - 1
            else:
                L, U = -(B53 - 1), -(B52 + 1)
            lo = max(lo, L - min(P, P + q))
            hi = min(hi, U - max(P, P + q))
            P += q
            j += 1
        G = c - 1075
        need = (G + 1 if even else G) if (hasx or even) else None
        if lo <= hi and abs(P) < B53:
            recs.append(('s', k, j, math.ldexp(P, G), math.ldexp(lo, G),
                         math.ldexp(hi, G), need))
        else:
            recs.append(('s', k, j, 0.0, math.inf, -math.inf, need))
        k = j
    return recs


def exact_sum(x, L=256, row=64, use_grid=True, guess_error=None):
    """The left-to-right sum of x by chunk programs and a checked walk;
    returns (sum, records, segments added value by value).  guess_error(j)
    may shift chunk j's guessed start (the result must not change)."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    P = (n + L - 1) // L
    sums = [float(np.sum(x[j * L:(j + 1) * L])) for j in range(P)]
    s, g, M, nrec, fails = 0.0, 0.0, 0.0, 0, 0
    for j in range(P):
        y = x[j * L:(j + 1) * L]
        gj = g + (guess_error(j) if guess_error else 0.0)
        recs = program(y, gj, max(M, abs(gj)), row, use_grid)
        M = max(M, abs(g))
        g += sums[j]
        if len(recs) > L // 2:            # no program: value by value
            for v in y:
                s = s + float(v)
            continue
        nrec += len(recs)
        for i, r in enumerate(recs):
            if r[0] == 'v':
                s = s + r[2]
                continue
            _, beg, end, add, lo, hi, need = r
            if lo <= s <= hi and (need is None or on_grid(s, need)):
                s = s + add
            else:
                fails += 1
                for v in y[beg:end]:
                    s = s + float(v)
    return s, nrec, fails
