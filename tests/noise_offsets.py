"""Second derivation of noise3d (opensimplex.py:266-759) for pinning the oracle.

The oracle (oracle/nx_oracle.c) transcribes the reference statement by
statement.  This module derives the same function differently: the region
logic is reduced to integer lattice offsets of the two extra vertices, every
contributing point is evaluated with the generic formula
    delta = d0 - o - (ox + oy + oz) / 3,   attn = 2 - |delta|^2,
and the arithmetic is vectorised numpy in FP64.  A misreading of one of the
reference's region branches in either derivation shows up as a selection
mismatch; an arithmetic slip shows up as a value mismatch.
"""
import numpy as np

S = 1.0 / 3.0
GRAD3 = np.array([
    -11, 4, 4, -4, 11, 4, -4, 4, 11,
    11, 4, 4, 4, 11, 4, 4, 4, 11,
    -11, -4, 4, -4, -11, 4, -4, -4, 11,
    11, -4, 4, 4, -11, 4, 4, -4, 11,
    -11, 4, -4, -4, 11, -4, -4, 4, -11,
    11, 4, -4, 4, 11, -4, 4, 4, -11,
    -11, -4, -4, -4, -11, -4, -4, -4, -11,
    11, -4, -4, 4, -11, -4, 4, -4, -11,
], dtype=np.float64).reshape(24, 3)

BASE = {
    0: [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
    1: [(1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)],
    2: [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
}


def _ext_offsets(xins, yins, zins):
    """Region and extra-vertex offsets for one point (scalar FP64 logic)."""
    in_sum = xins + yins + zins
    if in_sum <= 1:
        a_point, a_score, b_point, b_score = 1, xins, 2, yins
        if a_score >= b_score and zins > b_score:
            b_score, b_point = zins, 4
        elif a_score < b_score and zins > a_score:
            a_score, a_point = zins, 4
        wins = 1 - in_sum
        if wins > a_score or wins > b_score:
            c = b_point if b_score > a_score else a_point
            bits = [(c >> k) & 1 for k in range(3)]
            # the closest single axis c: both extras step along it, the
            # others get -1 on one coordinate each
            e0, e1 = [0, 0, 0], [0, 0, 0]
            for k in range(3):
                if bits[k]:
                    e0[k] = e1[k] = 1
            free = [k for k in range(3) if not bits[k]]
            e0[free[0]] = -1
            e1[free[1]] = -1
            return 0, tuple(e0), tuple(e1)
        c = a_point | b_point
        e0 = tuple((c >> k) & 1 for k in range(3))
        e1 = tuple(1 if (c >> k) & 1 else -1 for k in range(3))
        return 0, e0, e1
    if in_sum >= 2:
        a_point, a_score, b_point, b_score = 6, xins, 5, yins
        if a_score <= b_score and zins < b_score:
            b_score, b_point = zins, 3
        elif a_score > b_score and zins < a_score:
            a_score, a_point = zins, 3
        wins = 3 - in_sum
        if wins < a_score or wins < b_score:
            c = b_point if b_score < a_score else a_point
            bits = [(c >> k) & 1 for k in range(3)]
            e0, e1 = [0, 0, 0], [0, 0, 0]
            on = [k for k in range(3) if bits[k]]
            for k in on:
                e0[k] = e1[k] = 1
            e0[on[0]] = 2
            e1[on[1]] = 2
            return 1, tuple(e0), tuple(e1)
        c = a_point & b_point
        e0 = tuple((c >> k) & 1 for k in range(3))
        e1 = tuple(2 * ((c >> k) & 1) for k in range(3))
        return 1, e0, e1
    p1, p2, p3 = xins + yins, xins + zins, yins + zins
    if p1 > 1:
        a_score, a_point, a_far = p1 - 1, 3, True
    else:
        a_score, a_point, a_far = 1 - p1, 4, False
    if p2 > 1:
        b_score, b_point, b_far = p2 - 1, 5, True
    else:
        b_score, b_point, b_far = 1 - p2, 2, False
    if p3 > 1:
        score = p3 - 1
        if a_score <= b_score and a_score < score:
            a_point, a_far = 6, True
        elif a_score > b_score and b_score < score:
            b_point, b_far = 6, True
    else:
        score = 1 - p3
        if a_score <= b_score and a_score < score:
            a_point, a_far = 1, False
        elif a_score > b_score and b_score < score:
            b_point, b_far = 1, False

    def first_axis(mask):  # first axis whose bit is set
        return 0 if mask & 1 else (1 if mask & 2 else 2)

    def first_clear(mask):  # first axis whose bit is clear
        return 0 if not mask & 1 else (1 if not mask & 2 else 2)

    def minus_one_on(ax):
        return tuple(-1 if k == ax else 1 for k in range(3))

    def two_on(ax):
        return tuple(2 if k == ax else 0 for k in range(3))

    if a_far == b_far:
        if a_far:
            return 2, (1, 1, 1), two_on(first_axis(a_point & b_point))
        return 2, (0, 0, 0), minus_one_on(first_clear(a_point | b_point))
    c1, c2 = (a_point, b_point) if a_far else (b_point, a_point)
    return 2, minus_one_on(first_clear(c1)), two_on(first_axis(c2))


def noise3d(x, y, z, perm, pgi, with_sel=False):
    """Scalar noise3d by the offset formulation (FP64)."""
    so = (x + y + z) * (-1.0 / 6)
    xs, ys, zs = x + so, y + so, z + so
    xsb, ysb, zsb = int(np.floor(xs)), int(np.floor(ys)), int(np.floor(zs))
    sq = (xsb + ysb + zsb) * S
    xins, yins, zins = xs - xsb, ys - ysb, zs - zsb
    d0 = np.array([x - (xsb + sq), y - (ysb + sq), z - (zsb + sq)])
    region, e0, e1 = _ext_offsets(xins, yins, zins)
    pts = BASE[region] + [e0, e1]
    value = 0.0
    cs = 0
    for o in pts:
        o = np.array(o)
        g = pgi[(perm[(perm[(xsb + o[0]) & 255] + ysb + o[1]) & 255] + zsb + o[2]) & 255]
        cs = (cs * 31 + int(g)) & 0xFFFFFFFF
        d = d0 - o - S * o.sum()
        attn = 2.0 - d @ d
        if attn > 0:
            value += attn ** 4 * (GRAD3[g // 3] @ d)
    value /= 103
    if not with_sel:
        return value
    code = region
    for k in range(3):
        code |= (e0[k] + 1) << (2 + 2 * k)
        code |= (e1[k] + 1) << (8 + 2 * k)
    if cs >= 1 << 31:
        cs -= 1 << 32
    return value, np.array([xsb, ysb, zsb, code, cs], dtype=np.int64)
