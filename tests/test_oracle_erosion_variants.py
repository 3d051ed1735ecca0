"""Pin the oracle's erosion v1 and v4 (erosion.py:42-99, erosion.py:301-381)
on the CPU.

v1 moves each height by exactly 0.0005 per higher neighbour and -0.0005 per
lower one, reading only the old buffer.  v4 is v3's step with every
contribution clamped by max(x, 0): the pins are v4 == v3 (already pinned by
hand) where no clamp is active, and, for each clamp, the configuration that
makes it bite, where a dropped clamp changes the answer: a neighbour's
negative water acts exactly like zero water, a negative sediment sum leaves
height and sediment untouched, a negative water sum leaves the water
untouched.
"""
import numpy as np

H = 0.0005


def _graph():
    # a path 0 - 1 - 2 - 3 plus a pendant 4 on vertex 1; unit edge lengths
    verts = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.0, 0, 0], [3.0, 0, 0], [1.0, 1.0, 0]])
    nb = -np.ones((5, 6), np.int32)
    for a, b in ((0, 1), (1, 2), (2, 3), (1, 4)):
        nb[a, list(nb[a]).index(-1)] = b
        nb[b, list(nb[b]).index(-1)] = a
    return verts, nb


def test_v1_steps_by_hand(oracle):
    _, nb = _graph()
    h = np.array([1.0, 2.0, 2.0, 5.0, 0.0])
    w = np.zeros(5)
    oracle.erosion_iteration1(nb, h, w)
    # 0: one higher -> +H; 1: one lower (0), one equal (2), one lower (4) -> -2H;
    # 2: one equal, one higher -> +H; 3: one lower -> -H; 4: one higher -> +H
    assert w.tolist() == [1.0 + H, 2.0 - 2 * H, 2.0 + H, 5.0 - H, 0.0 + H]
    assert h.tolist() == [1.0, 2.0, 2.0, 5.0, 0.0]          # r_buff is only read


def test_v1_is_jacobi_and_ignores_pads(oracle):
    _, nb = _graph()
    # vertex 0 starts just above vertex 1 and drops below it in this step;
    # vertex 1 must still count it as higher (old heights only)
    h = np.array([2.0002, 2.0, 1.0, 0.0, 9.0])
    w = np.empty(5)
    oracle.erosion_iteration1(nb, h, w)
    assert w[0] == 2.0002 - H and w[0] < 2.0
    assert w[1] == 2.0 + H                    # +H (0, old) - H (2) + H (4)
    # pads anywhere in a row, real neighbours in the same order
    nb2 = -np.ones_like(nb)
    for v in range(5):
        real = [n for n in nb[v] if n != -1]
        slots = [1, 3, 4, 5][:len(real)]
        for n, k in zip(real, slots):
            nb2[v, k] = n
    w2 = np.empty(5)
    oracle.erosion_iteration1(nb2, h, w2)
    assert np.array_equal(w, w2)


def _v3(oracle, verts, nb, h, w, s):
    h, w, s = h.copy(), w.copy(), s.copy()
    oracle.erosion_iteration3(verts, nb, h, w, s)
    return h, w, s


def _v4(oracle, verts, nb, h, w, s):
    h, w, s = h.copy(), w.copy(), s.copy()
    oracle.erosion_iteration4(verts, nb, h, w, s)
    return h, w, s


def test_v4_equals_v3_where_no_clamp_is_active(oracle):
    verts, nb = _graph()
    # heights rising along the path: each vertex's sediment sum and water sum
    # stay positive (large own sediment and water)
    h = np.array([0.0, 1.0, 2.0, 3.0, 4.0])
    w = np.array([1.0, 1.0, 1.0, 1.0, 1.0])
    s = np.array([1.0, 1.0, 1.0, 1.0, 1.0]) * 1e-3
    a = _v3(oracle, verts, nb, h, w, s)
    b = _v4(oracle, verts, nb, h, w, s)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_v4_negative_neighbour_water_counts_as_zero(oracle):
    verts, nb = _graph()
    h = np.array([0.0, 1.0, 2.0, 3.0, 4.0])
    s = np.full(5, 1e-3)
    w_neg = np.array([1.0, -7.0, 1.0, 1.0, 1.0])
    w_zero = np.array([1.0, 0.0, 1.0, 1.0, 1.0])
    a = _v4(oracle, verts, nb, h, w_neg, s)
    b = _v4(oracle, verts, nb, h, w_zero, s)
    for v in (0, 2, 4):                       # the neighbours of vertex 1
        assert a[0][v] == b[0][v] and a[1][v] == b[1][v] and a[2][v] == b[2][v]
    c = _v3(oracle, verts, nb, h, w_neg, s)
    assert c[1][0] != a[1][0]                 # v3 does take the negative water


def test_v4_negative_sums_leave_fields(oracle):
    verts, nb = _graph()
    # vertex 3 is the highest at the end of the path: its only neighbour is
    # lower and very wet, so its sediment sum and water sum are negative
    h = np.array([0.0, 1.0, 2.0, 3.0, 0.5])
    w = np.array([0.0, 0.0, 50.0, 0.001, 0.0])
    s = np.zeros(5)
    a = _v4(oracle, verts, nb, h, w, s)
    assert a[0][3] == 3.0 and a[2][3] == 0.0  # height and sediment untouched
    assert a[1][3] == 0.001                   # water untouched
    b = _v3(oracle, verts, nb, h, w, s)
    assert b[0][3] > 3.0 and b[1][3] < 0.001  # v3 raises the height, drains the water
