"""Boundary tables beyond free slip for the BC-language tests: rows with
several neighbours, neighbours of the other component or two nodes away
(movemart.c:433-539 take any linear row V = c + sum k V_n), and
continuity-type rows (movemart.c:863-879, init.t3c's 'c' coefficient)."""
import numpy as np


def _append(hs, unknown, const, nbrs):
    """Point `unknown`'s bondm at a new row c + sum k V_n (nbrs: [(u, k)])."""
    n = np.zeros((1, 3), np.int64)
    v = np.zeros((1, 4))
    v[0, 0] = const
    for q, (u, k) in enumerate(nbrs):
        n[0, q] = u + 1
        v[0, q + 1] = k
    hs.bondn = np.vstack([hs.bondn, n])
    hs.bondv = np.vstack([hs.bondv, v])
    hs.bondm[unknown] = hs.bondn.shape[0] - 1


def general_rows(hs):
    """The lower free-slip rows Vx(m1,0) = Vx(m1,1) become mixtures: two
    neighbours of the same component, plus a small Vy term, plus (every
    third row) a neighbour two nodes in."""
    X, Y = hs.xnumx, hs.ynumy
    node = lambda a, b: a * Y + b
    rows = []
    for m1 in range(1, X - 2):
        u = 3 * node(m1, 0) + 1
        nb = [(3 * node(m1, 1) + 1, 0.7), (3 * node(m1 + 1, 1) + 1, 0.25), (3 * node(m1, 1) + 2, 0.05)]
        if m1 % 3 == 0:
            nb[1] = (3 * node(m1, 2) + 1, 0.25)
        _append(hs, u, 0.0, nb)
        rows.append((u, nb))
    return rows


def continuity_rows(hs):
    """An open right wall: the wall-normal Vx(x, m2) rows become the
    continuity equations of the cells next to the wall, whose pressures are
    fixed at 0 (the reference pairs a continuity-type Vx row with a P BC of
    the cell it names)."""
    X, Y = hs.xnumx, hs.ynumy
    x = X - 1
    node = lambda a, b: a * Y + b
    rows = []
    for m2 in range(1, Y - 3):
        u = 3 * node(x, m2) + 1
        _append(hs, 3 * node(x, m2 + 1), 0.0, [])
        _append(hs, u, 0.0, [(3 * node(x, m2 + 1) + 1, 1.1e30)])
        rows.append((u, x, m2 + 1))
    return rows


def chained_temperature_rows(hs):
    """An insulating left wall, T(0, m2) = T(1, m2), whose corner rows point
    at wall rows in turn, T(0, 0) = T(0, 1) and T(0, y) = T(0, y - 1): chains
    of two temperature BC rows (tkrecalc solves them with gausmat3,
    heatmart.c:758-826)."""
    X, Y = hs.xnumx, hs.ynumy
    N = X * Y
    node = lambda a, b: a * Y + b
    for m2 in range(1, Y - 1):
        _append(hs, 3 * N + node(0, m2), 0.0, [(3 * N + node(1, m2), 1.0)])
    _append(hs, 3 * N + node(0, 0), 0.0, [(3 * N + node(0, 1), 1.0)])
    _append(hs, 3 * N + node(0, Y - 1), 0.0, [(3 * N + node(0, Y - 2), 1.0)])


def cyclic_temperature_rows(hs):
    """Two temperature BC rows that reference each other."""
    X, Y = hs.xnumx, hs.ynumy
    N = X * Y
    a, b = 3 * N + 0 * Y + 0, 3 * N + 1 * Y + 0
    _append(hs, a, 0.0, [(b, 0.5)])
    _append(hs, b, 0.0, [(a, 0.5)])
