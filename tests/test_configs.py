"""Seeded setups: marker lattice order, jitter, materials and BC tables."""
import numpy as np

from i2elvis_planet_amd import configs


def test_c1_setup_layout():
    p, hs, cfg = configs.make("c1")
    X = Y = 51
    assert hs.marknum == 50 * 50 * 16 == 40000
    assert hs.gx[0] == 0 and hs.gx[-1] == 200e3 and np.all(np.diff(hs.gx) > 0)
    # lattice is cell-major: cell index of every marker is non-decreasing
    cell = (np.arange(hs.marknum) // 16)
    i = np.clip(np.searchsorted(hs.gx, hs.markx.astype(float), side="right") - 1, 0, X - 2)
    j = np.clip(np.searchsorted(hs.gy, hs.marky.astype(float), side="right") - 1, 0, Y - 2)
    # jitter is +-half a marker spacing, so markers stay within their cell
    assert np.array_equal(i * (Y - 1) + j, cell)
    r = np.hypot(hs.markx - 100e3, hs.marky - 100e3)
    assert set(np.unique(hs.markt[r < 49e3])) <= {configs.SILICATE, configs.IRON}
    assert np.all(hs.markt[r > 51e3] == configs.AIR)
    frac = np.mean(hs.markt[r < 49e3] == configs.IRON)
    assert 0.23 < frac < 0.27
    assert np.all(hs.markk == 150.0)


def test_setup_is_deterministic():
    a = configs.make_state("c1", seed=7)
    b = configs.make_state("c1", seed=7)
    c = configs.make_state("c1", seed=8)
    assert np.array_equal(a.markx, b.markx) and np.array_equal(a.markt, b.markt)
    assert not np.array_equal(a.markx, c.markx)


def test_free_slip_bc_tables():
    X = Y = 11
    N = X * Y
    bm, bn, bv = configs.free_slip_bc(X, Y)
    node = lambda i, j: i * Y + j
    # left wall Vx fixed to zero, top Vx mirrors its neighbour below
    b = bm[3 * node(0, 4) + 1]
    assert b and bn[b, 0] == 0 and bv[b, 0] == 0
    b = bm[3 * node(4, 0) + 1]
    assert bn[b, 0] == 3 * node(4, 1) + 1 + 1 and bv[b, 1] == 1.0
    # pressure anchor and mirrored corner cells
    assert bm[3 * node(3, 1)] and bn[bm[3 * node(3, 1)], 0] == 0
    assert bn[bm[3 * node(1, 1)], 0] == 3 * node(2, 1) + 1
    # every edge node has a fixed-temperature row, interior nodes none
    t = bm[3 * N:].reshape(X, Y)
    assert np.all(t[0, :] > 0) and np.all(t[-1, :] > 0) and np.all(t[:, 0] > 0) and np.all(t[:, -1] > 0)
    assert np.all(t[1:-1, 1:-1] == 0)
    assert np.all(bv[t[t > 0], 0] == 150.0)


def test_c3_blob_setup_small_sample():
    cfg = dict(configs.CONFIGS["c3"], n=101, size=120e3, radius=30e3)
    configs.CONFIGS["_b"] = cfg
    try:
        p, hs, _ = configs.make("_b")
    finally:
        del configs.CONFIGS["_b"]
    assert hs.marknum == 100 * 100 * 25
    assert (hs.markt == configs.IRON).sum() > 0
    assert p.melt_on == 1 and p.nubeg == 1e2
