"""The binary .prn checkpoint (i2e_prn_write / i2e_prn_read, host code of
libi2e): round trips, and the byte layout that the reference's loader
(loadmart.c:391-540) and its plotting script (support_files/plot2d.py) read.
CPU only."""
import struct

import numpy as np
import pytest

from i2elvis_planet_amd import configs, prn

YEAR = 3.15576e7


def state_with_fields(name="c1", seed=4):
    p, hs, _ = configs.make(name)
    rng = np.random.default_rng(seed)
    for f in ("vx", "vy", "pr", "exx", "eyy", "exy", "sxx", "syy", "sxy", "nu", "nd", "ro", "tk", "cp", "kt", "ht",
              "et"):
        getattr(hs, f)[:] = rng.standard_normal(hs.N) * 10.0 ** rng.integers(-12, 20)
    hs.markv[:] = rng.random(hs.marknum).astype(np.float32) * 1e20
    hs.marktmax[:] = (150 + 500 * rng.random(hs.marknum)).astype(np.float32)
    hs.markpor[:] = rng.random(hs.marknum).astype(np.float32) * 0.5
    hs.timesum = 1.234e11
    hs.gxkoef = 0.0421
    return p, hs


def f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def test_round_trip(tmp_path):
    p, hs = state_with_fields()
    path = tmp_path / "t.prn"
    prn.write(path, p, hs, n0=17)
    q, got, n0 = prn.read(path, params=p)
    assert n0 == 17
    assert got.marknum == hs.marknum and got.xnumx == hs.xnumx and got.ynumy == hs.ynumy
    # node fields, grid lines and BC values are float32 in the file
    for f in ("vx", "vy", "pr", "exx", "eyy", "exy", "sxx", "syy", "sxy", "nu", "nd", "ro", "tk", "cp", "kt", "ht",
              "et", "gx", "gy"):
        assert np.array_equal(getattr(got, f), f32(getattr(hs, f))), f
    assert np.array_equal(got.bondv[1:], f32(hs.bondv[1:]))
    # integers and float32 marker fields exactly
    assert np.array_equal(got.bondm, hs.bondm)
    assert np.array_equal(got.bondn[1:], hs.bondn[1:])
    for f in ("markx", "marky", "markk", "markv", "markpor", "marktmax", "markt"):
        assert np.array_equal(getattr(got, f), getattr(hs, f)), f
    assert got.timesum == pytest.approx(hs.timesum, rel=1e-15)
    assert got.gxkoef == hs.gxkoef
    # params carried by the file
    assert q.al2627_init == pytest.approx(p.al2627_init, rel=1e-15)
    assert (q.xsize, q.ysize, q.gykoef, q.tmp_ambient) == (p.xsize, p.ysize, p.gykoef, p.tmp_ambient)
    for t in range(p.ntypes):
        a, b = q.rock[t], p.rock[t]
        assert (a.n0, a.n1, a.nu0, a.dh, a.ro, a.cp, a.kt) == (b.n0, b.n1, b.nu0, b.dh, b.ro, b.cp, b.kt)
    # a second write of what was read gives the same bytes
    path2 = tmp_path / "t2.prn"
    prn.write(path2, q, got, n0=17)
    assert path.read_bytes() == path2.read_bytes()


def test_layout_matches_the_reference_reader(tmp_path):
    """Offsets and record sizes of plot2d.py:161-229 and the loader's read
    order: header values, the first node record, grid lines, the first BC row
    and the markers, with marknum + 1 marker records."""
    p, hs = state_with_fields(seed=9)
    path = tmp_path / "t.prn"
    prn.write(path, p, hs, n0=3)
    b = path.read_bytes()
    assert b[:4] == bytes([4, 8, 4, 8])
    X, Y, mx, my, M = struct.unpack_from("<5q", b, 4)
    assert (X, Y, mx, my, M) == (hs.xnumx, hs.ynumy, p.mnumx, p.mnumy, hs.marknum)
    xsize, ysize = struct.unpack_from("<2d", b, 44)
    assert (xsize, ysize) == (p.xsize, p.ysize)
    # GXKOEF, GYKOEF, tmp_ambient, timeexit, al2627*1e5, fe6056*1e8 after the
    # 4 + 5*8 + 2*8 + 3*8 + 8 + 8 + 8 + 8 + 8 + 4*8 = 164 header bytes
    gxk, gyk, tamb, texit, al, fe = struct.unpack_from("<6d", b, 164)
    assert (gxk, gyk, tamb, texit) == (hs.gxkoef, p.gykoef, p.tmp_ambient, 0.0)
    assert al == pytest.approx(p.al2627_init * 1e5, rel=1e-15) and fe == p.fe6056_init * 1e8
    rocknum, = struct.unpack_from("<i", b, 212)
    bondnum, = struct.unpack_from("<q", b, 216)
    n0, = struct.unpack_from("<i", b, 224)
    tyr, = struct.unpack_from("<d", b, 228)
    assert (rocknum, bondnum, n0) == (p.ntypes, hs.bondn.shape[0], 3)
    assert tyr == pytest.approx(hs.timesum / YEAR, rel=1e-15)
    # cursor positions as plot2d.py computes them
    curpos_nodes = 4 + 2 * 4 + 28 * 8 + rocknum * (8 * 24 + 4)
    curpos_grid = curpos_nodes + (4 * 20 + 8 * 4) * X * Y
    curpos_markers = curpos_grid + 4 * (X + Y) + (4 * 4 + 8 * 3) * (bondnum - 1)
    assert len(b) == curpos_markers + (9 * 4 + 3) * (M + 1)
    # rock table: markro (array 17 of 24) of type 6 after markim
    off_ro = curpos_nodes - rocknum * 8 * 24 + 16 * rocknum * 8
    assert struct.unpack_from("<d", b, off_ro + 6 * 8)[0] == p.rock[6].ro
    # node m = 5*Y + 7: pr vx vy, 3 BC ids, 14 floats, T BC id, cp kt ht
    m = 5 * Y + 7
    o = curpos_nodes + 112 * m
    prs, vx, vy = struct.unpack_from("<3f", b, o)
    assert (prs, vx, vy) == tuple(np.float32([hs.pr[m], hs.vx[m], hs.vy[m]]))
    assert struct.unpack_from("<3q", b, o + 12) == tuple(int(v) for v in hs.bondm[3 * m:3 * m + 3])
    f14 = struct.unpack_from("<14f", b, o + 36)
    want = [hs.exx[m], hs.eyy[m], hs.exy[m], hs.sxx[m], hs.syy[m], hs.sxy[m], hs.ro[m], hs.nu[m], hs.nd[m], 0, 0, 0,
            hs.et[m], hs.tk[m]]
    assert f14 == tuple(np.float32(want))
    assert struct.unpack_from("<q", b, o + 92)[0] == hs.bondm[3 * X * Y + m]
    assert struct.unpack_from("<3f", b, o + 100) == tuple(np.float32([hs.cp[m], hs.kt[m], hs.ht[m]]))
    # grid lines, then the first BC row
    assert struct.unpack_from("<%df" % X, b, curpos_grid) == tuple(np.float32(hs.gx))
    ob = curpos_grid + 4 * (X + Y)
    assert struct.unpack_from("<4f", b, ob) == tuple(np.float32(hs.bondv[1]))
    assert struct.unpack_from("<3q", b, ob + 16) == tuple(int(v) for v in hs.bondn[1])
    # marker 11 and the zero record after the last marker
    om = curpos_markers + 39 * 11
    rec = struct.unpack_from("<9f3b", b, om)
    assert rec[0] == hs.markx[11] and rec[1] == hs.marky[11] and rec[2] == hs.markk[11]
    assert rec[5] == hs.markpor[11] and rec[7] == hs.marktmax[11] and rec[11] == hs.markt[11]
    assert struct.unpack_from("<9f3b", b, curpos_markers + 39 * M) == (0.0,) * 9 + (0, 0, 0)


def test_reader_rejects_other_sizes(tmp_path):
    p, hs = state_with_fields()
    path = tmp_path / "t.prn"
    prn.write(path, p, hs)
    b = bytearray(path.read_bytes())
    b[1] = 4  # a 4-byte long
    bad = tmp_path / "bad.prn"
    bad.write_bytes(bytes(b))
    with pytest.raises(Exception, match="variable sizes"):
        prn.read(bad, params=p)
    short = tmp_path / "short.prn"
    short.write_bytes(path.read_bytes()[:-100])
    with pytest.raises(Exception, match="ends early"):
        prn.read(short, params=p)


def test_immobile_rock_types_round_trip(tmp_path):
    """markim follows the header as one int per rock type (loadmart.c:443,
    715): the params' immobile table is written there and read back."""
    p, hs = state_with_fields(seed=5)
    p.immobile[6] = 1
    path = tmp_path / "t.prn"
    prn.write(path, p, hs, n0=1)
    b = path.read_bytes()
    # markim follows the header: 164 bytes, 6 doubles, rocknum (4), bondnum
    # (8), n0 (4), timesum (8)
    off = 164 + 48 + 4 + 8 + 4 + 8
    assert struct.unpack_from("<i", b, off - 24)[0] == p.ntypes
    want = [0] * p.ntypes
    want[6] = 1
    assert list(struct.unpack_from("<%di" % p.ntypes, b, off)) == want
    q, _, _ = prn.read(path, params=p)
    assert list(q.immobile)[:p.ntypes] == want


def test_magnetisation_bytes(tmp_path):
    """markmg_old and markmg_time ride in the two bytes before the type, each
    cast to char as the saver does (loadmart.c:805-806): the state exactly,
    the time in Ma truncated toward zero and clamped to a byte."""
    p, hs = state_with_fields(seed=4)
    rng = np.random.default_rng(2)
    hs.markmg_old[:] = rng.integers(0, 4, hs.marknum)
    hs.markmg_time[:] = rng.uniform(-1.0, 5.0, hs.marknum)
    hs.markmg_time[:3] = (-1.0, 300.0, 2.99)
    path = tmp_path / "t.prn"
    prn.write(path, p, hs, n0=1)
    _, got, _ = prn.read(path, params=p)
    assert np.array_equal(got.markmg_old, hs.markmg_old)
    assert np.array_equal(got.markmg_time, np.clip(np.trunc(hs.markmg_time), -128, 127))
    assert tuple(got.markmg_time[:3]) == (-1.0, 127.0, 2.0)

