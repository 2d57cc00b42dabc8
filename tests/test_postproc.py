"""Post-processing in the oracle and the host writers (CPU): the
visualisation grid of i2jmart's saver (loadjmart.c:619-1197), the text files
saver() and markersave() write, and grain() (grain.c:4-156).

Pins: a uniform field comes back uniform (markers and the node fallback); a
lone marker reaches exactly the four pixels around it, with the bilinear
weights saver() gives them; a rock-type half-space splits at the boundary
pixel; the gridmod fill takes the nearest ring's type; the text writer's
run-length code expands back to the grid; markersave's four files.  grain():
growing over dt twice equals growing once over 2 dt (the law d^n + k t is a
semigroup), nothing grows before the impact or above the olivine liquidus,
the pallasite sector is the annulus sector the reference names.
"""
import os

import numpy as np
import pytest

from i2elvis_planet_amd import abi, configs, markersave, visual_write
from oracle import oracle as orc


def spec(t, xr=41, yr=31, gridmod=0, box=(0.0, 1.0, 0.0, 1.0)):
    return abi.VisualSpec(type=t, gridmod=gridmod, xresol=xr, yresol=yr, xbeg=box[0], xend=box[1], ybeg=box[2],
                          yend=box[3])


@pytest.fixture(scope="module")
def c1():
    p, hs, _ = configs.make("c1")
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    o.step()
    return p, hs


def ulps(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))


def test_uniform_temperature_comes_back_uniform(c1):
    p, hs = c1
    h = hs.copy()
    h.markk[:] = np.float32(500.0)
    h.tk[:] = 500.0
    val, _ = orc.Oracle(p, h).visual(spec(1))
    # float sums of w * 500 and of w over the ~10-40 markers of a pixel, each
    # addition rounding by up to half an ulp: within 32 ulps of 500, and
    # unbiased
    assert ulps(val, np.float32(500.0)).max() <= 32
    assert abs(float(np.mean(val.astype(np.float64))) - 500.0) < 2e-5


def test_lone_marker_reaches_its_four_pixels():
    p, hs, _ = configs.make("c1")
    h = hs.copy()
    L = p.xsize
    xr, yr = 11, 9
    dx, dy = L / (xr - 1), L / (yr - 1)
    # every marker outside the window but one, at (3.25, 2.6) pixel spacings
    h.markx[:] = np.float32(-1.0)
    x0, y0 = np.float32(3.25 * dx), np.float32(2.6 * dy)
    h.markx[0], h.marky[0], h.markt[0], h.markk[0] = x0, y0, 6, np.float32(700.0)
    val, _ = orc.Oracle(p, h).visual(spec(1, xr, yr))
    ea = (float(x0) - 0.0) / dx + 1.0
    m1 = int(ea)
    ea -= m1
    na = float(y0) / dy + 1.0
    m2 = int(na)
    na -= m2
    v = val.reshape(xr, yr)
    got = {(i, j) for i in range(xr) for j in range(yr) if v[i, j] == np.float32(700.0)}
    assert got == {(m1 - 1, m2 - 1), (m1 - 1, m2), (m1, m2 - 1), (m1, m2)}
    # every other pixel interpolates the node T (the fixture's 150 K)
    rest = [v[i, j] for i in range(xr) for j in range(yr) if (i, j) not in got]
    assert max(abs(float(r) - 150.0) for r in rest) < 1e-3
    # the same marker alone in a rock-type grid: exactly those pixels carry its type
    rk, _ = orc.Oracle(p, h.copy()).visual(spec(5, xr, yr))
    r = rk.reshape(xr, yr)
    assert {(i, j) for i in range(xr) for j in range(yr) if r[i, j] == 6} == got
    assert (r == -1).sum() == xr * yr - 4
    # gridmod 1: the ring around them takes the type too
    rk, _ = orc.Oracle(p, h.copy()).visual(spec(5, xr, yr, gridmod=1))
    r = rk.reshape(xr, yr)
    ring = {(i, j) for i in range(m1 - 2, m1 + 2) for j in range(m2 - 2, m2 + 2) if 0 <= i < xr and 0 <= j < yr}
    assert {(i, j) for i in range(xr) for j in range(yr) if r[i, j] == 6} == ring


def test_rock_half_space_splits_at_the_boundary():
    p, hs, _ = configs.make("c1")
    h = hs.copy()
    h.markt[:] = np.where(h.markx < np.float32(0.5 * p.xsize), 6, 7).astype(np.int8)
    xr, yr = 21, 21
    rk, _ = orc.Oracle(p, h).visual(spec(5, xr, yr))
    r = rk.reshape(xr, yr)
    xs = np.arange(xr) * p.xsize / (xr - 1)
    for i in range(xr):
        if xs[i] < 0.45 * p.xsize:
            assert (r[i] == 6).all(), i
        elif xs[i] > 0.55 * p.xsize:
            assert (r[i] == 7).all(), i


def test_refused_outputs():
    p, hs, _ = configs.make("c1")
    for t in (3, 9, 10, 11):
        with pytest.raises(RuntimeError, match="visual"):
            orc.Oracle(p, hs.copy()).visual(spec(t))


def parse(path, sp):
    toks = open(path).read().split()
    t, xr, yr = float(toks[0]), int(toks[1]), int(toks[2])
    assert (xr, yr) == (sp.xresol, sp.yresol)
    vals = toks[3:]
    if sp.type != 5:
        return t, np.array([float(v) for v in vals])
    out, k = [], 0
    while k < len(vals):
        if vals[k] == "-2":
            out += [int(vals[k + 2])] * int(vals[k + 1])
            k += 3
        else:
            out.append(int(vals[k]))
            k += 1
    return t, np.array(out)


def test_writer_round_trips(tmp_path, c1):
    p, hs = c1
    for t in (1, 5):
        sp = spec(t, 31, 23, gridmod=1)
        out, _ = orc.Oracle(p, hs.copy()).visual(sp)
        path = str(tmp_path / ("v%d.txt" % t))
        visual_write(path, sp, 3.15576e7 * 12.5, out)
        ty, got = parse(path, sp)
        assert ty == 12.5
        if t == 5:
            assert np.array_equal(got, out)
            assert " -2 " in open(path).read()  # runs of more than three were coded
        else:
            assert np.allclose(got, out, rtol=5e-5, atol=0)
            assert open(path).read().splitlines()[2].startswith("  1.")  # " % 5.4e" of a positive value


def test_markersave_files(tmp_path, c1):
    p, hs = c1
    melt = np.linspace(0, 1, hs.marknum)
    markersave(str(tmp_path), p, hs, melt)
    names = ["marker_melt.txt", "marker_comp.txt", "marker_magnet.txt", "marker_magnet_time.txt"]
    for n in names:
        lines = open(os.path.join(str(tmp_path), n)).read().splitlines()
        assert len(lines) == hs.marknum + 1
    first = open(os.path.join(str(tmp_path), "marker_comp.txt")).readline().split()
    assert float(first[0]) == pytest.approx(float(hs.markx[0]) / p.xsize, rel=1e-4)
    assert int(first[2]) == int(hs.markt[0])
    last = open(os.path.join(str(tmp_path), "marker_melt.txt")).read().splitlines()[-2].split()
    assert float(last[2]) == pytest.approx(1.0)


YEAR = 3.15576e7
MA = 1.000e+6 * 365.250 * 24.000 * 3600.000


def grain_state(n=2000, seed=3):
    p, hs, _ = configs.make("c1")
    rng = np.random.default_rng(seed)
    h = hs.copy()
    h.markt[:n] = rng.choice(np.array([5, 6, 25, 26, 7], np.int8), n)
    h.markk[:n] = rng.uniform(900.0, 2300.0, n).astype(np.float32)
    return p, h


def run_grain(p, h, gp, t, dt, gr):
    h.timesum, h.timestep = t, dt
    orc.Oracle(p, h).grain(gp, gr)


def test_grain_growth_is_a_semigroup():
    p, h = grain_state()
    gp = abi.GrainParams(growth_model=2, gr_init=1.0, start_time_ma=0.0, impact_time_ma=0.0)
    gr0 = np.full(h.marknum, 1.0, np.float32)
    a, b = gr0.copy(), gr0.copy()
    h1, h2 = h.copy(), h.copy()
    dt = 2.5e3 * YEAR
    run_grain(p, h1, gp, 5 * MA, dt, a)
    run_grain(p, h1, gp, 5 * MA + dt, dt, a)
    run_grain(p, h2, gp, 5 * MA, 2 * dt, b)
    grew = a != gr0
    assert grew.sum() > 100
    # float32 storage between the two halves: a few ulps
    assert np.max(np.abs(a.astype(np.float64) - b) / b) < 5e-7
    # above the olivine liquidus types 25/26 are reset, type 6 grows at any T
    hot = (h.markk[:2000] >= 2103.0) & np.isin(h.markt[:2000], [25, 26])
    assert np.all(a[:2000][hot] == np.float32(1.0))


def test_grain_waits_for_the_impact_and_emplaces_pallasite_in_the_sector():
    p, hs, _ = configs.make("c1")
    h = hs.copy()
    # markers on a ring of radius 150 km around the box centre, every degree
    # (at half degrees, off the sector's edges)
    n = 360
    ang = np.radians(np.arange(n) + 0.5)
    cx, cy = p.xsize / 2, p.ysize / 2
    h.markx[:n] = (cx + 150e3 * np.sin(ang)).astype(np.float32)
    h.marky[:n] = (cy - 150e3 * np.cos(ang)).astype(np.float32)
    h.markt[:n] = 6
    h.markk[:n] = np.float32(1200.0)
    gp = abi.GrainParams(growth_model=1, gr_init=3.0, start_time_ma=0.0, impact_time_ma=2.0)
    gr = np.full(h.marknum, 1.0, np.float32)
    before = h.copy()
    run_grain(p, h, gp, 1.0 * MA, 2.5e3 * YEAR, gr)  # before the impact: nothing happens
    assert np.all(gr == np.float32(1.0)) and np.array_equal(h.markt, before.markt)
    run_grain(p, h, gp, 2.1 * MA, 2.5e3 * YEAR, gr)  # inside the 0.45 Ma window
    # dist_x > 0, dist_y < 0 and asin(dist_x / r) in [35, 55] degrees: the
    # markers at 35..55 degrees from the box's upward direction (y grows down)
    deg = np.degrees(np.arcsin(np.sin(ang)))
    want = (np.sin(ang) > 0) & (-np.cos(ang) < 0) & (deg >= 35.0) & (deg <= 55.0)
    assert np.array_equal(h.markt[:n] == 5, want)
    assert want.sum() == 20
    assert np.all(gr[:n][want] > np.float32(3.0))  # reset to gr_init, then grown as FeS-olivine at 1200 K
