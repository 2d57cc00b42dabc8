"""GPU parity of the visualisation grid (i2e_visual, i2jmart's saver,
loadjmart.c:619-1197) against the oracle, through the C-ABI.

The device merges each pixel's markers in marker order and repeats saver()'s
float32 additions, so the grids are equal bit for bit for temperature, rock
types, density, conductivity and heat generation; log10 viscosity, log10
shear heating and the vorticity ratio pass through the device's log and pow,
which may round differently from glibc's: within 2 float ulps there.  The
context's markers are cell-sorted after a step, so these cases also check
that the pass follows the caller's marker order.
"""
import sys

import numpy as np
import pytest

from i2elvis_planet_amd import Context, abi, configs, default_solver
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

EXACT = (1, 5, 6, 7, 8)
ROUNDED = (2, 4, 12)


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def spec(t, xr, yr, gridmod=0, box=(0.0, 1.0, 0.0, 1.0)):
    return abi.VisualSpec(type=t, gridmod=gridmod, xresol=xr, yresol=yr, xbeg=box[0], xend=box[1], ybeg=box[2],
                          yend=box[3])


def ulps(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))


@pytest.fixture(scope="module")
def stepped():
    say("case visual c1 after two steps")
    p, hs, _ = configs.make("c1")
    o = default_solver()
    o.precision = 0
    ctx = Context(p, hs, solver=o)
    ctx.ronurecalc()
    ctx.step()
    ctx.step()
    got = ctx.download(hs.copy())
    yield p, ctx, got
    ctx.close()


@pytest.mark.parametrize("case", [
    dict(xr=81, yr=61),                                       # markstp 1
    dict(xr=17, yr=13, gridmod=2),                            # markstp 18: every 18th marker
    dict(xr=64, yr=50, gridmod=3, box=(0.2, 0.7, -0.1, 0.8)),  # a window (clamped below 0)
])
def test_visual_matches_oracle(stepped, case):
    p, ctx, got = stepped
    for t in EXACT + ROUNDED:
        sp = spec(t, case["xr"], case["yr"], case.get("gridmod", 0), case.get("box", (0.0, 1.0, 0.0, 1.0)))
        say("case visual type %d %s" % (t, case))
        gv, gmelt = ctx.visual(sp)
        ov, omelt = orc.Oracle(p, got.copy()).visual(sp)
        if t in EXACT:
            bad = int(np.sum(gv != ov))
            assert bad == 0, (t, bad)
        else:
            u = ulps(gv, ov)
            say("type %d: max ulps %d, differing pixels %d of %d" % (t, int(u.max()), int((u > 0).sum()), u.size))
            assert u.max() <= 2, t
        if t != 5:
            assert np.array_equal(gmelt, omelt)


def test_visual_rock_types_after_a_c3_like_mix(stepped):
    """Four rock types in stripes, sampled sparsely, with the gridmod fill."""
    p, ctx, got = stepped
    h = got.copy()
    h.markt[:] = (6 + (np.floor(h.markx / (p.xsize / 9)).astype(np.int64) % 2)).astype(np.int8)
    h.markt[h.marky < np.float32(0.2 * p.ysize)] = 0
    o = default_solver()
    o.precision = 0
    c2 = Context(p, h, solver=o)
    for xr, yr, gm in ((120, 90, 0), (31, 29, 4)):
        sp = spec(5, xr, yr, gm)
        gv, _ = c2.visual(sp)
        ov, _ = orc.Oracle(p, h.copy()).visual(sp)
        assert np.array_equal(gv, ov), (xr, yr, gm, int(np.sum(gv != ov)))
    c2.close()


YEAR = 3.15576e7
MA = 1.000e+6 * 365.250 * 24.000 * 3600.000


@pytest.mark.parametrize("model,t_ma", [(1, 1.0), (1, 2.1), (2, 0.5), (2, 2.3)])
def test_grain_matches_oracle(stepped, model, t_ma):
    """grain() on the device against the oracle, on the cell-sorted context
    (caller's order restored): marker types equal, grain sizes within one
    float ulp (pow, exp and asin of the device and of glibc)."""
    p, ctx, got = stepped
    say("case grain model %d t %.1f Ma" % (model, t_ma))
    rng = np.random.default_rng(7)
    h = got.copy()
    M = h.marknum
    h.markt[:] = rng.choice(np.array([0, 5, 6, 25, 26, 7], np.int8), M)
    h.markk[:] = rng.uniform(900.0, 2300.0, M).astype(np.float32)
    # a ring of markers through the pallasite sector (100-200 km from the centre)
    n = min(M, 720)
    ang = np.radians(np.arange(n) * 0.5 + 0.25)
    r = np.linspace(90e3, 210e3, n)
    h.markx[:n] = (p.xsize / 2 + r * np.sin(ang)).astype(np.float32)
    h.marky[:n] = (p.ysize / 2 - r * np.cos(ang)).astype(np.float32)
    h.timesum, h.timestep = t_ma * MA, 2.5e3 * YEAR
    gp = abi.GrainParams(growth_model=model, gr_init=1.5, start_time_ma=0.2, impact_time_ma=2.0)
    gr0 = rng.uniform(1.0, 50.0, M).astype(np.float32)
    o = default_solver()
    o.precision = 0
    c2 = Context(p, h, solver=o)
    c2.set_marker_sort(1)
    c2.sort_markers()
    c2.set_scalars(h.timestep, h.timesum)
    g_gpu = gr0.copy()
    c2.grain(gp, g_gpu)
    d = c2.download(h.copy())
    c2.close()
    ho = h.copy()
    g_orc = gr0.copy()
    orc.Oracle(p, ho).grain(gp, g_orc)
    assert np.array_equal(d.markt, ho.markt)
    u = ulps(g_gpu, g_orc)
    say("grain: changed %d, max ulps %d" % (int(np.sum(g_orc != gr0)), int(u.max())))
    assert u.max() <= 1
    if t_ma > 2.0:
        assert np.sum(ho.markt != h.markt) > 0  # the sector was emplaced
