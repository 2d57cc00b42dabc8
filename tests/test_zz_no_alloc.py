"""i2e_maxvelstep_noalloc (the no-allocation variant of i2e_maxvelstep)
against the context entry point, through the C-ABI: two calls in a row, the
second on a smaller grid with smaller velocities (a scratch left from the
first call would show), then a zero-velocity field and a rejected size."""
import ctypes as C
import sys

import numpy as np
import pytest

from i2elvis_planet_amd import Context, DeviceArray, configs, default_solver, load_library

pytestmark = pytest.mark.gpu


def case(n, scale, seed):
    cfg = dict(configs.CONFIGS["c1"], n=n, size=4e3 * (n - 1), radius=1e3 * (n - 1))
    p, hs = configs.make_params(cfg), configs.make_state(cfg, seed)
    rng = np.random.default_rng(seed)
    hs.vx[:] = scale * rng.standard_normal(hs.N)
    hs.vy[:] = 0.7 * scale * rng.standard_normal(hs.N)
    return p, hs


def noalloc(p, hs, dt0):
    L = load_library()
    vx, vy, ts = DeviceArray(hs.vx), DeviceArray(hs.vy), DeviceArray(np.array([dt0]))
    st = L.i2e_maxvelstep_noalloc(C.c_void_p(vx.ptr), C.c_void_p(vy.ptr), hs.xnumx, hs.ynumy, p.xsize, p.ysize,
                                  p.maxxystep, C.c_void_p(ts.ptr), None)
    assert st == 0, L.i2e_last_error()
    return ts.numpy()[0]


def reference(p, hs, dt0):
    o = default_solver()
    o.precision = 0
    ctx = Context(p, hs, solver=o)
    ctx.set_scalars(dt0, 0.0)
    ts = ctx.maxvelstep()
    ctx.close()
    return ts


def test_noalloc_maxvelstep_matches_the_context_entry_point():
    for n, scale, seed in ((61, 3e-8, 1), (41, 5e-9, 2)):
        print("case maxvelstep_noalloc", n, scale, file=sys.stderr, flush=True)
        p, hs = case(n, scale, seed)
        want = reference(p, hs, p.maxtmstep)
        got = noalloc(p, hs, p.maxtmstep)
        assert got == want and got < p.maxtmstep
    p, hs = case(41, 0.0, 3)
    assert noalloc(p, hs, p.maxtmstep) == p.maxtmstep == reference(p, hs, p.maxtmstep)


def test_noalloc_maxvelstep_rejects_sizes_beyond_its_maximum():
    L = load_library()
    ts = DeviceArray(np.array([1.0]))
    v = DeviceArray(16)
    st = L.i2e_maxvelstep_noalloc(C.c_void_p(v.ptr), C.c_void_p(v.ptr), 1, 16, 1.0, 1.0, 0.05, C.c_void_p(ts.ptr),
                                  None)
    assert st == 2 and b"outside" in L.i2e_last_error()
    st = L.i2e_maxvelstep_noalloc(C.c_void_p(v.ptr), C.c_void_p(v.ptr), 20000, 16, 1.0, 1.0, 0.05,
                                  C.c_void_p(ts.ptr), None)
    assert st == 2
    assert ts.numpy()[0] == 1.0
