"""GPU parity of libi2e (through its C-ABI) against the CPU oracle on the same
seeded inputs.  Integer/index results bit-exact, FP64 fields within stated
relative L2 tolerances."""
import math
import sys

import numpy as np
import pytest

from i2elvis_planet_amd import Context, DeviceArray, configs, default_solver

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / n if n > 0 else np.linalg.norm(a - b)


def fp64_opts(rtol=1e-12):
    o = default_solver()
    o.precision = 0
    o.stokes_rtol = rtol
    o.heat_rtol = 1e-13
    o.grav_rtol = 1e-13
    return o


def small_cfg(name, **kw):
    cfg = dict(configs.CONFIGS[name], **kw)
    key = "_t_" + name
    configs.CONFIGS[key] = cfg
    try:
        return configs.make(key)
    finally:
        del configs.CONFIGS[key]


def rotation_field(hs, om=2e-13, noise=0.0, seed=0):
    X, Y = hs.xnumx, hs.ynumy
    c = hs.gx[-1] / 2
    xm = (hs.gx[:-1] + hs.gx[1:]) / 2
    ym = (hs.gy[:-1] + hs.gy[1:]) / 2
    vx = np.zeros((X, Y))
    vy = np.zeros((X, Y))
    vx[:, :Y - 1] = -om * (ym[None, :] - c) + 0 * hs.gx[:, None]
    vy[:X - 1, :] = om * (xm[:, None] - c) + 0 * hs.gy[None, :]
    if noise:
        rng = np.random.default_rng(seed)
        vx[:, :Y - 1] *= 1 + noise * rng.standard_normal((X, Y - 1))
        vy[:X - 1, :] *= 1 + noise * rng.standard_normal((X - 1, Y))
    hs.vx[:] = vx.ravel()
    hs.vy[:] = vy.ravel()


# ------------------------------------------------------------------ markers
@pytest.mark.parametrize("name,markmod", [("c1", 2), ("c3", 2), ("c1", 1)])
def test_movemark_bitexact(orc, name, markmod):
    say("case movemark", name, "markmod", markmod)
    p, hs, _ = configs.make(name)
    p.markmod = markmod  # 1: Euler, 2: RK4 (markmart.c:74-97)
    rotation_field(hs, noise=0.3)
    rng = np.random.default_rng(3)
    hs.pr[:] = 1e6 * rng.random(hs.N)
    hs.ga[:] = rng.random(hs.N)
    hs.timestep = 0.05 * (hs.gx[1] - hs.gx[0]) / np.abs(hs.vx).max()
    ref = hs.copy()
    ctx = Context(p, hs, solver=fp64_opts())
    ctx.set_scalars(hs.timestep, 0.0)
    c0 = ctx.type_counts()
    ctx.movemark()
    got = ctx.download(hs.copy())
    gpu_counts = ctx.type_counts()
    ctx.close()
    assert np.array_equal(c0, orc.Oracle(p, hs).type_counts())
    orc.Oracle(p, ref).movemark()
    assert np.array_equal(got.markx, ref.markx)
    assert np.array_equal(got.marky, ref.marky)
    assert np.array_equal(got.markk, ref.markk)
    assert np.array_equal(got.markp, ref.markp)
    np.testing.assert_allclose(got.markv, ref.markv, rtol=2e-7)
    # cell indices and per-cell type counts, bit-exact: the GPU's own
    # histogram (k_type_counts) against the oracle's of the oracle's markers
    cb = orc.Oracle(p, ref).type_counts()
    assert np.array_equal(gpu_counts, cb)
    inside = (got.markx > 0) & (got.marky > 0) & (got.markx < p.xsize) & (got.marky < p.ysize)
    assert gpu_counts.sum() == np.count_nonzero(inside)


@pytest.mark.parametrize("name,kw", [("c1", {}), ("c2", dict(n=81, size=160e3, radius=40e3)),
                                     ("c3", dict(n=101, size=120e3, radius=30e3))])
def test_ronurecalc_parity(orc, name, kw):
    say("case ronurecalc", name)
    p, hs, _ = small_cfg(name, **kw)
    rng = np.random.default_rng(4)
    # temperatures across solidus/liquidus and the 700 K sintering switch
    hs.markk[:] = rng.uniform(150, 1900, hs.marknum).astype(np.float32)
    hs.pr[:] = 1e7 * rng.random(hs.N)
    hs.timestep = 7.9e10
    hs.timesum = 3e12
    ref = hs.copy()
    ctx = Context(p, hs, solver=fp64_opts())
    ctx.set_scalars(hs.timestep, hs.timesum)
    ctx.ronurecalc()
    got = ctx.download(hs.copy())
    ctx.close()
    orc.Oracle(p, ref).ronurecalc()
    for f in ("nu", "nd", "ro", "cp", "kt", "ht", "tk", "et"):
        assert rel(getattr(got, f), getattr(ref, f)) < 1e-10, f
    assert np.array_equal(got.markt, ref.markt)
    assert np.array_equal(got.markk, ref.markk)
    np.testing.assert_allclose(got.markpor, ref.markpor, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(got.marktmax, ref.marktmax, rtol=0)


def shuffle_markers(hs, rng):
    perm = rng.permutation(hs.marknum)
    for f in ("markx", "marky", "markk", "markv", "markp", "markpor", "marktmax", "markt", "markmg_old",
              "markmg_time"):
        a = getattr(hs, f)
        a[:] = a[perm]


@pytest.mark.parametrize("dynamo", [0, 1])
def test_ronurecalc_soft_turbulence_and_magnetisation(orc, dynamo):
    """ronurecalc with the soft-turbulence conductivity (markmart.c:508-610)
    and the magnetisation states (markmart.c:651-744) on: mixed silicate,
    molten and iron types over 150-1900 K (both Curie temperatures, both melt
    windows), random previous states and times, markers shuffled and then
    sorted by cell on the device (the temperature difference is taken
    against the caller's marker 0 wherever the sort put it)."""
    say("case ronurecalc soft_turb magnet dynamo", dynamo)
    p, hs, _ = small_cfg("c3", n=101, size=120e3, radius=30e3)
    p.soft_turb, p.core_form_time, p.magnet, p.dynamo = 1, 0.0, 1, dynamo
    p.fe_ts, p.fe_tl = 1400.0, 1450.0
    p.k_cutoff = 1e12
    p.meltable[5] = 1
    for t in (5, 6, 7, 17, 25, 26):
        p.rock[t].bb = 3e-5
    rng = np.random.default_rng(9)
    inside = hs.markt != configs.AIR
    pick = rng.choice(np.array([5, 6, 7, 17, 25, 26], np.int8), size=hs.marknum)
    hs.markt[inside] = pick[inside]
    hs.markk[:] = rng.uniform(150, 1900, hs.marknum).astype(np.float32)
    hs.markmg_old[:] = rng.integers(0, 4, hs.marknum)
    hs.markmg_time[:] = rng.uniform(0, 5, hs.marknum)
    hs.pr[:] = 1e7 * rng.random(hs.N)
    hs.ga[:] = rng.random(hs.N)
    hs.timestep = 7.9e10
    hs.timesum = 3e12
    shuffle_markers(hs, rng)
    ref = hs.copy()
    ctx = Context(p, hs, solver=fp64_opts())
    ctx.set_scalars(hs.timestep, hs.timesum)
    ctx.sort_markers()
    ctx.ronurecalc()
    got = ctx.download(hs.copy())
    ctx.close()
    orc.Oracle(p, ref).ronurecalc()
    assert np.array_equal(got.markt, ref.markt)
    assert np.array_equal(got.markk, ref.markk)
    assert np.array_equal(got.markmg_old, ref.markmg_old)
    assert np.array_equal(got.markmg_time, ref.markmg_time)
    for f in ("nu", "nd", "ro", "cp", "kt", "ht", "tk", "et"):
        assert rel(getattr(got, f), getattr(ref, f)) < 1e-10, f
    # the model changed the conductivity of molten and iron nodes
    p.soft_turb = 0
    off = hs.copy()
    orc.Oracle(p, off).ronurecalc()
    assert rel(off.kt, ref.kt) > 1e-3


@pytest.mark.parametrize("hydrate", [True, False])
def test_hydration_statistics_match_the_oracle(orc, hydrate):
    """impact()'s hydration and body-statistics sweep (impact.c:3022-3228) on
    the device state at 101^2 with every silicate type and peak temperatures
    across all thresholds, markers tile-sorted on the device: switched types,
    counts, fractions, maximum temperatures exact; the mean temperatures to
    1e-14 (sums in another order)."""
    say("case hydration", hydrate)
    p, hs, _ = small_cfg("c2", n=101, size=160e3, radius=40e3)
    rng = np.random.default_rng(12)
    hs.markt[:] = rng.choice(np.array([0, 5, 6, 7, 25, 26], np.int8), hs.marknum)
    hs.marktmax[:] = rng.uniform(100, 1900, hs.marknum).astype(np.float32)
    hs.markk[:] = rng.uniform(150, 1800, hs.marknum).astype(np.float32)
    hs.tk[:] = np.where(rng.random(hs.N) < 0.4, p.tmp_ambient, rng.uniform(150, 1700, hs.N))
    ref = hs.copy()
    ctx = Context(p, hs, solver=fp64_opts())
    ctx.sort_markers()
    g = ctx.hydration(t_sol=1416.193, t_liq=1973.0, hydrate=hydrate)
    got = ctx.download(hs.copy())
    ctx.close()
    o = orc.Oracle(p, ref).hydration(t_sol=1416.193, t_liq=1973.0, hydrate=hydrate)
    assert np.array_equal(got.markt, ref.markt)
    for k in ("count_tot", "count_toohot", "count_body_nodes"):
        assert getattr(g, k) == getattr(o, k), k
    for k in ("sol", "liq", "hydrous", "primitive", "n2co2", "cocl", "h2o", "phyllo1", "phyllo2", "phyllo3",
              "phyllo4", "perco", "melt1", "melt2"):
        a, b = getattr(g, k + "_frac"), getattr(o, k + "_frac")
        assert a == b or (math.isnan(a) and math.isnan(b)), k
    assert g.maxtk == o.maxtk and g.t_max_body == o.t_max_body
    assert g.meantk == pytest.approx(o.meantk, rel=1e-14)
    if hydrate:
        assert o.count_tot > 1000
        assert g.t_mean_body == pytest.approx(o.t_mean_body, rel=1e-14)
    else:
        assert math.isnan(g.t_mean_body) and math.isnan(o.t_mean_body)


@pytest.mark.parametrize("sort", [False, True])
def test_immobile_types_match_the_oracle(orc, sort):
    """Immobile rock types (markim): movemark leaves their markers where they
    are, bit-exact, and ronurecalc keeps their temperatures out of the node
    temperature (the scatter's IMM variant with its own weight sum), on the
    grid-stride and the tiled scatter."""
    say("case immobile, sorted", sort)
    p, hs, _ = small_cfg("c2", n=81, size=160e3, radius=40e3)
    p.immobile[configs.IRON] = 1
    rotation_field(hs, noise=0.3)
    rng = np.random.default_rng(15)
    hs.markk[:] = rng.uniform(150, 1500, hs.marknum).astype(np.float32)
    hs.pr[:] = 1e6 * rng.random(hs.N)
    hs.timestep = 0.05 * (hs.gx[1] - hs.gx[0]) / np.abs(hs.vx).max()
    hs.timesum = 3e12
    ref = hs.copy()
    ctx = Context(p, hs, solver=fp64_opts())
    ctx.set_scalars(hs.timestep, hs.timesum)
    if sort:
        ctx.sort_markers()
    ctx.movemark()
    ctx.ronurecalc()
    got = ctx.download(hs.copy())
    ctx.close()
    o = orc.Oracle(p, ref)
    o.movemark()
    o.ronurecalc()
    iron = ref.markt == configs.IRON
    assert iron.sum() > 100
    assert np.array_equal(got.markx, ref.markx) and np.array_equal(got.marky, ref.marky)
    assert np.array_equal(got.markx[iron], hs.markx[iron])
    assert np.array_equal(got.markt, ref.markt) and np.array_equal(got.markk, ref.markk)
    for f in ("nu", "nd", "ro", "cp", "kt", "ht", "tk", "et"):
        assert rel(getattr(got, f), getattr(ref, f)) < 1e-10, f


# ------------------------------------------------------------------ operators
def _dev(x):
    return DeviceArray(x)


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_operator_apply_matches_assembled_rows(orc, name):
    say("case operators", name)
    p, hs, _ = configs.make(name)
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    hs.timestep = p.maxtmstep
    o.gravity()
    ctx = Context(p, hs, solver=fp64_opts())
    ctx.set_scalars(hs.timestep, hs.timesum)
    rng = np.random.default_rng(5)
    # Stokes rows: scale unknowns like the solution (P ~ 1e7, v ~ 1e-10)
    M = o.stokes_matrix()
    x = rng.standard_normal(3 * hs.N)
    x[0::3] *= 1e7
    x[1::3] *= 1e-10
    x[2::3] *= 1e-10
    y_ref = M.matvec(x)
    xd, yd = _dev(x), _dev(np.zeros_like(x))
    ctx.apply("stokes", xd.ptr, yd.ptr)
    y = yd.numpy()
    # row-wise bound relative to sum |a_ij x_j|
    A = M.scipy()
    scale = abs(A) @ np.abs(x)
    assert np.all(np.abs(y - y_ref) <= 1e-13 * scale + 1e-300)
    # gravity rows
    G = orc.Matrix(hs.N, 12)
    import ctypes as C
    assert orc.lib().orc_gravity_build(C.byref(p), C.byref(hs.cstruct()), C.byref(G.c)) == 0
    xg = rng.standard_normal(hs.N)
    yd = _dev(np.zeros(hs.N))
    ctx.apply("gravity", _dev(xg).ptr, yd.ptr)
    scale = abs(G.scipy()) @ np.abs(xg)
    assert np.all(np.abs(yd.numpy() - G.matvec(xg)) <= 1e-13 * scale)
    # heat rows at this timestep
    H = orc.Matrix(hs.N, 12)
    sol1 = np.zeros(hs.N)
    assert orc.lib().orc_heat_build(C.byref(p), C.byref(hs.cstruct()), sol1.ctypes.data_as(C.POINTER(C.c_double)),
                                    C.byref(H.c)) == 0
    xh = 150 + rng.standard_normal(hs.N)
    yd = _dev(np.zeros(hs.N))
    ctx.apply("heat", _dev(xh).ptr, yd.ptr)
    scale = abs(H.scipy()) @ np.abs(xh)
    assert np.all(np.abs(yd.numpy() - H.matvec(xh)) <= 1e-13 * scale)
    ctx.close()


# ------------------------------------------------------------------ solves
def test_viterate_parity_fp64(orc):
    say("case viterate c1")
    p, hs, _ = configs.make("c1")
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    ctx = Context(p, hs, solver=fp64_opts(1e-13))
    hs.timestep = p.maxtmstep
    o.maxvelstep()
    ctx.set_scalars(hs.timestep, hs.timesum)
    ctx.viterate()
    st = ctx.stats()
    got = ctx.download(hs.copy())
    ctx.close()
    o.viterate()
    say("stokes its", st.stokes_iters, "relres", st.stokes_relres, "grav its", st.grav_iters)
    assert st.stokes_relres <= 1e-13
    assert got.gxkoef == pytest.approx(hs.gxkoef, rel=1e-13)
    for f in ("gp", "vx", "vy", "exx", "eyy", "exy", "sxy"):
        assert rel(getattr(got, f), getattr(hs, f)) < 1e-8, f
    # pressure is fixed only through the anchor cell: compare it in L2 too
    assert rel(got.pr, hs.pr) < 1e-8


@pytest.mark.parametrize("orth", [0, 1])
def test_viterate_parity_fp64_orthogonalisation(orc, orth):
    """Both Gram-Schmidt variants of the FGMRES solvers (lagged CGS2, CGS with
    DGKS reorthogonalisation) reach the direct solve on the 51^2 case."""
    say("case viterate c1 orth", orth)
    p, hs, _ = configs.make("c1")
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    opts = fp64_opts(1e-13)
    opts.orth = orth
    ctx = Context(p, hs, solver=opts)
    hs.timestep = p.maxtmstep
    o.maxvelstep()
    ctx.set_scalars(hs.timestep, hs.timesum)
    ctx.viterate()
    st = ctx.stats()
    got = ctx.download(hs.copy())
    ctx.close()
    o.viterate()
    say("stokes its", st.stokes_iters, "relres", st.stokes_relres)
    assert st.stokes_relres <= 1e-13
    for f in ("gp", "vx", "vy", "pr"):
        assert rel(getattr(got, f), getattr(hs, f)) < 1e-8, f


def test_titerate_parity_fp64(orc):
    say("case titerate c1")
    p, hs, _ = configs.make("c1")
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    hs.timestep = p.maxtmstep
    o.maxvelstep()
    o.viterate()
    hs.timesum = 5e12  # switches shear heating on
    ref = hs.copy()
    ctx = Context(p, hs, solver=fp64_opts())
    ctx.set_scalars(p.maxtmstep, hs.timesum)
    ctx.maxvelstep()
    ctx.titerate()
    got = ctx.download(hs.copy())
    ctx.close()
    oo = orc.Oracle(p, ref)
    ref.timestep = p.maxtmstep
    oo.maxvelstep()
    oo.titerate()
    assert got.timestep == ref.timestep
    assert rel(got.tk, ref.tk) < 1e-10
    np.testing.assert_allclose(got.markk, ref.markk, rtol=1e-6)


@pytest.mark.parametrize("name,nsteps,sort_every", [("c1", 3, 10), ("c1", 3, 1)])
def test_full_step_parity_fp64(orc, name, nsteps, sort_every):
    say("case steps", name, nsteps, "sort every", sort_every)
    p, hs, _ = configs.make(name)
    ref = hs.copy()
    ctx = Context(p, hs, solver=fp64_opts(1e-13))
    ctx.set_marker_sort(sort_every)
    ctx.ronurecalc()
    o = orc.Oracle(p, ref)
    o.ronurecalc()
    for k in range(nsteps):
        ctx.step()
        o.step()
        got = ctx.download(hs.copy())
        say("step", k, "dt", got.timestep, ref.timestep, "vx", rel(got.vx, ref.vx), "tk", rel(got.tk, ref.tk))
        assert got.timestep == pytest.approx(ref.timestep, rel=1e-10)
        for f in ("vx", "vy", "pr", "tk", "gp"):
            assert rel(getattr(got, f), getattr(ref, f)) < 1e-8, (k, f)
        assert np.array_equal(got.markt, ref.markt)
        assert np.array_equal(o.type_counts(), ctx.type_counts())
    ctx.close()


def test_full_step_parity_201_fp64(orc):
    """One full FP64 step of the 201^2 / 640k-marker setup (Arrhenius,
    porosity and sintering, shear heating) against the oracle's step: dt,
    velocity, pressure, temperature and gravity potential, marker types and
    per-cell type counts.  The 201^2 Stokes system stalls near 5e-12 in FP64:
    solved to 3e-11."""
    say("case step c2")
    p, hs, _ = configs.make("c2")
    ref = hs.copy()
    ctx = Context(p, hs, solver=fp64_opts(3e-11))
    ctx.ronurecalc()
    ctx.step()
    got = ctx.download(hs.copy())
    counts = ctx.type_counts()
    ctx.close()
    o = orc.Oracle(p, ref)
    o.ronurecalc()
    o.step()
    say("dt", got.timestep, ref.timestep, "vx", rel(got.vx, ref.vx), "tk", rel(got.tk, ref.tk))
    assert got.timestep == pytest.approx(ref.timestep, rel=1e-8)
    for f in ("vx", "vy", "pr", "tk", "gp"):
        assert rel(getattr(got, f), getattr(ref, f)) < 1e-8, f
    assert np.array_equal(got.markt, ref.markt)
    np.testing.assert_allclose(got.markpor, ref.markpor, rtol=1e-6, atol=1e-7)
    assert np.array_equal(counts, o.type_counts())


def test_marker_sort_is_invisible_at_the_boundary(orc):
    """The cell sort permutes the device markers only: after sorting, the
    device order is nondecreasing in the tile-major cell key, download returns the
    caller's order bit for bit, and advection from the sorted slots gives
    per marker exactly what it gives unsorted."""
    import ctypes as C
    from i2elvis_planet_amd import load_library
    say("case marker sort")
    p, hs, _ = configs.make("c1")
    rotation_field(hs, noise=0.3)
    rng = np.random.default_rng(9)
    perm = rng.permutation(hs.marknum)  # start from a shuffled order
    for f in ("markx", "marky", "markk", "markv", "markp", "markpor", "marktmax", "markt"):
        getattr(hs, f)[:] = getattr(hs, f)[perm]
    hs.timestep = 0.05 * (hs.gx[1] - hs.gx[0]) / np.abs(hs.vx).max()
    outs = []
    for sort in (False, True):
        ctx = Context(p, hs.copy(), solver=fp64_opts())
        ctx.set_scalars(hs.timestep, 0.0)
        if sort:
            ctx.sort_markers()
            L = load_library()
            xs = np.empty(hs.marknum, np.float32)
            ys = np.empty(hs.marknum, np.float32)
            orig = np.empty(hs.marknum, np.int32)
            for a, name in ((xs, "markx"), (ys, "marky"), (orig, "markorig")):
                assert L.i2e_copy_to_host(a.ctypes.data_as(C.c_void_p), C.c_void_p(ctx.device_ptr(name)),
                                          a.nbytes) == 0
            i = np.clip(np.searchsorted(hs.gx, xs.astype(np.float64), side="right") - 1, 0, hs.xnumx - 2)
            j = np.clip(np.searchsorted(hs.gy, ys.astype(np.float64), side="right") - 1, 0, hs.ynumy - 2)
            # tile-major key: 8 x 8 cells per tile, tiles and the cells in
            # a tile column-major
            T, ntj = 8, (hs.ynumy - 1 + 7) // 8
            key = (((i // T) * ntj + j // T) * T + i % T) * T + j % T
            assert np.all(np.diff(key) >= 0)
            assert np.array_equal(np.sort(orig), np.arange(hs.marknum))
            assert np.array_equal(xs, hs.markx[orig])
            back = ctx.download(hs.copy())
            assert np.array_equal(back.markx, hs.markx) and np.array_equal(back.markt, hs.markt)
        ctx.movemark()
        outs.append(ctx.download(hs.copy()))
        ctx.close()
    a, b = outs
    for f in ("markx", "marky", "markk", "markv", "markp", "markpor", "marktmax", "markt"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_type_counts_after_step_501(orc):
    """Per-cell marker-type counts after one full FP64 step of the 501^2 /
    6.25M-marker setup: the GPU histogram equals the oracle's after the
    oracle's own step (PARDISO or SuperLU solves), bit-exact."""
    say("case type counts after a 501^2 step")
    p, hs, _ = configs.make("c3")
    ref = hs.copy()
    # the 501^2 FP64 solve stagnates near 1.5e-10; 1e-9 leaves marker
    # displacements (~60 m) far below one float32 ulp of difference
    ctx = Context(p, hs, solver=fp64_opts(1e-9))
    ctx.ronurecalc()
    ctx.step()
    gpu_counts = ctx.type_counts()
    got = ctx.download(hs.copy())
    ctx.close()
    o = orc.Oracle(p, ref)
    o.ronurecalc()
    o.step()
    say("dt", got.timestep, ref.timestep, "vx", rel(got.vx, ref.vx))
    cb = o.type_counts()
    assert gpu_counts.sum() == cb.sum()
    assert np.array_equal(gpu_counts, cb)
    # dt = maxxystep dx / max|v| follows the solve's velocity maximum
    assert got.timestep == pytest.approx(ref.timestep, rel=1e-6)


def test_mixed_precision_error_vs_fp64(orc):
    say("case mixed vs fp64 c1")
    p, hs, _ = configs.make("c1")
    a = Context(p, hs.copy(), solver=fp64_opts(1e-8))
    om = default_solver()
    b = Context(p, hs.copy(), solver=om)
    for c in (a, b):
        c.ronurecalc()
        c.step()
    ga = a.download(hs.copy())
    gb = b.download(hs.copy())
    sb = b.stats()
    a.close()
    b.close()
    assert sb.stokes_relres <= 1e-8
    for f in ("vx", "vy", "pr", "tk"):
        e = rel(gb.__dict__[f], ga.__dict__[f])
        say("mixed vs fp64", f, e)
        assert e < 1e-5, f


@pytest.mark.parametrize("fine", [1, 2, 3, 4])
def test_mixed_viterate_fine_level_variants(orc, fine):
    """Mixed precision with each smoother of the finest velocity level
    (matrix-free rows, FP32 stencil, FP64 stencil on FP32 iterates, FP32
    stencil in difference form) reaches the metric's 1e-8 and the oracle's
    direct solve at 201^2."""
    say("case mixed viterate c2 fine_stencil", fine)
    p, hs, _ = configs.make("c2")
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    hs.timestep = p.maxtmstep
    o.maxvelstep()
    opts = default_solver()
    opts.fine_stencil = fine
    ctx = Context(p, hs, solver=opts)
    ctx.set_scalars(hs.timestep, hs.timesum)
    ctx.viterate()
    st = ctx.stats()
    got = ctx.download(hs.copy())
    ctx.close()
    o.viterate()
    say("stokes its", st.stokes_iters, "relres", st.stokes_relres)
    assert st.stokes_relres <= 1e-8
    for f in ("vx", "vy", "pr"):
        e = rel(getattr(got, f), getattr(hs, f))
        say(f, e)
        assert e < 1e-5, f


@pytest.mark.parametrize("precision", [0, 1])
def test_autotuned_tile_shapes(orc, precision):
    """With autotune on, every coarse velocity level smoothed by tiles gets
    one of the four candidate shapes, a second context of the same size
    reuses the choice, and the Stokes solve still reaches the oracle's
    direct solve at 201^2 (FP64 to 1e-8, mixed to 1e-5 relative L2)."""
    say("case autotune precision", precision)
    p, hs, _ = configs.make("c2")
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    hs.timestep = p.maxtmstep
    o.maxvelstep()
    # the 201^2 system stagnates near 5e-12 in FP64; 3e-11 is reached
    opts = fp64_opts(3e-11) if precision == 0 else default_solver()
    opts.autotune = 1
    cand = {0: {(16, 16), (16, 8), (32, 8), (32, 16)}, 1: {(32, 16), (32, 8), (64, 8), (32, 32)}}[precision]
    shapes = []
    for k in range(2):
        ctx = Context(p, hs.copy(), solver=opts)
        ctx.set_scalars(hs.timestep, hs.timesum)
        ctx.viterate()
        st = ctx.stats()
        got = ctx.download(hs.copy())
        shapes.append(ctx.tuned_shapes())
        ctx.close()
    say("shapes", shapes[0], "stokes its", st.stokes_iters, "relres", st.stokes_relres)
    assert len(shapes[0]) >= 2 and shapes[0] == shapes[1]
    assert all((sx, sy) in cand for _, sx, sy in shapes[0])
    o.viterate()
    assert st.stokes_relres <= (3e-11 if precision == 0 else 1e-8)
    for f in ("vx", "vy", "pr"):
        assert rel(getattr(got, f), getattr(hs, f)) < (1e-8 if precision == 0 else 1e-5), f


def test_graph_replay_is_refused():
    """Graph replay of the Stokes preconditioner is switched off (a profiled
    run ended in a host fault inside the graph launch whose cause is not
    established): the option is refused with I2E_ERR_UNSUPPORTED, and the
    context keeps its previous, working options."""
    import ctypes as C
    from i2elvis_planet_amd import load_library
    say("case graph replay refused")
    p, hs, _ = configs.make("c1")
    ctx = Context(p, hs, solver=fp64_opts())
    opts = fp64_opts()
    opts.graphs = 1
    L = load_library()
    assert L.i2e_set_solver(ctx._h, C.byref(opts)) == 5
    assert b"graph" in L.i2e_last_error()
    ctx.ronurecalc()
    ctx.step()
    assert ctx.stats().stokes_relres <= 1e-12
    ctx.close()


def test_residual_diagnostics_parity(orc):
    """i2e_stokes_residuals against the oracle's restatement of viterate's
    diagnostics (movemart.c:1049-1176) on the same perturbed solution."""
    say("case residual diagnostics c1")
    import ctypes as C
    p, hs, _ = configs.make("c1")
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    hs.timestep = p.maxtmstep
    o.maxvelstep()
    o.viterate()
    x = np.zeros(3 * hs.N)
    x[0::3], x[1::3], x[2::3] = hs.pr, hs.vx, hs.vy
    rng = np.random.default_rng(12)
    x[1::3] *= 1 + 1e-4 * rng.standard_normal(hs.N)
    o.L.orc_stokes_post(C.byref(p), o._s(), x.ctypes.data_as(C.POINTER(C.c_double)))
    ref = o.stokes_residuals()
    ctx = Context(p, hs, solver=fp64_opts())
    ctx.set_scalars(hs.timestep, hs.timesum)
    got = ctx.stokes_residuals()
    ctx.close()
    say("gpu", got["stokes"], got["cont"], got["bond"], "oracle", ref["stokes"], ref["cont"], ref["bond"])
    for k in ("stoknum", "contnum", "bondnum", "vxmin", "vxmax", "vymin", "vymax", "prmin", "prmax", "numin", "numax"):
        assert got[k] == ref[k], k
    for k in ("stokes", "cont", "bond"):
        assert got[k] == pytest.approx(ref[k], rel=1e-10), k


def test_batch_matches_sequential():
    """Ensemble stepping (bodies on separate streams driven from separate host
    threads) gives the results of stepping each body alone: 16 bodies of
    the ensemble's 101^2-301^2 grids, two FP64 steps each."""
    from i2elvis_planet_amd import step_batch
    say("case batch vs sequential, 16 bodies")
    bodies = configs.ensemble(16)
    seq, bat = [], []
    for p, hs, _ in bodies:
        ctx = Context(p, hs.copy(), solver=fp64_opts(1e-10))
        ctx.ronurecalc()
        ctx.step()
        ctx.step()
        seq.append(ctx.download(hs.copy()))
        ctx.close()
    ctxs = [Context(p, hs.copy(), solver=fp64_opts(1e-10)) for p, hs, _ in bodies]
    for c in ctxs:
        c.ronurecalc()
    step_batch(ctxs, 2)
    for c, (p, hs, _) in zip(ctxs, bodies):
        bat.append(c.download(hs.copy()))
        c.close()
    for a, b in zip(seq, bat):
        assert a.timestep == pytest.approx(b.timestep, rel=1e-12)
        for f in ("vx", "vy", "pr", "tk"):
            assert rel(b.__dict__[f], a.__dict__[f]) < 1e-9, f
        assert np.array_equal(a.markt, b.markt)
        np.testing.assert_allclose(b.markx, a.markx, rtol=1e-6)
        np.testing.assert_allclose(b.marky, a.marky, rtol=1e-6)


def test_batch_bodies_match_the_oracle(orc):
    """Three small ensemble bodies (41^2-61^2) stepped together, each against
    the oracle's steps of the same body: dt, velocity, pressure and
    temperature as in the single-body step parity."""
    from i2elvis_planet_amd import step_batch
    say("case batch vs oracle, 3 bodies")
    bodies = configs.ensemble(3, nmin=41, nmax=61)
    # the Arrhenius bodies stall between 1e-13 and 2e-12; 1e-11 is reached
    ctxs = [Context(p, hs.copy(), solver=fp64_opts(1e-11)) for p, hs, _ in bodies]
    for c in ctxs:
        c.ronurecalc()
    step_batch(ctxs, 2)
    for c, (p, hs, _) in zip(ctxs, bodies):
        got = c.download(hs.copy())
        c.close()
        ref = hs.copy()
        o = orc.Oracle(p, ref)
        o.ronurecalc()
        o.step()
        o.step()
        say("body", p.xnumx, "dt", got.timestep, ref.timestep, "vx", rel(got.vx, ref.vx), "pr", rel(got.pr, ref.pr))
        assert got.timestep == pytest.approx(ref.timestep, rel=1e-10)
        for f in ("vx", "vy", "pr", "tk", "gp"):
            assert rel(getattr(got, f), getattr(ref, f)) < 1e-8, f
        assert np.array_equal(got.markt, ref.markt)


def test_viterate_parity_201_fp64(orc):
    """viterate at 201^2 (Arrhenius, porosity) against the oracle's direct solve."""
    say("case viterate c2")
    p, hs, _ = configs.make("c2")
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    hs.timestep = p.maxtmstep
    o.maxvelstep()
    # the 201^2 system stagnates near 5e-12 in FP64; 3e-11 is reached.  A
    # solve stopped just under 1e-10 left sxy at 1-5e-8 of the direct solve,
    # depending on the Krylov path: the field bounds need the margin.
    ctx = Context(p, hs, solver=fp64_opts(3e-11))
    ctx.set_scalars(hs.timestep, hs.timesum)
    ctx.viterate()
    st = ctx.stats()
    got = ctx.download(hs.copy())
    ctx.close()
    o.viterate()
    say("stokes its", st.stokes_iters, "relres", st.stokes_relres)
    assert st.stokes_relres <= 3e-11
    for f in ("gp", "vx", "vy", "pr", "exx", "sxy"):
        e = rel(getattr(got, f), getattr(hs, f))
        say(f, e)
        assert e < 1e-8, f


@pytest.mark.parametrize("melt_eta,dt", [(1e15, None), (1e13, 1e7)])
def test_viterate_high_contrast_fp64(orc, melt_eta, dt):
    """The molten 501^2 geometry at 101^2 (same grid step and blob density,
    core at 1800 K) with the melt viscosity raised until the GPU solve
    converges: contrast 1e8 at dt = maxtmstep, 1e10 at dt = 1e7 s (at 1e10
    and dt = maxtmstep the stabilisation term dominates the melt rows and the
    solve diverges; from 1e14 it diverges at every dt).  FP64 to the metric's
    1e-8 against the oracle's rows solved directly after Ruiz equilibration
    (the unscaled LU is not accurate enough at these contrasts): vx, vy, P
    within 1e-6 relative L2."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    say("case high contrast melt_eta", melt_eta, "dt", dt)
    p, hs, cfg = configs.make_hot("c3", n=101, size=120e3, radius=30e3, nblob=16)
    p.melt_eta = melt_eta
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    hs.timestep = dt or p.maxtmstep
    ctx = Context(p, hs.copy(), solver=fp64_opts(1e-8))
    ctx.set_scalars(hs.timestep, hs.timesum)
    ctx.viterate()
    st = ctx.stats()
    got = ctx.download(hs.copy())
    ctx.close()
    o.gravity()
    M = o.stokes_matrix()
    A = M.scipy().tocsr()
    r = np.ones(A.shape[0])
    c = np.ones(A.shape[1])
    B = A.copy()
    for _ in range(10):
        rs = np.sqrt(abs(B).max(axis=1).toarray().ravel())
        cs = np.sqrt(abs(B).max(axis=0).toarray().ravel())
        rs[rs == 0] = 1
        cs[cs == 0] = 1
        B = sp.diags(1 / rs) @ B @ sp.diags(1 / cs)
        r /= rs
        c /= cs
    lu = spla.splu(B.tocsc())
    y = lu.solve(r * M.b)
    for _ in range(2):
        y = y + lu.solve(r * M.b - B @ y)
    o.stokes_post(c * y)
    say("stokes its", st.stokes_iters, "relres", st.stokes_relres)
    assert st.stokes_relres <= 1e-8
    for f in ("vx", "vy", "pr"):
        e = rel(getattr(got, f), getattr(hs, f))
        say(f, e)
        assert e < 1e-6, f


@pytest.mark.parametrize("kind", ["general", "continuity", "both"])
def test_viterate_general_bc_rows(orc, kind):
    """Velocity BC rows beyond free slip at 51^2: several neighbours, a
    neighbour of the other component or two nodes in (the operator applies
    them exactly, the multigrid keeps the strongest adjacent neighbour), and
    continuity-type rows of an open right wall (movemart.c:863-879), against
    the oracle's direct solve: vx, vy, P to 1e-8."""
    import sys as _s
    from harness.bc import continuity_rows, general_rows
    say("case viterate bc", kind)
    p, hs, _ = configs.make("c1")
    if kind in ("general", "both"):
        general_rows(hs)
    if kind in ("continuity", "both"):
        continuity_rows(hs)
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    hs.timestep = p.maxtmstep
    o.maxvelstep()
    ctx = Context(p, hs.copy(), solver=fp64_opts(1e-12))
    ctx.set_scalars(hs.timestep, hs.timesum)
    ctx.viterate()
    st = ctx.stats()
    got = ctx.download(hs.copy())
    ctx.close()
    o.viterate()
    say("stokes its", st.stokes_iters, "relres", st.stokes_relres)
    assert st.stokes_relres <= 1e-12
    for f in ("vx", "vy", "pr"):
        e = rel(getattr(got, f), getattr(hs, f))
        say(f, e)
        assert e < 1e-8, f


def test_chained_temperature_bc_rows(orc):
    """Temperature BC rows that point at other BC rows (an insulating wall
    with its corners pointing at the wall): ronurecalc's projection matches
    the oracle's direct solve of tkrecalc's system, and a full FP64 step at
    51^2 the oracle's step; rows that reference each other in a cycle are
    refused."""
    from i2elvis_planet_amd import I2EError
    from harness.bc import chained_temperature_rows, cyclic_temperature_rows
    say("case chained temperature BC")
    p, hs, _ = configs.make("c1")
    chained_temperature_rows(hs)
    rng = np.random.default_rng(21)
    hs.markk[:] = rng.uniform(150, 900, hs.marknum).astype(np.float32)
    ref = hs.copy()
    ctx = Context(p, hs, solver=fp64_opts(1e-13))
    ctx.ronurecalc()
    got = ctx.download(hs.copy())
    o = orc.Oracle(p, ref)
    o.ronurecalc()
    assert rel(got.tk, ref.tk) < 1e-12
    N, Y = hs.N, hs.ynumy
    assert got.tk[0] == got.tk[1] == got.tk[Y + 1]  # corner = wall = inner node
    ctx.step()
    o.step()
    got = ctx.download(hs.copy())
    ctx.close()
    for f in ("vx", "vy", "pr", "tk"):
        assert rel(getattr(got, f), getattr(ref, f)) < 1e-8, f
    p2, h2, _ = configs.make("c1")
    cyclic_temperature_rows(h2)
    with pytest.raises(I2EError, match="status 5.*cycle"):
        Context(p2, h2, solver=fp64_opts())


def test_step_501_properties():
    """One full step of the 501^2 / 6.25M-marker setup: the Stokes solution
    satisfies the assembled rows to the metric's tolerance when checked
    independently in FP64 (operator apply on the downloaded fields), the
    residual diagnostics are finite, markers keep their count and types
    outside melting, and temperatures stay within physical bounds."""
    say("case step 501 properties")
    p, hs, _ = configs.make("c3")
    ctx = Context(p, hs.copy())
    ctx.ronurecalc()
    # the step's Stokes solve, checked before titerate changes the timestep
    ctx.set_scalars(p.maxtmstep, 0.0)
    ctx.maxvelstep()
    ctx.viterate()
    st = ctx.stats()
    got = ctx.download(hs.copy())
    assert st.stokes_relres <= 1e-8
    # independent residual: b - K x with x the downloaded fields
    n = 3 * hs.N
    x = np.zeros(n)
    x[0::3], x[1::3], x[2::3] = got.pr, got.vx, got.vy
    xd, yd, bd = DeviceArray(x), DeviceArray(n), DeviceArray(n)
    ctx.stokes_rhs(bd.ptr)
    ctx.apply("stokes", xd.ptr, yd.ptr)
    r = bd.numpy() - yd.numpy()
    relres = np.linalg.norm(r) / np.linalg.norm(bd.numpy())
    say("independent relres", relres)
    assert relres <= 1.5e-8
    d = ctx.stokes_residuals()
    assert all(np.isfinite(v) for v in d.values())
    # rest of the step
    ctx.set_scalars(p.maxtmstep, 0.0)
    ctx.maxvelstep()
    ctx.titerate()
    ctx.movemark()
    ctx.ronurecalc()
    got = ctx.download(hs.copy())
    ctx.close()
    assert got.marknum == hs.marknum
    assert np.array_equal(np.sort(np.unique(got.markt)), np.sort(np.unique(hs.markt)))
    assert np.all(got.tk >= 149.0) and np.all(got.tk < 2500.0)


def test_restart_from_prn_continues_the_run(tmp_path):
    """Two steps, the binary .prn checkpoint written and read back into a new
    context, one more step in both: the restart rounds node fields to float32
    (as the reference's saver does), so fields agree to 1e-5 and the markers
    to float32 noise, with identical types."""
    from i2elvis_planet_amd import prn
    say("case prn restart c1")
    p, hs, _ = configs.make("c1")
    opts = fp64_opts(1e-10)
    ctx = Context(p, hs, solver=opts)
    ctx.ronurecalc()
    for _ in range(2):
        ctx.step()
    mid = ctx.download(hs.copy())
    path = tmp_path / "c1_2.prn"
    prn.write(path, p, mid, n0=2)
    q, back, n0 = prn.read(path, params=p)
    assert n0 == 2 and back.marknum == mid.marknum
    ctx.step()
    want = ctx.download(hs.copy())
    ctx.close()
    ctx2 = Context(q, back, solver=opts)
    ctx2.step()
    got = ctx2.download(back.copy())
    ctx2.close()
    for f in ("vx", "vy", "pr", "tk"):
        e = rel(getattr(got, f), getattr(want, f))
        say(f, "restart vs continued rel L2 %.2e" % e)
        assert e < 1e-5, f
    assert np.array_equal(got.markt, want.markt)
    assert np.max(np.abs(got.markx - want.markx)) <= 1e-6 * p.xsize
    assert np.max(np.abs(got.markk - want.markk)) <= 1e-3
