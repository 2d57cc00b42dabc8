"""Pins of the CPU oracle against closed forms, brute force and structural
identities of the reference discretisation (the reference ships no golden
vectors or tests for this path)."""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from i2elvis_planet_amd import configs
from i2elvis_planet_amd.state import HostState


def small_state(n=13, size=48e3, mnum=2, seed=5):
    cfg = dict(configs.CONFIGS["c1"], n=n, size=size, mnum=mnum, radius=size / 4)
    configs.CONFIGS["_t"] = cfg
    try:
        p, hs, _ = configs.make("_t", seed=seed)
    finally:
        del configs.CONFIGS["_t"]
    return p, hs


def node_xy(hs):
    X, Y = hs.xnumx, hs.ynumy
    return np.repeat(hs.gx, Y), np.tile(hs.gy, X)


# ------------------------------------------------------------------ search / interpolation
def test_m1serch_matches_bruteforce(orc):
    L = orc.lib()
    rng = np.random.default_rng(1)
    g = np.cumsum(np.r_[0.0, rng.uniform(0.5, 2.0, 20)])
    dp = g.ctypes.data_as(C.POINTER(C.c_double))
    pts = np.r_[rng.uniform(-1, g[-1] + 1, 400), g, g[-1]]
    for x in pts:
        # largest m with g[m] <= x, clamped to [0, n-2] (markmart.c:1051-1067)
        bf = max(0, min(len(g) - 2, int(np.searchsorted(g, x, side="right")) - 1))
        assert L.orc_m1serch(dp, len(g), float(x)) == bf


def _interp(orc, p, hs, x, y):
    out = np.zeros(2)
    cs = hs.cstruct()
    orc.lib().orc_allinterv(C.byref(p), C.byref(cs), x, y, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def test_allinterv_reproduces_affine_fields(orc):
    p, hs = small_state()
    X, Y = hs.xnumx, hs.ynumy
    xm = (hs.gx[:-1] + hs.gx[1:]) / 2
    ym = (hs.gy[:-1] + hs.gy[1:]) / 2
    # Vx at (gx[m1], ym[m2]); Vy at (xm[m1], gy[m2])
    fx = lambda x, y: 1.5e-9 + 2e-14 * x - 3e-14 * y
    fy = lambda x, y: -0.7e-9 + 1e-14 * x + 4e-14 * y
    vx = np.zeros((X, Y))
    vy = np.zeros((X, Y))
    vx[:, :Y - 1] = fx(hs.gx[:, None], ym[None, :])
    vy[:X - 1, :] = fy(xm[:, None], hs.gy[None, :])
    hs.vx[:] = vx.ravel()
    hs.vy[:] = vy.ravel()
    rng = np.random.default_rng(2)
    for x, y in zip(rng.uniform(0, p.xsize, 300), rng.uniform(0, p.ysize, 300)):
        a, b = _interp(orc, p, hs, x, y)
        assert abs(a - fx(x, y)) <= 1e-12 * abs(fx(x, y)) + 1e-24
        assert abs(b - fy(x, y)) <= 1e-12 * abs(fy(x, y)) + 1e-24


# ------------------------------------------------------------------ advection
def test_movemark_uniform_flow_is_exact_float32_update(orc):
    p, hs = small_state()
    hs.vx[:] = 2.5e-10
    hs.vy[:] = -1.25e-10
    hs.markk[:] = 0  # no viscosity update
    hs.timestep = 3.1e10
    x0, y0 = hs.markx.copy(), hs.marky.copy()
    o = orc.Oracle(p, hs)
    o.movemark()
    ex = (x0 + np.float32(hs.timestep * 2.5e-10)).astype(np.float32)
    ey = (y0 + np.float32(hs.timestep * -1.25e-10)).astype(np.float32)
    inside = (ex >= 0) & (ey >= 0) & (ex <= p.xsize) & (ey <= p.ysize)
    assert np.array_equal(hs.markx[inside], ex[inside])
    assert np.array_equal(hs.marky[inside], ey[inside])


def test_movemark_rk4_solid_body_rotation(orc):
    p, hs = small_state(n=21, size=100e3, mnum=2)
    X, Y = hs.xnumx, hs.ynumy
    c = p.xsize / 2
    om = 1e-12
    xm = (hs.gx[:-1] + hs.gx[1:]) / 2
    ym = (hs.gy[:-1] + hs.gy[1:]) / 2
    vx = np.zeros((X, Y))
    vy = np.zeros((X, Y))
    vx[:, :Y - 1] = -om * (ym[None, :] - c) + 0 * hs.gx[:, None]
    vy[:X - 1, :] = om * (xm[:, None] - c) + 0 * hs.gy[None, :]
    hs.vx[:] = vx.ravel()
    hs.vy[:] = vy.ravel()
    hs.markk[:] = 0
    dt = 0.05 / om
    hs.timestep = dt
    x0 = hs.markx.astype(np.float64) - c
    y0 = hs.marky.astype(np.float64) - c
    orc.Oracle(p, hs).movemark()
    th = om * dt
    xe = x0 * math.cos(th) - y0 * math.sin(th)
    ye = x0 * math.sin(th) + y0 * math.cos(th)
    r = np.hypot(x0, y0)
    keep = r < 0.4 * p.xsize
    err = np.hypot(hs.markx[keep] - c - xe[keep], hs.marky[keep] - c - ye[keep])
    # RK4 local error th^5/120 * r plus float32 rounding of the coordinates
    bound = (th ** 5 / 120 + 1e-15) * r[keep] + 2 * np.spacing(np.float32(p.xsize))
    assert np.all(err <= bound)


def test_movemark_flags_markers_leaving_the_grid(orc):
    p, hs = small_state()
    hs.vx[:] = 1e-9
    hs.vy[:] = 0
    hs.timestep = 0.3 * p.xsize / 1e-9
    orc.Oracle(p, hs).movemark()
    out = hs.markx > p.xsize
    assert out.any()
    assert np.all(hs.markk[out] == 0)
    assert np.all(hs.markk[~out] > 0)


# ------------------------------------------------------------------ marker -> node
def test_ronurecalc_constant_properties_and_partition_of_unity(orc):
    p, hs = small_state()
    hs.markt[:] = configs.SILICATE
    hs.markpor[:] = 0
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    # every node touched by a marker takes the constant exactly (ratio of sums)
    assert np.allclose(hs.ro, 3500.0, rtol=1e-15, atol=0)
    assert np.allclose(hs.cp, 1000.0, rtol=1e-15, atol=0)
    X, Y = hs.xnumx, hs.ynumy
    nu = hs.nu.reshape(X, Y)
    # shear viscosity: silicate 1e20 inside radmax, 1e22 outside
    assert np.all((np.isclose(nu, 1e20, rtol=1e-13)) | (np.isclose(nu, 1e22, rtol=1e-13)) |
                  ((nu > 1e20) & (nu < 1e22)))


def test_ronurecalc_sticky_air_and_heating(orc):
    p, hs = small_state(n=21, size=48e3)
    hs.markt[hs.markt == configs.IRON] = configs.SILICATE
    hs.markk[:] = np.float32(400.0)
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    air = hs.markt == configs.AIR
    assert np.all(hs.markk[air] == np.float32(p.tmp_ambient))
    assert np.all(hs.markk[~air] == np.float32(400.0))
    # nodes fully inside silicate carry H = H0 * rho (t = 0)
    X, Y = hs.xnumx, hs.ynumy
    ht, ro = hs.ht.reshape(X, Y), hs.ro.reshape(X, Y)
    pure = np.isclose(ro, 3500.0, rtol=1e-14)
    assert pure.any()
    assert np.allclose(ht[pure], configs.H0 * 3500.0, rtol=1e-12)


def test_type_counts_match_histogram(orc):
    p, hs, _ = configs.make("c1")
    cnt = orc.Oracle(p, hs).type_counts()
    X, Y = hs.xnumx, hs.ynumy
    xs = hs.markx.astype(np.float64)
    ys = hs.marky.astype(np.float64)
    i = np.clip(np.searchsorted(hs.gx, xs, side="right") - 1, 0, X - 2)
    j = np.clip(np.searchsorted(hs.gy, ys, side="right") - 1, 0, Y - 2)
    ok = (hs.markx > 0) & (hs.marky > 0) & (xs < p.xsize) & (ys < p.ysize)
    ref = np.zeros(((X - 1) * (Y - 1), 64), np.int64)
    np.add.at(ref, ((i * (Y - 1) + j)[ok], hs.markt[ok].astype(np.int64)), 1)
    assert np.array_equal(cnt, ref)
    assert cnt.sum() == ok.sum()


def test_viscalc_closed_forms(orc):
    p = configs.make_params(configs.CONFIGS["c2"])
    L = orc.lib()
    for T in (150.0, 900.0, 1000.0, 1300.0):
        v = L.orc_viscalc(C.byref(p), T, 0.0, configs.SILICATE, 0.0)
        e = min(configs.E_ACT / (8.314 * T), 150.0)
        ref = min(max(configs.NU0_ARRH * math.exp(e), 1e17), 1e23)
        assert v == pytest.approx(ref, rel=1e-14)
    assert L.orc_viscalc(C.byref(p), 500.0, 0.0, configs.AIR, 0.0) == pytest.approx(1e18, rel=1e-15)
    assert L.orc_viscalc(C.byref(p), 2000.0, 0.0, configs.MOLTEN_SILICATE, 0.0) == 1e2


# ------------------------------------------------------------------ gravity
def test_gravity_uniform_disk(orc):
    p, hs, _ = configs.make("c1")
    o = orc.Oracle(p, hs)
    hs.ro[:] = 3000.0
    o.gravity()
    G = 6.67384e-11
    R = p.gykoef
    # surface gravity of the disk (movemart.c:650) from the exact mean density
    assert hs.gxkoef == pytest.approx(4.0 / 3.0 * math.pi * R * 3000.0 * G, rel=1e-12)
    gx, gy = node_xy(hs)
    r = np.hypot(gx - p.xsize / 2, gy - p.ysize / 2)
    inner = (r < 0.6 * R) & (r > 0.2 * R)
    # inside a uniform disk the 8/3 pi G rho source (movemart.c:702) gives the
    # spherical g(r) = 4/3 pi G rho r.  |g| as the reference computes it
    # (movemart.c:775-781): x-derivative centred on the node, "y" term is
    # the diagonal difference (gp[m+Y+1]-gp[m])/(2 dy) = (phi_x + phi_y)/2
    k = 4.0 / 3.0 * math.pi * G * 3000.0
    px, py = k * (gx - p.xsize / 2), k * (gy - p.ysize / 2)
    ga_ref = np.hypot(px, (px + py) / 2)
    assert np.allclose(hs.ga[inner], ga_ref[inner], rtol=0.02)


# ------------------------------------------------------------------ Stokes structure
def test_stokes_matrix_structure(orc):
    """dt = 0 removes the stabilisation: the velocity block must then be
    symmetric on a uniform grid, and the pressure columns of the momentum
    rows (-dP/dx) must equal the
    transpose of the (mukoef-scaled) divergence rows, interior to interior."""
    p, hs = small_state(n=15, size=56e3)
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    hs.timestep = 0.0
    o.gravity()
    K = o.stokes_matrix().scipy().tocsr()
    n = K.shape[0]
    N = n // 3
    bm = hs.bondm[:3 * N]
    X, Y = hs.xnumx, hs.ynumy
    m1 = (np.arange(n) // 3) // Y
    m2 = (np.arange(n) // 3) % Y
    comp = np.arange(n) % 3
    ghost = ((comp == 0) & ((m1 == 0) | (m2 == 0))) | ((comp == 1) & (m2 == Y - 1)) | ((comp == 2) & (m1 == X - 1))
    free = (bm == 0) & ~ghost
    v = np.where(free & (comp > 0))[0]
    q = np.where(free & (comp == 0))[0]
    A = K[v][:, v].toarray()
    assert np.allclose(A, A.T, rtol=1e-13, atol=1e-13 * np.abs(A).max())
    Gm = K[v][:, q].toarray()
    D = K[q][:, v].toarray()
    mu = p.nubeg * p.strmin
    assert np.allclose(Gm, D.T / mu, rtol=1e-13, atol=1e-13 * np.abs(Gm).max())


def test_stokes_linear_fields_give_pressure_gradient(orc):
    """With constant viscosity and dt = 0, a linear divergence-free velocity
    field has zero viscous force, so interior x/y rows reduce to -dP/dx,dy."""
    p, hs = small_state(n=15, size=56e3)
    hs.nu[:] = 1e20
    hs.nd[:] = 1e20
    hs.ro[:] = 3000.0
    hs.timestep = 0.0
    o = orc.Oracle(p, hs)
    M = o.stokes_matrix()
    X, Y = hs.xnumx, hs.ynumy
    xm = (hs.gx[:-1] + hs.gx[1:]) / 2
    ym = (hs.gy[:-1] + hs.gy[1:]) / 2
    x = np.zeros(3 * X * Y)
    a, b, cpx, cpy = 1e-14, 2e-14, 3.0, -5.0
    for i in range(X):
        for j in range(Y):
            m = i * Y + j
            if i and j:
                x[3 * m] = cpx * xm[i - 1] + cpy * ym[j - 1]
            if j < Y - 1:
                x[3 * m + 1] = a * hs.gx[i] + b * ym[j]
            if i < X - 1:
                x[3 * m + 2] = -a * hs.gy[j] + 2e-14 * xm[i]
    y = M.matvec(x)
    bm = hs.bondm
    for i in range(1, X - 1):
        for j in range(1, Y - 2):
            m = i * Y + j
            if not bm[3 * m + 1]:
                assert y[3 * m + 1] == pytest.approx(-cpx, rel=1e-9, abs=1e-9)
            if i < X - 2 and not bm[3 * m + 2]:
                assert y[3 * m + 2] == pytest.approx(-cpy, rel=1e-9, abs=1e-9)
            if not bm[3 * m]:
                assert abs(y[3 * m]) < 1e-9 * p.nubeg * p.strmin * 1e-14


# ------------------------------------------------------------------ heat
def test_heat_fourier_mode_decay(orc):
    """Implicit Euler on the 5-point operator: a discrete sine mode decays by
    exactly 1/(1 + kappa dt lambda_h)."""
    p, hs = small_state(n=21, size=100e3)
    X, Y = hs.xnumx, hs.ynumy
    hs.kt[:] = 3.0
    hs.ro[:] = 3000.0
    hs.cp[:] = 1000.0
    hs.ht[:] = 0.0
    L = p.xsize
    gx, gy = node_xy(hs)
    amp = 50.0
    hs.tk[:] = 150.0 + amp * np.sin(np.pi * gx / L) * np.sin(np.pi * gy / L)
    hs.timestep = 3e12
    sol1 = np.zeros(X * Y)
    M = orc.Matrix(X * Y, 12)
    st = orc.lib().orc_heat_build(C.byref(p), C.byref(hs.cstruct()), sol1.ctypes.data_as(C.POINTER(C.c_double)),
                                  C.byref(M.c))
    assert st == 0
    T1 = spla.spsolve(M.scipy().tocsc(), M.b)
    h = hs.gx[1] - hs.gx[0]
    kap = 3.0 / (3000.0 * 1000.0)
    lam = 2 * (4 / h ** 2) * np.sin(np.pi * h / (2 * L)) ** 2
    inner = (gx > 0) & (gx < L) & (gy > 0) & (gy < L)
    ratio = (T1[inner] - 150.0) / (hs.tk[inner] - 150.0)
    # grid lines are float32-rounded: allow their relative spacing error
    assert np.allclose(ratio, 1.0 / (1.0 + kap * hs.timestep * lam), rtol=1e-6)


def test_heat_uniform_source_energy(orc):
    """Uniform H in a block with zero-flux free interior: dT = H dt / (rho cp)."""
    p, hs = small_state(n=21, size=100e3)
    X, Y = hs.xnumx, hs.ynumy
    hs.kt[:] = 3.0
    hs.ro[:] = 3000.0
    hs.cp[:] = 1000.0
    hs.ht[:] = 3000.0 * 1e-7
    hs.tk[:] = 150.0
    hs.timestep = 1e9
    N = X * Y
    sol1 = np.zeros(N)
    tk_ts = C.c_double()
    o = orc.Oracle(p, hs)
    cs = hs.cstruct()
    assert orc.lib().orc_heat_prepare(C.byref(p), C.byref(cs), sol1.ctypes.data_as(C.POINTER(C.c_double)),
                                      C.byref(tk_ts)) == 0
    interior = hs.bondm[3 * N:] == 0
    assert np.allclose(sol1[interior], 1e-7 / 1000.0, rtol=1e-14)
    # tktimestep = maxtkstep / max rate (heatmart.c:378-386)
    assert tk_ts.value == pytest.approx(p.maxtkstep / (1e-7 / 1000.0), rel=1e-12)


def test_direct_solvers_agree_on_stokes(orc):
    """The oracle's Stokes solution does not depend on which exact sparse LU
    solves it (PARDISO as gausmart.c:203-260 calls it, or SuperLU): the
    residual of both is at round-off and the fields agree far below the
    tolerances of the GPU parity tests."""
    import sys
    p, hs, _ = configs.make("c1")
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    hs.timestep = p.maxtmstep
    o.maxvelstep()
    o.gravity()
    M = o.stokes_matrix()
    A = M.scipy().tocsc()
    xs = orc.lu_solve(A, M.b)
    assert np.linalg.norm(A @ xs - M.b) <= 1e-12 * np.linalg.norm(M.b)
    if "torch" in sys.modules or orc._mkl_lib() is None:
        pytest.skip("standalone MKL runtime unusable in this process")
    xp = orc.pardiso_solve(M.scipy(), M.b)
    assert np.linalg.norm(A @ xp - M.b) <= 1e-12 * np.linalg.norm(M.b)
    for k in (1, 2):  # velocities
        assert np.linalg.norm(xp[k::3] - xs[k::3]) <= 1e-9 * np.linalg.norm(xs[k::3])
    assert np.linalg.norm(xp[0::3] - xs[0::3]) <= 1e-9 * np.linalg.norm(xs[0::3])


def test_residual_diagnostics_match_assembled_rows(orc):
    """viterate's error-form residuals (movemart.c:1049-1176), computed from the
    post-solve stresses, equal the residual of the assembled rows: momentum
    rows divided by the mean nodal viscosity, velocity BC rows by mukoef.  Two
    independent code paths (stresses vs CSR coefficients) must agree."""
    p, hs, _ = configs.make("c1")
    o = orc.Oracle(p, hs)
    o.ronurecalc()
    hs.timestep = p.maxtmstep
    o.maxvelstep()
    o.viterate()
    exact = o.stokes_residuals()
    assert exact["stokes"] < 1e-10 and exact["bond"] < 1e-12
    M = o.stokes_matrix()
    A = M.scipy()
    x = np.zeros(3 * hs.N)
    x[0::3], x[1::3], x[2::3] = hs.pr, hs.vx, hs.vy
    rng = np.random.default_rng(11)
    x[1::3] *= 1 + 1e-3 * rng.standard_normal(hs.N)
    x[2::3] *= 1 + 1e-3 * rng.standard_normal(hs.N)
    o.L.orc_stokes_post(C.byref(p), o._s(), x.ctypes.data_as(C.POINTER(C.c_double)))
    got = o.stokes_residuals()
    x[0::3], x[1::3], x[2::3] = hs.pr, hs.vx, hs.vy  # reload conventions (ghosts) as the post step left them
    r = A @ x - M.b
    X = Y = p.xnumx
    m = np.arange(hs.N)
    m1, m2 = m // Y, m % Y
    bm = hs.bondm[:3 * hs.N]
    mu = p.nubeg * p.strmin
    sx = (m2 < Y - 1) & (bm[3 * m + 1] == 0)
    sy = (m1 < X - 1) & (bm[3 * m + 2] == 0)
    nux = (hs.nu[m[sx]] + hs.nu[m[sx] + 1]) / 2
    nuy = (hs.nu[m[sy]] + hs.nu[m[sy] + Y]) / 2
    es = np.concatenate([r[3 * m[sx] + 1] / nux, r[3 * m[sy] + 2] / nuy])
    bx = (m2 < Y - 1) & (bm[3 * m + 1] != 0)
    by = (m1 < X - 1) & (bm[3 * m + 2] != 0)
    eb = np.concatenate([r[3 * m[bx] + 1], r[3 * m[by] + 2]]) / mu
    vxr = hs.vx[m2 < Y - 1]
    vyr = hs.vy[m1 < X - 1]
    vrange = max(vxr.max() - vxr.min(), vyr.max() - vyr.min())
    mindx = min(p.xsize / (X - 1), p.ysize / (Y - 1))
    assert got["stoknum"] == es.size and got["bondnum"] == eb.size
    assert got["stokes"] == pytest.approx(np.sqrt(np.mean(es ** 2)) / (vrange / mindx ** 2), rel=1e-9)
    assert got["bond"] == pytest.approx(np.sqrt(np.mean(eb ** 2)) / vrange, rel=1e-9)
    assert got["stokes"] > 1e3 * exact["stokes"]


def test_peak_t_objective_closed_forms(orc):
    """The adjoint's objective: mass-weighted mean marktmax of silicate markers
    inside the box.  Closed forms on two-valued fields: equal porosity cancels
    from the weights; unequal porosity weights by (1 - phi); iron, air and
    markers outside the box do not count."""
    cfg = dict(configs.CONFIGS["c4"], n=31)
    p, hs = configs.make_params(cfg), configs.make_state(cfg)
    sil = hs.markt == configs.SILICATE
    left = hs.markx < cfg["size"] / 2
    nL, nR = int(np.sum(sil & left)), int(np.sum(sil & ~left))
    assert nL > 10 and nR > 10
    hs.marktmax[:] = np.float32(1234.0)
    assert orc.peak_t_objective(p, hs) == pytest.approx(1234.0, rel=1e-14)
    a, b = 600.0, 900.0
    hs.marktmax[:] = np.where(left, np.float32(a), np.float32(b))
    hs.markpor[sil] = np.float32(0.375)
    assert orc.peak_t_objective(p, hs) == pytest.approx((nL * a + nR * b) / (nL + nR), rel=1e-13)
    fL, fR = 0.25, 0.5
    hs.markpor[sil] = np.where(left[sil], np.float32(fL), np.float32(fR))
    want = (nL * (1 - fL) * a + nR * (1 - fR) * b) / (nL * (1 - fL) + nR * (1 - fR))
    assert orc.peak_t_objective(p, hs) == pytest.approx(want, rel=1e-13)
    hs.marktmax[~sil] = np.float32(5000.0)
    assert orc.peak_t_objective(p, hs) == pytest.approx(want, rel=1e-13)
    i = int(np.flatnonzero(sil & left)[0])
    hs.marktmax[i] = np.float32(1e6)
    hs.markx[i] = np.float32(-1.0)
    wantL = ((nL - 1) * (1 - fL) * a + nR * (1 - fR) * b) / ((nL - 1) * (1 - fL) + nR * (1 - fR))
    assert orc.peak_t_objective(p, hs) == pytest.approx(wantL, rel=1e-13)
