"""Seeded synthetic planetesimal setups (the initial conditions in2mart.c
would write for them) and their run parameters.

The reference's initial-condition generator (in2mart.c) is outside the time
step; what is kept from it is the layout semantics the step depends on:

* grid lines are a cumulative sum of the uniform step with the last line reset
  to the box size (in2mart.c:135-138, 981-984), rounded to float32 as the
  binary save/load does (loadmart.c:502-507, 774-779);
* markers sit on a per-cell lattice at half-spacing offsets, cell by cell with
  x outer and y inner (in2mart.c:153-182) -- so they start sorted by cell --
  and get a uniform jitter (in2mart.c:188-197; here numpy's PCG64 instead of
  glibc rand());
* boundary conditions are the free-slip box, pressure anchor and fixed
  ambient temperature of init.t3c:49-86, in the bondm/bondn/bondv tables.

Rock constants follow the init.t3c ROCKS table where the configs do not say
otherwise (air: cp 3000, k 2000; silicate/iron: cp 1000, k 3).
"""
import math

import numpy as np

from . import abi
from .state import HostState

YEAR = 3.15576e7          # seconds per year, i2mart.c:69 / loadmart.c:46
T_AMBIENT = 150.0
H0 = 1.5e-7               # W/kg, 26Al heating at t = 0
# al2627_init giving init_al = H0 in markmart.c:770
AL2627_FOR_H0 = H0 * 3.264e13 / (1.89716e23 * 4.999e-13)
E_ACT = 300e3             # J/mol
ETA_REF_T = 1000.0        # Arrhenius prefactor chosen so eta(1000 K) = 1e20
NU0_ARRH = 1e20 / math.exp(E_ACT / (8.314 * ETA_REF_T))

AIR, SILICATE, IRON, MOLTEN_SILICATE = 0, 6, 7, 26

CONFIGS = {
    # 51x51, 200 km box, 4x4 markers/cell, disk R=50 km, constant viscosity
    "c1": dict(n=51, size=200e3, mnum=4, radius=50e3, steps=10, rheology="const", porosity=False,
               iron="random", melt=False, nubeg=1e17, nucontr=1e18, densimod=0),
    # 201x201, 400 km, 4x4/cell (640k), R=100 km, Arrhenius 300 kJ/mol, porosity U[0.3,0.5]
    "c2": dict(n=201, size=400e3, mnum=4, radius=100e3, steps=100, rheology="arrhenius", porosity=True,
               iron="random", melt=False, nubeg=1e17, nucontr=1e18, densimod=1),
    # 501x501, 600 km, 5x5/cell (6.25M), R=150 km, Arrhenius + linear melting, 400 iron blobs
    "c3": dict(n=501, size=600e3, mnum=5, radius=150e3, steps=200, rheology="arrhenius", porosity=False,
               iron="blobs", melt=True, nubeg=1e2, nucontr=1e21, densimod=1),
    # adjoint case: the 201x201 setup over 50 steps with the porosity held at
    # its initial U[0.3,0.5] values (no sintering), no shear heating and no
    # heating-rate limit on the time step (titerate's dt = maxtkstep/max|dT/dt|
    # would make dt, and with it every trajectory, depend on the heating)
    "c4": dict(n=201, size=400e3, mnum=4, radius=100e3, steps=50, rheology="arrhenius", porosity=True,
               porosity_frozen=True, frictyn=1e30, maxtkstep=1e30, iron="random", melt=False, nubeg=1e17,
               nucontr=1e18, densimod=1),
}


def make_params(cfg):
    e_act = cfg.get("e_act", E_ACT)
    nu0_arrh = 1e20 / math.exp(e_act / (8.314 * ETA_REF_T))  # eta(1000 K) = 1e20 for any E
    p = abi.Params()
    n = cfg["n"]
    p.xnumx = p.ynumy = n
    p.mnumx = p.mnumy = cfg["mnum"]
    p.xsize = p.ysize = cfg["size"]
    p.maxxystep = 0.05
    p.maxtkstep = cfg.get("maxtkstep", 1e4)
    p.maxtmstep = 2.5e3 * YEAR
    p.nubeg = cfg["nubeg"]
    p.nuend = 1e23
    p.nucontr = cfg["nucontr"]
    # mukoef = nubeg*strmin scales continuity/BC rows (movemart.c:792); mode.t3c
    # gives 1e17*0.1 = 1e16.  Configs that lower nubeg to admit 1e2 Pa s melt keep
    # the same row scale, otherwise the assembled system is badly scaled.
    p.strmin = 1e16 / cfg["nubeg"]
    p.heatdif = 1.0
    p.frictyn = cfg.get("frictyn", 1.0)
    p.adiabyn = 0.0
    p.heatmod = 0
    p.densimod = cfg["densimod"]
    p.outgrid = 2
    p.markmod = 2
    p.gykoef = cfg["radius"]
    p.tmp_ambient = T_AMBIENT
    p.al2627_init = AL2627_FOR_H0
    p.fe6056_init = 0.0
    p.h0_scale = cfg.get("h0", H0) / H0
    p.longlived = 0.0
    p.por_init = 0.4 if cfg["porosity"] else 0.0
    p.k_cutoff = 1e6
    # The configs put sticky air (1e18 Pa s) everywhere outside the disk; the
    # reference's forced 1e22 beyond 0.95*xsize/2 (markmart.c:268, 340) is
    # kept as a parameter but placed outside the box.  With it, the mean
    # pressure of the stiff corners holding the anchor cell is so weakly
    # determined that two direct solvers disagree on P at 2.6e-7 (201^2 case).
    p.radmax_frac = 10.0
    p.eta_outer = 1e22
    p.melt_ts, p.melt_tl, p.melt_switch, p.melt_eta, p.melt_hlat = 1360.0, 1850.0, 0.4, 1e2, 4e5
    p.melt_on = 1 if cfg["melt"] else 0
    p.porosity_on = 1 if cfg["porosity"] and not cfg.get("porosity_frozen") else 0
    p.ntypes = abi.MAXTYPES

    def rock(t, ro, nu0, dh=0.0, cp=1000.0, kt=3.0, n0=1e17, n1=1e23):
        r = p.rock[t]
        r.nu0, r.dh, r.dv, r.ss, r.mm = nu0, dh, 0.0, 0.0, 1.0
        r.n0, r.n1 = n0, n1
        r.ro, r.aa, r.bb, r.cp, r.kt, r.kf, r.kp = ro, 0.0, 0.0, cp, kt, 0.0, 0.0

    for t in range(abi.MAXTYPES):  # inert defaults so no type reads zeros
        rock(t, 3500.0, 1e20)
    rock(AIR, 1.0, 1e18, cp=3000.0, kt=2000.0)
    if cfg["rheology"] == "const":
        rock(SILICATE, 3500.0, 1e20)
        rock(IRON, 7800.0, 1e20)
    else:
        rock(SILICATE, 3500.0, nu0_arrh, dh=e_act)
        rock(IRON, 7800.0, nu0_arrh, dh=e_act)
    rock(MOLTEN_SILICATE, 3500.0, nu0_arrh, dh=e_act, n0=1e2)
    for t in (5, 6, 25, 26):
        p.is_silicate[t] = 1
    for t in (7, 8, 9, 10, 17, 18, 19):
        p.is_iron[t] = 1
    for t in (5, 6):
        p.porous[t] = 1
    p.meltable[SILICATE] = 1
    return p


def grid_lines(n, size):
    """Cumulative-sum grid lines, endpoint reset, float32-rounded."""
    step = size / float(n - 1)
    g = np.zeros(n)
    for i in range(1, n):  # sequential sum, as in2mart.c:136
        g[i] = g[i - 1] + step
    g[n - 1] = size
    return g.astype(np.float32).astype(np.float64)


def free_slip_bc(X, Y, t_ambient=T_AMBIENT):
    """bondm/bondn/bondv for init.t3c:49-86 (x = X-1, y = Y-1)."""
    N = X * Y
    bondm = np.zeros(4 * N, np.int64)
    rows_n = [[0, 0, 0]]
    rows_v = [[0.0, 0.0, 0.0, 0.0]]

    def add(unknown, const, nbr=None, k=0.0):
        rows_v.append([const, k, 0.0, 0.0])
        rows_n.append([nbr + 1 if nbr is not None else 0, 0, 0])
        bondm[unknown] = len(rows_v) - 1

    node = lambda m1, m2: m1 * Y + m2
    x, y = X - 1, Y - 1
    # pressure: anchor cell (3,1) and the four corner cells mirror a neighbour
    add(3 * node(3, 1) + 0, 0.0)
    add(3 * node(1, 1) + 0, 0.0, 3 * node(2, 1) + 0, 1.0)
    add(3 * node(1, y) + 0, 0.0, 3 * node(2, y) + 0, 1.0)
    add(3 * node(x, 1) + 0, 0.0, 3 * node(x - 1, 1) + 0, 1.0)
    add(3 * node(x, y) + 0, 0.0, 3 * node(x - 1, y) + 0, 1.0)
    for m1 in range(1, x):            # upper / lower free slip
        add(3 * node(m1, 0) + 1, 0.0, 3 * node(m1, 1) + 1, 1.0)
        add(3 * node(m1, y - 1) + 1, 0.0, 3 * node(m1, y - 2) + 1, 1.0)
    for m1 in range(0, x):
        add(3 * node(m1, 0) + 2, 0.0)
        add(3 * node(m1, y) + 2, 0.0)
    for m2 in range(0, y):            # left / right walls
        add(3 * node(0, m2) + 1, 0.0)
        add(3 * node(x, m2) + 1, 0.0)
    for m2 in range(1, y):
        add(3 * node(0, m2) + 2, 0.0, 3 * node(1, m2) + 2, 1.0)
        add(3 * node(x - 1, m2) + 2, 0.0, 3 * node(x - 2, m2) + 2, 1.0)
    for m2 in range(1, y):            # fixed ambient temperature on every edge
        add(3 * N + node(0, m2), t_ambient)
        add(3 * N + node(x, m2), t_ambient)
    for m1 in range(0, x + 1):
        add(3 * N + node(m1, 0), t_ambient)
        add(3 * N + node(m1, y), t_ambient)
    return bondm, np.array(rows_n, np.int64), np.array(rows_v)


def make_state(name, seed=20261203):
    """Seeded initial state of config `name` (or of a config dict: markers,
    grid, BCs; node fields zero except T = ambient).  Node properties come
    from a first ronurecalc."""
    cfg = CONFIGS[name] if isinstance(name, str) else name
    n, size, mnum, R = cfg["n"], cfg["size"], cfg["mnum"], cfg["radius"]
    X = Y = n
    cells = (X - 1) * (Y - 1)
    M = cells * mnum * mnum
    bondm, bondn, bondv = free_slip_bc(X, Y)
    hs = HostState(X, Y, M, bondnum=bondn.shape[0])
    hs.bondm[:], hs.bondn, hs.bondv = bondm, bondn, bondv
    hs.gx[:] = grid_lines(X, size)
    hs.gy[:] = grid_lines(Y, size)
    xstp = size / float(X - 1)
    dx = xstp / float(mnum)
    # lattice positions, cell-major (x outer, y inner), then the two in-cell loops
    m1 = np.repeat(np.arange(X - 1), (Y - 1) * mnum * mnum)
    m2 = np.tile(np.repeat(np.arange(Y - 1), mnum * mnum), X - 1)
    m4 = np.tile(np.repeat(np.arange(mnum), mnum), cells)
    m5 = np.tile(np.arange(mnum), cells * mnum)
    x = m1 * xstp + dx / 2.0
    for k in range(1, mnum):
        x = np.where(m4 >= k, x + dx, x)
    y = m2 * xstp + dx / 2.0
    for k in range(1, mnum):
        y = np.where(m5 >= k, y + dx, y)
    rng = np.random.Generator(np.random.PCG64(seed))
    x = x + rng.uniform(-0.5, 0.5, M) * dx
    y = y + rng.uniform(-0.5, 0.5, M) * dx
    hs.markx[:] = x.astype(np.float32)
    hs.marky[:] = y.astype(np.float32)
    cx = hs.markx.astype(np.float64) - size / 2.0
    cy = hs.marky.astype(np.float64) - size / 2.0
    r = np.sqrt(cx * cx + cy * cy)
    inside = r < R
    t = np.where(inside, SILICATE, AIR).astype(np.int8)
    rng2 = np.random.Generator(np.random.PCG64(seed + 1))
    if cfg["iron"] == "random":
        t[inside & (rng2.random(M) < 0.25)] = IRON
    else:
        sigma, nblob = 2e3, cfg.get("nblob", 400)
        rb = (R - 3 * sigma) * np.sqrt(rng2.random(nblob))
        th = 2 * np.pi * rng2.random(nblob)
        bx, by = size / 2 + rb * np.cos(th), size / 2 + rb * np.sin(th)
        prob = np.zeros(M)
        per_cell = mnum * mnum
        h = int(math.ceil(3 * sigma / xstp)) + 1
        for b in range(nblob):
            i0, j0 = int(bx[b] // xstp), int(by[b] // xstp)
            ii = np.arange(max(i0 - h, 0), min(i0 + h + 1, X - 1))
            jj = np.arange(max(j0 - h, 0), min(j0 + h + 1, Y - 1))
            cid = (ii[:, None] * (Y - 1) + jj[None, :]).ravel()
            idx = (cid[:, None] * per_cell + np.arange(per_cell)[None, :]).ravel()
            d2 = (hs.markx[idx] - bx[b]) ** 2 + (hs.marky[idx] - by[b]) ** 2
            np.maximum.at(prob, idx, np.exp(-d2 / sigma ** 2))
        u = rng2.random(M)
        t[inside & (u < prob)] = IRON
    hs.markt[:] = t
    hs.markk[:] = np.float32(T_AMBIENT)
    hs.marktmax[:] = np.float32(T_AMBIENT)
    if cfg["porosity"]:
        por = rng2.uniform(0.3, 0.5, M).astype(np.float32)
        hs.markpor[:] = np.where(t == SILICATE, por, np.float32(0.0))
    hs.tk[:] = T_AMBIENT
    hs.timestep = 0.0
    hs.timesum = 0.0
    hs.gxkoef = 1.0
    return hs


def make(name, seed=20261203):
    cfg = CONFIGS[name]
    return make_params(cfg), make_state(name, seed), cfg


def ensemble(nbody=64, seed=20261203, nmin=101, nmax=301):
    """The batched-ensemble setup: bodies with radius uniform in [10, 150] km,
    grids scaled with the radius from nmin^2 to nmax^2 nodes (odd sizes, box
    4R), 4x4 markers per cell, H0 log-uniform in [1e-8, 3e-7] W/kg and
    activation energy uniform in [250, 350] kJ/mol (Arrhenius, 25 % iron).
    Returns a list of (params, state, cfg)."""
    rng = np.random.Generator(np.random.PCG64(seed + 7))
    out = []
    for k in range(nbody):
        R = rng.uniform(10e3, 150e3)
        n = int(round(nmin + (R - 10e3) / 140e3 * (nmax - nmin))) | 1
        cfg = dict(n=n, size=4 * R, mnum=4, radius=R, steps=50, rheology="arrhenius", porosity=False,
                   iron="random", melt=False, nubeg=1e17, nucontr=1e18, densimod=1,
                   h0=math.exp(rng.uniform(math.log(1e-8), math.log(3e-7))), e_act=rng.uniform(250e3, 350e3))
        out.append((make_params(cfg), make_state(cfg, seed + 100 * k), cfg))
    return out


def heat_state(hs, cfg, t_core=1800.0, t_surface=T_AMBIENT):
    """The molten regime of the 501^2 setup, seeded directly: marker
    temperatures follow T(r) = t_surface + (t_core - t_surface)(1 - (r/R)^2)
    inside the disk (R the body radius) and stay ambient outside.  With the
    linear melting of 1360-1850 K the core inside r = 0.385 R (t_core 1800 K)
    exceeds the 0.4 melt fraction (eta -> 1e2 Pa s), a partially molten shell
    follows, and the cold rim stays at the 1e23 Pa s clip: the 1e21
    viscosity contrast of the run's hot half (the 26Al heating raises the
    interior by ~12 K per 2500-yr step, so the solidus is crossed after ~100
    steps).  marktmax is raised to match; node fields follow from the next
    ronurecalc.  Modifies and returns hs."""
    size, R = cfg["size"], cfg["radius"]
    cx = hs.markx.astype(np.float64) - size / 2.0
    cy = hs.marky.astype(np.float64) - size / 2.0
    r2 = (cx * cx + cy * cy) / (R * R)
    T = np.where(r2 < 1.0, t_surface + (t_core - t_surface) * (1.0 - r2), T_AMBIENT)
    hs.markk[:] = T.astype(np.float32)
    hs.marktmax[:] = np.maximum(hs.marktmax, hs.markk)
    return hs


def make_hot(name="c3", seed=20261203, t_core=1800.0, **over):
    """make(name) with the molten-core temperature field of heat_state; keyword
    arguments override the config (e.g. n, size, radius for smaller grids)."""
    cfg = dict(CONFIGS[name], **over)
    p = make_params(cfg)
    hs = make_state(cfg, seed)
    heat_state(hs, cfg, t_core=t_core)
    return p, hs, cfg
