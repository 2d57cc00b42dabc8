"""GPU tests of the coupled peak-temperature adjoint
(i2e_adjoint_peak_t_coupled) through the C-ABI: the gradient of J, the
mass-weighted mean peak temperature of the silicate markers, through the
whole recorded step -- thermal chain, marker advection, Stokes and gravity
solves, viscosity scatter and viscalc -- against central finite differences
of the full time step, against the thermal-chain adjoint where the flow does
not depend on the parameter, and FP64 against mixed precision.

What limits the finite differences here (each test states its tolerance):
marker temperatures and positions are float32 (headmart.c:122-127), so J(p)
is a staircase of ulp-sized steps under its smooth part; the viscosity
clips (nubeg, nuend) put kinks into J wherever a marker's Arrhenius
viscosity crosses them; and a warm body's flow amplifies both over 50 steps.
The activation-energy derivative is the smallest (a 1e-3 relative change of
E moves J by ~1e-7 of itself), so its differences take relative steps of
1e-3..2e-3 and a body whose viscosities stay above the lower clip; the
porosity shift is 4e-3 of the porosity.  h0_scale keeps the relative step
1e-4.
"""
import copy
import sys

import numpy as np
import pytest

from i2elvis_planet_amd import Context, I2EError, configs, default_solver

pytestmark = pytest.mark.gpu

EPS = 1e-4


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def opts(precision=0, rtol=1e-11):
    o = default_solver()
    o.precision = precision
    o.stokes_rtol = rtol
    o.heat_rtol = 1e-13
    o.grav_rtol = 1e-13
    return o


def scaled(p, h0=1.0, s=1.0):
    q = copy.deepcopy(p)
    q.h0_scale = p.h0_scale * h0
    for t in range(len(q.rock)):
        q.rock[t].dh = p.rock[t].dh * s
    return q


def run_full(ctx, p, hs, K):
    ctx.set_params(p)
    ctx.upload(hs)
    ctx.ronurecalc()
    for _ in range(K):
        ctx.step()
    return ctx.peak_t()


def shift_porosity(hs, d):
    """hs with every positive porosity shifted by d, and the shift float32
    storage applied to nearly all of them (as in test_gpu_adjoint)."""
    h = hs.copy()
    live = h.markpor > 0
    h.markpor[live] = (h.markpor[live].astype(np.float64) + d).astype(np.float32)
    eff = h.markpor[live].astype(np.float64) - hs.markpor[live].astype(np.float64)
    vals, counts = np.unique(eff, return_counts=True)
    assert counts.max() >= 0.999 * eff.size, (vals, counts)
    return h, float(vals[np.argmax(counts)])


def warm_c4(t_core=1100.0, nubeg=None):
    """The adjoint case (201^2, 640k markers, 50 steps, porosity held) with a
    warm interior: T(r) = 150 + (t_core - 150)(1 - (r/R)^2) inside the body
    (continuous at the cold surface), porosities in (0.41, 0.49) where the
    porous-conductivity closure is smooth, markv = 0 (the first ronurecalc
    evaluates viscalc).  With E = 300 kJ/mol and eta(1000 K) = 1e20 the inner
    half sits between the clips and the activation energy reaches the flow;
    the 26Al heating (~8 K per step) takes the centre to the 1e17 lower clip
    over the 50 steps.  maxxystep 1 (a marker may cross one cell per step)
    keeps maxvelstep from shortening the 2500-yr step: the gradient holds
    the time steps, and so must the finite differences.  nubeg (with strmin
    keeping the row scale nubeg strmin = 1e16) lowers the lower clip."""
    p, hs, cfg = configs.make("c4")
    configs.heat_state(hs, cfg, t_core=t_core, t_surface=configs.T_AMBIENT)
    hs.markv[:] = 0
    live = hs.markpor > 0
    hs.markpor[live] = np.random.default_rng(5).uniform(0.41, 0.49, int(live.sum())).astype(np.float32)
    p.maxxystep = 1.0
    if nubeg is not None:
        p.nubeg = nubeg
        p.strmin = 1e16 / nubeg
        for t in range(len(p.rock)):
            p.rock[t].n0 = min(p.rock[t].n0, nubeg)
    return p, hs, cfg


def coupled_at(ctx, o, rtol=3e-10):
    """The coupled adjoint with its transposed Stokes solves to rtol: on the
    warm 201^2 case they stall near 1e-10 (FGMRES(200) restarts; the
    stabilisation term makes the transposed rows less symmetric than the
    forward preconditioner assumes), where the forward solves reach the
    1e-11 of the record.  The gradient agrees to 9 digits between 1e-8 and
    1e-10 here."""
    r = o.stokes_rtol
    o.stokes_rtol = rtol
    ctx.set_solver(o)
    try:
        return ctx.adjoint_peak_t_coupled()
    finally:
        o.stokes_rtol = r
        ctx.set_solver(o)


def central(f, h):
    return (f(h) - f(-h)) / (2 * h)


def recorded(p, hs, K, precision=0):
    o = opts(precision)
    ctx = Context(p, hs, solver=o)
    ctx.record(K)
    J = ctx.peak_t()
    adj = coupled_at(ctx, o)
    return ctx, J, adj


@pytest.fixture(scope="module")
def warm():
    say("case coupled adjoint warm c4 record")
    p, hs, cfg = warm_c4()
    K = cfg["steps"]
    ctx, J, adj = recorded(p, hs, K)
    say("warm c4: J %.12f coupled adjoint %s" % (J, adj))
    yield dict(p=p, hs=hs, K=K, ctx=ctx, J=J, adj=adj)
    ctx.close()


def test_coupled_objective_and_solves(warm):
    a = warm["adj"]
    assert a["J"] == warm["J"] and a["steps"] == warm["K"]
    assert a["stokes_relres_max"] <= 3e-10 and a["heat_relres_max"] <= 1e-13 and a["grav_relres_max"] <= 1e-13


def test_coupled_h0_matches_fd(warm):
    """h0_scale reaches J through the heating and, with T-dependent
    viscosities, through the flow: central difference with the relative step
    1e-4 of the full step, within 2e-5 (the float32 staircase of a 50-step
    warm run; 1e-5 where the flow is frozen, test_gpu_adjoint)."""
    say("case coupled adjoint fd h0")
    p, hs, K, ctx = warm["p"], warm["hs"], warm["K"], warm["ctx"]
    fd = central(lambda e: run_full(ctx, scaled(p, h0=1 + e), hs, K), EPS) / p.h0_scale
    ad = warm["adj"]["dJ_dh0"]
    say("dJ/dh0: adjoint %.12e fd %.12e rel %.2e" % (ad, fd, abs(ad - fd) / abs(fd)))
    ctx.set_params(p)
    assert abs(ad - fd) <= 2e-5 * abs(fd)


def test_coupled_phi0_matches_fd(warm):
    """Porosity reaches J through density, heat capacity, radiogenic heat,
    the porous conductivity and the marker mass, and also through buoyancy
    and gravity into the flow: central difference of the full step with a
    shift of 2e-3 (4e-3 of the mean porosity), within 2e-5.  Smaller shifts
    sink into the float32 staircase (at 1e-4 of the porosity the difference
    scatters by ~1e-4 around the adjoint's value)."""
    say("case coupled adjoint fd phi0")
    p, hs, K, ctx = warm["p"], warm["hs"], warm["K"], warm["ctx"]
    hp, dp = shift_porosity(hs, 2e-3)
    hm, dm = shift_porosity(hs, -2e-3)
    fd = (run_full(ctx, p, hp, K) - run_full(ctx, p, hm, K)) / (dp - dm)
    ad = warm["adj"]["dJ_dphi0"]
    say("dJ/dphi0: adjoint %.12e fd %.12e rel %.2e" % (ad, fd, abs(ad - fd) / abs(fd)))
    assert abs(ad - fd) <= 2e-5 * abs(fd)


def test_coupled_energy_matches_fd_above_the_lower_clip():
    """dJ/ds, s a factor on every activation energy, with the lower clip at
    1e10 Pa s so that no viscosity reaches it in the 50 steps (J then has no
    clip kinks): central differences with relative steps 1e-3 and 2e-3
    agree with each other to ~6e-3 (the staircase), and the adjoint with
    each to 1e-2."""
    say("case coupled adjoint fd E (nubeg 1e10)")
    p, hs, cfg = warm_c4(nubeg=1e10)
    K = cfg["steps"]
    ctx, J, adj = recorded(p, hs, K)
    ad = adj["dJ_dEscale"]
    f = lambda e: run_full(ctx, scaled(p, s=1 + e), hs, K)
    fd1, fd2 = central(f, 1e-3), central(f, 2e-3)
    ctx.close()
    say("dJ/ds: adjoint %.10e fd(1e-3) %.10e rel %.2e fd(2e-3) %.10e rel %.2e"
        % (ad, fd1, abs(ad - fd1) / abs(fd1), fd2, abs(ad - fd2) / abs(fd2)))
    assert abs(ad) > 1e-2  # the activation energy reaches J
    assert abs(ad - fd1) <= 1e-2 * abs(fd1)
    assert abs(ad - fd2) <= 1e-2 * abs(fd2)


def test_coupled_energy_matches_fd_with_the_clips(warm):
    """The same derivative in the case itself (lower clip 1e17, reached at
    the centre during the run): the clip kinks spread the finite differences
    of steps 1e-3..4e-3 over ~4 %; the adjoint lies within 5e-2 of each."""
    say("case coupled adjoint fd E")
    p, hs, K, ctx = warm["p"], warm["hs"], warm["K"], warm["ctx"]
    ad = warm["adj"]["dJ_dEscale"]
    f = lambda e: run_full(ctx, scaled(p, s=1 + e), hs, K)
    for h in (1e-3, 2e-3):
        fd = central(f, h)
        say("dJ/ds: adjoint %.10e fd(%.0e) %.10e rel %.2e" % (ad, h, fd, abs(ad - fd) / abs(fd)))
        assert abs(ad - fd) <= 5e-2 * abs(fd)
    ctx.set_params(p)


def test_coupled_positions_match_fd():
    """dJ/d(initial marker positions) over 5 steps of the warm case: J of the
    full run with every marker displaced by eps sin(pi x/L) sin(pi y/L) in x
    (then in y), eps = 5 m, against the adjoint's sum.  The float32
    positions (ulp 1.6 cm at 200 km) and temperatures limit this difference
    to a few percent (over 50 steps, to ~15 %); the bound of 5e-2 still
    catches a missing term or a wrong sign."""
    say("case coupled adjoint fd positions")
    p, hs, cfg = warm_c4()
    K = 5
    ctx, J, adj = recorded(p, hs, K)
    gx, gy = ctx.adjoint_marker_positions(hs.markx.size)
    L = p.xsize
    x, y = hs.markx.astype(np.float64), hs.marky.astype(np.float64)
    u = np.sin(np.pi * x / L) * np.sin(np.pi * y / L)
    for comp, g in ((0, gx), (1, gy)):
        Js, shifts = [], []
        for e in (5.0, -5.0):
            h = hs.copy()
            if comp == 0:
                h.markx[:] = (x + e * u).astype(np.float32)
                shifts.append(h.markx.astype(np.float64) - x)
            else:
                h.marky[:] = (y + e * u).astype(np.float32)
                shifts.append(h.marky.astype(np.float64) - y)
            Js.append(run_full(ctx, p, h, K))
        pred = float(np.dot(g, shifts[0] - shifts[1]))
        fd = Js[0] - Js[1]
        say("positions comp %d: adjoint %.6e fd %.6e rel %.2e" % (comp, pred, fd, abs(pred - fd) / abs(fd)))
        assert abs(pred - fd) <= 5e-2 * abs(fd)
    ctx.close()


def test_coupled_equals_thermal_adjoint_when_the_flow_is_frozen():
    """The cold case as configured (150 K, every silicate and iron viscosity
    at its 1e23 clip): the activation energy reaches nothing, and h0_scale
    reaches the flow nowhere, so the coupled adjoint's dJ/ds is exactly 0
    and its dJ/dh0 equals the thermal-chain adjoint's."""
    say("case coupled adjoint cold c4")
    p, hs, cfg = configs.make("c4")
    o = opts()
    ctx = Context(p, hs, solver=o)
    ctx.record(10)
    th = ctx.adjoint_peak_t()
    co = coupled_at(ctx, o)
    ctx.close()
    say("thermal %s\ncoupled %s" % (th, co))
    assert co["J"] == th["J"]
    assert co["dJ_dEscale"] == 0.0
    assert abs(co["dJ_dh0"] - th["dJ_dh0"]) <= 1e-9 * abs(th["dJ_dh0"])


def test_coupled_mixed_matches_fp64(warm):
    say("case coupled adjoint mixed")
    p, hs, K = warm["p"], warm["hs"], 20
    ctx64, J64, a64 = recorded(p, hs, K)
    ctx64.close()
    ctx32, J32, a32 = recorded(p, hs, K, precision=1)
    ctx32.close()
    say("fp64 %s\nmixed %s" % (a64, a32))
    assert abs(a32["J"] - a64["J"]) <= 1e-9 * abs(a64["J"])
    for k in ("dJ_dh0", "dJ_dphi0"):
        assert abs(a32[k] - a64[k]) <= 1e-6 * abs(a64[k]), k
    assert abs(a32["dJ_dEscale"] - a64["dJ_dEscale"]) <= 1e-4 * abs(a64["dJ_dEscale"])


def test_coupled_refuses_velocity_limited_steps():
    """A maxxystep small enough that maxvelstep sets dt: that min is not
    differentiated, so the adjoint refuses the record."""
    say("case coupled adjoint maxvelstep refusal")
    p, hs, cfg = warm_c4()
    p.maxxystep = 1e-4
    o = opts()
    ctx = Context(p, hs, solver=o)
    ctx.record(3)
    with pytest.raises(I2EError, match="maxvelstep"):
        ctx.adjoint_peak_t_coupled()
    with pytest.raises(I2EError, match="coupled adjoint"):
        ctx.adjoint_marker_positions(hs.markx.size)
    ctx.close()
