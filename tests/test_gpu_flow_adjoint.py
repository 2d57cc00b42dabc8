"""GPU tests of the flow adjoint (i2e_sink_rate / i2e_adjoint_sink_rate) and
of the transposed Stokes operator, through the C-ABI.

Pins: the dot-product identity <K x, y> = <x, K^T y> on the assembled rows
(free-slip box, and general and continuity-type BC rows); the objective
against the oracle's restatement on the GPU's velocities; the adjoint's
activation-energy derivative against central finite differences of the
GPU's own ronurecalc + viterate (relative step 1e-4); mixed precision
against FP64; the refusals.
"""
import copy
import sys

import numpy as np
import pytest

from i2elvis_planet_amd import Context, DeviceArray, I2EError, configs, default_solver
from harness import bc as bch

pytestmark = pytest.mark.gpu

EPS = 1e-4


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def opts(precision=0, rtol=1e-10):
    o = default_solver()
    o.precision = precision
    o.stokes_rtol = rtol
    o.heat_rtol = 1e-13
    o.grav_rtol = 1e-13
    return o


def scaled_e(p, s):
    q = copy.deepcopy(p)
    for t in range(len(q.rock)):
        q.rock[t].dh = p.rock[t].dh * s
    return q


def hot_c2(n=201, t_core=1100.0, t_surface=800.0):
    """The 201^2 Arrhenius setup with a warm body: T(r) from t_core at the
    centre to t_surface at the rim, so that all its silicate and iron sit
    between the viscosity clips (eta(1000 K) = 1e20, E = 300 kJ/mol:
    8.5e18 to 2.3e22 Pa s) and the activation energy reaches the whole flow.
    markv = 0: ronurecalc evaluates viscalc."""
    p, hs, cfg = configs.make("c2")
    configs.heat_state(hs, cfg, t_core=t_core, t_surface=t_surface)
    hs.markv[:] = 0
    return p, hs, cfg


DT = 1e9  # s: a fixed step, so the time step does not follow the velocities
# Stokes tolerance of the finite-difference cases: J moves by ~1e-4 of itself
# per 1e-4 change of s times E/RT ~ 40, so solver noise of 1e-11 in J is
# ~1e-7 of the difference; the central difference's own O((eps E/RT)^2 / 6)
# error is ~3e-6 at 900 K.
RTOL_FD = 1e-11


def solve_flow(ctx, p, hs):
    """ronurecalc + viterate at the fixed step, as the adjoint sees them."""
    ctx.set_params(p)
    ctx.upload(hs)
    ctx.ronurecalc()
    ctx.set_scalars(DT, 0.0)
    ctx.viterate()
    return ctx.sink_rate()


def adjoint_at(ctx, o, rtol=1e-10):
    """The adjoint with its transposed solve to rtol (it stalls near 3e-11 on
    the warm 201^2 case, where the forward solve reaches 1e-11)."""
    r = o.stokes_rtol
    o.stokes_rtol = rtol
    ctx.set_solver(o)
    try:
        return ctx.adjoint_sink_rate()
    finally:
        o.stokes_rtol = r
        ctx.set_solver(o)


# ------------------------------------------------------------------ operator transpose
def dot_test(ctx, N, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(3 * N)
    y = rng.standard_normal(3 * N)
    dx, dy, kx, kty = DeviceArray(x), DeviceArray(y), DeviceArray(3 * N), DeviceArray(3 * N)
    ctx.apply("stokes", dx.ptr, kx.ptr)
    ctx.apply("stokes_t", dy.ptr, kty.ptr)
    Kx, Kty = kx.numpy(), kty.numpy()
    for d in (dx, dy, kx, kty):
        d.free()
    lhs, rhs = float(np.dot(Kx, y)), float(np.dot(x, Kty))
    scale = float(np.sum(np.abs(Kx * y)))
    return lhs, rhs, scale


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_stokes_transpose_dot_product(name):
    say("case stokes transpose", name)
    p, hs, _ = configs.make(name)
    ctx = Context(p, hs, solver=opts())
    ctx.ronurecalc()
    ctx.viterate()
    for seed in (1, 2):
        lhs, rhs, scale = dot_test(ctx, hs.N, seed)
        say("%s: <Kx,y> %.16e <x,K^T y> %.16e scale %.3e" % (name, lhs, rhs, scale))
        assert abs(lhs - rhs) <= 1e-14 * scale
    ctx.close()


@pytest.mark.parametrize("rows", ["general", "continuity"])
def test_stokes_transpose_with_general_bc_rows(rows):
    """BC rows with neighbours anywhere (k_stokes_apply_t_bc) and the
    continuity-type rows (gathered with the cell shift)."""
    say("case stokes transpose bc rows", rows)
    p, hs, _ = configs.make("c1")
    (bch.general_rows if rows == "general" else bch.continuity_rows)(hs)
    ctx = Context(p, hs, solver=opts())
    ctx.ronurecalc()
    lhs, rhs, scale = dot_test(ctx, hs.N, 3)
    say("%s: <Kx,y> %.16e <x,K^T y> %.16e scale %.3e" % (rows, lhs, rhs, scale))
    assert abs(lhs - rhs) <= 1e-14 * scale
    ctx.close()


# ------------------------------------------------------------------ objective and gradient
@pytest.fixture(scope="module")
def hot():
    say("case flow adjoint hot c2")
    p, hs, cfg = hot_c2()
    o = opts(rtol=RTOL_FD)
    ctx = Context(p, hs, solver=o)
    J = solve_flow(ctx, p, hs)
    adj = adjoint_at(ctx, o)
    got = ctx.download(hs.copy())
    say("J %.12e adjoint %s" % (J, adj))
    yield dict(p=p, hs=hs, ctx=ctx, J=J, adj=adj, got=got)
    ctx.close()


def test_sink_rate_matches_oracle(hot, orc):
    """The objective restated by the oracle on the GPU's velocities and step."""
    p, got = hot["p"], hot["got"]
    got.timestep = DT
    Jo = orc.Oracle(p, got).sink_rate()
    say("J gpu %.15e oracle %.15e" % (hot["J"], Jo))
    # J nearly cancels (iron moves both ways around the centre): summation
    # order (GPU blocks vs. the oracle's loop) shows at ~1e-12 of it
    assert abs(hot["J"] - Jo) <= 1e-11 * abs(Jo)
    assert hot["adj"]["J"] == hot["J"]


def central(ctx, p, hs, eps):
    return (solve_flow(ctx, scaled_e(p, 1 + eps), hs) - solve_flow(ctx, scaled_e(p, 1 - eps), hs)) / (2 * eps)


def test_flow_adjoint_energy_matches_fd(hot):
    """dJ/ds, every activation energy scaled by s, against central
    differences of ronurecalc + viterate + the objective at s = 1 +- 1e-4:
    within 1e-5.  That difference carries its own O((h E/RT)^2 / 6) term,
    ~2e-6 here (E/RT ~ 33-45); Richardson's (4 D(h) - D(2h)) / 3 with
    h = 1e-4 cancels it (the next term, O((h E/RT)^4), is ~1e-9), leaving
    the forward solves' iteration error: within 2e-6."""
    p, hs, ctx = hot["p"], hot["hs"], hot["ctx"]
    ad = hot["adj"]["dJ_dEscale"]
    fd = central(ctx, p, hs, EPS)
    fd2 = central(ctx, p, hs, 2 * EPS)
    rich = (4 * fd - fd2) / 3
    say("dJ/ds: adjoint %.12e fd(1e-4) %.12e rel %.2e  fd(2e-4) %.12e  richardson %.12e rel %.2e "
        "(J %.6e, iters %d, relres %.2e)" % (ad, fd, abs(ad - fd) / abs(fd), fd2, rich, abs(ad - rich) / abs(rich),
                                            hot["J"], hot["adj"]["stokes_iters"], hot["adj"]["stokes_relres"]))
    assert abs(fd) > abs(hot["J"])  # the energy reaches the whole flow here
    assert abs(ad - fd) <= 1e-5 * abs(fd)
    assert abs(ad - rich) <= 2e-6 * abs(rich)


def test_flow_adjoint_mixed_matches_fp64(hot):
    say("case flow adjoint mixed")
    p, hs = hot["p"], hot["hs"]
    ctx = Context(p, hs, solver=opts(precision=1, rtol=1e-10))
    solve_flow(ctx, p, hs)
    adj = ctx.adjoint_sink_rate()
    ctx.close()
    a64 = hot["adj"]
    say("mixed %s" % adj)
    assert abs(adj["J"] - a64["J"]) <= 1e-8 * abs(a64["J"])
    assert abs(adj["dJ_dEscale"] - a64["dJ_dEscale"]) <= 1e-7 * abs(a64["dJ_dEscale"])


def test_flow_adjoint_euler_matches_fd():
    """markmod 1 (one velocity evaluation) on a smaller warm case."""
    say("case flow adjoint euler")
    p, hs, cfg = hot_c2()
    p.markmod = 1
    o = opts(rtol=RTOL_FD)
    ctx = Context(p, hs, solver=o)
    J = solve_flow(ctx, p, hs)
    ad = adjoint_at(ctx, o)["dJ_dEscale"]
    rich = (4 * central(ctx, p, hs, EPS) - central(ctx, p, hs, 2 * EPS)) / 3
    ctx.close()
    say("euler dJ/ds: adjoint %.12e richardson %.12e rel %.2e J %.6e" % (ad, rich, abs(ad - rich) / abs(rich), J))
    assert abs(ad - rich) <= 2e-6 * abs(rich)


def test_flow_adjoint_refusals():
    say("case flow adjoint refusals")
    p, hs, _ = configs.make("c1")
    ctx = Context(p, hs, solver=opts())
    ctx.ronurecalc()
    with pytest.raises(I2EError, match="no Stokes solution"):
        ctx.adjoint_sink_rate()
    ctx.viterate()
    ctx.titerate()  # changes the marker temperatures the viscosities came from
    with pytest.raises(I2EError, match="no Stokes solution"):
        ctx.adjoint_sink_rate()
    q = copy.deepcopy(p)
    q.rock[6].ss, q.rock[6].mm = 1e6, 3.0
    ctx.set_params(q)
    ctx.upload(hs)
    ctx.ronurecalc()
    ctx.viterate()
    with pytest.raises(I2EError, match="power"):
        ctx.adjoint_sink_rate()
    ctx.close()
