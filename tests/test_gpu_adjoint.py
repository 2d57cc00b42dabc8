"""GPU tests of the peak-temperature adjoint (i2e_record / i2e_replay /
i2e_adjoint_peak_t) and of the transposed heat operator, through the C-ABI.

Pins: the dot-product identity <A x, y> = <x, A^T y>; central finite
differences of the full time step (h0_scale, activation energy) and of the
frozen-trajectory replay (porosity shift) with relative steps 1e-4; central
finite differences of the CPU oracle's step; FP64 vs mixed-precision solves.
"""
import copy
import sys

import numpy as np
import pytest

from i2elvis_planet_amd import Context, DeviceArray, I2EError, configs, default_solver

pytestmark = pytest.mark.gpu

EPS = 1e-4  # relative finite-difference step
# Marker temperatures are float32 (headmart.c:122-127): every step rounds them,
# which puts a noise floor of ~4e-6 (201^2, 50 steps) to ~4e-5 (51^2, 10
# steps) under a 1e-4 central difference of J.  Where a test asks for 1e-5 on
# a smaller case it steps by 1e-3 instead; J is linear in h0_scale and smooth
# in the porosity there, so the larger step adds no truncation error of note.
EPS_SMALL_CASE = 1e-3


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def opts(precision=0):
    o = default_solver()
    o.precision = precision
    o.stokes_rtol = 1e-10
    o.heat_rtol = 1e-13
    o.grav_rtol = 1e-13
    return o


def params_with(p, **kw):
    q = copy.deepcopy(p)
    for k, v in kw.items():
        setattr(q, k, v)
    return q


def run_full(ctx, p, hs, nsteps):
    """J after ronurecalc + nsteps full steps from hs under params p."""
    ctx.set_params(p)
    ctx.upload(hs)
    ctx.ronurecalc()
    for _ in range(nsteps):
        ctx.step()
    return ctx.peak_t()


def run_replay(ctx, hs, nsteps):
    ctx.upload(hs)
    ctx.replay(nsteps)
    return ctx.peak_t()


def shift_porosity(hs, d):
    """hs with every positive porosity shifted by d, and the shift the float32
    storage actually applied: porosities of one binade share one ulp, so the
    rounded shift is the same for nearly all of them and differs from d by up
    to half an ulp -- 3e-5 of d = 4.5e-4, enough to spoil a finite
    difference.  The few markers at a binade edge (porosity near 0.5) get a
    shift one ulp off, a 1e-3 relative change on under 1e-3 of the markers."""
    h = hs.copy()
    live = h.markpor > 0
    h.markpor[live] = (h.markpor[live].astype(np.float64) + d).astype(np.float32)
    eff = h.markpor[live].astype(np.float64) - hs.markpor[live].astype(np.float64)
    vals, counts = np.unique(eff, return_counts=True)
    assert counts.max() >= 0.999 * eff.size, (vals, counts)
    return h, float(vals[np.argmax(counts)])


def replay_fd_phi0(ctx, hs, K, d):
    hp, dp = shift_porosity(hs, d)
    hm, dm = shift_porosity(hs, -d)
    return (run_replay(ctx, hp, K) - run_replay(ctx, hm, K)) / (dp - dm)


# ------------------------------------------------------------------ operator transpose
def test_heat_transpose_dot_product():
    """<A x, y> = <x, A^T y> for titerate's rows.  The fixed-T rows are
    identity rows, so y is scaled down there (1e-13) to keep them from
    hiding the equation rows; a second pass with dt = 1e30 leaves only the
    conductivity stencil."""
    say("case heat transpose c1")
    p, hs, _ = configs.make("c1")
    ctx = Context(p, hs, solver=opts())
    ctx.ronurecalc()
    ctx.step()  # leaves titerate's node fields and dt in the context
    dt, ts, _ = ctx.scalars()
    N = hs.N
    bc = hs.bondm[3 * N:] != 0
    rng = np.random.default_rng(11)
    for step in (dt, 1e30):
        ctx.set_scalars(step, ts)
        x = rng.standard_normal(N) * 300.0
        y = rng.standard_normal(N) * np.where(bc, 1e-13, 1.0)
        dx, dy, ax, aty = DeviceArray(x), DeviceArray(y), DeviceArray(N), DeviceArray(N)
        ctx.apply("heat", dx.ptr, ax.ptr)
        ctx.apply("heat_t", dy.ptr, aty.ptr)
        Ax, Aty = ax.numpy(), aty.numpy()
        lhs, rhs = float(np.dot(Ax, y)), float(np.dot(x, Aty))
        scale = float(np.sum(np.abs(Ax * y)))
        say("dt %.3e: <Ax,y> %.16e  <x,A^T y> %.16e  scale %.3e" % (step, lhs, rhs, scale))
        assert abs(lhs - rhs) <= 1e-13 * scale
        for d in (dx, dy, ax, aty):
            d.free()
    ctx.close()


# ------------------------------------------------------------------ the 201^2 adjoint case
@pytest.fixture(scope="module")
def c4():
    say("case adjoint c4 record")
    p, hs, cfg = configs.make("c4")
    K = cfg["steps"]
    ctx = Context(p, hs, solver=opts())
    ctx.record(K)
    J = ctx.peak_t()
    adj = ctx.adjoint_peak_t()
    say("c4: J %.12f  adjoint %s" % (J, adj))
    yield dict(p=p, hs=hs, K=K, ctx=ctx, J=J, adj=adj)
    ctx.close()


def test_adjoint_objective_matches_oracle(c4, orc):
    got = c4["ctx"].download(c4["hs"].copy())
    Jo = orc.peak_t_objective(c4["p"], got)
    say("J gpu %.15f oracle objective %.15f" % (c4["J"], Jo))
    assert abs(c4["J"] - Jo) <= 1e-13 * Jo
    assert c4["adj"]["J"] == c4["J"]
    assert c4["J"] > 150.0 + 1.0  # the run heated the silicate


def test_adjoint_h0_matches_full_step_fd(c4):
    """The flow of this case does not depend on h0_scale (density independent
    of T, viscosities at their upper clip), so the frozen-trajectory gradient
    is the full-model gradient."""
    say("case adjoint c4 fd h0")
    p, hs, K, ctx = c4["p"], c4["hs"], c4["K"], c4["ctx"]
    h0 = p.h0_scale
    Jp = run_full(ctx, params_with(p, h0_scale=h0 * (1 + EPS)), hs, K)
    Jm = run_full(ctx, params_with(p, h0_scale=h0 * (1 - EPS)), hs, K)
    ctx.set_params(p)
    fd = (Jp - Jm) / (2 * EPS * h0)
    ad = c4["adj"]["dJ_dh0"]
    say("dJ/dh0: adjoint %.12e  fd %.12e  rel %.2e" % (ad, fd, abs(ad - fd) / abs(fd)))
    assert abs(ad - fd) <= 1e-5 * abs(fd)


def test_activation_energy_does_not_reach_J(c4):
    """The adjoint carries no activation-energy path: in this case every
    silicate and iron viscosity stays at its 1e23 clip, so J(E) is flat --
    the full step at E(1 +- 1e-4) must give the same J up to summation noise."""
    say("case adjoint c4 fd E")
    p, hs, K, ctx = c4["p"], c4["hs"], c4["K"], c4["ctx"]
    J = []
    for s in (1 + EPS, 1 - EPS):
        q = copy.deepcopy(p)
        for t in range(len(q.rock)):
            q.rock[t].dh = p.rock[t].dh * s
        J.append(run_full(ctx, q, hs, K))
    ctx.set_params(p)
    say("J(E+) %.15f J(E-) %.15f" % tuple(J))
    assert abs(J[0] - J[1]) <= 1e-12 * c4["J"]


def test_replay_reproduces_record(c4):
    say("case adjoint c4 replay")
    Jr = run_replay(c4["ctx"], c4["hs"], c4["K"])
    say("J record %.15f replay %.15f" % (c4["J"], Jr))
    assert abs(Jr - c4["J"]) <= 1e-12 * c4["J"]


def test_adjoint_phi0_matches_replay_fd(c4):
    """Porosity reaches J through density (heat capacity, radiogenic heat,
    marker mass) and the porous-conductivity closure; the replay holds the
    recorded trajectories, so its finite difference is what the adjoint
    differentiates.  The closure jumps by 0.09 % at porosity 0.4
    (markmart.c:614-638): markers crossing it add an O(1e-4) share that no
    derivative sees, hence the wider tolerance."""
    say("case adjoint c4 fd phi0")
    hs, K, ctx = c4["hs"], c4["K"], c4["ctx"]
    fd = replay_fd_phi0(ctx, hs, K, EPS * 0.4)
    ad = c4["adj"]["dJ_dphi0"]
    say("dJ/dphi0: adjoint %.12e  fd %.12e  rel %.2e" % (ad, fd, abs(ad - fd) / abs(fd)))
    assert abs(ad - fd) <= 5e-4 * abs(fd)


def test_adjoint_phi0_smooth_branch_fd():
    """Same finite difference with every porosity in (0.41, 0.49), where the
    conductivity closure is smooth, over 20 steps: the adjoint must match to
    1e-5."""
    say("case adjoint c4 smooth-porosity fd phi0")
    p, hs, cfg = configs.make("c4")
    K = 20
    live = hs.markpor > 0
    rng = np.random.default_rng(5)
    hs.markpor[live] = rng.uniform(0.41, 0.49, int(live.sum())).astype(np.float32)
    ctx = Context(p, hs, solver=opts())
    ctx.record(K)
    ad = ctx.adjoint_peak_t()["dJ_dphi0"]
    fd = replay_fd_phi0(ctx, hs, K, EPS * 0.45)
    ctx.close()
    say("smooth dJ/dphi0: adjoint %.12e  fd %.12e  rel %.2e" % (ad, fd, abs(ad - fd) / abs(fd)))
    assert abs(ad - fd) <= 1e-5 * abs(fd)


def test_adjoint_mixed_matches_fp64(c4):
    say("case adjoint c4 mixed")
    p, hs, K = c4["p"], c4["hs"], c4["K"]
    ctx = Context(p, hs, solver=opts(precision=1))
    ctx.record(K)
    adj = ctx.adjoint_peak_t()
    ctx.close()
    a64 = c4["adj"]
    say("mixed %s" % adj)
    assert abs(adj["J"] - a64["J"]) <= 1e-9 * a64["J"]
    assert abs(adj["dJ_dh0"] - a64["dJ_dh0"]) <= 1e-7 * abs(a64["dJ_dh0"])
    assert abs(adj["dJ_dphi0"] - a64["dJ_dphi0"]) <= 1e-7 * abs(a64["dJ_dphi0"])


# ------------------------------------------------------------------ against the oracle's step
def test_adjoint_h0_matches_oracle_fd(orc):
    """51x51 case, 10 full steps (shear heating on, constant viscosity, so the
    flow does not depend on h0_scale): the GPU adjoint against central
    differences of the CPU oracle's time step."""
    say("case adjoint c1 vs oracle fd")
    p, hs, cfg = configs.make("c1")
    p.maxtkstep = 1e30  # no heating-rate limit on dt (see the c4 configuration)
    K = 10
    ctx = Context(p, hs, solver=opts())
    ctx.record(K)
    adj = ctx.adjoint_peak_t()
    ctx.close()
    J = []
    for s in (1 + EPS_SMALL_CASE, 1 - EPS_SMALL_CASE):
        q = params_with(p, h0_scale=p.h0_scale * s)
        h = hs.copy()
        o = orc.Oracle(q, h)
        o.ronurecalc()
        for _ in range(K):
            o.step()
        J.append(orc.peak_t_objective(q, h))
    fd = (J[0] - J[1]) / (2 * EPS_SMALL_CASE * p.h0_scale)
    say("c1 dJ/dh0: adjoint %.12e  oracle fd %.12e  rel %.2e" % (adj["dJ_dh0"], fd,
                                                              abs(adj["dJ_dh0"] - fd) / abs(fd)))
    assert abs(adj["dJ_dh0"] - fd) <= 1e-5 * abs(fd)


def test_adjoint_refuses_heating_limited_time_steps():
    """With the reference's maxtkstep (1e4 K) titerate shortens dt as the body
    heats; the gradient through that limit is not built, so the adjoint and
    the replay must refuse the record instead of returning a wrong value."""
    say("case adjoint c1 heating-limited dt")
    p, hs, _ = configs.make("c1")
    p.frictyn = 1e30  # no shear heating, which replay refuses on its own
    ctx = Context(p, hs, solver=opts())
    ctx.record(4)
    with pytest.raises(I2EError, match="heating-rate limit"):
        ctx.adjoint_peak_t()
    with pytest.raises(I2EError, match="heating-rate limit"):
        ctx.replay(4)
    ctx.close()


def test_torch_autograd_wrapper():
    """The adjoint wrapped as a torch.autograd.Function (tests/harness), run in
    a child process so that torch never shares this process with libi2e:
    autograd's gradient is the adjoint's, and dJ/dh0 matches a central
    difference through the same wrapper."""
    import json
    import os
    import subprocess
    import sys
    say("case adjoint torch.autograd wrapper (child process)")
    here = os.path.dirname(os.path.abspath(__file__))
    r = subprocess.run([sys.executable, "-u", os.path.join(here, "harness", "adjoint_autograd.py")],
                       capture_output=True, text=True, timeout=200)
    say(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    out = json.loads(r.stdout.strip().splitlines()[-1])
    # the same recorded run (node sums by atomics: equal to round-off)
    assert abs(out["grad_h0"] - out["direct_h0"]) <= 1e-12 * abs(out["direct_h0"])
    assert abs(out["grad_phi0"] - out["direct_phi0"]) <= 1e-12 * abs(out["direct_phi0"])
    assert abs(out["grad_h0"] - out["fd_h0"]) <= 1e-5 * abs(out["fd_h0"])
