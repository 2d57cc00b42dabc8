"""The coarse velocity levels in one cooperative launch per V-cycle
(coarse_coop = 1) against one launch per pass (coarse_coop = 0, the
default), through the C-ABI.  Both do the same arithmetic on the same 32x16 tiles
(block Gauss-Seidel with the tile halo frozen for a pass), the two-level
tail and the dense coarsest solve in the same order, so the V-cycle, the
Stokes preconditioner and a whole mixed-precision viterate agree bit for
bit; the launch counts show which path ran.
"""
import sys

import numpy as np
import pytest

from i2elvis_planet_amd import Context, DeviceArray, configs, default_solver

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def opts(coop):
    o = default_solver()
    o.coarse_coop = coop
    return o


def vcycle_out(ctx, f):
    df, du = DeviceArray(f), DeviceArray(f.size)
    ctx.kernel_timer("vcoarse")
    ctx.vcycle(df.ptr, du.ptr)
    n = ctx.kernel_timer_read()[0]
    ctx.kernel_timer(None, 0)
    u = du.numpy()
    df.free()
    du.free()
    return u, n


def precond_out(ctx, r):
    dr, dz = DeviceArray(r), DeviceArray(r.size)
    ctx.stokes_precond(dr.ptr, dz.ptr)
    z = dz.numpy()
    dr.free()
    dz.free()
    return z


@pytest.mark.parametrize("name", ["c2", "c3"])
def test_coop_vcycle_and_precond_bitwise(name):
    say("case coop vcycle", name)
    p, hs, _ = configs.make(name)
    ctx = Context(p, hs, solver=opts(1))
    ctx.ronurecalc()
    rng = np.random.default_rng(11)
    f = rng.standard_normal(2 * hs.N)
    r = rng.standard_normal(3 * hs.N)
    u1, n1 = vcycle_out(ctx, f)
    z1 = precond_out(ctx, r)
    u1b, _ = vcycle_out(ctx, f)  # the argument upload is cached between calls
    ctx.set_solver(opts(0))
    u0, n0 = vcycle_out(ctx, f)
    z0 = precond_out(ctx, r)
    ctx.close()
    say("%s: cooperative launches %d / %d, max |du| %.3e, max |dz| %.3e" %
        (name, n1, n0, np.max(np.abs(u1 - u0)), np.max(np.abs(z1 - z0))))
    assert n1 == 1 and n0 == 0
    assert np.all(np.isfinite(u1)) and np.max(np.abs(u1)) > 0
    assert np.array_equal(u1, u1b)
    assert np.array_equal(u1, u0)
    assert np.array_equal(z1, z0)


def test_coop_viterate_bitwise():
    """Whole mixed-precision Stokes solves at 501^2 (FGMRES with the V-cycle
    in every iteration), twice in a row (the second from the extrapolated
    guess), from one set of node fields: the same iterations, residuals,
    velocities and pressures to the bit.  (The fields come from one
    ronurecalc and are uploaded again for each mode: the marker scatter sums
    with atomics, so two ronurecalc calls may differ in the last bits.)"""
    say("case coop viterate c3")
    p, hs, _ = configs.make("c3")
    ctx = Context(p, hs, solver=opts(1))
    ctx.ronurecalc()
    snap = ctx.download(hs.copy())
    out = []
    for coop in (1, 0):
        ctx.set_solver(opts(coop))
        ctx.upload(snap)
        ctx.set_scalars(p.maxtmstep, 0.0)
        ctx.viterate()
        st1 = ctx.stats()
        ctx.viterate()
        st2 = ctx.stats()
        got = ctx.download(hs.copy())
        out.append((st1.stokes_iters, st1.stokes_relres, st2.stokes_iters, st2.stokes_relres, got))
        say("coop %d: iters %d relres %.3e; again iters %d relres %.3e" % ((coop,) + out[-1][:4]))
    ctx.close()
    a, b = out
    assert a[1] <= 1e-8 and a[3] <= 1e-8
    assert a[:4] == b[:4]
    for fld in ("vx", "vy", "pr"):
        assert np.array_equal(getattr(a[4], fld), getattr(b[4], fld)), fld
