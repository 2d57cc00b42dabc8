"""torch.autograd.Function around the peak-temperature adjoint (test harness
only: the library never imports torch).

    J = PeakT.apply(h0_scale, phi0_shift)     # CPU scalar tensors
    J.backward()                              # gradients from i2e_adjoint_peak_t

forward() runs ronurecalc + K full steps of libi2e from the given state
(i2e_record) and returns J; backward() returns dJ/dh0_scale and dJ/dphi0
from the recorded run.  Run as a script it checks autograd's gradient
against the adjoint called directly and against central differences, and
prints one JSON line.  libi2e is loaded before torch, so that torch's
bundled HIP runtime is not loaded next to the system one; torch only ever
sees CPU tensors.
"""
import copy
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import i2elvis_planet_amd as i2e  # noqa: E402

i2e.load_library()
import torch  # noqa: E402


class PeakT(torch.autograd.Function):
    """J(h0_scale, phi0): mass-weighted mean peak temperature of silicate markers."""

    @staticmethod
    def forward(ctx, h0_scale, phi0, run):
        p = copy.deepcopy(run["params"])
        p.h0_scale = float(h0_scale)
        hs = run["state"].copy()
        live = hs.markpor > 0
        hs.markpor[live] = (hs.markpor[live].astype(np.float64) + float(phi0)).astype(np.float32)
        c = run["ctx"]
        c.set_params(p)
        c.upload(hs)
        c.record(run["steps"])
        adj = c.adjoint_peak_t()
        ctx.grads = (adj["dJ_dh0"], adj["dJ_dphi0"])
        return torch.tensor(adj["J"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        gh, gp = ctx.grads
        return g * torch.tensor(gh, dtype=torch.float64), g * torch.tensor(gp, dtype=torch.float64), None


def main():
    from i2elvis_planet_amd import Context, configs, default_solver
    cfg = dict(configs.CONFIGS["c4"], n=81, steps=10)
    p, hs = configs.make_params(cfg), configs.make_state(cfg)
    o = default_solver()
    o.precision = 0
    o.stokes_rtol = 1e-10
    o.heat_rtol = 1e-13
    o.grav_rtol = 1e-13
    ctx = Context(p, hs, solver=o)
    run = {"params": p, "state": hs, "ctx": ctx, "steps": cfg["steps"]}
    h0 = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    ph = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    J = PeakT.apply(h0, ph, run)
    J.backward()
    direct = ctx.adjoint_peak_t()  # the same recorded run, called without torch
    # central difference in h0_scale through the full step (linear in h0 here)
    e = 1e-3
    Jp = PeakT.apply(torch.tensor(1.0 + e, dtype=torch.float64), torch.tensor(0.0, dtype=torch.float64), run)
    Jm = PeakT.apply(torch.tensor(1.0 - e, dtype=torch.float64), torch.tensor(0.0, dtype=torch.float64), run)
    fd = float(Jp - Jm) / (2 * e)
    ctx.close()
    out = {"J": float(J), "grad_h0": float(h0.grad), "grad_phi0": float(ph.grad),
           "direct_h0": direct["dJ_dh0"], "direct_phi0": direct["dJ_dphi0"], "fd_h0": fd}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
