"""Benchmark: I2ELVIS time steps per second at 501x501 nodes, 6.25M markers.

One step = the reference's i2mart.c:60-153 sequence (maxvelstep, viterate
with gravity + Stokes, maxvelstep, titerate, movemark, ronurecalc) on the
501x501 / 25 markers-per-cell planetesimal (radius 150 km, Arrhenius rheology,
linear melting, 400 iron blobs) with every array resident in HBM.  The
mixed-precision solver (FP32 multigrid inside FP64 FGMRES) is the headline;
the same steps with the FP64 multigrid run from the same state for the
mixed/FP64 ratio.  Both solve Stokes to ||b - Ax|| / ||b|| <= 1e-8.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the roofline kernel, measured by two rocprofv3 --pmc
# runs (FETCH_SIZE, WRITE_SIZE) and scaled by the same counters on a plain
# 1 GiB copy; written by `bench.py --pmc-summary` (counters cannot be read
# inside the timed run).
TRAFFIC_FILES = {
    "vsmoothc": os.path.join(ROOT, "profiles", "k_tsmooth_f32_501x501_pmc_traffic.md"),
    "gsdot": os.path.join(ROOT, "profiles", "k_gsdot2_501x501_pmc_traffic.md"),
    "p2g": os.path.join(ROOT, "profiles", "k_ronu_markers_501x501_pmc_traffic.md"),
    "advect": os.path.join(ROOT, "profiles", "k_movemark_501x501_pmc_traffic.md"),
}
# rocprof kernel-name prefix and, where only some launches count, the grid
# size (threads) of the launches bench.py times for each roofline kernel
PMC_MATCH = {
    "vsmoothc": ("void i2e::k_tsmooth<float, float, 2, false, false, 32, 16>(", None),
    "gsdot": ("i2e::k_gsdot2(", "stokes"),
    "p2g": ("void i2e::k_ronu_markers<true>(", None),
    "advect": ("i2e::k_movemark(", None),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def physical_cores():
    """Physical cores of this host from lscpu (sockets x cores per socket)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        f = dict(l.split(":", 1) for l in out.splitlines() if ":" in l)
        return int(f["Socket(s)"].strip()) * int(f["Core(s) per socket"].strip())
    except Exception:
        return os.cpu_count() or 1


def oracle_step_phases(orc):
    """One oracle step (i2mart.c:60-153 order) with wall-clock seconds per
    phase: gravity, Stokes assembly + direct solve, heat, advection,
    marker-to-node."""
    hs, p = orc.hs, orc.p
    t = {}
    c = time.perf_counter
    hs.timestep = p.maxtmstep
    orc.maxvelstep()
    a = c()
    orc.gravity()
    t["gravity"] = c() - a
    a = c()
    M = orc.stokes_matrix()
    x = orc.solve_matrix(M)
    orc.stokes_post(x)
    t["stokes"] = c() - a
    hs.timestep = p.maxtmstep
    orc.maxvelstep()
    a = c()
    if hs.timestep:
        orc.titerate()
    t["heat"] = c() - a
    a = c()
    if hs.timestep:
        orc.movemark()
    t["advection"] = c() - a
    a = c()
    orc.ronurecalc()
    t["marker_to_node"] = c() - a
    hs.timesum += hs.timestep
    return t


def cpu_baseline(name="c3"):
    """The CPU oracle (serial C marker loops and node assembly, PARDISO
    direct solves factorised every step as the reference does) on the same
    501x501 / 6.25M-marker input: one full step with the sparse solves on
    all physical cores, one on 2 and one on 1 thread (the reference's
    production settings), each from the same state, split by phase."""
    from i2elvis_planet_amd import configs
    from oracle import oracle
    p, hs, _ = configs.make(name)
    oracle.Oracle(p, hs).ronurecalc()
    cores = physical_cores()
    runs = {}
    # the reference takes PARDISO's answer unchecked (gausmart.c:174-318);
    # the oracle's own check, which falls back to SuperLU when the residual
    # is above 1e-10, is off for the timing
    solver0, oracle.SOLVER = oracle.SOLVER, "pardiso"
    try:
        for nt in (cores, 2, 1):
            if nt in runs:
                continue
            oracle.set_solver_threads(nt)
            orc = oracle.Oracle(p, hs.copy())
            a = time.perf_counter()
            ph = oracle_step_phases(orc)
            runs[nt] = (time.perf_counter() - a, ph)
            log("cpu oracle step, %d thread(s): %.1f s %s" % (nt, runs[nt][0],
                                                              {k: round(v, 2) for k, v in ph.items()}))
    finally:
        oracle.SOLVER = solver0
    dt, ph = runs[cores]
    return {"value": 1.0 / dt, "unit": "timesteps/s", "cores": cores, "kind": "port",
            "sample": "1 full step of the %s setup after its initial ronurecalc: marker loops and assembly "
                      "serial C, sparse solves PARDISO with %d threads, its answer taken unchecked as the "
                      "reference takes it (%.1f s)" % (WORKLOADS[name][1].split(";")[0], cores, dt),
            "s_per_step_by_threads": {str(k): round(v[0], 3) for k, v in sorted(runs.items())},
            "phase_s": {k: round(v, 3) for k, v in ph.items()}}


# algorithmic-byte roofline kernels (HBM bound), selectable with --roofline-kernel
ROOFLINE_KERNELS = {
    "gsdot": "k_gsdot2 (fused Gram-Schmidt projection pass of the Stokes FGMRES)",
    "gsaxpy": "k_gsaxpy2 (fused Gram-Schmidt update pass of the Stokes FGMRES)",
    "vsmoothc": "k_tsmooth<float,float,2> (tiled velocity smoothing pass, 32x16 tiles, Galerkin levels "
                "251^2..17^2, mixed mode)",
    "vsmooth0": "k_tsmooth<float,float,1,diff> (tiled velocity smoothing pass, finest level, FP32 "
                "difference-form stencil, mixed mode)",
    "stokes_apply": "k_stokes_apply (matrix-free Stokes operator)",
    "advect": "k_movemark (RK4 marker advection)",
    "p2g": "k_ronu_markers + k_ronu_nodes (marker-to-node scatter and normalisation)",
    "heatmark": "k_heat_markers_a + k_heat_nodes_norm + k_heat_markers_b (marker temperature update)",
}


# launches sampled by the kernel timer: every 7th coarse smoothing pass (10
# per Stokes iteration, so the sample cycles through all levels); the others
# run once per iteration and are all timed
ROOFLINE_EVERY = {"vsmoothc": 7, "vsmooth0": 3, "precond": 1}


def stokes_gs_grid(n_nodes):
    """Grid size (threads) of a Gram-Schmidt launch of the Stokes FGMRES:
    ceil(3N / 512) blocks of 128 threads."""
    return (3 * n_nodes + 511) // 512 * 128


def read_traffic(rk):
    try:
        for line in open(TRAFFIC_FILES[rk]):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except (OSError, KeyError):
        pass
    return None


def pmc_summary(fetch_dir, write_dir, rk="vsmoothc", n_nodes=501 * 501, copy_bytes=None):
    """Markdown summary of two rocprofv3 --pmc runs of
    `bench.py --no-cpu --steps 1 --warmup 0 --copy-calibration 1024`:
    FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of the launches of roofline
    kernel `rk` that bench.py times, scaled by what the same counters read
    for the plain copy."""
    import csv
    prefix, only = PMC_MATCH[rk]

    def load(d):
        rows = list(csv.DictReader(open(os.path.join(d, "run_counter_collection.csv"))))
        cal = [float(r["Counter_Value"]) * 1024 for r in rows if "k_dev_copy" in r["Kernel_Name"]]
        ker = [float(r["Counter_Value"]) * 1024 for r in rows if r["Kernel_Name"].startswith(prefix) and
               (only is None or int(r["Grid_Size"]) == stokes_gs_grid(n_nodes))]
        return cal, ker
    kernel = prefix.rstrip("(")
    cf, kf = load(fetch_dir)
    cw, kw = load(write_dir)
    if not (cf and cw and kf and kw):
        raise SystemExit("pmc summary: calibration copy or kernel dispatches missing")
    nb = copy_bytes
    fetch_factor = nb / (sum(cf) / len(cf))
    write_factor = nb / (sum(cw) / len(cw))
    f_raw, w_raw = sum(kf) / len(kf), sum(kw) / len(kw)
    traffic = f_raw * fetch_factor + w_raw * write_factor
    lines = [
        "calibration copy: %d bytes read + %d written per dispatch; FETCH_SIZE read %.4g B (x%.4f), "
        "WRITE_SIZE %.4g B (x%.4f), means of %d / %d dispatches" % (nb, nb, sum(cf) / len(cf), fetch_factor,
                                                                   sum(cw) / len(cw), write_factor, len(cf), len(cw)),
        "%s%s: %d / %d dispatches; raw FETCH %.4g B, raw WRITE %.4g B per dispatch"
        % (kernel, " (Stokes-size, grid %d threads)" % stokes_gs_grid(n_nodes) if only else "", len(kf), len(kw),
           f_raw, w_raw),
        "traffic_bytes_per_launch: %.6g" % traffic,
    ]
    return "\n".join(lines)


KERNEL_TABLE = ("advect", "p2g", "heatmark", "stokes_apply", "gsdot", "gsaxpy", "vsmooth0", "vsmoothc")


def timed_steps(ctx, o, snap, mode, steps, timer=None):
    """Upload snap, run `steps` steps in `mode` (mixed / fp64); returns
    (seconds, per-step stats, timer reading or None)."""
    o.precision = 1 if mode == "mixed" else 0
    ctx.set_solver(o)
    ctx.upload(snap)
    ctx.sync()
    if timer:
        ctx.kernel_timer(timer[0], 1 << 17, timer[1])
    iters = []
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.step()
        st = ctx.stats()
        iters.append((st.stokes_iters, st.stokes_relres, st.grav_iters, st.heat_iters, st.ms_stokes, st.ms_heat,
                      st.ms_grav, st.ms_markers))
    ctx.sync()
    el = time.perf_counter() - t0
    reading = None
    if timer:
        reading = ctx.kernel_timer_read()
        ctx.kernel_timer(None, 0)
    return el, iters, reading


def kernel_table(ctx, o, snap):
    """Every kernel unit with an algorithmic byte count, timed on each launch
    with HIP events over one mixed step, outside the timed region.  Units of
    one kernel get the events stamped by their own dispatch (the kernel's
    duration as a kernel trace reports it); p2g and heatmark span several
    kernels and are bracketed by an event pair (~3 us more)."""
    out = {}
    for k in KERNEL_TABLE:
        _, _, (n, ms, alg, rep) = timed_steps(ctx, o, snap, "mixed", 1, (k, 1))
        if not n:
            continue
        gbs = alg / (ms * 1e-3) / 1e9
        out[k] = {"launches": n, "avg_us": round(ms * 1e3 / n, 2), "alg_bytes_per_launch": round(alg / n),
                  "alg_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                  "repr_bytes_per_launch": round(rep / n), "repr_GBs": round(rep / (ms * 1e-3) / 1e9, 1)}
    return out


def hot_attempt(ctx, o, hs_cold, cfg, direct_steps=2):
    """The molten regime (core at 1800 K, melt above 0.4 at 1e2 Pa s against
    the 1e23 clip): one Stokes solve attempted per precision with the
    multigrid preconditioner, capped at 300 Krylov iterations (reports what
    it reached); then the banded LU of the equilibrated rows
    (stokes_direct = 2): the first molten solve to an equilibrated residual
    of 3e-12 (one FGMRES iteration at 501^2; its floor lies near 1e-12), and
    direct_steps full steps at 1e-5, timed: once the melt limits the step,
    the residual one application of fresh factors leaves, which further
    iterations do not lower, rises as the step shrinks and the grid grows
    (~5e-10 at 101^2, ~2e-6 at 501^2); each step's is reported."""
    from i2elvis_planet_amd import I2EError, configs
    hot = configs.heat_state(hs_cold.copy(), cfg, t_core=1800.0)
    out = {}
    for mode in ("mixed", "fp64"):
        o.precision = 1 if mode == "mixed" else 0
        o2 = type(o).from_buffer_copy(o)
        o2.maxit = 300
        ctx.set_solver(o2)
        ctx.upload(hot)
        ctx.ronurecalc()
        st0 = ctx.stats()
        ctx.set_scalars(hot.timestep or 7.8894e10, hot.timesum)
        ctx.maxvelstep()
        t0 = time.perf_counter()
        try:
            ctx.viterate()
            conv = True
        except I2EError:
            conv = False
        ctx.sync()
        st = ctx.stats()
        rr = st.stokes_relres if math.isfinite(st.stokes_relres) else None  # NaN: a non-finite iterate
        out[mode] = {"converged": conv, "stokes_iters": st.stokes_iters, "stokes_relres": rr,
                     "ms": round((time.perf_counter() - t0) * 1e3, 1), "eta_min": st0.numin, "eta_max": st0.numax}
        log("hot %s: converged %s, %d its, relres %r" % (mode, conv, st.stokes_iters, rr))
    o2 = type(o).from_buffer_copy(o)
    o2.precision, o2.stokes_direct, o2.stokes_rtol, o2.maxit, o2.restart = 0, 2, 3e-12, 20, 20
    ctx.set_solver(o2)
    ctx.upload(hot)
    ctx.ronurecalc()
    ctx.set_scalars(hot.timestep or 7.8894e10, hot.timesum)
    ctx.maxvelstep()
    d = {}
    try:
        ctx.sync()
        t0 = time.perf_counter()
        ctx.viterate()
        ctx.sync()
        st = ctx.stats()
        kl, ku, nbytes = ctx.band_info()
        d["first_solve"] = {"ms": round((time.perf_counter() - t0) * 1e3, 1), "stokes_iters": st.stokes_iters,
                            "equilibrated_relres": st.stokes_relres, "band_kl": kl, "band_ku": ku,
                            "band_GB": round(nbytes / 1e9, 2)}
        o2.stokes_rtol = 1e-5
        ctx.set_solver(o2)
        steps = []
        for k in range(direct_steps):
            t0 = time.perf_counter()
            ctx.step()
            ctx.sync()
            st = ctx.stats()
            steps.append({"ms": round((time.perf_counter() - t0) * 1e3, 1), "ms_stokes": round(st.ms_stokes, 1),
                          "stokes_iters": st.stokes_iters, "equilibrated_relres": st.stokes_relres,
                          "dt_s": ctx.scalars()[0]})
        d["steps"] = steps
    except I2EError as e:
        d["error"] = str(e)
    out["direct_equilibrated"] = d
    log("hot direct: %s" % d)
    ctx.set_solver(o)
    return out


def valu_summary(pmc_dir, rk):
    """VALU activity of roofline kernel `rk` from one rocprofv3 --pmc run of
    SQ_ACTIVE_INST_VALU SQ_INSTS_VALU SQ_WAVE_CYCLES GRBM_GUI_ACTIVE: VALU
    busy = SQ_ACTIVE_INST_VALU * 4 / SIMDs / (GRBM_GUI_ACTIVE / XCDs), per
    dispatch, averaged (4 SIMDs per CU, 256 CUs; SQ_ACTIVE_INST_VALU counts
    quad-cycles; GRBM_GUI_ACTIVE is summed over the 8 XCDs: 8.3x the
    kernel's duration in 2.4 GHz cycles)."""
    import csv
    prefix, _ = PMC_MATCH[rk]
    rows = [r for r in csv.DictReader(open(os.path.join(pmc_dir, "run_counter_collection.csv")))
            if r["Kernel_Name"].startswith(prefix)]
    if not rows:
        raise SystemExit("valu summary: no %s dispatches" % prefix)
    per = {}
    for r in rows:
        per.setdefault(r["Dispatch_Id"], {})[r["Counter_Name"]] = float(r["Counter_Value"])
    simds = 4 * 256
    busy, ipw = [], []
    for d in per.values():
        if d.get("GRBM_GUI_ACTIVE") and "SQ_ACTIVE_INST_VALU" in d:
            busy.append(100.0 * d["SQ_ACTIVE_INST_VALU"] * 4 / simds / (d["GRBM_GUI_ACTIVE"] / 8))
        if d.get("SQ_WAVE_CYCLES") and "SQ_INSTS_VALU" in d:
            ipw.append(d["SQ_INSTS_VALU"] / d["SQ_WAVE_CYCLES"])
    return "\n".join([
        "%s: %d dispatches" % (prefix.rstrip("("), len(per)),
        "VALU busy (SQ_ACTIVE_INST_VALU*4/SIMDs/(GRBM_GUI_ACTIVE/8)): %.1f %%" % (sum(busy) / max(len(busy), 1)),
        "VALU instructions per wave cycle (SQ_INSTS_VALU/SQ_WAVE_CYCLES): %.3f" % (sum(ipw) / max(len(ipw), 1)),
    ])


def kernel_stats_table(stats_csv, top=30):
    """Markdown table of a rocprofv3 --stats kernel summary (run_kernel_stats.csv)."""
    import csv
    rows = list(csv.DictReader(open(stats_csv)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    out = ["| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---:|---:|---:|---:|---:|---:|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:top]:
        name = r["Name"].replace("void ", "").replace("i2e::", "").split("(")[0]
        out.append("| %s | %s | %.1f | %.1f | %.1f | %.1f | %.2f |" % (
            name, r["Calls"], float(r["TotalDurationNs"]) / 1e6, float(r["AverageNs"]) / 1e3,
            float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3, 100.0 * float(r["TotalDurationNs"]) / tot))
    out.append("")
    out.append("All kernels: %.1f ms over %d kernel names." % (tot / 1e6, len(rows)))
    return "\n".join(out)


WORKLOADS = {
    "c3": ("timesteps/s @501\u00b2, 6.3M markers: mixed vs FP64 GPU at Stokes rel.resid 1e-8",
           "501x501 nodes, 25 markers/cell (6.25M), R=150 km, Arrhenius 300 kJ/mol, linear melting 1360-1850 K, "
           "400 iron blobs; one i2mart step"),
    "c2": ("timesteps/s @201\u00b2, 640k markers: mixed vs FP64 GPU at Stokes rel.resid 1e-8",
           "201x201 nodes, 16 markers/cell (640k), R=100 km, Arrhenius 300 kJ/mol, porosity U[0.3,0.5]; "
           "one i2mart step"),
    "c1": ("timesteps/s @51\u00b2, 40k markers: mixed vs FP64 GPU at Stokes rel.resid 1e-8",
           "51x51 nodes, 16 markers/cell (40k), R=50 km, constant viscosity, 25 % iron; one i2mart step"),
}


def run(steps, warmup, skip_cpu, rk, minimal=False, name="c3"):
    import numpy as np
    from i2elvis_planet_amd import Context, configs, default_solver

    p, hs, cfg = configs.make(name)
    cold = hs.copy()
    log("setup generated: %d markers" % hs.marknum)
    o = default_solver()
    o.precision = 1
    ctx = Context(p, hs, solver=o)
    ctx.ronurecalc()
    for k in range(warmup):
        ctx.step()
        log("warmup step %d done" % k)
    ctx.sync()
    snap = ctx.download(hs.copy())
    res = {}
    roof = None
    ends = {}
    for mode in ("mixed", "fp64"):
        timer = (rk, ROOFLINE_EVERY.get(rk, 1)) if mode == "mixed" else None
        el, iters, reading = timed_steps(ctx, o, snap, mode, steps, timer)
        if reading:
            roof = reading
        res[mode] = (steps / el, el * 1e3 / steps, iters)
        ends[mode] = ctx.download(snap.copy())  # after the timed region
        log("%s: %.3f steps/s, %.1f ms/step; stokes its %s" % (mode, steps / el, el * 1e3 / steps,
                                                                [i[0] for i in iters]))
    # each field of the mixed run against the FP64 run after the same steps
    # from the same state (relative L2; both solve Stokes to 1e-8)
    field_err = {}
    for f in ("vx", "vy", "pr", "tk", "gp"):
        a, b = getattr(ends["mixed"], f), getattr(ends["fp64"], f)
        field_err[f] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    log("mixed vs fp64 after %d steps: %s" % (steps, {k: "%.1e" % v for k, v in field_err.items()}))
    # outside the timed region: per-kernel figures, the pre-melt regime with
    # a hot interior, the molten regime, standalone kernel timings
    kt, warm_res, hot = {}, None, None
    if not minimal and name == "c3":
        kt = kernel_table(ctx, o, snap)
        warm = configs.heat_state(cold.copy(), cfg, t_core=1340.0)
        ctx.upload(warm)
        ctx.ronurecalc()
        ctx.sync()
        warm_snap = ctx.download(warm.copy())
        warm_res = {}
        for mode in ("mixed", "fp64"):
            el, iters, _ = timed_steps(ctx, o, warm_snap, mode, steps)
            warm_res[mode] = (steps / el, [i[0] for i in iters], [i[1] for i in iters])
            log("warm %s: %.3f steps/s; stokes its %s" % (mode, steps / el, warm_res[mode][1]))
        hot = hot_attempt(ctx, o, cold, cfg)
    ctx.upload(snap)
    ctx.sort_markers()  # as the steps run them: grouped by cell tile
    ms_adv, ms_p2g, ms_sap = ctx.time_kernels(5)
    mixed_sps, mixed_ms, it = res["mixed"]
    fp64_sps = res["fp64"][0]
    launches, ms_tot, bytes_tot, repr_tot = roof
    avg_ms = ms_tot / max(launches, 1)
    achieved = bytes_tot / max(ms_tot, 1e-30) * 1e3 / 1e9
    line = {
        "metric": WORKLOADS[name][0],
        "value": mixed_sps,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": mixed_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (f32 multigrid preconditioner)",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[name][1], "stokes_rtol": 1e-8, "precision": "mixed"},
        "fp64_value": fp64_sps,
        "mixed_over_fp64": mixed_sps / fp64_sps,
        "stokes_iters_mixed": [i[0] for i in it],
        "stokes_iters_fp64": [i[0] for i in res["fp64"][2]],
        "stokes_relres_mixed": [i[1] for i in it],
        "mixed_vs_fp64_rel_l2": field_err,
        "heat_iters_mixed": [i[3] for i in it],
        "grav_iters_mixed": [i[2] for i in it],
        "phase_ms_mixed": {"stokes": float(np.mean([i[4] for i in it])), "heat": float(np.mean([i[5] for i in it])),
                           "gravity": float(np.mean([i[6] for i in it])),
                           "markers": float(np.mean([i[7] for i in it]))},
        "warm_core_1340K": None if warm_res is None else {
            "mixed": warm_res["mixed"][0], "fp64": warm_res["fp64"][0],
            "mixed_over_fp64": warm_res["mixed"][0] / warm_res["fp64"][0],
            "stokes_iters_mixed": warm_res["mixed"][1], "stokes_iters_fp64": warm_res["fp64"][1],
            "stokes_relres_mixed": warm_res["mixed"][2]},
        "hot_core_1800K": hot,
        "kernels": kt,
        "kernels_standalone_ms": {"advect_rk4": round(ms_adv, 4), "p2g_scatter": round(ms_p2g, 4),
                                  "stokes_apply_fp64": round(ms_sap, 4)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(rk),
                     "kernel": ROOFLINE_KERNELS[rk], "sampled_every": ROOFLINE_EVERY.get(rk, 1),
                     "launches": launches, "avg_us": avg_ms * 1e3,
                     "bytes_per_launch": bytes_tot / max(launches, 1),
                     "repr_bytes_per_launch": repr_tot / max(launches, 1),
                     "repr_GBs": repr_tot / max(ms_tot, 1e-30) * 1e3 / 1e9},
        "cpu_baseline": None,
    }
    ctx.close()
    if not skip_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(name)
        except Exception as e:  # reported, never fatal to the GPU number
            log("cpu baseline failed: %r" % (e,))
    return line


def ensemble(nbody, steps, warmup):
    """Aggregate timesteps/s of an ensemble of independent bodies (radius
    10-150 km, grids 101^2-301^2, 16 markers/cell, H0 and activation energy
    drawn per body), every body from the same state both ways: all bodies
    stepped together (i2e_step_batch: one host thread and HIP stream per
    body, their launches overlapping on the GPU) against the same bodies
    stepped one after the other on the same GPU.  Mixed precision at Stokes
    rel. residual 1e-8 with autotuned coarse-level tile shapes."""
    from i2elvis_planet_amd import Context, configs, default_solver, step_batch
    bodies = configs.ensemble(nbody)
    sizes = [b[2]["n"] for b in bodies]
    o = default_solver()
    o.autotune = 1
    log("ensemble of %d bodies generated (grids %d^2..%d^2, %d markers)" % (
        nbody, min(sizes), max(sizes), sum(b[1].marknum for b in bodies)))
    ctxs = [Context(p, hs, solver=o) for p, hs, _ in bodies]
    for c in ctxs:
        c.ronurecalc()
    step_batch(ctxs, warmup)
    for c in ctxs:
        c.sync()
    snaps = [c.download(b[1].copy()) for c, b in zip(ctxs, bodies)]
    tuned = {}
    for c in ctxs:
        for lv, sx, sy in c.tuned_shapes():
            tuned["%d^2" % lv] = "%dx%d" % (sx, sy)
    log("tuned tile shapes: %s" % tuned)
    t0 = time.perf_counter()
    step_batch(ctxs, steps)
    for c in ctxs:
        c.sync()
    tb = time.perf_counter() - t0
    log("batched: %d bodies x %d steps in %.2f s" % (nbody, steps, tb))
    iters_b = [c.stats().stokes_iters for c in ctxs]
    for c, sn in zip(ctxs, snaps):
        c.upload(sn)
    for c in ctxs:
        c.sync()
    t0 = time.perf_counter()
    for c in ctxs:
        for _ in range(steps):
            c.step()
        c.sync()
    ts = time.perf_counter() - t0
    log("sequential: %.2f s" % ts)
    for c in ctxs:
        c.close()
    return {"metric": "ensemble aggregate timesteps/s", "value": nbody * steps / tb, "unit": "timesteps/s",
            "n_gpus": 1, "n_bodies": nbody, "steps": steps, "warmup": warmup, "higher_is_better": True,
            "batched": nbody * steps / tb, "sequential": nbody * steps / ts, "speedup": ts / tb,
            "dtype": "f64 (f32 multigrid preconditioner)", "data": "synthetic",
            "tuned_shapes": tuned, "last_stokes_iters": iters_b,
            "config": {"workload": "%d bodies, radius U[10,150] km, grids 101^2-301^2 (%d markers in all), "
                                   "16 markers/cell, H0 log-U[1e-8,3e-7] W/kg, E U[250,350] kJ/mol, mixed "
                                   "precision at Stokes rel. residual 1e-8"
                                   % (nbody, sum(b[1].marknum for b in bodies))}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--minimal", action="store_true",
                    help="only the timed steps and the roofline kernel (for counter runs)")
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS),
                    help="setup of the timed steps: c3 (501^2, the metric's), c2 (201^2) or c1 (51^2)")
    ap.add_argument("--roofline-kernel", default="vsmoothc", choices=sorted(ROOFLINE_KERNELS),
                    help="kernel of the roofline entry (default: the one with the most GPU time)")
    ap.add_argument("--ensemble", type=int, default=0, help="run the ensemble measurement with this many bodies")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device three times (PMC calibration)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) and exit")
    ap.add_argument("--valu-summary", metavar="PMC_DIR",
                    help="summarise a rocprofv3 --pmc run of the VALU counters for --roofline-kernel and exit")
    ap.add_argument("--stats-table", metavar="KERNEL_STATS_CSV",
                    help="print a rocprofv3 --stats kernel summary as a markdown table and exit")
    a = ap.parse_args()
    if a.pmc_summary:
        print(pmc_summary(*a.pmc_summary, rk=a.roofline_kernel, copy_bytes=a.copy_calibration * 2 ** 20 or 2 ** 30))
        return
    if a.valu_summary:
        print(valu_summary(a.valu_summary, a.roofline_kernel))
        return
    if a.stats_table:
        print(kernel_stats_table(a.stats_table))
        return
    if a.copy_calibration:
        from i2elvis_planet_amd import DeviceArray
        n = a.copy_calibration * 2 ** 20 // 8
        src, dst = DeviceArray(n), DeviceArray(n)
        for _ in range(3):
            dst.copy_from(src)
        src.free()
        dst.free()
        log("calibration copy: 3 x %d bytes" % (8 * n))
    if a.gpus != 1:
        print("bench.py runs one GPU only: the multi-GPU decomposition of this step is not built "
              "(see the sharding notes); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if a.ensemble:
        print(json.dumps(ensemble(a.ensemble, a.steps, a.warmup)), flush=True)
        return
    line = run(a.steps, a.warmup, a.no_cpu, a.roofline_kernel, a.minimal, a.workload)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
