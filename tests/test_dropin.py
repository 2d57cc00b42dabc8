"""The drop-in shim (include/i2e_dropin.h, csrc/dropin.c): the reference's
step functions by name over globals named and typed as headmart.c defines
them, driven by a C stand-in of the reference's main() loop
(tests/dropin/ref_driver.c, i2mart.c:60-164 order), against the oracle's
step on the same .prn state."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from i2elvis_planet_amd import configs, prn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "i2elvis_planet_amd")


def build_driver(tmp_path):
    exe = str(tmp_path / "ref_driver")
    cmd = ["gcc", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-o", exe,
           os.path.join(ROOT, "tests", "dropin", "ref_driver.c"), os.path.join(PKG, "csrc", "dropin.c"),
           "-L", PKG, "-l:libi2e.so", "-Wl,-rpath," + PKG]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


MODE_KEYS = ("maxxystep", "maxtkstep", "maxtmstep", "nubeg", "nuend", "nucontr", "strmin", "heatdif", "frictyn",
             "adiabyn", "por_init", "k_cutoff", "heatmod", "densimod", "outgrid", "markmod")


def write_case(tmp_path, name="c1", rtol=(1e-13, 1e-13, 1e-13)):
    p, hs, _ = configs.make(name)
    for t in (6, 7, 26):
        p.rock[t].bb = 3e-5  # thermal expansion drives the soft-turbulence flux
    path = tmp_path / "state.prn"
    prn.write(path, p, hs, n0=0)
    q, back, _ = prn.read(path, params=p)
    mode = tmp_path / "mode.txt"
    lines = ["%s %r" % (k, float(getattr(p, k))) for k in MODE_KEYS]
    # what the reference fixes in code, set by the hook to the config's values
    lines += ["radmax_frac %r" % p.radmax_frac, "longlived %r" % p.longlived, "h0_scale %r" % p.h0_scale,
              "porosity_on %d" % p.porosity_on, "stokes_rtol %r" % rtol[0], "heat_rtol %r" % rtol[1],
              "grav_rtol %r" % rtol[2]]
    # the soft-turbulence conductivity and the magnetisation run as in the
    # reference (the shim switches both on); the oracle gets the same
    q.soft_turb, q.magnet, q.core_form_time, q.dynamo = 1, 1, 0.0, 1
    q.fe_ts, q.fe_tl = 400.0, 450.0
    lines += ["core_form_time 0", "dynamo 1", "fe_ts 400", "fe_tl 450"]
    mode.write_text("\n".join(lines) + "\n")
    return q, back, path, mode


def read_out(path, N, M):
    b = open(path, "rb").read()
    f = {}
    o = 0
    for k in ("vx", "vy", "pr", "tk", "gp"):
        f[k] = np.frombuffer(b, np.float64, N, o)
        o += 8 * N
    for k in ("markx", "marky", "markk"):
        f[k] = np.frombuffer(b, np.float32, M, o)
        o += 4 * M
    f["markt"] = np.frombuffer(b, np.int8, M, o)
    o += M
    f["markmg_old"] = np.frombuffer(b, np.int8, M, o)
    o += M
    f["markmg_time"] = np.frombuffer(b, np.float64, M, o)
    o += 8 * M
    f["marknum"], f["timestep"], f["timesum"], f["gxkoef"] = struct.unpack_from("<q3d", b, o)
    assert o + 32 == len(b)
    return f


def test_driver_builds_and_reports_a_missing_gpu(tmp_path):
    """The shim compiles against the reference-typed globals and links with
    libi2e.so; without a visible GPU its first call exits with
    I2E_ERR_NODEVICE (6) and says why, as the reference exits on a failed
    solve (gausmart.c:271-304)."""
    exe = build_driver(tmp_path)
    if os.environ.get("I2E_GPU_PRESENT") or os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible here")
    q, back, state, mode = write_case(tmp_path)
    r = subprocess.run([exe, str(state), str(mode), "1", str(tmp_path / "o.bin")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 6, r.stderr
    assert "ronurecalc" in r.stderr and "device" in r.stderr


@pytest.mark.gpu
def test_dropin_steps_match_the_oracle(orc, tmp_path):
    """Three steps of the 51^2 setup through maxvelstep/viterate/titerate/
    movemark/ronurecalc called by name from C, against the oracle's steps:
    dt to 1e-10, node fields to 1e-8, marker types and per-cell type counts
    exact."""
    print("case dropin c1 3 steps", file=sys.stderr, flush=True)
    exe = build_driver(tmp_path)
    q, back, state, mode = write_case(tmp_path)
    out = tmp_path / "o.bin"
    r = subprocess.run([exe, str(state), str(mode), "3", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ref = back.copy()
    o = orc.Oracle(q, ref)
    o.ronurecalc()
    for _ in range(3):
        o.step()
    got = read_out(out, ref.N, ref.marknum)
    assert got["marknum"] == ref.marknum
    assert got["timestep"] == pytest.approx(ref.timestep, rel=1e-10)
    assert got["timesum"] == pytest.approx(ref.timesum, rel=1e-10)
    for k in ("vx", "vy", "pr", "tk", "gp"):
        a, b = got[k], getattr(ref, k)
        e = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(k, "rel L2 %.2e" % e, file=sys.stderr)
        assert e < 1e-8, k
    assert np.array_equal(got["markt"], ref.markt)
    assert np.array_equal(got["markmg_old"], ref.markmg_old)
    assert np.array_equal(got["markmg_time"], ref.markmg_time)
    g = ref.copy()
    g.markx[:], g.marky[:], g.markt[:] = got["markx"], got["marky"], got["markt"]
    assert np.array_equal(orc.Oracle(q, g).type_counts(), o.type_counts())
