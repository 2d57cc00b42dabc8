"""The C-ABI library loads without a GPU, exports every entry point that
include/i2e.h declares, and the ctypes structs match the C layouts."""
import ctypes as C
import os
import re
import subprocess

import pytest

from i2elvis_planet_amd import abi, load_library, I2EError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "i2e.h")


def declared_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(i2e_[a-z_0-9]+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_entry_point():
    L = load_library()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "i2e.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(i2e_params),"
        " sizeof(i2e_host_state), sizeof(i2e_solver_opts), sizeof(i2e_stats), sizeof(i2e_rock),"
        " offsetof(i2e_params, rock), offsetof(i2e_host_state, bondm), sizeof(i2e_resid),"
        " sizeof(i2e_adjoint_out), offsetof(i2e_adjoint_out, steps), sizeof(i2e_init_info),"
        " sizeof(i2e_body_stats), offsetof(i2e_body_stats, count_toohot), sizeof(i2e_coupled_adjoint_out),"
        " offsetof(i2e_coupled_adjoint_out, steps)); return 0;}\n")
    exe = tmp_path / "sz"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)])
    got = [int(v) for v in subprocess.check_output([str(exe)]).split()]
    want = [C.sizeof(abi.Params), C.sizeof(abi.HostStateC), C.sizeof(abi.SolverOpts), C.sizeof(abi.Stats),
            C.sizeof(abi.Rock), abi.Params.rock.offset, abi.HostStateC.bondm.offset, C.sizeof(abi.Resid),
            C.sizeof(abi.AdjointOut), abi.AdjointOut.steps.offset, C.sizeof(abi.InitInfo), C.sizeof(abi.BodyStats),
            abi.BodyStats.count_toohot.offset, C.sizeof(abi.CoupledAdjointOut), abi.CoupledAdjointOut.steps.offset]
    assert got == want


def test_default_solver_options():
    from i2elvis_planet_amd import default_solver
    o = default_solver()
    assert o.precision == abi.PREC_MIXED
    assert o.stokes_rtol == 1e-8
    assert o.restart > 0 and o.maxit >= o.restart


def test_create_without_device_fails_loudly():
    from i2elvis_planet_amd import configs, Context
    p, hs, _ = configs.make("c1")
    try:
        ctx = Context(p, hs)
    except I2EError as e:
        assert "device" in str(e).lower()
        return
    ctx.close()
    pytest.skip("a GPU is visible")


def test_dropin_source_defines_every_function_its_header_declares(tmp_path):
    """include/i2e_dropin.h declares the reference-named step functions
    (headmart.c:321-366); csrc/dropin.c, compiled on its own, defines each of
    them (the globals it binds stay undefined until the caller links them)."""
    hdr = open(os.path.join(ROOT, "include", "i2e_dropin.h")).read()
    names = sorted(set(re.findall(r"^\s*void\s+\*?\s*([a-z_0-9]+)\s*\(", hdr, re.M)))
    assert {"maxvelstep", "viterate", "titerate", "movemark", "ronurecalc"} <= set(names)
    obj = tmp_path / "dropin.o"
    subprocess.check_call(["gcc", "-c", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-o",
                           str(obj), os.path.join(ROOT, "i2elvis_planet_amd", "csrc", "dropin.c")])
    syms = subprocess.check_output(["nm", str(obj)], text=True).split("\n")
    defined = {l.split()[-1] for l in syms if len(l.split()) == 3 and l.split()[1] == "T"}
    undefined = {l.split()[-1] for l in syms if len(l.split()) == 2 and l.split()[0] == "U"}
    for n in names:
        assert n in defined, n
    assert {"vx", "markx", "bondm", "timestep", "marknum"} <= undefined
