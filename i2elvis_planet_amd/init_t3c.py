"""The initial state of a reference run from an init.t3c file (in2mart's
semantics, in2mart.c:23-1453), through libi2e's C-ABI (host code only: no
GPU is needed).

    info = read_info(path)
    params, hs = generate(path, params=None)

generate() fills the grid, rock table, radiogenic ratios, ambient
temperature, GYKOEF and por_init from the file and keeps every other
parameter of `params` (time-step and solver controls, which the reference
reads from mode.t3c).  Node fields other than T and the velocity BC values
are zero: a first ronurecalc computes them, as in2mart does before saving.
"""
import copy
import ctypes as C

from . import abi
from .state import HostState


def _lib():
    from . import load_library
    return load_library()


def read_info(path):
    from . import _chk
    info = abi.InitInfo()
    _chk(_lib().i2e_init_read(str(path).encode(), C.byref(info)), "i2e_init_read")
    return info


def generate(path, params=None):
    from . import _chk
    info = read_info(path)
    p = copy.deepcopy(params) if params is not None else abi.Params()
    hs = HostState(info.xnumx, info.ynumy, info.marknum, bondnum=info.bondnum)
    cs = hs.cstruct()
    _chk(_lib().i2e_init_generate(str(path).encode(), C.byref(p), C.byref(cs)), "i2e_init_generate")
    hs.pull_scalars(cs)
    return p, hs


def read_mode(path, file_index=0, params=None):
    """Run controls of a mode.t3c file for output file `file_index` (time-step
    limits, rheology bounds, heat options); returns (params, cyc0max)."""
    from . import _chk
    p = copy.deepcopy(params) if params is not None else abi.Params()
    n = C.c_int(0)
    _chk(_lib().i2e_mode_read(str(path).encode(), int(file_index), C.byref(p), C.byref(n)), "i2e_mode_read")
    return p, n.value
