"""The reference's binary .prn checkpoint (saver/loader, loadmart.c:391-540
and 671-809) for a HostState, through libi2e's C-ABI (host code only: no GPU
is needed).

    write(path, params, hs, n0=0)
    params, hs, n0 = read(path, params=None)

read() fills the grid, rock table, radiogenic ratios, ambient temperature
and GYKOEF from the file and keeps every other parameter of `params` (time
step and solver controls, closures), which the reference reads from mode.t3c.
"""
import copy
import ctypes as C

from . import abi
from .state import HostState


def _lib():
    from . import load_library
    return load_library()


def _chk(st, what):
    from . import _chk as chk
    chk(st, what)


def write(path, params, hs, n0=0):
    cs = hs.cstruct()
    _chk(_lib().i2e_prn_write(str(path).encode(), C.byref(params), C.byref(cs), int(n0)), "i2e_prn_write")


def header(path):
    h = abi.PrnHeader()
    _chk(_lib().i2e_prn_read_header(str(path).encode(), C.byref(h)), "i2e_prn_read_header")
    return h


def read(path, params=None):
    h = header(path)
    p = copy.deepcopy(params) if params is not None else abi.Params()
    hs = HostState(h.xnumx, h.ynumy, h.marknum, bondnum=h.bondnum)
    cs = hs.cstruct()
    n0 = C.c_int(0)
    _chk(_lib().i2e_prn_read(str(path).encode(), C.byref(p), C.byref(cs), C.byref(n0)), "i2e_prn_read")
    hs.pull_scalars(cs)
    return p, hs, n0.value
