"""MI355X-native I2ELVIS marker-in-cell time step.

Python mirror of the reference's step interface (headmart.c:321-366, driven by
i2mart.c:60-153): a `Context` owns one device copy of the model state and
exposes maxvelstep / viterate / titerate / movemark / ronurecalc / step with
the reference's semantics, calling libi2e.so through its C-ABI
(include/i2e.h).  There is no CPU fallback: without the built library or a
visible GPU every call raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import Params, SolverOpts, Stats
from .state import HostState

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libi2e.so")

_EXPORTS = [
    "i2e_create", "i2e_destroy", "i2e_set_solver", "i2e_default_solver", "i2e_upload", "i2e_download",
    "i2e_maxvelstep", "i2e_viterate", "i2e_titerate", "i2e_movemark", "i2e_ronurecalc", "i2e_step",
    "i2e_get_scalars", "i2e_set_scalars", "i2e_get_stats", "i2e_sync", "i2e_last_error",
    "i2e_stokes_apply_dev", "i2e_gravity_apply_dev", "i2e_heat_apply_dev", "i2e_device_ptr", "i2e_time_kernels",
    "i2e_stokes_rhs_dev", "i2e_stokes_precond_dev", "i2e_vcycle_dev",
    "i2e_dev_alloc", "i2e_dev_free", "i2e_copy_to_dev", "i2e_copy_to_host",
    "i2e_kernel_timer", "i2e_kernel_timer_read", "i2e_stokes_residuals", "i2e_band_info", "i2e_step_batch",
    "i2e_dev_copy", "i2e_prn_write", "i2e_prn_read_header", "i2e_prn_read", "i2e_set_params", "i2e_record", "i2e_replay", "i2e_peak_t", "i2e_adjoint_peak_t", "i2e_adjoint_peak_t_coupled", "i2e_adjoint_marker_positions", "i2e_visual", "i2e_visual_write", "i2e_markersave", "i2e_grain", "i2e_heat_apply_t_dev", "i2e_sink_rate", "i2e_adjoint_sink_rate", "i2e_stokes_apply_t_dev",
    "i2e_type_counts", "i2e_maxvelstep_noalloc", "i2e_kernel_timer_repr", "i2e_set_marker_sort",
    "i2e_sort_markers", "i2e_init_read", "i2e_init_generate", "i2e_mode_read", "i2e_tuned_shapes", "i2e_hydration",
]

_lib = None


class I2EError(RuntimeError):
    pass


def visual_write(path, spec, timesum, out):
    """saver()'s text file for a grid from Context.visual (loadjmart.c:681, 1130-1165)."""
    load_library()
    a = np.ascontiguousarray(out)
    _chk(_lib.i2e_visual_write(os.fsencode(path), C.byref(spec), float(timesum),
                               None if spec.type == 5 else a.ctypes.data, a.ctypes.data if spec.type == 5 else None),
         "visual_write")


def markersave(directory, params, hs, melt=None):
    """markersave()'s four marker text files (loadjmart.c:3-53)."""
    load_library()
    cs = hs.cstruct()
    m = None if melt is None else np.ascontiguousarray(melt, np.float64)
    _chk(_lib.i2e_markersave(os.fsencode(directory), C.byref(params), C.byref(cs),
                             None if m is None else m.ctypes.data), "markersave")


def load_library(path=LIB_PATH):
    """dlopen libi2e.so and declare its signatures (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise I2EError("libi2e.so not built: run `python -m i2elvis_planet_amd.build`")
    L = C.CDLL(path)
    vp = C.c_void_p
    P = C.POINTER(Params)
    S = C.POINTER(abi.HostStateC)
    dp = C.POINTER(C.c_double)
    sig = {
        "i2e_create": ([P, C.c_int64, C.POINTER(vp)], C.c_int),
        "i2e_destroy": ([vp], C.c_int),
        "i2e_set_solver": ([vp, C.POINTER(SolverOpts)], C.c_int),
        "i2e_default_solver": ([C.POINTER(SolverOpts)], C.c_int),
        "i2e_upload": ([vp, S], C.c_int),
        "i2e_download": ([vp, S], C.c_int),
        "i2e_maxvelstep": ([vp, dp], C.c_int),
        "i2e_viterate": ([vp, C.c_int], C.c_int),
        "i2e_titerate": ([vp, C.c_int], C.c_int),
        "i2e_movemark": ([vp], C.c_int),
        "i2e_ronurecalc": ([vp], C.c_int),
        "i2e_step": ([vp], C.c_int),
        "i2e_get_scalars": ([vp, dp, dp, dp], C.c_int),
        "i2e_set_scalars": ([vp, C.c_double, C.c_double], C.c_int),
        "i2e_get_stats": ([vp, C.POINTER(Stats)], C.c_int),
        "i2e_sync": ([vp], C.c_int),
        "i2e_last_error": ([], C.c_char_p),
        "i2e_stokes_apply_dev": ([vp, vp, vp], C.c_int),
        "i2e_gravity_apply_dev": ([vp, vp, vp], C.c_int),
        "i2e_heat_apply_dev": ([vp, vp, vp], C.c_int),
        "i2e_device_ptr": ([vp, C.c_char_p], vp),
        "i2e_time_kernels": ([vp, C.c_int, dp, dp, dp], C.c_int),
        "i2e_stokes_rhs_dev": ([vp, vp], C.c_int),
        "i2e_stokes_precond_dev": ([vp, vp, vp], C.c_int),
        "i2e_vcycle_dev": ([vp, vp, vp], C.c_int),
        "i2e_kernel_timer": ([vp, C.c_char_p, C.c_int, C.c_int], C.c_int),
        "i2e_stokes_residuals": ([vp, C.POINTER(abi.Resid)], C.c_int),
        "i2e_step_batch": ([C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)], C.c_int),
        "i2e_kernel_timer_read": ([vp, C.POINTER(C.c_int64), dp, dp], C.c_int),
        "i2e_kernel_timer_repr": ([vp, dp], C.c_int),
        "i2e_band_info": ([vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), dp], C.c_int),
        "i2e_set_marker_sort": ([vp, C.c_int], C.c_int),
        "i2e_sort_markers": ([vp], C.c_int),
        "i2e_tuned_shapes": ([vp, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                              C.POINTER(C.c_int)], C.c_int),
        "i2e_hydration": ([vp, C.c_double, C.c_double, C.c_int, C.POINTER(abi.BodyStats)], C.c_int),
        "i2e_init_read": ([C.c_char_p, C.POINTER(abi.InitInfo)], C.c_int),
        "i2e_init_generate": ([C.c_char_p, P, S], C.c_int),
        "i2e_mode_read": ([C.c_char_p, C.c_int, P, C.POINTER(C.c_int)], C.c_int),
        "i2e_dev_alloc": ([C.c_size_t, C.POINTER(vp)], C.c_int),
        "i2e_dev_free": ([vp], C.c_int),
        "i2e_copy_to_dev": ([vp, vp, C.c_size_t], C.c_int),
        "i2e_copy_to_host": ([vp, vp, C.c_size_t], C.c_int),
        "i2e_dev_copy": ([vp, vp, C.c_size_t], C.c_int),
        "i2e_prn_write": ([C.c_char_p, P, S, C.c_int], C.c_int),
        "i2e_prn_read_header": ([C.c_char_p, C.POINTER(abi.PrnHeader)], C.c_int),
        "i2e_prn_read": ([C.c_char_p, P, S, C.POINTER(C.c_int)], C.c_int),
        "i2e_set_params": ([vp, P], C.c_int),
        "i2e_record": ([vp, C.c_int], C.c_int),
        "i2e_replay": ([vp, C.c_int], C.c_int),
        "i2e_peak_t": ([vp, dp], C.c_int),
        "i2e_adjoint_peak_t": ([vp, C.POINTER(abi.AdjointOut)], C.c_int),
        "i2e_adjoint_peak_t_coupled": ([vp, C.POINTER(abi.CoupledAdjointOut)], C.c_int),
        "i2e_adjoint_marker_positions": ([vp, vp, vp], C.c_int),
        "i2e_visual": ([vp, C.POINTER(abi.VisualSpec), vp, vp, vp], C.c_int),
        "i2e_visual_write": ([C.c_char_p, C.POINTER(abi.VisualSpec), C.c_double, vp, vp], C.c_int),
        "i2e_markersave": ([C.c_char_p, C.POINTER(abi.Params), C.POINTER(abi.HostStateC), vp], C.c_int),
        "i2e_grain": ([vp, C.POINTER(abi.GrainParams), vp], C.c_int),
        "i2e_sink_rate": ([vp, C.POINTER(C.c_double)], C.c_int),
        "i2e_adjoint_sink_rate": ([vp, C.POINTER(abi.FlowAdjointOut)], C.c_int),
        "i2e_stokes_apply_t_dev": ([vp, vp, vp], C.c_int),
        "i2e_heat_apply_t_dev": ([vp, vp, vp], C.c_int),
        "i2e_type_counts": ([vp, C.c_int, C.POINTER(C.c_int32)], C.c_int),
        "i2e_maxvelstep_noalloc": ([vp, vp, C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, vp, vp],
                                   C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _chk(st, what):
    if st != 0:
        msg = _lib.i2e_last_error().decode(errors="replace")
        raise I2EError("%s failed (status %d): %s" % (what, st, msg))


def default_solver():
    o = SolverOpts()
    _chk(load_library().i2e_default_solver(C.byref(o)), "i2e_default_solver")
    return o


class DeviceArray:
    """float64 device buffer allocated through libi2e's own HIP runtime."""

    def __init__(self, n_or_array):
        L = load_library()
        a = None if np.isscalar(n_or_array) else np.ascontiguousarray(n_or_array, np.float64)
        self.n = int(n_or_array) if a is None else a.size
        self._p = C.c_void_p()
        _chk(L.i2e_dev_alloc(8 * self.n, C.byref(self._p)), "i2e_dev_alloc")
        self.upload(np.zeros(self.n) if a is None else a)

    @property
    def ptr(self):
        return self._p.value

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        if a.size != self.n:
            raise ValueError("size %d != %d" % (a.size, self.n))
        _chk(_lib.i2e_copy_to_dev(self._p, a.ctypes.data_as(C.c_void_p), 8 * self.n), "i2e_copy_to_dev")

    def numpy(self):
        out = np.empty(self.n)
        _chk(_lib.i2e_copy_to_host(out.ctypes.data_as(C.c_void_p), self._p, 8 * self.n), "i2e_copy_to_host")
        return out

    def copy_from(self, other):
        """Device-to-device copy by libi2e's plain streaming kernel."""
        if other.n != self.n:
            raise ValueError("size %d != %d" % (other.n, self.n))
        _chk(_lib.i2e_dev_copy(self._p, other._p, 8 * self.n), "i2e_dev_copy")

    def free(self):
        if self._p:
            _lib.i2e_dev_free(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """Device-resident model state plus the reference's step functions."""

    def __init__(self, params, hs, max_markers=None, solver=None):
        L = load_library()
        self.p = params
        self._h = C.c_void_p()
        M = hs.marknum if max_markers is None else int(max_markers)
        _chk(L.i2e_create(C.byref(params), M, C.byref(self._h)), "i2e_create")
        if solver is not None:
            self.set_solver(solver)
        self.upload(hs)

    def close(self):
        if self._h:
            _lib.i2e_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_solver(self, o):
        _chk(_lib.i2e_set_solver(self._h, C.byref(o)), "i2e_set_solver")

    def upload(self, hs):
        self._cs = hs.cstruct()
        _chk(_lib.i2e_upload(self._h, C.byref(self._cs)), "i2e_upload")
        self.M = int(hs.marknum)

    def download(self, hs):
        cs = hs.cstruct()
        _chk(_lib.i2e_download(self._h, C.byref(cs)), "i2e_download")
        hs.pull_scalars(cs)
        return hs

    # the reference's step functions ------------------------------------------------
    def maxvelstep(self):
        ts = C.c_double(0)
        _chk(_lib.i2e_maxvelstep(self._h, C.byref(ts)), "maxvelstep")
        return ts.value

    def viterate(self, n0=0):
        _chk(_lib.i2e_viterate(self._h, n0), "viterate")

    def titerate(self, n0=0):
        _chk(_lib.i2e_titerate(self._h, n0), "titerate")

    def movemark(self):
        _chk(_lib.i2e_movemark(self._h), "movemark")

    def ronurecalc(self):
        _chk(_lib.i2e_ronurecalc(self._h), "ronurecalc")

    def step(self):
        _chk(_lib.i2e_step(self._h), "step")

    def sync(self):
        _chk(_lib.i2e_sync(self._h), "sync")

    # scalars / diagnostics -----------------------------------------------------------
    def scalars(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _chk(_lib.i2e_get_scalars(self._h, C.byref(a), C.byref(b), C.byref(c)), "get_scalars")
        return a.value, b.value, c.value

    def set_scalars(self, timestep, timesum):
        _chk(_lib.i2e_set_scalars(self._h, timestep, timesum), "set_scalars")

    def stats(self):
        s = Stats()
        _chk(_lib.i2e_get_stats(self._h, C.byref(s)), "get_stats")
        return s

    def device_ptr(self, name):
        p = _lib.i2e_device_ptr(self._h, name.encode())
        if not p:
            raise I2EError("unknown device array %r" % name)
        return p

    def apply(self, which, x_dev, y_dev):
        f = {"stokes": _lib.i2e_stokes_apply_dev, "gravity": _lib.i2e_gravity_apply_dev,
             "heat": _lib.i2e_heat_apply_dev, "heat_t": _lib.i2e_heat_apply_t_dev,
             "stokes_t": _lib.i2e_stokes_apply_t_dev}[which]
        _chk(f(self._h, C.c_void_p(x_dev), C.c_void_p(y_dev)), which + "_apply")

    def stokes_rhs(self, b_dev):
        _chk(_lib.i2e_stokes_rhs_dev(self._h, C.c_void_p(b_dev)), "stokes_rhs")

    def stokes_precond(self, r_dev, z_dev):
        _chk(_lib.i2e_stokes_precond_dev(self._h, C.c_void_p(r_dev), C.c_void_p(z_dev)), "stokes_precond")

    def vcycle(self, f_dev, u_dev):
        _chk(_lib.i2e_vcycle_dev(self._h, C.c_void_p(f_dev), C.c_void_p(u_dev)), "vcycle")

    def band_info(self):
        """(kl, ku, bytes) of the last banded Stokes factorisation (stokes_direct = 1)."""
        kl, ku, b = C.c_int32(), C.c_int32(), C.c_double()
        _chk(_lib.i2e_band_info(self._h, C.byref(kl), C.byref(ku), C.byref(b)), "band_info")
        return kl.value, ku.value, b.value

    def stokes_residuals(self):
        """viterate's residual diagnostics (movemart.c:1049-1176) as a dict."""
        r = abi.Resid()
        _chk(_lib.i2e_stokes_residuals(self._h, C.byref(r)), "stokes_residuals")
        return {k: getattr(r, k) for k, _ in abi.Resid._fields_}

    def kernel_timer(self, kernel, max_launches=65536, every=1):
        """Time every `every`-th launch of `kernel` from now on with HIP events (None stops)."""
        _chk(_lib.i2e_kernel_timer(self._h, (kernel or "").encode(), max_launches, every), "kernel_timer")

    def kernel_timer_read(self):
        """(launches, total ms, total algorithmic bytes, total representation
        bytes) since kernel_timer()."""
        n, ms, b, r = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        _chk(_lib.i2e_kernel_timer_read(self._h, C.byref(n), C.byref(ms), C.byref(b)), "kernel_timer_read")
        _chk(_lib.i2e_kernel_timer_repr(self._h, C.byref(r)), "kernel_timer_repr")
        return n.value, ms.value, b.value, r.value

    def set_params(self, params):
        _chk(_lib.i2e_set_params(self._h, C.byref(params)), "set_params")
        self.p = params

    # adjoint of the peak-temperature objective -------------------------------------
    def record(self, nsteps):
        """ronurecalc + nsteps steps from the uploaded state, history kept on the GPU."""
        _chk(_lib.i2e_record(self._h, nsteps), "record")

    def replay(self, nsteps):
        """ronurecalc + the thermal part of nsteps recorded steps (recorded positions and dt)."""
        _chk(_lib.i2e_replay(self._h, nsteps), "replay")

    def peak_t(self):
        """Mass-weighted mean marktmax of the silicate markers inside the grid."""
        J = C.c_double()
        _chk(_lib.i2e_peak_t(self._h, C.byref(J)), "peak_t")
        return J.value

    def adjoint_peak_t(self):
        """J and dJ/d(h0_scale, phi0) through the recorded thermal chain (call after record)."""
        o = abi.AdjointOut()
        _chk(_lib.i2e_adjoint_peak_t(self._h, C.byref(o)), "adjoint_peak_t")
        return {k: getattr(o, k) for k, _ in abi.AdjointOut._fields_}

    def adjoint_peak_t_coupled(self):
        """J and dJ/d(h0_scale, phi0, activation-energy factor) through the
        whole recorded step, flow and advection included (call after record;
        the context's node fields are left at the first recorded step's)."""
        o = abi.CoupledAdjointOut()
        _chk(_lib.i2e_adjoint_peak_t_coupled(self._h, C.byref(o)), "adjoint_peak_t_coupled")
        return {k: getattr(o, k) for k, _ in abi.CoupledAdjointOut._fields_ if k != "pad"}

    def adjoint_marker_positions(self, n):
        """dJ/dx, dJ/dy of the n initial marker positions (caller's order)
        from the last adjoint_peak_t_coupled."""
        gx = np.zeros(n)
        gy = np.zeros(n)
        _chk(_lib.i2e_adjoint_marker_positions(self._h, gx.ctypes.data, gy.ctypes.data), "adjoint_marker_positions")
        return gx, gy

    # post-processing: i2jmart's visualisation grid ---------------------------------
    def visual(self, spec):
        """One output of i2jmart's saver on the device (loadjmart.c:619-1197):
        returns (values float32, or rock types int32 for type 5; melt fraction
        per marker)."""
        n = spec.xresol * spec.yresol
        out = np.zeros(n, np.int32 if spec.type == 5 else np.float32)
        melt = np.zeros(max(self.M, 1))
        vp_val = None if spec.type == 5 else out.ctypes.data
        vp_rock = out.ctypes.data if spec.type == 5 else None
        _chk(_lib.i2e_visual(self._h, C.byref(spec), vp_val, vp_rock, melt.ctypes.data), "visual")
        return out, melt[:self.M]

    def grain(self, gp, markgr):
        """grain() on the device (grain.c:4-156): markgr (float32, one per
        marker, caller's order) updated in place; marker types may change."""
        assert markgr.dtype == np.float32 and markgr.flags.c_contiguous and markgr.size == self.M
        _chk(_lib.i2e_grain(self._h, C.byref(gp), markgr.ctypes.data), "grain")

    # flow adjoint: one Stokes solve and the advection that reads it ---------------
    def sink_rate(self):
        """Mass-weighted mean outward RK4 velocity of the iron markers (m/s)."""
        J = C.c_double()
        _chk(_lib.i2e_sink_rate(self._h, C.byref(J)), "sink_rate")
        return J.value

    def adjoint_sink_rate(self):
        """sink_rate and its derivative in a common factor on every activation
        energy, through the last Stokes solve (call right after viterate)."""
        o = abi.FlowAdjointOut()
        _chk(_lib.i2e_adjoint_sink_rate(self._h, C.byref(o)), "adjoint_sink_rate")
        return {k: getattr(o, k) for k, _ in abi.FlowAdjointOut._fields_}

    def set_marker_sort(self, every):
        """Sort the markers by cell every `every` steps of step() (0: never)."""
        _chk(_lib.i2e_set_marker_sort(self._h, int(every)), "set_marker_sort")

    def hydration(self, t_sol=1393.811, t_liq=1973.0, hydrate=True):
        """impact()'s per-step hydration and body statistics on the device
        state (impact.c:3022-3228): switches rock+ice / hydrous silicate
        marker types by their peak temperature and returns the counts and
        fractions (abi.BodyStats).  t_sol, t_liq: impact()'s peridotite
        solidus and liquidus; hydrate: growth_model == 0 && por_init <= 0.001."""
        o = abi.BodyStats()
        _chk(_lib.i2e_hydration(self._h, float(t_sol), float(t_liq), int(bool(hydrate)), C.byref(o)), "hydration")
        return o

    def tuned_shapes(self):
        """[(level nodes along x, tile width, tile height)] of the coarse
        velocity levels smoothed by tiles (the autotuner's choice)."""
        n = C.c_int()
        _chk(_lib.i2e_tuned_shapes(self._h, 0, None, None, None, C.byref(n)), "tuned_shapes")
        k = n.value
        lv, sx, sy = (np.zeros(k, np.int32) for _ in range(3))
        i32 = C.POINTER(C.c_int32)
        _chk(_lib.i2e_tuned_shapes(self._h, k, lv.ctypes.data_as(i32), sx.ctypes.data_as(i32),
                                   sy.ctypes.data_as(i32), C.byref(n)), "tuned_shapes")
        return [(int(a), int(b), int(c)) for a, b, c in zip(lv, sx, sy)]

    def sort_markers(self):
        """Group the device markers by cell now (download keeps the caller's order)."""
        _chk(_lib.i2e_sort_markers(self._h), "sort_markers")

    def type_counts(self, ntypes=64):
        """Per-cell marker-type histogram of the device markers, shape (cells, ntypes)."""
        X, Y = self.p.xnumx, self.p.ynumy
        out = np.zeros((X - 1) * (Y - 1) * ntypes, np.int32)
        _chk(_lib.i2e_type_counts(self._h, ntypes, out.ctypes.data_as(C.POINTER(C.c_int32))), "type_counts")
        return out.reshape((X - 1) * (Y - 1), ntypes)

    def time_kernels(self, reps=5):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _chk(_lib.i2e_time_kernels(self._h, reps, C.byref(a), C.byref(b), C.byref(c)), "time_kernels")
        return a.value, b.value, c.value


def step_batch(contexts, nsteps=1, nthreads=0):
    """nsteps of step() on every context, bodies overlapping on the GPU
    (one host thread per body unless nthreads says otherwise)."""
    L = load_library()
    n = len(contexts)
    arr = (C.c_void_p * n)(*[c._h.value for c in contexts])
    st = (C.c_int * n)()
    _chk(L.i2e_step_batch(arr, n, nsteps, nthreads, st), "step_batch")


__all__ = ["step_batch", "Context", "DeviceArray", "HostState", "Params", "SolverOpts", "Stats", "default_solver", "load_library", "I2EError"]
