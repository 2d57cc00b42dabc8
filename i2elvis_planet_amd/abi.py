"""ctypes mirrors of the structs declared in include/i2e.h.

The layouts must match the C header field for field; tests/test_abi.py checks
the sizes against the compiled library.
"""
import ctypes as C

MAXTYPES = 64

I2E_OK = 0
I2E_ERR_ARG = 1
I2E_ERR_CAPACITY = 2
I2E_ERR_HIP = 3
I2E_ERR_NOCONV = 4
I2E_ERR_UNSUPPORTED = 5
I2E_ERR_NODEVICE = 6

PREC_FP64 = 0
PREC_MIXED = 1
SCHUR_BFBT = 0
SCHUR_MASS = 1
SCHUR_WBFBT = 2
ORTH_LAGGED = 0
ORTH_DGKS = 1
FINE_AUTO, FINE_ROWS, FINE_STENCIL, FINE_STENCIL64, FINE_DIFF32 = 0, 1, 2, 3, 4
GUESS_LAST, GUESS_EXTRAP = 0, 1

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i8p = C.POINTER(C.c_int8)
_i64p = C.POINTER(C.c_int64)


class Rock(C.Structure):
    """One row of the reference's ROCKS table (init.t3c, headmart.c:127-131)."""
    _fields_ = [(n, C.c_double) for n in
                ("nu0", "dh", "dv", "ss", "mm", "n0", "n1", "ro", "aa", "bb", "cp", "kt", "kf", "kp")]


class Params(C.Structure):
    _fields_ = [
        ("xnumx", C.c_int64), ("ynumy", C.c_int64),
        ("mnumx", C.c_int64), ("mnumy", C.c_int64),
        ("xsize", C.c_double), ("ysize", C.c_double),
        ("maxxystep", C.c_double), ("maxtkstep", C.c_double), ("maxtmstep", C.c_double),
        ("nubeg", C.c_double), ("nuend", C.c_double), ("nucontr", C.c_double), ("strmin", C.c_double),
        ("heatdif", C.c_double), ("frictyn", C.c_double), ("adiabyn", C.c_double),
        ("heatmod", C.c_int32), ("densimod", C.c_int32),
        ("outgrid", C.c_int32), ("markmod", C.c_int32),
        ("gykoef", C.c_double), ("tmp_ambient", C.c_double),
        ("al2627_init", C.c_double), ("fe6056_init", C.c_double),
        ("h0_scale", C.c_double), ("longlived", C.c_double),
        ("por_init", C.c_double), ("k_cutoff", C.c_double),
        ("radmax_frac", C.c_double), ("eta_outer", C.c_double),
        ("melt_ts", C.c_double), ("melt_tl", C.c_double), ("melt_switch", C.c_double),
        ("melt_eta", C.c_double), ("melt_hlat", C.c_double),
        ("melt_on", C.c_int32), ("porosity_on", C.c_int32),
        ("ntypes", C.c_int32), ("pad0", C.c_int32),
        ("rock", Rock * MAXTYPES),
        ("is_silicate", C.c_uint8 * MAXTYPES),
        ("is_iron", C.c_uint8 * MAXTYPES),
        ("porous", C.c_uint8 * MAXTYPES),
        ("meltable", C.c_uint8 * MAXTYPES),
        ("core_form_time", C.c_double), ("fe_ts", C.c_double), ("fe_tl", C.c_double),
        ("soft_turb", C.c_int32), ("magnet", C.c_int32), ("dynamo", C.c_int32), ("pad1", C.c_int32),
        ("immobile", C.c_uint8 * MAXTYPES),
    ]


class HostStateC(C.Structure):
    _fields_ = [
        ("gx", _dp), ("gy", _dp),
        ("vx", _dp), ("vy", _dp), ("pr", _dp), ("gp", _dp), ("ga", _dp),
        ("exx", _dp), ("eyy", _dp), ("exy", _dp), ("sxx", _dp), ("syy", _dp), ("sxy", _dp),
        ("nu", _dp), ("nd", _dp), ("ro", _dp), ("tk", _dp), ("cp", _dp), ("kt", _dp), ("ht", _dp), ("et", _dp),
        ("marknum", C.c_int64),
        ("markx", _fp), ("marky", _fp), ("markk", _fp), ("markv", _fp), ("markp", _fp),
        ("markpor", _fp), ("marktmax", _fp),
        ("markt", _i8p),
        ("bondm", _i64p), ("bondn", _i64p), ("bondv", _dp), ("bondnum", C.c_int64),
        ("timestep", C.c_double), ("timesum", C.c_double), ("gxkoef", C.c_double),
        ("markmg_old", _i8p), ("markmg_time", _dp),
    ]


class SolverOpts(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("restart", C.c_int32), ("maxit", C.c_int32), ("nsmooth", C.c_int32),
        ("stokes_rtol", C.c_double), ("heat_rtol", C.c_double), ("grav_rtol", C.c_double),
        ("schur_scale", C.c_double), ("omega", C.c_double),
        ("coarse_max", C.c_int32), ("schur", C.c_int32),
        ("graphs", C.c_int32), ("nsmooth_coarse", C.c_int32), ("orth", C.c_int32), ("fine_stencil", C.c_int32),
        ("stokes_guess", C.c_int32), ("nsmooth_cell", C.c_int32), ("autotune", C.c_int32), ("coarse_coop", C.c_int32), ("stokes_direct", C.c_int32), ("pad_opts", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("stokes_iters", C.c_int32), ("heat_iters", C.c_int32), ("grav_iters", C.c_int32), ("pad", C.c_int32),
        ("stokes_relres", C.c_double), ("heat_relres", C.c_double), ("grav_relres", C.c_double),
        ("numin", C.c_double), ("numax", C.c_double),
        ("ms_grav", C.c_double), ("ms_stokes", C.c_double), ("ms_heat", C.c_double), ("ms_markers", C.c_double),
    ]


NODE_FIELDS = ("vx", "vy", "pr", "gp", "ga", "exx", "eyy", "exy", "sxx", "syy", "sxy",
               "nu", "nd", "ro", "tk", "cp", "kt", "ht", "et")
MARKER_FIELDS_F32 = ("markx", "marky", "markk", "markv", "markp", "markpor", "marktmax")


class Resid(C.Structure):
    """i2e_resid: viterate's residual diagnostics (movemart.c:1049-1176)."""
    _fields_ = [(k, C.c_double) for k in ("stokes", "cont", "bond", "stoknum", "contnum", "bondnum", "vxmin", "vxmax",
                                          "vymin", "vymax", "prmin", "prmax", "numin", "numax")]


class AdjointOut(C.Structure):
    """i2e_adjoint_out: peak-temperature objective and its gradient."""
    _fields_ = [("J", C.c_double), ("dJ_dh0", C.c_double), ("dJ_dphi0", C.c_double),
                ("heat_relres_max", C.c_double), ("steps", C.c_int32), ("heat_iters", C.c_int32)]


class CoupledAdjointOut(C.Structure):
    """i2e_coupled_adjoint_out: peak-temperature objective and its gradient
    through the whole step (flow included)."""
    _fields_ = [("J", C.c_double), ("dJ_dh0", C.c_double), ("dJ_dphi0", C.c_double), ("dJ_dEscale", C.c_double),
                ("heat_relres_max", C.c_double), ("stokes_relres_max", C.c_double), ("grav_relres_max", C.c_double),
                ("heat_iters", C.c_int64), ("stokes_iters", C.c_int64), ("grav_iters", C.c_int64),
                ("steps", C.c_int32), ("pad", C.c_int32)]


class FlowAdjointOut(C.Structure):
    """i2e_flow_adjoint_out: iron sinking rate and its activation-energy derivative."""
    _fields_ = [("J", C.c_double), ("dJ_dEscale", C.c_double), ("stokes_relres", C.c_double),
                ("stokes_iters", C.c_int64)]


class InitInfo(C.Structure):
    """i2e_init_info: sizes of the state an init.t3c file generates."""
    _fields_ = [("xnumx", C.c_int64), ("ynumy", C.c_int64), ("mnumx", C.c_int64), ("mnumy", C.c_int64),
                ("marknum", C.c_int64), ("bondnum", C.c_int64), ("rocknum", C.c_int32), ("nonstab", C.c_int32),
                ("timesum", C.c_double)]


class PrnHeader(C.Structure):
    """i2e_prn_header: sizes and scalars of a binary .prn file."""
    _fields_ = [("xnumx", C.c_int64), ("ynumy", C.c_int64), ("mnumx", C.c_int64), ("mnumy", C.c_int64),
                ("marknum", C.c_int64), ("bondnum", C.c_int64), ("rocknum", C.c_int32), ("n0", C.c_int32),
                ("xsize", C.c_double), ("ysize", C.c_double), ("por_init", C.c_double), ("gxkoef", C.c_double),
                ("gykoef", C.c_double), ("tmp_ambient", C.c_double), ("al2627_init", C.c_double),
                ("fe6056_init", C.c_double), ("timesum", C.c_double)]


class VisualSpec(C.Structure):
    """i2e_visual_spec: one output of i2jmart's saver (type = amir.t3c's letter
    code, the window as fractions of the box, the pixel resolution)."""
    _fields_ = [("type", C.c_int32), ("gridmod", C.c_int32), ("xresol", C.c_int32), ("yresol", C.c_int32),
                ("xbeg", C.c_double), ("xend", C.c_double), ("ybeg", C.c_double), ("yend", C.c_double)]


class GrainParams(C.Structure):
    """i2e_grain_params: grain() controls (growth_model, gr_init in
    micrometres, start_time and impact_time[1] in Ma after CAIs)."""
    _fields_ = [("growth_model", C.c_int32), ("pad", C.c_int32), ("gr_init", C.c_double),
                ("start_time_ma", C.c_double), ("impact_time_ma", C.c_double)]


class BodyStats(C.Structure):
    """i2e_body_stats (include/i2e.h): impact()'s hydration and body
    statistics (impact.c:3022-3228)."""
    _fields_ = [(n, C.c_double) for n in (
        "sol_frac", "liq_frac", "hydrous_frac", "primitive_frac", "n2co2_frac", "cocl_frac", "h2o_frac",
        "phyllo1_frac", "phyllo2_frac", "phyllo3_frac", "phyllo4_frac", "perco_frac", "melt1_frac", "melt2_frac",
        "maxtk", "t_max_body", "meantk", "t_mean_body")] + [
        ("count_toohot", C.c_int64), ("count_tot", C.c_int64), ("count_body_nodes", C.c_int64)]

