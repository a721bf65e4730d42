"""ctypes mirror of ``include/ouxsbli.h`` (the C ABI of libouxsbli).

Only layout and constants live here; the loader is ``_native.py``.
"""
import ctypes

OX_ABI_VERSION = 2

OX_OK = 0
OX_EINVAL = -1
OX_ENOMEM = -2
OX_EDEVICE = -3
OX_EUNSUPPORTED = -4
OX_ESTATE = -5
OX_EBUSY = -6
OX_ENOTFOUND = -7

SCHEME = {"KEEP": 0, "SLAU": 1, "Roe": 2, "Hybrid": 3}
VISC = {"Euler": 0, "NS": 1, "LES": 2}
TVD = {"none": 0, "tvd": 1, "hybrid": 2}
SLAU_VARIANT = {"SLAU": 0, "HRSLAU2": 1}
STENCIL = {"Central": 0, "ME4Base": 1}
PRECISION = {"fp64": 0, "fp32_bulk": 1}
BC = {"periodic": 0, "extrapolate": 1, "slip": 2}  # OX_BC_* of one direction


class OxConfig(ctypes.Structure):
    """``ox_config`` -- one case's compile-time switches and constants."""

    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("order", ctypes.c_int32),
        ("scheme", ctypes.c_int32),
        ("visc", ctypes.c_int32),
        ("visc_order", ctypes.c_int32),
        ("visc_stencil", ctypes.c_int32),
        ("tvd", ctypes.c_int32),
        ("slau_variant", ctypes.c_int32),
        ("rk", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("bc_x", ctypes.c_int32),
        ("bc_y", ctypes.c_int32),
        ("bc_z", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("gamma", ctypes.c_double),
        ("R", ctypes.c_double),
        ("Pr", ctypes.c_double),
        ("Prt", ctypes.c_double),
        ("threshold", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("Lx", ctypes.c_double),
        ("Ly", ctypes.c_double),
        ("Lz", ctypes.c_double),
    ]


OX2D_BC_OBLIQUE_SHOCK = 1


class Ox2dConfig(ctypes.Structure):
    """``ox2d_config`` -- one 2-D case with the oblique-shock boundaries."""

    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("order", ctypes.c_int32),
        ("scheme", ctypes.c_int32),
        ("tvd", ctypes.c_int32),
        ("slau_variant", ctypes.c_int32),
        ("rk", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("bc", ctypes.c_int32),
        ("top_split", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("gamma", ctypes.c_double),
        ("R", ctypes.c_double),
        ("threshold", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("Lx", ctypes.c_double),
        ("Ly", ctypes.c_double),
        ("inflow", ctypes.c_double * 4),
        ("post", ctypes.c_double * 4),
    ]


OX_IC_TAYLOR_GREEN = 1
OX_IC_SHEAR_LAYER = 2
OX_IC_SPECTRAL = 3
OX_IC_ETGV = 4
OX_IC_KHI = 5


class OxIc(ctypes.Structure):
    """``ox_ic`` -- an initial condition built on the device."""

    _fields_ = [
        ("kind", ctypes.c_int32),
        ("kmax", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("L0", ctypes.c_double),
        ("V0", ctypes.c_double),
        ("rho0", ctypes.c_double),
        ("p0", ctypes.c_double),
        ("T0", ctypes.c_double),
        ("delta", ctypes.c_double),
        ("noise", ctypes.c_double),
        ("modes", ctypes.POINTER(ctypes.c_double)),
        ("u2", ctypes.c_double),
        ("rho2", ctypes.c_double),
        ("amp", ctypes.c_double),
    ]


class OxDiag(ctypes.Structure):
    """``ox_diag`` -- interior diagnostics."""

    _fields_ = [
        ("ke", ctypes.c_double),
        ("entropy", ctypes.c_double),
        ("enstrophy", ctypes.c_double),
        ("max_wave_speed", ctypes.c_double),
        ("mass", ctypes.c_double),
        ("energy", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class OxKernelTime(ctypes.Structure):
    """``ox_kernel_time`` -- accumulated device time of one kernel."""

    _fields_ = [
        ("name", ctypes.c_char * 32),
        ("total_ms", ctypes.c_double),
        ("launches", ctypes.c_int64),
    ]


class OxBatchConfig(ctypes.Structure):
    """``ox_batch_config``."""

    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("max_cases", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("autotune", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("hbm_budget", ctypes.c_uint64),
    ]


class OxBatchStats(ctypes.Structure):
    """``ox_batch_stats``."""

    _fields_ = [
        ("active", ctypes.c_int32),
        ("done", ctypes.c_int32),
        ("groups", ctypes.c_int32),
        ("graph_captures", ctypes.c_int32),
        ("steps", ctypes.c_int64),
        ("graph_replays", ctypes.c_int64),
        ("bytes_in_use", ctypes.c_uint64),
        ("cell_stages", ctypes.c_double),
        ("device_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# Every symbol include/ouxsbli.h declares (checked by tests/test_abi.py).
EXPORTS = (
    "ox_create",
    "ox_destroy",
    "ox_set_state",
    "ox_get_state",
    "ox_set_qj",
    "ox_get_qj",
    "ox_step",
    "ox_rhs",
    "ox_diagnostics",
    "ox_diagnostics_noalloc",
    "ox_real4_sum",
    "ox_check_state",
    "ox_synchronize",
    "ox_last_step_ms",
    "ox_set_profiling",
    "ox_kernel_times",
    "ox_footprint",
    "ox_autotune",
    "ox_set_segments",
    "ox_batch_create",
    "ox_batch_destroy",
    "ox_batch_admit",
    "ox_batch_tune",
    "ox_batch_step",
    "ox_batch_run",
    "ox_batch_done",
    "ox_batch_retire",
    "ox_batch_query",
    "ox2d_create",
    "ox2d_destroy",
    "ox2d_set_state",
    "ox2d_get_state",
    "ox2d_step",
    "ox2d_last_step_ms",
    "ox2d_footprint",
    "ox_batch_admit2d",
    "ox_init_state",
    "ox_batch_admit_ic",
    "ox2d_init_state",
    "ox_host_alloc",
    "ox_host_free",
    "ox_last_error",
    "ox_version",
)
