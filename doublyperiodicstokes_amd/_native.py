"""ctypes binding of ``lib/libdpstokes_hip.so`` (C ABI in include/dpstokes_hip.h).

The library is the only compute path: if it is missing or fails to load, every
entry point raises ``RuntimeError``; there is no CPU fallback.
"""
import ctypes
import os

from .build import lib_path

DPS_OK = 0
DPS_ERR_ARG = -1
DPS_ERR_STATE = -2
DPS_ERR_RANGE = -3
DPS_ERR_HIP = -4
DPS_ERR_FFT = -5
DPS_ERR_ALLOC = -6

DPS_MODE_NOWALL = 1
DPS_MODE_BOTTOM = 2
DPS_MODE_SLIT = 3
DPS_MODE_PERIODIC = 4
DPS_PREC_DOUBLE = 0
DPS_PREC_MIXED = 1
DPS_SOLVER_AUTO = 0
DPS_SOLVER_LANE = 1
DPS_F32 = 0
DPS_F64 = 1
DPS_NOALLOC_SCRATCH_BYTES = 268435456

EXPORTS = ("dps_create", "dps_destroy", "dps_set_positions", "dps_update_positions",
           "dps_check_positions", "dps_mdot",
           "dps_mdot_adjoint", "dps_mdot_adjoint_noalloc", "dps_batch_create", "dps_batch_destroy", "dps_batch_mdot", "dps_batch_tune",
           "dps_batch_set_chunks",
           "dps_stage_spread",
           "dps_stage_solve", "dps_stage_interp", "dps_get_info", "dps_set_profiling",
           "dps_get_timings", "dps_get_precision", "dps_last_error")


class DpsParams(ctypes.Structure):
    _fields_ = [("viscosity", ctypes.c_double),
                ("Lx", ctypes.c_double), ("Ly", ctypes.c_double),
                ("zmin", ctypes.c_double), ("zmax", ctypes.c_double),
                ("w", ctypes.c_int), ("w_d", ctypes.c_int),
                ("alpha", ctypes.c_double), ("alpha_d", ctypes.c_double),
                ("beta", ctypes.c_double), ("beta_d", ctypes.c_double),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("mode", ctypes.c_int), ("precision", ctypes.c_int),
                ("solver", ctypes.c_int),
                ("reg_kernel", ctypes.c_int)]


class DpsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


_lib = None


def load():
    """Load the shared library once; raise RuntimeError when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError("libdpstokes_hip.so is not built (expected at %s); run "
                           "`python -m doublyperiodicstokes_amd.build`" % path)
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    L.dps_create.argtypes = [ctypes.POINTER(DpsParams), ctypes.c_int, ctypes.POINTER(vp)]
    L.dps_destroy.argtypes = [vp]
    L.dps_set_positions.argtypes = [vp, vp, ctypes.c_int, vp]
    L.dps_update_positions.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_int, vp]
    L.dps_check_positions.argtypes = [vp, vp]
    L.dps_mdot.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int, vp]
    L.dps_mdot_adjoint.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int, vp]
    L.dps_mdot_adjoint_noalloc.argtypes = L.dps_mdot_adjoint.argtypes
    pp = ctypes.POINTER(vp)
    L.dps_batch_create.argtypes = [pp, ctypes.c_int, ctypes.c_int, pp]
    L.dps_batch_destroy.argtypes = [vp]
    L.dps_batch_mdot.argtypes = [vp, pp, pp, pp, pp, ctypes.c_int, vp]
    L.dps_batch_tune.argtypes = [vp, pp, pp, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_int)]
    L.dps_batch_set_chunks.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int, vp]
    L.dps_stage_spread.argtypes = [vp, vp, vp, ctypes.c_int, vp, vp]
    L.dps_stage_solve.argtypes = [vp, vp, vp, vp]
    L.dps_stage_interp.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp]
    L.dps_get_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.dps_set_profiling.argtypes = [vp, ctypes.c_int]
    L.dps_get_timings.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.dps_get_precision.argtypes = [vp]
    L.dps_get_precision.restype = ctypes.c_char_p
    L.dps_last_error.argtypes = []
    L.dps_last_error.restype = ctypes.c_char_p
    for name in EXPORTS:
        if name not in ("dps_get_precision", "dps_last_error"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(rc):
    """Raise the Python exception matching a C status code."""
    if rc == DPS_OK:
        return
    msg = load().dps_last_error().decode(errors="replace")
    if rc == DPS_ERR_ARG:
        raise ValueError(msg)
    if rc == DPS_ERR_RANGE:
        raise ValueError(msg)
    raise DpsError(rc, msg)
