"""ctypes binding of libfdlbm.so (declarations: include/fdlbm.h).

The library is loaded from this package's ``_lib`` folder only.  There is no
fallback: if it is missing or cannot be loaded the import of the solver
raises, so nothing ever runs on a silent CPU path.
"""
import ctypes
import os

_LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")
LIB_PATH = os.path.join(_LIBDIR, "libfdlbm.so")

FD_OK = 0


class FdParams(ctypes.Structure):
    """Mirror of ``fd_params`` in include/fdlbm.h."""
    _fields_ = [
        ("H", ctypes.c_int32), ("W", ctypes.c_int32),
        ("tau", ctypes.c_double), ("eta_n", ctypes.c_double), ("M", ctypes.c_double),
        ("a", ctypes.c_double), ("kappa", ctypes.c_double), ("gamma", ctypes.c_double),
        ("psi_wall", ctypes.c_double), ("u0", ctypes.c_double), ("rho_out", ctypes.c_double),
        ("outlet_f3_coef", ctypes.c_double), ("gx", ctypes.c_double), ("gy", ctypes.c_double),
        ("y_bc", ctypes.c_int32), ("x_bc", ctypes.c_int32), ("profile", ctypes.c_int32),
        ("outlet", ctypes.c_int32), ("init_mode", ctypes.c_int32),
        ("wall_lo", ctypes.c_int32), ("wall_hi", ctypes.c_int32),
        ("psi_visc_bound", ctypes.c_double),
    ]


EXPORTS = {
    # name: (restype, argtypes)
    "fd_create": (ctypes.c_int, [ctypes.POINTER(FdParams), ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.POINTER(ctypes.c_void_p)]),
    "fd_init": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fd_init_fields": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 4),
    "fd_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "fd_get_fusion": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_int32)]),
    "fd_use_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "fd_steps_done": (ctypes.c_int64, [ctypes.c_void_p]),
    "fd_solid_inert": (ctypes.c_int, [ctypes.c_void_p]),
    "fd_get_fields": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 4),
    "fd_get_fields_window": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32] + [ctypes.c_void_p] * 4),
    "fd_get_populations": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fd_set_populations": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fd_phase_mass": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "fd_phase_mass_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "fd_watchdog": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "fd_state_bytes": (ctypes.c_int64, [ctypes.c_void_p]),
    "fd_save_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "fd_load_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "fd_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "fd_fields_device": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 4),
    "fd_init_fields_device": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 4),
    "fd_state_device_bytes": (ctypes.c_int64, [ctypes.c_void_p]),
    "fd_save_state_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "fd_load_state_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64]),
    "fd_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "fd_last_step_ms": (ctypes.c_double, [ctypes.c_void_p]),
    "fd_set_tile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "fd_set_fusion": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "fd_copy_probe": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]),
    "fd_destroy": (None, [ctypes.c_void_p]),
    "fd_ens_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "fd_ens_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "fd_ens_step_sequential": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "fd_ens_set_tile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "fd_ens_set_fusion": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "fd_ens_use_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "fd_ens_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "fd_ens_last_step_ms": (ctypes.c_double, [ctypes.c_void_p]),
    "fd_ens_destroy": (None, [ctypes.c_void_p]),
    "fd_last_error": (ctypes.c_char_p, []),
}

_lib = None


def load():
    """Load libfdlbm.so (once) and declare every export; raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libfdlbm.so is not built (%s); run `python -m fingering_dynamics_amd.build`"
                           % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class FdError(RuntimeError):
    """A libfdlbm call returned a negative status."""

    def __init__(self, code, msg):
        super().__init__("libfdlbm error %d: %s" % (code, msg))
        self.code = code


def check(rc):
    if rc != FD_OK:
        raise FdError(rc, _lib.fd_last_error().decode(errors="replace"))
    return rc


def ptr(arr):
    """Data pointer of a C-contiguous numpy array (or None)."""
    if arr is None:
        return None
    assert arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(ctypes.c_void_p)
