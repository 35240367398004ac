"""ctypes binding of libcsrst.so (the C ABI of include/csrst_hip.h).

The library is loaded from this package directory.  Nothing here falls back to
a CPU implementation: if the library or a GPU is missing, calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CSRST_LIB selects another build of the library (for comparing variants)
LIBPATH = os.environ.get("CSRST_LIB") or os.path.join(HERE, "libcsrst.so")
_LIB = None

ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "EWINDOW", -5: "EGEOM", -6: "ESTATE"}

# every symbol declared in include/csrst_hip.h
SYMBOLS = [
    "csrst_version", "csrst_last_error", "csrst_device_count",
    "csrst_create", "csrst_destroy", "csrst_geometry", "csrst_boundary_cells", "csrst_set_mask",
    "csrst_set_boundary_cells", "csrst_set_stream", "csrst_init_bunch",
    "csrst_restore_bunch",
    "csrst_step", "csrst_first_step", "csrst_second_step", "csrst_move_window", "csrst_copy_fields", "csrst_substep",
    "csrst_download_field", "csrst_upload_field", "csrst_download_particles", "csrst_download_charges",
    "csrst_info",
    "csrst_wake", "csrst_wake_reset", "csrst_is_inside_bend", "csrst_diag", "csrst_beam_stats", "csrst_field_energy",
    "csrst_line_fields", "csrst_gauss_residual",
    "csrst_mw_create", "csrst_mw_destroy", "csrst_mw_init_bunch", "csrst_mw_step", "csrst_mw_push", "csrst_mw_set_stream",
    "csrst_mw_place",
    "csrst_mw_upload_fields",
    "csrst_mw_download_fields", "csrst_mw_download_particles", "csrst_mw_beam_stats", "csrst_mw_info", "csrst_mw_wake",
    "csrst_mw_timing",
    "csrst_poisson_solve", "csrst_poisson_compute_field", "csrst_mw_poisson",
    "csrst_sg_gather", "csrst_sg_scatter", "csrst_sf_current_density", "csrst_sf_current_density_noalloc",
    "csrst_sf_charge_density",
    "csrst_sf_charge_density_diff",
    "csrst_warmstart_create", "csrst_warmstart_destroy", "csrst_warmstart_num_params",
    "csrst_warmstart_get_weights", "csrst_warmstart_set_weights", "csrst_warmstart_infer",
    "csrst_warmstart_gradients", "csrst_warmstart_train_step", "csrst_warmstart_solve",
    "csrst_warmstart_train_iterates",
]


class CsrstError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class Params(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("lenBefore", C.c_double), ("lenAfter", C.c_double),
                ("width", C.c_double), ("phiDeg", C.c_double), ("ekinMeV", C.c_double), ("bz", C.c_double),
                ("order", C.c_int), ("s2bug", C.c_int), ("statModulus", C.c_int)]


class MWParams(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("lenBefore", C.c_double), ("lenAfter", C.c_double), ("width", C.c_double),
                ("phiDeg", C.c_double), ("ekinMeV", C.c_double), ("bz", C.c_double),
                ("order", C.c_int), ("precision", C.c_int), ("sortEvery", C.c_int),
                ("statModulus", C.c_int), ("cfl", C.c_double), ("pml", C.c_int), ("wakeGrid", C.c_int)]


def lib():
    """Load libcsrst.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise ImportError("libcsrst.so is not built; run __graft_entry__.build() or csrst_amd/build.py")
        L = C.CDLL(LIBPATH)
        L.csrst_version.restype = C.c_char_p
        L.csrst_last_error.restype = C.c_char_p
        L.csrst_last_error.argtypes = [C.c_void_p]
        _LIB = L
    return _LIB


def has_gpu():
    n = C.c_int(0)
    lib().csrst_device_count(C.byref(n))
    return n.value > 0


def check(rc, handle=None):
    if rc != 0:
        msg = lib().csrst_last_error(handle)
        raise CsrstError(rc, (msg or b"").decode())


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def u8ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))
