"""Cut-domain Poisson solve of the initialisation (csrst_poisson_solve).

The operator is PoissonSolver::cutoffStencil (PoissonSolver.cpp:678-766); the
solver is FP64 multigrid on the GPU (y-line relaxation, x-semicoarsening).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

FIXED, UNKNOWN, KNOWN = 0, 1, 2


def classes(flag, active):
    """Node classes from cutoffStencil's two masks ([ny, nx]): stencil rows are
    unknowns; flagged identity rows are known values; the rest are fixed."""
    flag = np.asarray(flag, bool)
    active = np.asarray(active, bool)
    cls = np.zeros(flag.shape, np.uint8)
    cls[flag & ~active] = KNOWN
    cls[active] = UNKNOWN
    return cls


def solve(hx, hy, cls, b, phi0=None, tol=1e-12, maxit=200):
    """Solve A phi = b on an [ny, nx] node grid.  hx is the gamma-scaled x
    spacing.  Returns (phi, cycles, relative residual)."""
    cls = np.ascontiguousarray(cls, np.uint8)
    b = np.ascontiguousarray(b, np.float64)
    if cls.shape != b.shape or b.ndim != 2:
        raise ValueError("cls and b must be [ny, nx] arrays of one shape")
    ny, nx = b.shape
    phi = np.zeros_like(b) if phi0 is None else np.array(phi0, np.float64, copy=True, order="C")
    cyc = C.c_int(0)
    res = C.c_double(0.0)
    check(_lib.lib().csrst_poisson_solve(C.c_int(nx), C.c_int(ny), C.c_double(hx), C.c_double(hy),
                                         cls.ctypes.data_as(C.POINTER(C.c_uint8)), _lib.dptr(b), _lib.dptr(phi),
                                         C.c_double(tol), C.c_int(maxit), C.byref(cyc), C.byref(res)))
    return phi, cyc.value, res.value
