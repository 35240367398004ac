"""Python front end of the gfx950 PIC step, mirroring the reference's
interfaces for this path (TETMSolver, PartBunch, Bend, the Commands input).

The objects hold a handle of libcsrst.so; all numerical work happens in the
HIP kernels behind the C ABI (include/csrst_hip.h).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr, u8ptr

C_LIGHT = 299792458.0
M_E_GEV = 0.51099906e-03      # Physics.cpp:51
M_E_BEND_EV = 0.511e6         # Bend.cpp:66


@dataclass
class Commands:
    """The subset of main.cpp:111-136 that shapes the time step."""
    numgridpx: int = 100          # -x
    numgridpy: int = 100          # -y
    numsteps: int = 10            # -N
    statmodulus: int = 100000     # --statmodulus
    phi: float = 0.0              # -P [deg]
    width: float = 0.10           # -w [m]
    length: float = 0.10          # -l [m] (after the bend)
    lengthbefore: float = 0.10    # --lengthbefore [m]
    Ekin: float = 250.0           # -E [MeV]
    Bz: float = 0.10              # -B [T]
    order: int = 0                # IntOp_t: 0 IntCIC, 1 IntS1, 2 IntS2
    s2bug: bool = True            # ScatterGatherImpl.hh:467,525 as written

    def params(self):
        return _lib.Params(self.numgridpx, self.numgridpy, self.lengthbefore, self.length, self.width, self.phi,
                           self.Ekin, self.Bz, self.order, int(self.s2bug), self.statmodulus)


def bend_radius(ekin_mev, bz):
    """Bend radius of Bend.cpp:76-78 (electron mass 0.511 MeV)."""
    ek = ekin_mev * 1e6
    return np.sqrt(ek * ek + 2.0 * ek * M_E_BEND_EV) / (C_LIGHT * abs(bz))


class TwoGridStep:
    """The reference's two-grid step: global chamber grid + bunch window.

    Methods map onto TETMSolver::updateField_firstStep/secondStep
    (TETMSolver.hh:110-266) and the loop body of main.cpp:390-506.
    """

    def __init__(self, cmd: Commands):
        self.cmd = cmd
        self._p = cmd.params()
        h = C.c_void_p()
        L = _lib.lib()
        check(L.csrst_create(C.byref(self._p), C.byref(h)))
        self.h = h
        g = np.zeros(7)
        check(L.csrst_geometry(self.h, dptr(g)), self.h)
        self.hr = g[0:2].copy()
        self.origin = g[2:4].copy()
        self.dt = g[4]
        self.R, self.phi = g[5], g[6]
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().csrst_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # -- set-up ------------------------------------------------------------
    def boundary_cells(self):
        L = _lib.lib()
        cnt = C.c_int(0)
        check(L.csrst_boundary_cells(self.h, None, None, None, 0, C.byref(cnt)), self.h)
        n = cnt.value
        mask = np.zeros((self.cmd.numgridpy, self.cmd.numgridpx), np.uint8)
        ij = np.zeros((max(n, 1), 2), np.int32)
        data = np.zeros((max(n, 1), 5))
        check(L.csrst_boundary_cells(self.h, u8ptr(mask), iptr(ij), dptr(data), n, C.byref(cnt)), self.h)
        return mask, ij[:n], data[:n]

    def set_mask(self, mask):
        """The inside mask [Ny, Nx] of the masked sweeps (before init_bunch)."""
        m = np.ascontiguousarray(mask, np.uint8)
        if m.shape != (self.cmd.numgridpy, self.cmd.numgridpx):
            raise ValueError("mask must be [Ny, Nx]")
        check(_lib.lib().csrst_set_mask(self.h, u8ptr(m)), self.h)

    def set_boundary_cells(self, ij, data):
        """SCBoundaryCondition's cell list (SCBoundaryCondition.hh:39-84): ij
        [n, 2] cell indices, data [n, 5] = lambda0, lambda1, area, lambdaN0,
        lambdaN1 (before init_bunch)."""
        ij = np.ascontiguousarray(ij, np.int32).reshape(-1, 2)
        data = np.ascontiguousarray(data, np.float64).reshape(-1, 5)
        if len(ij) != len(data):
            raise ValueError("ij and data differ in length")
        check(_lib.lib().csrst_set_boundary_cells(self.h, iptr(ij), dptr(data), C.c_int(len(ij))), self.h)

    def set_stream(self, stream):
        """Queue the step's GPU work on a caller hipStream_t handle (None: the
        context's own stream)."""
        check(_lib.lib().csrst_set_stream(self.h, C.c_void_p(stream) if stream else None), self.h)

    def init_bunch(self, R, P, Q):
        R = np.ascontiguousarray(R, np.float64).reshape(-1, 2)
        P = np.ascontiguousarray(P, np.float64).reshape(-1, 2)
        Q = np.ascontiguousarray(Q, np.float64)
        self.n = len(R)
        check(_lib.lib().csrst_init_bunch(self.h, dptr(R), dptr(P), dptr(Q), C.c_int(self.n)), self.h)

    def restore(self, E, H, R, oldR, P, Q, step=0):
        """Restart (LoadH5Part::restoreFields + restoreBunch, main.cpp:147-153,
        226-228): global E and H ([2, Ny+2, Nx+2] planes), then the bunch with
        its oldR (no makeReady); the step counter starts at step."""
        self.set_field("E", E)
        self.set_field("H", H)
        R, O, P = (np.ascontiguousarray(a, np.float64).reshape(-1, 2) for a in (R, oldR, P))
        Q = np.ascontiguousarray(Q, np.float64)
        self.n = len(R)
        check(_lib.lib().csrst_restore_bunch(self.h, dptr(R), dptr(O), dptr(P), dptr(Q), C.c_int(self.n),
                                             C.c_int(step)), self.h)

    def computeField(self, tol=1e-12, maxit=500):
        """PoissonSolver::computeField (PoissonSolver.cpp:85-201) on the GPU:
        the three multigrid solves into the global E and H, resetOutsideDomain
        and the window copy (main.cpp:229-341).  Call after init_bunch, before
        the first step.  Returns the V-cycles and relative residual per solve."""
        cyc = (C.c_int * 3)()
        res = (C.c_double * 3)()
        check(_lib.lib().csrst_poisson_compute_field(self.h, C.c_double(tol), C.c_int(maxit), cyc, res), self.h)
        return dict(cycles=list(cyc), resid=list(res))

    # -- stepping ------------------------------------------------------------
    def step(self, n=1):
        check(_lib.lib().csrst_step(self.h, C.c_int(n)), self.h)

    def updateField_firstStep(self):
        check(_lib.lib().csrst_first_step(self.h), self.h)

    def updateField_secondStep(self):
        check(_lib.lib().csrst_second_step(self.h), self.h)

    def substep(self, op):
        """One part of the step (codes of csrst_substep), for bisection against the oracle."""
        check(_lib.lib().csrst_substep(self.h, C.c_int(op)), self.h)

    def moveMesh(self):
        mi, mj = C.c_int(), C.c_int()
        check(_lib.lib().csrst_move_window(self.h, C.byref(mi), C.byref(mj)), self.h)
        return mi.value, mj.value

    def copyFields(self, copyCurrent=True):
        """copyFields (main.cpp:534-548): the window's E, H (and J) overwrite
        the global fields under it; done before every field plot."""
        check(_lib.lib().csrst_copy_fields(self.h, C.c_int(int(bool(copyCurrent)))), self.h)

    # -- state ------------------------------------------------------------------
    def info(self):
        a = np.zeros(9)
        check(_lib.lib().csrst_info(self.h, dptr(a)), self.h)
        return dict(dt=a[0], origin=(a[1], a[2]), pos=(int(a[3]), int(a[4])), W=int(a[5]), H=int(a[6]),
                    steps=int(a[7]), nsamp=int(a[8]))

    _CODES = {"E": 0, "H": 1, "mE": 2, "mH": 3, "mJ": 4, "J": 5}

    def field(self, which):
        """[2, m+2, n+2] over [-1, n] x [-1, m]; 'E','H','J' global, 'mE','mH','mJ' window."""
        L = _lib.lib()
        cnt = C.c_int(0)
        code = self._CODES[which]
        check(L.csrst_download_field(self.h, code, None, 0, C.byref(cnt)), self.h)
        out = np.zeros(cnt.value)
        check(L.csrst_download_field(self.h, code, dptr(out), cnt.value, C.byref(cnt)), self.h)
        if which in ("E", "H", "J"):
            n, m = self.cmd.numgridpx, self.cmd.numgridpy
        else:
            inf = self.info()
            n, m = inf["W"], inf["H"]
        return out.reshape(2, m + 2, n + 2)

    def set_field(self, which, arr):
        a = np.ascontiguousarray(arr, np.float64).ravel()
        check(_lib.lib().csrst_upload_field(self.h, self._CODES[which], dptr(a), C.c_int(a.size)), self.h)

    def particles(self):
        R = np.zeros((self.n, 2))
        P = np.zeros((self.n, 2))
        O = np.zeros((self.n, 2))
        check(_lib.lib().csrst_download_particles(self.h, dptr(R), dptr(P), dptr(O), C.c_int(self.n)), self.h)
        return R, P, O

    def wake_reset(self):
        """IntegrateElongCommand::reset: the wake accumulator to zero."""
        check(_lib.lib().csrst_wake_reset(self.h), self.h)

    def isInsideBend(self, R):
        """Bend::isInsideBend (Bend.hh:307-323)."""
        r = np.ascontiguousarray(R, np.float64)
        out = C.c_int(0)
        check(_lib.lib().csrst_is_inside_bend(self.h, dptr(r), C.byref(out)), self.h)
        return bool(out.value)

    def charges(self):
        Q = np.zeros(self.n)
        check(_lib.lib().csrst_download_charges(self.h, dptr(Q), C.c_int(self.n)), self.h)
        return Q

    def wake(self):
        L = _lib.lib()
        cnt = C.c_int(0)
        check(L.csrst_wake(self.h, None, 0, C.byref(cnt)), self.h)
        out = np.zeros(max(cnt.value, 1))
        check(L.csrst_wake(self.h, dptr(out), cnt.value, C.byref(cnt)), self.h)
        return out[:cnt.value]

    def diag(self):
        out = np.zeros(6)
        check(_lib.lib().csrst_diag(self.h, dptr(out)), self.h)
        return out

    def lineFields(self):
        """SaveEMFieldsCommand::execute (SaveEMFields.cpp:34-160) on the global
        fields: rows [n, 7] = s, E along P, E across P, Hz, J along P, x, y."""
        L = _lib.lib()
        cnt = C.c_int(0)
        check(L.csrst_line_fields(self.h, None, 0, C.byref(cnt)), self.h)
        out = np.zeros(7 * max(cnt.value, 1))
        check(L.csrst_line_fields(self.h, dptr(out), cnt.value, C.byref(cnt)), self.h)
        return out[:7 * cnt.value].reshape(cnt.value, 7)

    def gaussResidual(self):
        """div E - rho/eps0 on the window's vertices, [H, W] (csrst_gauss_residual)."""
        L = _lib.lib()
        cnt = C.c_int(0)
        check(L.csrst_gauss_residual(self.h, None, 0, C.byref(cnt)), self.h)
        out = np.zeros(max(cnt.value, 1))
        check(L.csrst_gauss_residual(self.h, dptr(out), cnt.value, C.byref(cnt)), self.h)
        inf = self.info()
        return out[:cnt.value].reshape(inf["H"], inf["W"])

    def fieldEnergySums(self):
        """(sum E.E, sum (Hzx+Hzy)^2) x cell area over global + window - overlap,
        the sums of CalcEnergyCommand::execute (CalcEnergy.cpp:60-124)."""
        out = np.zeros(2)
        check(_lib.lib().csrst_field_energy(self.h, dptr(out)), self.h)
        return out

    def calcBeamParameters(self):
        """PartBunch::calcBeamParameters (PartBunch.cpp:584-657) of the current bunch,
        as a dict keyed by csrst_amd.stats.FIELDS."""
        from .stats import N_STATS, as_dict
        out = np.zeros(N_STATS)
        check(_lib.lib().csrst_beam_stats(self.h, dptr(out)), self.h)
        return as_dict(out)
