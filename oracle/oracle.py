"""ctypes front end of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the
CPU-baseline leg of bench.py.  The product package never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)


def _ptr(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class SimParams(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int),
                ("lenBefore", C.c_double), ("lenAfter", C.c_double), ("width", C.c_double),
                ("phiDeg", C.c_double), ("ekinMeV", C.c_double), ("bz", C.c_double),
                ("order", C.c_int), ("s2bug", C.c_int), ("statModulus", C.c_int)]


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        # CSRST_ORACLE_LIB loads another build of the same source (the mutation
        # tests use it for deliberately broken copies)
        path = os.environ.get("CSRST_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_q_rsqrt.restype = C.c_float
        L.orc_q_rsqrt.argtypes = [C.c_float]
        L.orc_crossings.restype = C.c_int
        L.orc_bend_create.restype = C.c_void_p
        L.orc_bend_create.argtypes = [C.c_double] * 6 + [C.c_int, C.c_int]
        for f in ("orc_bend_destroy", "orc_bend_info", "orc_bend_mask", "orc_bend_num_bcells",
                  "orc_bend_bcells", "orc_bend_external_field", "orc_bend_is_inside"):
            getattr(L, f).argtypes = None
        L.orc_bend_is_inside.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_sim_create.restype = C.c_void_p
        L.orc_sim_create.argtypes = [C.POINTER(SimParams)]
        L.orc_poisson_solve.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p]
        L.orc_sim_poisson.argtypes = [C.c_void_p, C.c_double]
        _LIB = L
    return _LIB


def q_rsqrt(x):
    return lib().orc_q_rsqrt(C.c_float(x))


def crossings(oldR, newR, dx, origin, dual=False):
    frac = np.zeros(10)
    sub = np.zeros(20)
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (oldR, newR, dx, origin)]
    n = lib().orc_crossings(int(dual), *[_ptr(v) for v in a], _ptr(frac), _ptr(sub))
    return frac[:n].copy(), sub[:2 * n].reshape(n, 2).copy()


def gather_edge(order, comp, patch, origin, dx, R, s2bug=True):
    patch = np.ascontiguousarray(patch, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 2)
    out = np.zeros(len(R))
    pny, pnx = patch.shape
    lib().orc_gather_edge(C.c_int(order), C.c_int(comp), C.c_int(int(s2bug)), _ptr(patch), C.c_int(pnx),
                          C.c_int(pny), _ptr(np.asarray(origin, np.float64)), _ptr(np.asarray(dx, np.float64)),
                          _ptr(R), C.c_int(len(R)), _ptr(out))
    return out


def gather_cell(order, patch, origin, dx, R, s2bug=True):
    patch = np.ascontiguousarray(patch, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 2)
    out = np.zeros(len(R))
    pny, pnx = patch.shape
    lib().orc_gather_cell(C.c_int(order), C.c_int(int(s2bug)), _ptr(patch), C.c_int(pnx), C.c_int(pny),
                          _ptr(np.asarray(origin, np.float64)), _ptr(np.asarray(dx, np.float64)),
                          _ptr(R), C.c_int(len(R)), _ptr(out))
    return out


def scatter_edge(order, comp, shape, origin, dx, R, val):
    patch = np.zeros(shape)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 2)
    val = np.ascontiguousarray(val, dtype=np.float64)
    lib().orc_scatter_edge(C.c_int(order), C.c_int(comp), _ptr(patch), C.c_int(shape[1]), C.c_int(shape[0]),
                           _ptr(np.asarray(origin, np.float64)), _ptr(np.asarray(dx, np.float64)),
                           _ptr(R), _ptr(val), C.c_int(len(R)))
    return patch


def scatter_cell(order, shape, origin, dx, R, val):
    patch = np.zeros(shape)
    R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1, 2)
    val = np.ascontiguousarray(val, dtype=np.float64)
    lib().orc_scatter_cell(C.c_int(order), _ptr(patch), C.c_int(shape[1]), C.c_int(shape[0]),
                           _ptr(np.asarray(origin, np.float64)), _ptr(np.asarray(dx, np.float64)),
                           _ptr(R), _ptr(val), C.c_int(len(R)))
    return patch


def deposit(order, comp, box, origin, dx, dt, oldR, R, Q):
    """box = (i0, j0, nx, ny) absolute node indices; returns J[ny, nx]."""
    i0, j0, nx, ny = box
    J = np.zeros((ny, nx))
    oldR = np.ascontiguousarray(oldR, np.float64).reshape(-1, 2)
    R = np.ascontiguousarray(R, np.float64).reshape(-1, 2)
    Q = np.ascontiguousarray(Q, np.float64)
    rc = lib().orc_deposit(C.c_int(order), C.c_int(comp), _ptr(J), C.c_int(i0), C.c_int(j0), C.c_int(nx),
                           C.c_int(ny), _ptr(np.asarray(origin, np.float64)), _ptr(np.asarray(dx, np.float64)),
                           C.c_double(dt), _ptr(oldR), _ptr(R), _ptr(Q), C.c_int(len(R)))
    if rc != 0:
        raise ValueError("deposit left the patch")
    return J


def push_patch(order, bend, dt, ex, ey, hz, origin, dx, R, P, s2bug=False, drift_back=True):
    """The particle loop of PartBunch::push_particles on given fields (one
    patch [ny, nx] for Ex, Ey and Hz with node (0, 0) at origin), plus the
    -dt/2 drift after it.  Returns (R, P, oldR, Rmid), Rmid the positions
    after the push before the drift back (the end of the deposit segment)."""
    ex, ey, hz = (np.ascontiguousarray(a, np.float64) for a in (ex, ey, hz))
    R = np.array(R, np.float64).reshape(-1, 2)
    P = np.array(P, np.float64).reshape(-1, 2)
    oldR = np.zeros_like(R)
    Rmid = np.zeros_like(R)
    ny, nx = ex.shape
    rc = lib().orc_push_patch(C.c_int(order), C.c_int(int(s2bug)), C.c_void_p(bend.h), C.c_double(dt), _ptr(ex),
                              _ptr(ey), _ptr(hz), C.c_int(nx), C.c_int(ny), _ptr(np.asarray(origin, np.float64)),
                              _ptr(np.asarray(dx, np.float64)), _ptr(R), _ptr(P), _ptr(oldR), _ptr(Rmid),
                              C.c_int(len(R)), C.c_int(int(drift_back)))
    if rc != 0:
        raise ValueError("oracle push failed")
    return R, P, oldR, Rmid


BEAM_FIELDS = ("Ekin", "rrms_x", "rrms_y", "prms_x", "prms_y", "emit_x", "emit_y", "rmean_x", "rmean_y",
               "pmean_x", "pmean_y", "Dy", "DDy", "dE", "rmin_x", "rmin_y", "rmax_x", "rmax_y", "rprms_x",
               "rprms_y")


def beam_parameters(R, P):
    """PartBunch::calcMoments + calcBeamParameters (PartBunch.cpp:584-756), as a
    dict keyed like csrst_amd.stats.FIELDS."""
    R = np.ascontiguousarray(R, np.float64).reshape(-1, 2)
    P = np.ascontiguousarray(P, np.float64).reshape(-1, 2)
    out = np.zeros(len(BEAM_FIELDS))
    if lib().orc_beam_parameters(_ptr(R), _ptr(P), C.c_int(len(R)), _ptr(out)) != 0:
        raise ValueError("bad beam")
    return {k: float(out[i]) for i, k in enumerate(BEAM_FIELDS)}


def charge_density(order, box, origin, dx, R, Q):
    i0, j0, nx, ny = box
    rho = np.zeros((ny, nx))
    R = np.ascontiguousarray(R, np.float64).reshape(-1, 2)
    Q = np.ascontiguousarray(Q, np.float64)
    rc = lib().orc_charge_density(C.c_int(order), _ptr(rho), C.c_int(i0), C.c_int(j0), C.c_int(nx), C.c_int(ny),
                                  _ptr(np.asarray(origin, np.float64)), _ptr(np.asarray(dx, np.float64)),
                                  _ptr(R), _ptr(Q), C.c_int(len(R)))
    if rc != 0:
        raise ValueError("charge density left the patch")
    return rho


def poisson_solve(hx, hy, flag, active, b):
    """Direct solve of the cutoffStencil system (PoissonSolver.cpp:678-766);
    flag, active, b are [ny, nx]; identity rows return phi = b."""
    flag = np.ascontiguousarray(flag, np.uint8)
    active = np.ascontiguousarray(active, np.uint8)
    b = np.ascontiguousarray(b, np.float64)
    ny, nx = b.shape
    phi = np.zeros_like(b)
    rc = lib().orc_poisson_solve(nx, ny, hx, hy, flag.ctypes.data, active.ctypes.data, b.ctypes.data,
                                 phi.ctypes.data)
    if rc != 0:
        raise RuntimeError("oracle poisson solve failed (%d)" % rc)
    return phi


class Bend:
    """Bend geometry (Bend.cpp) on an Nx x Ny global vertex mesh."""

    def __init__(self, lenBefore, lenAfter, width, phiDeg, ekinMeV, bz, Nx, Ny):
        L = lib()
        self.h = L.orc_bend_create(lenBefore, lenAfter, width, phiDeg, ekinMeV, bz, Nx, Ny)
        if not self.h:
            raise RuntimeError("bend geometry failed")
        self.Nx, self.Ny = Nx, Ny
        info = np.zeros(10)
        L.orc_bend_info(C.c_void_p(self.h), _ptr(info))
        self.hr = info[0:2].copy()
        self.origin = info[2:4].copy()
        self.R, self.phi = info[4], info[5]
        self.center = info[6:8].copy()
        self.domain = info[8:10].copy()

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_bend_destroy(C.c_void_p(self.h))
            self.h = None

    def is_inside(self, x, y):
        return bool(lib().orc_bend_is_inside(C.c_void_p(self.h), x, y))

    def mask(self):
        m = np.zeros((self.Ny, self.Nx), np.uint8)
        lib().orc_bend_mask(C.c_void_p(self.h), _ptr(m, C.c_uint8))
        return m

    def bcells(self):
        n = lib().orc_bend_num_bcells(C.c_void_p(self.h))
        ij = np.zeros((n, 2), np.int32)
        data = np.zeros((n, 5))
        lib().orc_bend_bcells(C.c_void_p(self.h), _ptr(ij, C.c_int), _ptr(data))
        return ij, data

    def poisson_classes(self):
        """(flag, active, maxX) of PoissonSolver::cutoffStencil on this grid, [Ny, Nx] uint8."""
        f = np.zeros((self.Ny, self.Nx), np.uint8)
        a = np.zeros((self.Ny, self.Nx), np.uint8)
        mx = lib().orc_poisson_classes(C.c_void_p(self.h), _ptr(f, C.c_uint8), _ptr(a, C.c_uint8))
        return f, a, mx

    def external_field(self, R):
        R = np.ascontiguousarray(R, np.float64).reshape(-1, 2)
        out = np.zeros(len(R))
        lib().orc_bend_external_field(C.c_void_p(self.h), _ptr(R), C.c_int(len(R)), _ptr(out))
        return out


def copy_window(G, Wf, pos):
    """Communicator::copy(from=window, to=global, start=-ddim.min)
    (Communicator.hh:212-268): window cell (i, j) lands on global cell
    (pos[0] + i, pos[1] + j) wherever that lies in the global domain; guard
    layers stay.  G, Wf: [2, m+2, n+2] arrays over [-1, n] x [-1, m].
    Returns the new global array."""
    G = np.array(G, copy=True)
    ny, nx = G.shape[1] - 2, G.shape[2] - 2
    h, w = Wf.shape[1] - 2, Wf.shape[2] - 2
    for j in range(h):
        gj = pos[1] + j
        if not 0 <= gj < ny:
            continue
        for i in range(w):
            gi = pos[0] + i
            if 0 <= gi < nx:
                G[:, gj + 1, gi + 1] = Wf[:, j + 1, i + 1]
    return G


class Sim:
    """Two-grid PIC time step of the reference (main.cpp:390-506), FP64 on the CPU."""

    def __init__(self, Nx, Ny, lenBefore, lenAfter, width, phiDeg, ekinMeV, bz, order=0, s2bug=True,
                 statModulus=1):
        self.p = SimParams(Nx, Ny, lenBefore, lenAfter, width, phiDeg, ekinMeV, bz, order, int(s2bug), statModulus)
        self.h = lib().orc_sim_create(C.byref(self.p))
        if not self.h:
            raise RuntimeError("oracle sim create failed")
        self.n = 0

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(C.c_void_p(self.h))
            self.h = None

    def init(self, R, P, Q):
        R = np.ascontiguousarray(R, np.float64).reshape(-1, 2)
        P = np.ascontiguousarray(P, np.float64).reshape(-1, 2)
        Q = np.ascontiguousarray(Q, np.float64)
        self.n = len(R)
        rc = lib().orc_sim_init(C.c_void_p(self.h), _ptr(R), _ptr(P), _ptr(Q), C.c_int(self.n))
        if rc != 0:
            raise RuntimeError("oracle init failed (%d)" % rc)

    def restore(self, E, H, R, oldR, P, Q, step=0):
        """The restart of main.cpp:147-153,226-228: global E and H, then the
        bunch with its oldR (LoadH5Part::restoreBunch, no makeReady)."""
        self.set_field("E", E)
        self.set_field("H", H)
        R, O, P = (np.ascontiguousarray(a, np.float64).reshape(-1, 2) for a in (R, oldR, P))
        Q = np.ascontiguousarray(Q, np.float64)
        self.n = len(R)
        rc = lib().orc_sim_restore(C.c_void_p(self.h), _ptr(R), _ptr(O), _ptr(P), _ptr(Q), C.c_int(self.n),
                                   C.c_int(step))
        if rc != 0:
            raise RuntimeError("oracle restore failed (%d)" % rc)

    def poisson(self, gamma=0.0):
        """PoissonSolver::computeField + resetOutsideDomain + window copy (after init)."""
        rc = lib().orc_sim_poisson(C.c_void_p(self.h), C.c_double(gamma))
        if rc != 0:
            raise RuntimeError("oracle poisson failed (%d)" % rc)

    def substep(self, op):
        if lib().orc_sim_substep(C.c_void_p(self.h), C.c_int(op)) != 0:
            raise RuntimeError("oracle substep %d failed" % op)

    def step(self, n=1):
        rc = lib().orc_sim_step(C.c_void_p(self.h), C.c_int(n))
        if rc != 0:
            raise RuntimeError("oracle step failed")

    def info(self):
        a = np.zeros(9)
        lib().orc_sim_info(C.c_void_p(self.h), _ptr(a))
        return dict(dt=a[0], origin=(a[1], a[2]), pos=(int(a[3]), int(a[4])), W=int(a[5]), H=int(a[6]),
                    steps=int(a[7]), nsamp=int(a[8]))

    def field(self, which):
        """which: 'E', 'H' (global), 'mE', 'mH', 'mJ' (window), 'J' (global).
        Returns array [2, n+2, m+2] over [-1, n] x [-1, m] (component, j, i)."""
        code = {"E": 0, "H": 1, "mE": 2, "mH": 3, "mJ": 4, "J": 5}[which]
        inf = self.info()
        if which in ("E", "H", "J"):
            n, m = self.p.Nx, self.p.Ny
        else:
            n, m = inf["W"], inf["H"]
        out = np.zeros((m + 2) * (n + 2) * 2)
        lib().orc_sim_get_field(C.c_void_p(self.h), C.c_int(code), _ptr(out))
        return out.reshape(m + 2, n + 2, 2).transpose(2, 0, 1).copy()

    def set_field(self, which, arr):
        code = {"E": 0, "H": 1, "mE": 2, "mH": 3, "mJ": 4, "J": 5}[which]
        a = np.ascontiguousarray(np.asarray(arr, np.float64).transpose(1, 2, 0))
        lib().orc_sim_set_field(C.c_void_p(self.h), C.c_int(code), _ptr(a))

    def field_steps(self, n=1):
        rc = lib().orc_sim_field_steps(C.c_void_p(self.h), C.c_int(n))
        if rc != 0:
            raise RuntimeError("oracle field steps failed")

    def particles(self):
        R = np.zeros((self.n, 2))
        P = np.zeros((self.n, 2))
        O = np.zeros((self.n, 2))
        lib().orc_sim_get_particles(C.c_void_p(self.h), _ptr(R), _ptr(P), _ptr(O))
        return R, P, O

    def wake(self):
        n = self.info()["nsamp"]
        out = np.zeros(max(n, 1))
        lib().orc_sim_get_wake(C.c_void_p(self.h), _ptr(out))
        return out[:n]

    def diag(self):
        out = np.zeros(6)
        lib().orc_sim_diag(C.c_void_p(self.h), _ptr(out))
        return out

    def copy_fields(self, copyCurrent=True):
        """copyFields (main.cpp:534-548) on the oracle's state: window E, H
        (and J) into the global fields, through copy_window below."""
        pos = self.info()["pos"]
        for g, w in (("E", "mE"), ("H", "mH")) + ((("J", "mJ"),) if copyCurrent else ()):
            self.set_field(g, copy_window(self.field(g), self.field(w), pos))

    def line_fields(self, cap=100000):
        """SaveEMFieldsCommand::execute (SaveEMFields.cpp:34-160) on the global
        fields: rows [n, 7] = s, E along P, E across P, Hz, J along P, x, y."""
        out = np.zeros(7 * cap)
        n = lib().orc_sim_line_fields(C.c_void_p(self.h), _ptr(out), C.c_int(cap))
        if n < 0:
            raise ValueError("line longer than cap")
        return out[:7 * n].reshape(n, 7)

    def field_energy_sums(self):
        """CalcEnergyCommand sums (CalcEnergy.cpp:60-124): sum E.E and sum
        (Hzx+Hzy)^2 over global + window - overlap, times the cell area."""
        out = np.zeros(2)
        lib().orc_sim_field_energy(C.c_void_p(self.h), _ptr(out))
        return out
