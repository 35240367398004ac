"""Production moving-window PIC step (csrst_mw_* of include/csrst_hip.h).

One W x H grid that follows the bunch and carries the conformal chamber walls
of the bend; FP64, or FP32 fields with int32-cell + FP32-offset positions.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dptr

C_LIGHT = 299792458.0


class MovingWindowStep:
    def __init__(self, W, H, dx, dy, lenBefore, lenAfter, width, phiDeg, ekinMeV, bz, order=2, fp32=True,
                 sortEvery=20, statModulus=100, cfl=1.0, pml=0, wakeGrid=False):
        self.W, self.H = W, H
        self._p = _lib.MWParams(W, H, dx, dy, lenBefore, lenAfter, width, phiDeg, ekinMeV, bz, order,
                                1 if fp32 else 0, sortEvery, statModulus, cfl, pml, int(wakeGrid))
        h = C.c_void_p()
        check(_lib.lib().csrst_mw_create(C.byref(self._p), C.byref(h)))
        self.h = h
        self.n = 0
        self.fp32 = fp32

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().csrst_mw_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init_bunch(self, R, P, Q):
        R = np.ascontiguousarray(R, np.float64).reshape(-1, 2)
        P = np.ascontiguousarray(P, np.float64).reshape(-1, 2)
        Q = np.ascontiguousarray(Q, np.float64)
        self.n = len(R)
        check(_lib.lib().csrst_mw_init_bunch(self.h, dptr(R), dptr(P), dptr(Q), C.c_int(self.n)), self.h)

    def poisson(self, tol=1e-12, maxit=500):
        """Poisson field start on the window (after init_bunch, before the
        first step); returns the V-cycles and relative residual per solve."""
        cyc = (C.c_int * 3)()
        res = (C.c_double * 3)()
        check(_lib.lib().csrst_mw_poisson(self.h, C.c_double(tol), C.c_int(maxit), cyc, res), self.h)
        return dict(cycles=list(cyc), resid=list(res))

    def place(self, wx0, wy0):
        self.n = 0
        check(_lib.lib().csrst_mw_place(self.h, C.c_int(wx0), C.c_int(wy0)), self.h)

    def step(self, n=1):
        check(_lib.lib().csrst_mw_step(self.h, C.c_int(n)), self.h)

    def push(self, comp):
        """One push with the deposit of J[comp] on the current fields, then the
        drift back by half a step (csrst_mw_push)."""
        check(_lib.lib().csrst_mw_push(self.h, C.c_int(comp)), self.h)

    def timing(self):
        out = np.zeros(5)
        check(_lib.lib().csrst_mw_timing(self.h, dptr(out), C.c_int(5)), self.h)
        return dict(push_ms=out[0], sweep_ms=out[1], move_ms=out[2], sort_ms=out[3])

    def info(self):
        a = np.zeros(10)
        check(_lib.lib().csrst_mw_info(self.h, dptr(a)), self.h)
        return dict(dt=a[0], origin=(a[1], a[2]), pos=(int(a[3]), int(a[4])), W=int(a[5]), H=int(a[6]),
                    steps=int(a[7]), nsamp=int(a[8]), n=int(a[9]))

    def fields(self):
        """[6, H+2, W+2]: Ex, Ey, Hzx, Hzy (= Hz/2 each), Jx, Jy in logical order, guards included."""
        L = _lib.lib()
        cnt = C.c_int(0)
        check(L.csrst_mw_download_fields(self.h, None, 0, C.byref(cnt)), self.h)
        out = np.zeros(cnt.value)
        check(L.csrst_mw_download_fields(self.h, dptr(out), cnt.value, C.byref(cnt)), self.h)
        return out.reshape(6, self.H + 2, self.W + 2)

    def set_fields(self, arr):
        a = np.ascontiguousarray(arr, np.float64).ravel()
        check(_lib.lib().csrst_mw_upload_fields(self.h, dptr(a), C.c_int(a.size)), self.h)

    def particles(self):
        R = np.zeros((self.n, 2))
        P = np.zeros((self.n, 2))
        check(_lib.lib().csrst_mw_download_particles(self.h, dptr(R), dptr(P), C.c_int(self.n)), self.h)
        return R, P

    def calcBeamParameters(self):
        """PartBunch::calcBeamParameters (PartBunch.cpp:584-657) of the bunch, as a
        dict keyed by csrst_amd.stats.FIELDS."""
        from .stats import N_STATS, as_dict
        out = np.zeros(N_STATS)
        check(_lib.lib().csrst_mw_beam_stats(self.h, dptr(out)), self.h)
        return as_dict(out)

    def wake(self):
        L = _lib.lib()
        cnt = C.c_int(0)
        check(L.csrst_mw_wake(self.h, None, 0, C.byref(cnt)), self.h)
        out = np.zeros(max(cnt.value, 1))
        check(L.csrst_mw_wake(self.h, dptr(out), cnt.value, C.byref(cnt)), self.h)
        return out[:cnt.value]
