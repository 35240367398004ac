"""ShapeFunction and ScatterGatherImpl on device arrays (csrst_sg_*, csrst_sf_*
of include/csrst_hip.h), with the reference's names.

`DeviceArray` is a minimal owner of device memory through the HIP runtime
(hipMalloc / hipMemcpy / hipFree), enough to hand the kernels their device
pointers from numpy arrays.  The functions below copy numpy inputs to the
device, run the kernel and copy the result back.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL("libamdhip64.so")
    return _HIP


class DeviceArray:
    """Device copy of a float64 numpy array (or an uninitialised one of a shape)."""

    def __init__(self, host=None, shape=None):
        self.host = None if host is None else np.ascontiguousarray(host, np.float64)
        self.shape = self.host.shape if host is not None else tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        self.ptr = C.c_void_p()
        if hip().hipMalloc(C.byref(self.ptr), C.c_size_t(max(self.nbytes, 8))) != 0:
            raise MemoryError("hipMalloc of %d bytes failed" % self.nbytes)
        if self.host is not None and self.nbytes:
            if hip().hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), 1) != 0:
                raise RuntimeError("hipMemcpy to the device failed")

    def get(self):
        out = np.empty(self.shape, np.float64)
        if self.nbytes and hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2) != 0:
            raise RuntimeError("hipMemcpy to the host failed")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        self.free()


class Stream:
    """A HIP stream of the caller's (hipStreamCreate), for csrst_set_stream."""

    def __init__(self):
        self.h = C.c_void_p()
        if hip().hipStreamCreate(C.byref(self.h)) != 0:
            raise RuntimeError("hipStreamCreate failed")

    @property
    def handle(self):
        return self.h.value

    def close(self):
        if self.h:
            hip().hipStreamSynchronize(self.h)
            hip().hipStreamDestroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()


def _v2(a):
    return (C.c_double * 2)(float(a[0]), float(a[1]))


def _pairs(R):
    return np.ascontiguousarray(np.asarray(R, np.float64).reshape(-1, 2))


KIND = {"Ex": 0, "Ey": 1, "cell": 2}


def gather(order, kind, field, origin, dx, R, s2bug=True):
    """ScatterGatherImpl<IntOp,double>::gather (ScatterGatherImpl.hh:27-64)
    of a node patch field[ny, nx] with node (0, 0) at origin."""
    field = np.ascontiguousarray(field, np.float64)
    R = _pairs(R)
    f, r, out = DeviceArray(field), DeviceArray(R), DeviceArray(shape=(len(R),))
    check(_lib.lib().csrst_sg_gather(C.c_int(order), C.c_int(int(s2bug)), C.c_int(KIND.get(kind, kind)), f.ptr,
                                     C.c_int(field.shape[1]), C.c_int(field.shape[0]), _v2(origin), _v2(dx), r.ptr,
                                     C.c_int(len(R)), out.ptr, None))
    return out.get()


def scatter(order, kind, shape, origin, dx, R, val, field=None):
    """ScatterGatherImpl<IntOp,double>::scatter: adds val times the weights
    into field (zeros of `shape` if not given)."""
    R = _pairs(R)
    base = np.zeros(shape) if field is None else np.asarray(field, np.float64)
    f, r, v = DeviceArray(base), DeviceArray(R), DeviceArray(np.asarray(val, np.float64))
    check(_lib.lib().csrst_sg_scatter(C.c_int(order), C.c_int(KIND.get(kind, kind)), f.ptr, C.c_int(base.shape[1]),
                                      C.c_int(base.shape[0]), _v2(origin), _v2(dx), r.ptr, v.ptr, C.c_int(len(R)),
                                      None))
    return f.get()


def getCurrentDensity(order, comp, box, origin, dx, dt, oldR, R, Q, noalloc=False):
    """ShapeFunction<IntOp>::getCurrentDensity{X,Y} onto the node patch
    box = (i0, j0, nx, ny) of the mesh with node 0 at origin; returns J[ny, nx].
    noalloc: through csrst_sf_current_density_noalloc (static scratch)."""
    i0, j0, nx, ny = box
    O, Rn = _pairs(oldR), _pairs(R)
    J = DeviceArray(shape=(ny, nx))
    o, r, q = DeviceArray(O), DeviceArray(Rn), DeviceArray(np.asarray(Q, np.float64))
    fn = _lib.lib().csrst_sf_current_density_noalloc if noalloc else _lib.lib().csrst_sf_current_density
    check(fn(C.c_int(order), C.c_int(comp), J.ptr, C.c_int(i0), C.c_int(j0), C.c_int(nx), C.c_int(ny), _v2(origin),
             _v2(dx), C.c_double(dt), o.ptr, r.ptr, q.ptr, C.c_int(len(Rn)), None))
    return J.get()


def getChargeDensity(order, box, origin, dx, R, Q, oldR=None):
    """ShapeFunction<IntOp>::getChargeDensity (with oldR: getChargeDensityDiff,
    rho(R) - rho(oldR)) onto the node patch box = (i0, j0, nx, ny)."""
    i0, j0, nx, ny = box
    Rn = _pairs(R)
    rho = DeviceArray(shape=(ny, nx))
    r, q = DeviceArray(Rn), DeviceArray(np.asarray(Q, np.float64))
    L = _lib.lib()
    if oldR is None:
        check(L.csrst_sf_charge_density(C.c_int(order), rho.ptr, C.c_int(i0), C.c_int(j0), C.c_int(nx), C.c_int(ny),
                                        _v2(origin), _v2(dx), r.ptr, q.ptr, C.c_int(len(Rn)), None))
    else:
        o = DeviceArray(_pairs(oldR))
        check(L.csrst_sf_charge_density_diff(C.c_int(order), rho.ptr, C.c_int(i0), C.c_int(j0), C.c_int(nx),
                                             C.c_int(ny), _v2(origin), _v2(dx), r.ptr, o.ptr, q.ptr,
                                             C.c_int(len(Rn)), None))
    return rho.get()
