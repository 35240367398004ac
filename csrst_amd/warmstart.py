"""Learned warm start for the Poisson multigrid (csrst_warmstart_*).

A 4-level U-Net (32 base channels, BF16 on the MFMA units) maps a charge
density to a potential; the multigrid iterates from that guess to its
tolerance.  `layers()` gives the parameter layout: per convolution the
weights [Cout][3][3][Cin] then the biases [Cout].
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check


def layers(base=32):
    b = base
    io = [(1, b, 1), (b, b, 1), (b, 2 * b, 1), (2 * b, 2 * b, 1), (2 * b, 4 * b, 1), (4 * b, 4 * b, 1),
          (4 * b, 8 * b, 1), (8 * b, 8 * b, 1), (8 * b, 4 * b, 1), (4 * b, 4 * b, 1), (4 * b, 2 * b, 1),
          (2 * b, 2 * b, 1), (2 * b, b, 1), (b, b, 1), (b, 1, 0)]
    out, off = [], 0
    for cin, cout, relu in io:
        out.append(dict(cin=cin, cout=cout, relu=relu, woff=off, boff=off + cout * 9 * cin))
        off += cout * 9 * cin + cout
    return out, off


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class WarmStart:
    def __init__(self, H, W, max_batch=16):
        h = C.c_void_p()
        check(_lib.lib().csrst_warmstart_create(C.c_int(H), C.c_int(W), C.c_int(max_batch), C.byref(h)))
        self.h, self.H, self.W = h, H, W
        n = C.c_int(0)
        check(_lib.lib().csrst_warmstart_num_params(self.h, C.byref(n)), self.h)
        self.nparam = n.value

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().csrst_warmstart_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def weights(self):
        w = np.zeros(self.nparam, np.float32)
        check(_lib.lib().csrst_warmstart_get_weights(self.h, _fp(w), C.c_int(self.nparam)), self.h)
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        check(_lib.lib().csrst_warmstart_set_weights(self.h, _fp(w), C.c_int(w.size)), self.h)

    def infer(self, rho, scale=1.0):
        rho = np.ascontiguousarray(rho, np.float32).reshape(-1, self.H, self.W)
        out = np.zeros_like(rho)
        check(_lib.lib().csrst_warmstart_infer(self.h, _fp(rho), _fp(out), C.c_int(len(rho)), C.c_float(scale)),
              self.h)
        return out

    def gradients(self, rho, phi, scale=1.0):
        rho = np.ascontiguousarray(rho, np.float32).reshape(-1, self.H, self.W)
        phi = np.ascontiguousarray(phi, np.float32).reshape(rho.shape)
        g = np.zeros(self.nparam, np.float32)
        loss = C.c_float(0)
        check(_lib.lib().csrst_warmstart_gradients(self.h, _fp(rho), _fp(phi), C.c_int(len(rho)), C.c_float(scale),
                                                   C.byref(loss), _fp(g), C.c_int(self.nparam)), self.h)
        return loss.value, g

    def train_step(self, rho, phi, lr=1e-3, scale=1.0):
        rho = np.ascontiguousarray(rho, np.float32).reshape(-1, self.H, self.W)
        phi = np.ascontiguousarray(phi, np.float32).reshape(rho.shape)
        loss = C.c_float(0)
        check(_lib.lib().csrst_warmstart_train_step(self.h, _fp(rho), _fp(phi), C.c_int(len(rho)), C.c_float(scale),
                                                    C.c_float(lr), C.byref(loss)), self.h)
        return loss.value

    def _grid(self, cls):
        cls = np.ascontiguousarray(cls, np.uint8)
        if cls.shape != (self.H, self.W):
            raise ValueError("node classes must be [%d, %d]" % (self.H, self.W))
        return cls

    def solve(self, hx, hy, cls, b, phi0=None, tol=1e-10, maxit=200, mode=1):
        """Poisson solve on the network's grid (csrst_warmstart_solve): mode 0
        plain V-cycles, 1 the network's step before every V-cycle, 2 before the
        first only, 3 the step over the network's direction and the previous
        change, 4 over the previous change alone.  Returns (phi, cycles,
        relative residual)."""
        cls = self._grid(cls)
        b = np.ascontiguousarray(b, np.float64)
        if b.shape != cls.shape:
            raise ValueError("b must be [%d, %d]" % (self.H, self.W))
        phi = np.zeros_like(b) if phi0 is None else np.array(phi0, np.float64, copy=True, order="C")
        cyc, res = C.c_int(0), C.c_double(0.0)
        check(_lib.lib().csrst_warmstart_solve(self.h, C.c_double(hx), C.c_double(hy),
                                               cls.ctypes.data_as(C.POINTER(C.c_uint8)), _lib.dptr(b), _lib.dptr(phi),
                                               C.c_double(tol), C.c_int(maxit), C.c_int(mode), C.byref(cyc),
                                               C.byref(res)), self.h)
        return phi, cyc.value, res.value

    def train_iterates(self, hx, hy, cls, b, kmax=12, inloop=False, lr=1e-4):
        """Adam steps on (residual, error) pairs of the solver's own iterates
        for the right-hand sides b[n, H, W] (csrst_warmstart_train_iterates).
        Returns (mean loss, pairs used)."""
        cls = self._grid(cls)
        b = np.ascontiguousarray(b, np.float64).reshape(-1, self.H, self.W)
        loss, pairs = C.c_float(0), C.c_int(0)
        check(_lib.lib().csrst_warmstart_train_iterates(self.h, C.c_double(hx), C.c_double(hy),
                                                        cls.ctypes.data_as(C.POINTER(C.c_uint8)), _lib.dptr(b),
                                                        C.c_int(len(b)), C.c_int(kmax), C.c_int(int(inloop)),
                                                        C.c_float(lr), C.byref(loss), C.byref(pairs)), self.h)
        return loss.value, pairs.value
