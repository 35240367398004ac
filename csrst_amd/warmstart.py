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
