"""CPU (numpy, float64) restatement of the warm-start U-Net of
csrst_amd/csrc/csrst_unet.hip, forward and backward, used to check the GPU
kernels.  With `bf16=True` it rounds what the GPU stores in BF16 (input,
weights, activations, gradients) at the same points, so the remaining
differences are those of FP32 accumulation order."""
import numpy as np

from csrst_amd.warmstart import layers


def to_bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


class Ref:
    def __init__(self, params, base=32, bf16=True):
        self.L, n = layers(base)
        assert params.size == n
        self.r = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
        self.Wt, self.bias = [], []
        for l in self.L:
            w = params[l["woff"]:l["woff"] + l["cout"] * 9 * l["cin"]].reshape(l["cout"], 9, l["cin"])
            self.Wt.append(self.r(w))
            self.bias.append(np.asarray(params[l["boff"]:l["boff"] + l["cout"]], np.float64))

    @staticmethod
    def _conv(x, w, b=None, flip=False):
        """3x3, zero padding 1; x [B,H,W,Cin], w [Cout,9,Cin]."""
        B, H, W, _ = x.shape
        xp = np.zeros((B, H + 2, W + 2, x.shape[3]))
        xp[:, 1:-1, 1:-1] = x
        out = np.zeros((B, H, W, w.shape[0]))
        for t in range(9):
            ky, kx = divmod(t, 3)
            if flip:
                ky, kx = 2 - ky, 2 - kx
            out += xp[:, ky:ky + H, kx:kx + W, :] @ w[:, t, :].T
        if b is not None:
            out += b
        return out

    def forward(self, rho, scale=1.0):
        r = self.r
        x0 = r(np.asarray(rho, np.float64)[..., None] * np.float32(scale))
        A = [x0]
        self.A = A

        def conv(i, x):
            l = self.L[i]
            y = self._conv(x, self.Wt[i], self.bias[i])
            if l["relu"]:
                y = np.maximum(y, 0)
            return y

        def pool(x):
            return r(0.25 * (x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2]))

        def up(x):
            return x.repeat(2, axis=1).repeat(2, axis=2)

        A.append(r(conv(0, A[0])))
        A.append(r(conv(1, A[1])))
        A.append(pool(A[2]))
        A.append(r(conv(2, A[3])))
        A.append(r(conv(3, A[4])))
        A.append(pool(A[5]))
        A.append(r(conv(4, A[6])))
        A.append(r(conv(5, A[7])))
        A.append(pool(A[8]))
        A.append(r(conv(6, A[9])))
        A.append(r(conv(7, A[10])))
        A.append(r(conv(8, A[11])))
        A.append(r(up(A[12]) + A[8]))
        A.append(r(conv(9, A[13])))
        A.append(r(conv(10, A[14])))
        A.append(r(up(A[15]) + A[5]))
        A.append(r(conv(11, A[16])))
        A.append(r(conv(12, A[17])))
        A.append(r(up(A[18]) + A[2]))
        A.append(r(conv(13, A[19])))
        out = conv(14, A[20])          # fp32 output
        A.append(r(out))
        return out[..., 0]

    def backward(self, out, phi):
        """Gradients of mean((out - phi)^2) for every parameter, in the
        layout of csrst_amd.warmstart.layers()."""
        r, A = self.r, self.A
        n = out.size
        g = np.zeros(sum(l["cout"] * 9 * l["cin"] + l["cout"] for l in self.L))
        D = [None] * 22
        D[21] = r(2.0 * (out - phi)[..., None] / n)

        def back(li, xin, yout, need_dx=True):
            l = self.L[li]
            dpre = D[yout]
            if l["relu"]:
                dpre = np.where(A[yout] > 0, dpre, 0.0)
            B, H, W, _ = dpre.shape
            xp = np.zeros((B, H + 2, W + 2, A[xin].shape[3]))
            xp[:, 1:-1, 1:-1] = A[xin]
            dw = np.zeros((l["cout"], 9, l["cin"]))
            for t in range(9):
                ky, kx = divmod(t, 3)
                dw[:, t, :] = np.einsum("bhwo,bhwi->oi", dpre, xp[:, ky:ky + H, kx:kx + W, :])
            g[l["woff"]:l["woff"] + dw.size] = dw.ravel()
            g[l["boff"]:l["boff"] + l["cout"]] = dpre.sum(axis=(0, 1, 2))
            if need_dx:
                wt = np.transpose(self.Wt[li], (2, 1, 0))           # [cin, 9, cout]
                D[xin] = r(self._conv(dpre, wt, flip=True))

        def upT(d):
            return d[:, 0::2, 0::2] + d[:, 0::2, 1::2] + d[:, 1::2, 0::2] + d[:, 1::2, 1::2]

        def poolT(d):
            return 0.25 * d.repeat(2, axis=1).repeat(2, axis=2)

        back(14, 20, 21)
        back(13, 19, 20)
        D[18] = r(upT(D[19]))
        back(12, 17, 18)
        back(11, 16, 17)
        D[15] = r(upT(D[16]))
        back(10, 14, 15)
        back(9, 13, 14)
        D[12] = r(upT(D[13]))
        back(8, 11, 12)
        back(7, 10, 11)
        back(6, 9, 10)
        D[8] = r(r(poolT(D[9])) + D[13])
        back(5, 7, 8)
        back(4, 6, 7)
        D[5] = r(r(poolT(D[6])) + D[16])
        back(3, 4, 5)
        back(2, 3, 4)
        D[2] = r(r(poolT(D[3])) + D[19])
        back(1, 1, 2)
        back(0, 0, 1, need_dx=False)
        return g
