"""CPU checks of the U-Net restatement used to test the GPU warm start: the
analytic backward against central differences of the forward (float64, no
BF16 rounding), on a tiny grid with a narrow network."""
import numpy as np

from tests.unet_ref import Ref, to_bf16
from csrst_amd.warmstart import layers


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], np.float64)
    r = to_bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 3.0e-3) < 3.0e-3 * 2 ** -8


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    base = 4
    L, n = layers(base)
    p = rng.normal(0, 0.3, n)
    rho = rng.normal(size=(1, 16, 16))
    phi = rng.normal(size=(1, 16, 16))
    ref = Ref(p, base=base, bf16=False)
    out = ref.forward(rho)
    g = ref.backward(out, phi)

    def loss(q):
        rr = Ref(q, base=base, bf16=False)
        o = rr.forward(rho)
        return np.mean((o - phi) ** 2)

    idx = np.concatenate([rng.choice(n, 24, replace=False), [L[0]["boff"], L[14]["boff"], L[7]["woff"] + 5]])
    for k in idx:
        e = np.zeros(n)
        e[k] = 1e-6
        fd = (loss(p + e) - loss(p - e)) / 2e-6
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])
