"""GPU warm start (BF16 MFMA U-Net) against the numpy restatement, and a
short training run that must lower the loss and the multigrid V-cycles."""
import numpy as np
import pytest

from csrst_amd import poisson as P
from csrst_amd.warmstart import WarmStart, layers
from tests.unet_ref import Ref

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("H,W,B", [(16, 24, 2), (32, 48, 2)])
def test_forward_and_gradients_match_cpu(gpu, H, W, B):
    """At 32 x 48 x 2 the top level holds 3072 pixels, more than one
    1024-pixel chunk of the weight-gradient kernel: its partial sums from
    several chunks are added."""
    rng = np.random.default_rng(1)
    u = WarmStart(H, W, max_batch=B)
    p = u.weights()
    p[np.concatenate([np.arange(l["boff"], l["boff"] + l["cout"]) for l in layers()[0]])] = \
        rng.normal(0, 0.05, sum(l["cout"] for l in layers()[0]))
    u.set_weights(p)
    rho = rng.normal(size=(B, H, W)).astype(np.float32)
    phi = rng.normal(size=(B, H, W)).astype(np.float32)
    out = u.infer(rho)
    ref = Ref(u.weights())
    want = ref.forward(rho)
    e = rel(out, want)
    print("forward relative error %.2e" % e)
    assert e < 2e-2
    loss, g = u.gradients(rho, phi)
    want_out = ref.forward(rho)
    gw = ref.backward(want_out, phi)
    assert abs(loss - np.mean((want_out - phi) ** 2)) < 2e-2 * loss
    cos = float(g @ gw / (np.linalg.norm(g) * np.linalg.norm(gw)))
    print("gradient cosine %.6f, relative error %.2e" % (cos, rel(g, gw)))
    assert cos > 0.999
    for i, l in enumerate(layers()[0]):
        sl = slice(l["woff"], l["boff"] + l["cout"])
        e = rel(g[sl], gw[sl])
        assert e < 6e-2, (i, e)


def make_sample(rng, H, W, h):
    """1-4 Gaussians (sigma 40-500 um) summing to one charge unit, on an
    H x W grid of spacing h inside a rectangular chamber (edges grounded)."""
    y, x = np.mgrid[0:H, 0:W]
    rho = np.zeros((H, W))
    k = rng.integers(1, 5)
    amps = rng.random(k)
    amps /= amps.sum()
    for a in amps:
        s = rng.uniform(40e-6, 500e-6) / h
        cx, cy = rng.uniform(0.25 * W, 0.75 * W), rng.uniform(0.25 * H, 0.75 * H)
        rho += a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s)) / (2 * np.pi * s * s)
    return rho


def test_training_saves_vcycles(gpu):
    rng = np.random.default_rng(2)
    H, W, h = 64, 128, 80e-6
    cls = np.zeros((H, W), np.uint8)
    cls[1:-1, 1:-1] = P.UNKNOWN
    data = []
    for _ in range(96):
        rho = make_sample(rng, H, W, h)
        rho[cls != P.UNKNOWN] = 0.0             # grounded walls: phi = 0 there
        phi, cyc, res = P.solve(h, h, cls, rho, tol=1e-12, maxit=100)
        data.append((rho, phi))
    # inputs scaled to peak 1, targets to unit RMS (the problem is linear)
    scale = 1.0 / max(np.abs(r).max() for r, _ in data)
    pscale = 1.0 / np.sqrt(np.mean([np.mean(p * p) for _, p in data]))
    u = WarmStart(H, W, max_batch=8)
    train, test = data[:80], data[80:]
    losses = []
    for epoch in range(40):
        for b in range(0, len(train), 8):
            rb = np.stack([r for r, _ in train[b:b + 8]])
            pb = np.stack([p for _, p in train[b:b + 8]]) * pscale
            losses.append(u.train_step(rb, pb, lr=2e-4, scale=scale))
    first, last = np.mean(losses[:10]), np.mean(losses[-10:])
    print("loss %.3e -> %.3e" % (first, last))
    assert last < 0.1 * first
    cold, warm = [], []
    for rho, _ in test:
        guess = u.infer(rho[None], scale=scale)[0].astype(np.float64) / pscale
        _, c0, _ = P.solve(h, h, cls, rho, tol=1e-10, maxit=200)
        _, c1, _ = P.solve(h, h, cls, rho, phi0=guess, tol=1e-10, maxit=200)
        cold.append(c0)
        warm.append(c1)
    print("V-cycles to 1e-10: cold %.2f, warm %.2f" % (np.mean(cold), np.mean(warm)))
    assert np.mean(warm) < np.mean(cold)


def test_first_adam_step_is_a_sign_step(gpu):
    """Adam's first step moves every parameter by lr sign(g) (m = g, v = g^2)."""
    rng = np.random.default_rng(3)
    H, W = 16, 24
    u = WarmStart(H, W, max_batch=2)
    rho = rng.normal(size=(2, H, W)).astype(np.float32)
    phi = rng.normal(size=(2, H, W)).astype(np.float32)
    w0 = u.weights()
    _, g = u.gradients(rho, phi)
    u.train_step(rho, phi, lr=1e-3)
    d = u.weights() - w0
    sel = np.abs(g) > 1e-3 * np.abs(g).max()
    print("moved %d of %d parameters; max |step| %.3e" % ((d != 0).sum(), d.size, np.abs(d).max()))
    assert np.allclose(d[sel], -1e-3 * np.sign(g[sel]), rtol=1e-3, atol=1e-7)


def test_training_matches_cpu_trajectory(gpu):
    """Ten Adam steps on one batch, GPU against the numpy restatement doing the
    same updates from the same start: the losses follow each other and fall."""
    rng = np.random.default_rng(4)
    H, W = 16, 32
    u = WarmStart(H, W, max_batch=4)
    y, x = np.mgrid[0:H, 0:W]
    rho = np.stack([np.exp(-((x - rng.uniform(8, 24)) ** 2 + (y - rng.uniform(4, 12)) ** 2) / 4.0)
                    for _ in range(4)]).astype(np.float32)
    phi = np.stack([np.exp(-np.sqrt((x - 16.0) ** 2 + (y - 8.0) ** 2) / 6.0)] * 4).astype(np.float32)
    phi /= np.sqrt(np.mean(phi ** 2))
    p = u.weights().astype(np.float64)
    m, v = np.zeros_like(p), np.zeros_like(p)
    gpu_l, cpu_l = [], []
    for t in range(1, 11):
        gpu_l.append(u.train_step(rho, phi, lr=1e-3))
        ref = Ref(p.astype(np.float32))
        out = ref.forward(rho)
        cpu_l.append(float(np.mean((out - phi) ** 2)))
        g = ref.backward(out, phi)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        p -= 1e-3 * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
    print("gpu losses", ["%.4f" % l for l in gpu_l])
    print("cpu losses", ["%.4f" % l for l in cpu_l])
    assert abs(gpu_l[0] - cpu_l[0]) < 2e-2 * cpu_l[0]
    assert cpu_l[-1] < 0.9 * cpu_l[0]
    assert gpu_l[-1] < 0.9 * gpu_l[0]


def test_solve_modes_and_iterate_training(gpu):
    """csrst_warmstart_solve: mode 0 is the multigrid of csrst_poisson_solve
    (same V-cycles, same solution bit for bit); training on the solver's own
    iterates lowers the loss; with the network's step before every V-cycle
    (mode 1), before the first (mode 2), with the previous change as a second
    direction (mode 3) and the previous change alone (mode 4) the solve
    reaches the same tolerance and the same potential."""
    rng = np.random.default_rng(5)
    H, W, h = 64, 128, 80e-6
    cls = np.zeros((H, W), np.uint8)
    cls[1:-1, 1:-1] = P.UNKNOWN

    def sample():
        rho = make_sample(rng, H, W, h)
        rho[cls != P.UNKNOWN] = 0.0
        return rho

    u = WarmStart(H, W, max_batch=4)
    rho = sample()
    phi0, c0, r0 = P.solve(h, h, cls, rho, tol=1e-10, maxit=200)
    phi_m0, cm0, rm0 = u.solve(h, h, cls, rho, tol=1e-10, maxit=200, mode=0)
    assert cm0 == c0 and rm0 == r0
    assert np.array_equal(phi_m0, phi0)
    losses = []
    for it in range(40):
        loss, n = u.train_iterates(h, h, cls, np.stack([sample(), sample()]), kmax=8, inloop=it >= 20, lr=3e-4)
        assert n > 0
        losses.append(loss)
    print("iterate-pair loss %.3f -> %.3f" % (np.mean(losses[:5]), np.mean(losses[-5:])))
    assert np.mean(losses[-5:]) < 0.8 * np.mean(losses[:5])
    for mode in (1, 2, 3, 4):
        phi, c, r = u.solve(h, h, cls, rho, tol=1e-10, maxit=200, mode=mode)
        print("mode %d: %d V-cycles (plain %d), residual %.1e" % (mode, c, c0, r))
        assert r <= 1e-10
        assert np.abs(phi - phi0).max() <= 1e-8 * np.abs(phi0).max()
    with pytest.raises(ValueError):
        u.solve(h, h, cls[:-1], rho)
    u.close()
