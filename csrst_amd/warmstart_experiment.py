"""Learned warm start experiment: multigrid-generated (rho, phi) pairs, GPU
training of the BF16 U-Net, and the V-cycles multigrid needs to reach a
residual of 1e-10 from zero and from the network's guess on held-out seeds.

Densities: 1-4 Gaussians (sigma 20-500 um, random offsets and amplitudes
summing to one unit of charge) in a rectangular chamber whose walls are the
grid edges.  The operator is the isotropic 5-point Laplacian (the Poisson
start of a bunch at rest in its own frame, gamma hx = hy): the case where
multigrid needs the most cycles.

    python -m csrst_amd.warmstart_experiment --H 128 --W 256 --train 2000 --test 200

prints one JSON line.
"""
import argparse
import json
import sys
import time

import numpy as np


def sample(rng, H, W, h):
    y, x = np.mgrid[0:H, 0:W]
    rho = np.zeros((H, W))
    k = int(rng.integers(1, 5))
    amps = rng.random(k)
    amps /= amps.sum()
    for a in amps:
        s = rng.uniform(20e-6, 500e-6) / h
        s = max(s, 0.5)
        cx, cy = rng.uniform(0.2 * W, 0.8 * W), rng.uniform(0.2 * H, 0.8 * H)
        chirp = rng.uniform(-0.3, 0.3)
        xs = (x - cx) * (1 + chirp * (y - cy) / H)
        rho += a * np.exp(-(xs ** 2 + (y - cy) ** 2) / (2 * s * s)) / (2 * np.pi * s * s)
    return rho


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--H", type=int, default=128)
    ap.add_argument("--W", type=int, default=256)
    ap.add_argument("--h", type=float, default=80e-6)
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=200)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--seconds", type=float, default=150.0, help="training time budget")
    ap.add_argument("--lr", type=float, default=2e-4)
    ap.add_argument("--seed", type=int, default=7)
    a = ap.parse_args()
    sys.path.insert(0, __file__.rsplit("/csrst_amd/", 1)[0])
    from csrst_amd import poisson as P
    from csrst_amd.warmstart import WarmStart
    rng = np.random.default_rng(a.seed)
    H, W, h = a.H, a.W, a.h
    cls = np.zeros((H, W), np.uint8)
    cls[1:-1, 1:-1] = P.UNKNOWN
    t0 = time.time()
    rhos, phis, cold = [], [], []
    for k in range(a.train + a.test):
        rho = sample(rng, H, W, h)
        rho[cls != P.UNKNOWN] = 0.0             # grounded walls: identity rows give phi = 0
        phi, c, res = P.solve(h, h, cls, rho, tol=1e-10, maxit=300)
        rhos.append(rho.astype(np.float32))
        phis.append(phi.astype(np.float32))
        cold.append(c)
    tgen = time.time() - t0
    print("warmstart: %d pairs in %.1f s" % (len(rhos), tgen), file=sys.stderr, flush=True)
    rscale = 1.0 / max(float(np.abs(r).max()) for r in rhos[:a.train])
    pscale = 1.0 / float(np.sqrt(np.mean([np.mean(p.astype(np.float64) ** 2) for p in phis[:a.train]])))
    u = WarmStart(H, W, max_batch=a.batch)
    t0 = time.time()
    order = np.arange(a.train)
    steps, losses = 0, []
    while time.time() - t0 < a.seconds:
        rng.shuffle(order)
        for b in range(0, a.train - a.batch + 1, a.batch):
            idx = order[b:b + a.batch]
            lr = a.lr * (0.3 if time.time() - t0 > 0.7 * a.seconds else 1.0)
            loss = u.train_step(np.stack([rhos[i] for i in idx]), np.stack([phis[i] for i in idx]) * pscale,
                                lr=lr, scale=rscale)
            losses.append(loss)
            steps += 1
            if time.time() - t0 >= a.seconds:
                break
        print("warmstart: %d steps, loss %.3e" % (steps, np.mean(losses[-20:])), file=sys.stderr, flush=True)
    ttrain = time.time() - t0
    warm, relerr, cold_t = [], [], []
    for k in range(a.train, a.train + a.test):
        guess = u.infer(rhos[k][None], scale=rscale)[0].astype(np.float64) / pscale
        relerr.append(float(np.linalg.norm(guess - phis[k]) / np.linalg.norm(phis[k])))
        _, c1, res1 = P.solve(h, h, cls, rhos[k].astype(np.float64), phi0=guess, tol=1e-10, maxit=300)
        warm.append(c1)
        cold_t.append(cold[k])
    rec = {"grid": [H, W], "h_m": h, "train_pairs": a.train, "test_pairs": a.test, "train_steps": steps,
           "train_seconds": ttrain, "final_loss": float(np.mean(losses[-50:])),
           "guess_rel_error_median": float(np.median(relerr)),
           "vcycles_cold_mean": float(np.mean(cold_t)), "vcycles_warm_mean": float(np.mean(warm)),
           "vcycle_ratio": float(np.mean(cold_t) / max(np.mean(warm), 1e-9)), "tol": 1e-10}
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
