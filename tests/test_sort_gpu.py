"""The production step's particle sort (every sortEvery steps).

- It only permutes the particles: FP64 runs that sort every step and every
  25 steps give the same fields, wake and beam to round-off.
- Its order keeps runs of consecutive particles within a small part of one
  cell (tiles over the bunch's cell box, then the cell, then Z-ordered bins
  in the cell), which the push relies on for its register accumulation.
"""
import numpy as np
import pytest

from csrst_amd.mw import MovingWindowStep
from csrst_amd import gaussian_bunch
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GEOM = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=0.1704)
NX, NY = 1100, 120


def make(sort_every, fp32=False, n=200000, steps=0):
    R, P, Q = gaussian_bunch(n, (0.03, 0.0064), 0.4e-3, 0.3e-3, 100.0, 1.0, seed=3, energy_spread=1e-3)
    b = O.Bend(Nx=NX, Ny=NY, **GEOM)
    m = MovingWindowStep(512, 100, b.hr[0], b.hr[1], order=2, fp32=fp32, sortEvery=sort_every, statModulus=1,
                         **GEOM)
    m.init_bunch(R, P, Q)
    if steps:
        m.step(steps)
    return m, b


def test_sort_is_a_permutation(gpu):
    steps = 50
    m1, _ = make(1, steps=steps)
    m25, _ = make(25, steps=steps)
    assert m1.info()["pos"] == m25.info()["pos"]
    w1, w25 = m1.wake(), m25.wake()
    assert np.abs(w1 - w25).max() <= 1e-12 * np.abs(w25).max()
    f1, f25 = m1.fields(), m25.fields()
    for c in range(4):
        assert np.abs(f1[c] - f25[c]).max() <= 1e-12 * np.abs(f25[c]).max(), "field %d" % c
    R1, P1 = m1.particles()
    R25, P25 = m25.particles()
    # the same particles in another order: each coordinate's sorted values
    for A, B in ((R1, R25), (P1, P25)):
        for c in range(2):
            assert np.abs(np.sort(A[:, c]) - np.sort(B[:, c])).max() <= 1e-12 * np.abs(B[:, c]).max()
    m1.close()
    m25.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_sorted_runs_are_compact(gpu, fp32):
    m, b = make(20, fp32=fp32)
    R, _ = m.particles()
    h = np.array(b.hr)
    o = np.array(m.info()["origin"])
    u = (R - o) / h                     # positions in cells from the window's node (0, 0)
    n4 = (len(u) // 4) * 4
    runs = u[:n4].reshape(-1, 4, 2)
    cells = np.floor(runs)
    same_cell = np.all(cells == cells[:, :1, :], axis=(1, 2))
    spread = (runs.max(axis=1) - runs.min(axis=1)).max(axis=1)
    # hundreds of particles per cell at the core: 4 consecutive particles
    # share a cell and lie within a few hundredths of a cell of each other
    # (ordered by half-cells alone they would spread over about half a cell)
    assert same_cell.mean() > 0.99, same_cell.mean()
    assert np.median(spread) < 0.05, np.median(spread)
    assert np.quantile(spread, 0.9) < 0.2, np.quantile(spread, 0.9)
    m.close()
