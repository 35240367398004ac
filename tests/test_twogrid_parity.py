"""Two-grid step on the GPU against the CPU FP64 oracle: bit-identical fields,
particles and window position after every step (small grids, seeded bunches).
"""
import numpy as np
import pytest

from csrst_amd import Commands, TwoGridStep, gaussian_bunch
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GEOM = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=0.1704)


def make(order, Nx=1100, Ny=120, n=2000, seed=3, s2bug=True, stat=1, center=(0.02, 0.0064)):
    cmd = Commands(numgridpx=Nx, numgridpy=Ny, lengthbefore=GEOM["lenBefore"], length=GEOM["lenAfter"],
                   width=GEOM["width"], phi=GEOM["phiDeg"], Ekin=GEOM["ekinMeV"], Bz=GEOM["bz"], order=order,
                   s2bug=s2bug, statmodulus=stat)
    R, P, Q = gaussian_bunch(n, center, 1e-3, 0.5e-3, 100.0, 1.0, seed=seed)
    g = TwoGridStep(cmd)
    g.init_bunch(R, P, Q)
    o = O.Sim(Nx, Ny, order=order, s2bug=s2bug, statModulus=stat, **GEOM)
    o.init(R, P, Q)
    return g, o


def same(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def compare(g, o, tag):
    gi, oi = g.info(), o.info()
    assert gi["pos"] == oi["pos"] and gi["W"] == oi["W"] and gi["H"] == oi["H"], tag
    assert gi["origin"] == oi["origin"], tag
    for f in ("E", "H", "mE", "mH", "mJ"):
        a, b = g.field(f), o.field(f)
        if not same(a, b):
            d = np.abs(a - b)
            idx = np.unravel_index(np.argmax(d), d.shape)
            pytest.fail("%s: field %s differs, max |d| %.3e at %s (gpu %r, cpu %r)" % (tag, f, d.max(), idx, a[idx], b[idx]))
    Rg, Pg, _ = g.particles()
    Ro, Po, _ = o.particles()
    assert same(Rg, Ro), "%s: positions differ (max %.3e)" % (tag, np.abs(Rg - Ro).max())
    assert same(Pg, Po), "%s: momenta differ (max %.3e)" % (tag, np.abs(Pg - Po).max())


@pytest.mark.parametrize("order", [0, 1, 2])
def test_init_matches(gpu, order):
    g, o = make(order)
    compare(g, o, "init")


@pytest.mark.parametrize("order", [0, 1, 2])
def test_steps_bit_identical(gpu, order):
    g, o = make(order)
    for k in range(12):
        g.step(1)
        o.step(1)
        compare(g, o, "order %d step %d" % (order, k))


def test_wake_and_diag(gpu):
    g, o = make(0, stat=1)
    g.step(15)
    o.step(15)
    wg, wo = g.wake(), o.wake()
    assert len(wg) == len(wo)
    scale = np.abs(wo).max()
    assert scale > 0
    assert np.abs(wg - wo).max() <= 1e-12 * scale
    dg, do = g.diag(), o.diag()
    assert np.allclose(dg[:5], do[:5], rtol=1e-13, atol=0)
    assert abs(dg[5] - do[5]) <= 1e-10 * abs(do[5])


SUBSTEPS = [0, 2, 5, 8, 3, 6, 9, 2, 5, 8, 11, 1, 4, 7, 10, 11, 12, 13, 14]


@pytest.mark.parametrize("order", [0, 2])
def test_substeps_bit_identical(gpu, order):
    """The step taken apart (csrst_substep) against the oracle's parts, every
    field compared bit for bit after every part of two steps."""
    g, o = make(order, n=1000, stat=1)
    for step in range(2):
        for op in SUBSTEPS:
            g.substep(op)
            o.substep(op)
            for f in ("E", "H", "mE", "mH", "mJ"):
                assert same(g.field(f), o.field(f)), "step %d, part %d: field %s differs" % (step, op, f)
    assert g.info()["pos"] == o.info()["pos"]


def test_copy_fields_bit_identical(gpu):
    """copyFields (main.cpp:534-548) through the C ABI against the oracle's
    Communicator::copy restatement (window E, H and optionally J into the
    global fields), and the steps after it, which start from the changed
    global fields."""
    g, o = make(1, n=1500, seed=9)
    for _ in range(3):
        g.step(1)
        o.step(1)
    for cur in (False, True):
        before = g.field("E")
        g.copyFields(cur)
        o.copy_fields(cur)
        assert not same(before, g.field("E")), "the copy left the global E unchanged"
        compare(g, o, "copy (J %s)" % cur)
        assert same(g.field("J"), o.field("J")), "global J after the copy"
        for k in range(2):
            g.step(1)
            o.step(1)
            compare(g, o, "step %d after copy (J %s)" % (k, cur))


def test_restore_bit_identical(gpu):
    """The restart (LoadH5Part::restoreFields + restoreBunch, main.cpp:147-153,
    226-228) through the C ABI against the oracle: global E and H, the bunch
    with its oldR (no makeReady) and the step counter, which phases the wake
    samples; then steps on both."""
    stat = 3
    g, o = make(2, n=1500, seed=21, stat=stat)
    for _ in range(4):
        g.step(1)
        o.step(1)
    g.copyFields(True)
    E, H = g.field("E"), g.field("H")
    R, P, Rold = g.particles()
    Q = g.charges()
    g2, o2 = make(2, n=10, seed=1, stat=stat)       # fresh steppers, then restored
    g2.restore(E, H, R, Rold, P, Q, step=4)
    o2.restore(E, H, R, Rold, P, Q, step=4)
    assert g2.info()["steps"] == 4 and o2.info()["steps"] == 4
    compare(g2, o2, "restore")
    for k in range(5):
        g2.step(1)
        o2.step(1)
        compare(g2, o2, "step %d after restore" % k)
    assert same(g2.wake(), o2.wake()) and np.abs(g2.wake()).max() > 0
