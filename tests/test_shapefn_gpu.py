"""ShapeFunction and ScatterGatherImpl through the C ABI (csrst_sg_*,
csrst_sf_*) against the CPU oracle, and the caller-geometry entries of the
two-grid step (csrst_set_mask, csrst_set_boundary_cells, csrst_set_stream).

The stateless kernels keep the reference's per-particle summation order, so
every result is compared bit for bit.  The charge-density difference is
checked against the discrete continuity equation with the GPU's own
currents (dt div J = -(rho(R) - rho(oldR)) to round-off).
"""
import numpy as np
import pytest

from csrst_amd import Commands, TwoGridStep, gaussian_bunch, shapefn as SF
from csrst_amd._lib import CsrstError
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("kind", ["Ex", "Ey", "cell"])
@pytest.mark.parametrize("s2bug", [False, True])
def test_gather_bitwise(gpu, order, kind, s2bug):
    rng = np.random.default_rng(order * 7 + len(kind))
    field = rng.normal(size=(30, 40))
    org, dx = (0.013, -0.021), (0.7, 1.3)
    R = np.stack([rng.uniform(org[0] + 4 * dx[0], org[0] + 34 * dx[0], 500),
                  rng.uniform(org[1] + 4 * dx[1], org[1] + 24 * dx[1], 500)], 1)
    got = SF.gather(order, kind, field, org, dx, R, s2bug=s2bug)
    if kind == "cell":
        ref = O.gather_cell(order, field, org, dx, R, s2bug)
    else:
        ref = O.gather_edge(order, 0 if kind == "Ex" else 1, field, org, dx, R, s2bug)
    assert same(got, ref)


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("kind", ["Ex", "Ey", "cell"])
def test_scatter_bitwise(gpu, order, kind):
    rng = np.random.default_rng(order * 5 + len(kind))
    org, dx = (0.0, 0.0), (1.0, 1.0)
    R = rng.uniform(5, 15, (2000, 2))          # many particles per node: the order of the sums matters
    val = rng.normal(size=2000)
    got = SF.scatter(order, kind, (24, 24), org, dx, R, val)
    if kind == "cell":
        ref = O.scatter_cell(order, (24, 24), org, dx, R, val)
    else:
        ref = O.scatter_edge(order, 0 if kind == "Ex" else 1, (24, 24), org, dx, R, val)
    assert same(got, ref)


def trajectories(n, seed, scale=(0.7, 1.3)):
    rng = np.random.default_rng(seed)
    oldR = rng.uniform(8, 20, (n, 2)) * np.array(scale)
    R = oldR + rng.uniform(-0.99, 0.99, (n, 2)) * np.array(scale)
    return oldR, R, rng.uniform(-1, 1, n)


@pytest.mark.parametrize("order", [0, 1, 2])
def test_current_and_charge_bitwise(gpu, order):
    dx, dt = (0.7, 1.3), 0.37
    box, org = (-6, -6, 40, 40), (0.0, 0.0)
    oldR, R, q = trajectories(3000, order + 3)
    for comp in (0, 1):
        got = SF.getCurrentDensity(order, comp, box, org, dx, dt, oldR, R, q)
        ref = O.deposit(order, comp, box, org, dx, dt, oldR, R, q)
        assert same(got, ref), "comp %d" % comp
    assert same(SF.getChargeDensity(order, box, org, dx, R, q), O.charge_density(order, box, org, dx, R, q))


@pytest.mark.parametrize("order", [0, 1, 2])
def test_charge_diff_continuity(gpu, order):
    dx, dt = (0.7, 1.3), 0.37
    box, org = (-6, -6, 40, 40), (0.0, 0.0)
    oldR, R, q = trajectories(3000, order + 17)
    diff = SF.getChargeDensity(order, box, org, dx, R, q, oldR=oldR)
    r1 = O.charge_density(order, box, org, dx, R, q)
    r0 = O.charge_density(order, box, org, dx, oldR, q)
    scale = np.abs(r1).max()
    assert np.abs(diff - (r1 - r0)).max() < 1e-13 * scale
    Jx = SF.getCurrentDensity(order, 0, box, org, dx, dt, oldR, R, q)
    Jy = SF.getCurrentDensity(order, 1, box, org, dx, dt, oldR, R, q)
    div = np.zeros_like(diff)
    div[:, 1:] += (Jx[:, 1:] - Jx[:, :-1]) / dx[0]
    div[1:, :] += (Jy[1:, :] - Jy[:-1, :]) / dx[1]
    assert np.abs(diff + dt * div).max() < 1e-12 * scale


def test_patch_overflow_is_reported(gpu):
    oldR, R, q = trajectories(50, 1)
    with pytest.raises(CsrstError) as e:
        SF.getCurrentDensity(2, 0, (10, 10, 5, 5), (0.0, 0.0), (0.7, 1.3), 0.37, oldR, R, q)
    assert e.value.code == -4


GEOM = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=0.1704)


def stepper():
    cmd = Commands(numgridpx=1100, numgridpy=120, lengthbefore=GEOM["lenBefore"], length=GEOM["lenAfter"],
                   width=GEOM["width"], phi=GEOM["phiDeg"], Ekin=GEOM["ekinMeV"], Bz=GEOM["bz"], order=1,
                   statmodulus=1)
    return TwoGridStep(cmd)


def test_caller_geometry_and_stream(gpu):
    """The Bend's own mask and cells handed back through csrst_set_mask and
    csrst_set_boundary_cells, and the work moved to a caller stream, leave
    the steps bit-identical to the oracle; a changed cell list changes them."""
    from tests.test_twogrid_parity import compare
    R, P, Q = gaussian_bunch(1500, (0.046, 0.0064), 1e-3, 0.5e-3, 100.0, 1.0, seed=5)
    g = stepper()
    mask, ij, data = g.boundary_cells()
    g.set_mask(mask)
    g.set_boundary_cells(ij, data)
    stream = SF.Stream()
    g.set_stream(stream.handle)
    g.init_bunch(R, P, Q)
    o = O.Sim(1100, 120, order=1, statModulus=1, **GEOM)
    o.init(R, P, Q)
    for k in range(30):
        g.step(1)
        o.step(1)
    compare(g, o, "caller geometry, step 30")
    g.set_stream(None)
    g.step(1)
    o.step(1)
    compare(g, o, "own stream again")
    # cut cells given half their area: the fields at the walls change
    g2 = stepper()
    data2 = data.copy()
    data2[:, 2] *= 0.5
    g2.set_boundary_cells(ij, data2)
    g2.init_bunch(R, P, Q)
    g2.step(31)
    assert not same(g2.field("H"), g.field("H"))
    with pytest.raises(CsrstError):
        g2.set_mask(mask)                     # the geometry is fixed once the bunch is in
    g3 = stepper()
    bad = ij.copy()
    bad[0] = bad[1]
    with pytest.raises(CsrstError):
        g3.set_boundary_cells(bad, data)      # a cell listed twice
    stream.close()


def test_empty_bunch_and_one_node_patch(gpu):
    """No particles: the deposits and the charge density are zero on the
    patch, a scatter leaves the field as it was, a gather returns nothing.
    The smallest patch that holds an S0 (CIC) charge, the 2 x 2 vertices of
    its cell, gets the oracle's values."""
    org, dx, dt = (0.0, 0.0), (0.7, 1.3), 0.37
    none = np.zeros((0, 2))
    for order in (0, 1, 2):
        for comp in (0, 1):
            J = SF.getCurrentDensity(order, comp, (0, 0, 7, 5), org, dx, dt, none, none, np.zeros(0))
            assert J.shape == (5, 7) and not J.any()
        assert not SF.getChargeDensity(order, (0, 0, 7, 5), org, dx, none, np.zeros(0)).any()
        field = np.arange(35.0).reshape(5, 7)
        assert same(SF.scatter(order, "cell", (5, 7), org, dx, none, np.zeros(0), field=field), field)
        assert SF.gather(order, "Ex", field, org, dx, none).shape == (0,)
    R, q = np.array([[3.2 * dx[0], 2.1 * dx[1]]]), np.array([2.0])
    rho = SF.getChargeDensity(0, (3, 2, 2, 2), org, dx, R, q)
    assert rho.shape == (2, 2) and same(rho, O.charge_density(0, (3, 2, 2, 2), org, dx, R, q))
    with pytest.raises(CsrstError):
        SF.getChargeDensity(0, (3, 2, 1, 1), org, dx, R, q)      # the stencil leaves a 1 x 1 patch
