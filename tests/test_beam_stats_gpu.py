"""GPU beam parameters (csrst_beam_stats, csrst_mw_beam_stats) against the
oracle's restatement of PartBunch::calcMoments / calcBeamParameters
(PartBunch.cpp:584-756) on the same particles."""
import numpy as np
import pytest

from csrst_amd import Commands, TwoGridStep, gaussian_bunch
from csrst_amd.mw import MovingWindowStep
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GEOM = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=0.1704)

# the longitudinal momentum spread (PartBunch.cpp:625) cancels sum p^2 against
# N pmean^2 and loses ~(|p|/prms)^2 of relative precision: looser there
TOL = dict(prms_x=1e-8, dE=1e-8, emit_x=1e-6, rprms_x=1e-8, Dy=1e-9, DDy=1e-9)


def check(g, o, tag):
    worst = 0.0
    for k, v in o.items():
        tol = TOL.get(k, 1e-12)
        d = abs(g[k] - v)
        scale = max(abs(v), 1e-300)
        worst = max(worst, d / scale / tol)
        assert d <= tol * scale, "%s: %s gpu %r oracle %r" % (tag, k, g[k], v)
    print("%s: worst difference %.2f of its tolerance" % (tag, worst))


def test_two_grid_bunch_parameters(gpu):
    cmd = Commands(numgridpx=1100, numgridpy=120, lengthbefore=GEOM["lenBefore"], length=GEOM["lenAfter"],
                   width=GEOM["width"], phi=GEOM["phiDeg"], Ekin=GEOM["ekinMeV"], Bz=GEOM["bz"], order=1,
                   statmodulus=1)
    R, P, Q = gaussian_bunch(4000, (0.02, 0.0064), 1e-3, 0.5e-3, 100.0, 1.0, seed=8, energy_spread=1e-3)
    g = TwoGridStep(cmd)
    g.init_bunch(R, P, Q)
    for it in range(3):
        Rg, Pg, _ = g.particles()
        check(g.calcBeamParameters(), O.beam_parameters(Rg, Pg), "two-grid step %d" % it)
        g.step(1)


@pytest.mark.parametrize("fp32", [False, True])
def test_moving_window_bunch_parameters(gpu, fp32):
    b = O.Bend(Nx=1100, Ny=120, **GEOM)
    R, P, Q = gaussian_bunch(20000, (0.03, 0.0064), 0.4e-3, 0.3e-3, 100.0, 1.0, seed=9, energy_spread=1e-3)
    m = MovingWindowStep(512, 100, b.hr[0], b.hr[1], order=2, fp32=fp32, sortEvery=5, statModulus=0, **GEOM)
    m.init_bunch(R, P, Q)
    for it in range(3):
        Rm, Pm = m.particles()
        check(m.calcBeamParameters(), O.beam_parameters(Rm, Pm), "moving window fp32=%s step %d" % (fp32, 4 * it))
        m.step(4)


def test_field_energy_sums_match_oracle(gpu):
    """csrst_field_energy against the oracle's CalcEnergyCommand sums over the
    global grid plus the window minus the overlap (CalcEnergy.cpp:60-124)."""
    cmd = Commands(numgridpx=1100, numgridpy=120, lengthbefore=GEOM["lenBefore"], length=GEOM["lenAfter"],
                   width=GEOM["width"], phi=GEOM["phiDeg"], Ekin=GEOM["ekinMeV"], Bz=GEOM["bz"], order=1,
                   statmodulus=1)
    R, P, Q = gaussian_bunch(2000, (0.02, 0.0064), 1e-3, 0.5e-3, 100.0, 1.0, seed=8)
    g = TwoGridStep(cmd)
    g.init_bunch(R, P, Q)
    g.computeField()
    o = O.Sim(1100, 120, order=1, statModulus=1, **GEOM)
    o.init(R, P, Q)
    o.poisson()
    for it in range(3):
        a, b = g.fieldEnergySums(), o.field_energy_sums()
        print("step %d: E.E %.12e / %.12e, H^2 %.12e / %.12e" % (it, a[0], b[0], a[1], b[1]))
        assert np.all(np.abs(a - b) <= 1e-12 * np.abs(b))
        g.step(1)
        o.step(1)
