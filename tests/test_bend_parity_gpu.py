"""GPU against the CPU oracle inside the bend.

* Two-grid engine, bit for bit, with the bunch starting 4 mm before the bend
  and stepping 10 mm into it: the external field acts on most particles, the
  bunch turns, particles cross cell rows and the fields reach the chamber's
  cut cells (Bend.cpp:865-907, TETMSolver.cpp:87-127).
* The reference's CPU-sized workload (the 10 deg, R = 1 m bend on its
  ~50 um global grid, 1e5 particles, first-order shape function, -I start):
  200 steps into the bend, fields, particles and the CSR wake of
  IntegrateElong.cpp:38-89 against the oracle.
* The production moving-window engine on the same workload and steps (its
  window carries the chamber walls; wake sampled at the window's grid
  points as IntegrateElong does), FP64 and FP32, with the error against the
  oracle stated per result.
* One production push (gather, kick, rotation in the bend field, deposit) on
  identical uploaded fields against the oracle's push loop.
"""
import numpy as np
import pytest

from csrst_amd import Commands, TwoGridStep, gaussian_bunch
from csrst_amd.mw import MovingWindowStep
from oracle import oracle as O
from tests.test_twogrid_parity import compare, same

pytestmark = pytest.mark.gpu

GEOM = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=0.1704)
Z0 = np.sqrt(1.256637061e-06 / 8.854187817e-12)


def twogrid_pair(order, Nx, Ny, R, P, Q, stat):
    cmd = Commands(numgridpx=Nx, numgridpy=Ny, lengthbefore=GEOM["lenBefore"], length=GEOM["lenAfter"],
                   width=GEOM["width"], phi=GEOM["phiDeg"], Ekin=GEOM["ekinMeV"], Bz=GEOM["bz"], order=order,
                   statmodulus=stat)
    g = TwoGridStep(cmd)
    g.init_bunch(R, P, Q)
    o = O.Sim(Nx, Ny, order=order, statModulus=stat, **GEOM)
    o.init(R, P, Q)
    return g, o


@pytest.mark.parametrize("order", [0, 1, 2])
def test_twogrid_through_bend_entry(gpu, order):
    Nx, Ny = 1100, 120
    b = O.Bend(Nx=Nx, Ny=Ny, **GEOM)
    R, P, Q = gaussian_bunch(2000, (0.046, 0.0064), 1e-3, 0.5e-3, 100.0, 1.0, seed=4)
    g, o = twogrid_pair(order, Nx, Ny, R, P, Q, stat=1)
    R0, _, _ = g.particles()
    print("order %d: bunch centre 0.046 m, bend entry 0.05 m" % order, flush=True)
    for k in range(70):
        g.step(1)
        o.step(1)
        if k % 10 == 9:
            compare(g, o, "order %d step %d" % (order, k))
    assert same(g.wake(), o.wake())
    R1, P1, _ = g.particles()
    # what the run went through
    assert (b.external_field(R1) != 0).mean() > 0.5
    assert np.mean(P1[:, 1] / P1[:, 0]) > 5e-3
    rows0 = np.floor((R0[:, 1] - b.origin[1]) / b.hr[1])
    rows1 = np.floor((R1[:, 1] - b.origin[1]) / b.hr[1])
    assert (rows0 != rows1).sum() > 50
    ij, _ = b.bcells()
    H = g.field("H")
    assert np.abs((H[0] + H[1])[ij[:, 1] + 1, ij[:, 0] + 1]).max() > 0


# the reference's CPU-sized workload: the 10 deg, R = 1 m bend on a ~50 um
# global grid (4694 x 593: hy >= hx, which the production window requires),
# 1e5 macro-particles, first-order shape function
C1 = dict(Nx=4694, Ny=593, n=100_000, order=1, steps=200, centre=(0.045, 0.0064))


@pytest.fixture(scope="module")
def c1_oracle():
    R, P, Q = gaussian_bunch(C1["n"], C1["centre"], 1e-3, 0.5e-3, 100.0, 1.0, seed=1)
    o = O.Sim(C1["Nx"], C1["Ny"], order=C1["order"], statModulus=1, **GEOM)
    o.init(R, P, Q)
    Rs, Ps, Os = o.particles()              # sorted (PartBunch::sortParticles), oldR from makeReady
    pos = [o.info()["pos"]]
    for _ in range(C1["steps"]):
        o.step(1)
        pos.append(o.info()["pos"])
    Rf, Pf, Of = o.particles()
    return dict(inputs=(R, P, Q), start=(Rs, Ps, Os), sim=o, pos=pos, R=Rf, P=Pf, wake=o.wake(), info=o.info(),
                fields={f: o.field(f) for f in ("E", "H", "mE", "mH", "mJ")})


def test_c1_workload_twogrid_parity(gpu, c1_oracle):
    R, P, Q = c1_oracle["inputs"]
    cmd = Commands(numgridpx=C1["Nx"], numgridpy=C1["Ny"], lengthbefore=GEOM["lenBefore"], length=GEOM["lenAfter"],
                   width=GEOM["width"], phi=GEOM["phiDeg"], Ekin=GEOM["ekinMeV"], Bz=GEOM["bz"], order=C1["order"],
                   statmodulus=1)
    g = TwoGridStep(cmd)
    g.init_bunch(R, P, Q)
    print("c1 workload: %d steps on the GPU" % C1["steps"], flush=True)
    g.step(C1["steps"])
    o = c1_oracle
    assert g.info()["pos"] == o["info"]["pos"]
    for f, ref in o["fields"].items():
        assert same(g.field(f), ref), "field %s differs after %d steps" % (f, C1["steps"])
    Rg, Pg, _ = g.particles()
    assert same(Rg, o["R"]) and same(Pg, o["P"])
    wg, wo = g.wake(), o["wake"]
    scale = np.abs(wo).max()
    assert scale > 0 and np.abs(wg - wo).max() <= 1e-12 * scale
    b = O.Bend(Nx=C1["Nx"], Ny=C1["Ny"], **GEOM)
    assert (b.external_field(Rg) != 0).mean() > 0.9       # the bunch is in the bend


def mw_for_c1(fp32):
    """The production window for the comparison: 512 x 320 cells, tall enough
    that the chamber's cross-section stays inside it while the bunch turns
    (a window edge that cuts the chamber reflects like a wall)."""
    b = O.Bend(Nx=C1["Nx"], Ny=C1["Ny"], **GEOM)
    return MovingWindowStep(512, 320, b.hr[0], b.hr[1], order=C1["order"], fp32=fp32, sortEvery=0, statModulus=1,
                            cfl=1.0, pml=0, wakeGrid=True, **GEOM)


@pytest.mark.parametrize("fp32,tol", [(False, dict(R=1e-8, P=1e-8, E=1e-8, wake=1e-10)),
                                      (True, dict(R=5e-5, P=2e-5, E=1e-4, wake=1e-6))])
def test_c1_workload_production_vs_oracle(gpu, c1_oracle, fp32, tol):
    """The production window (walls inside it, merged Hz, FMA arithmetic,
    split positions) against the oracle over the same 200 steps.  Errors:
    positions in cells, momentum changes relative to the largest change,
    fields over the oracle's bunch window relative to their maximum, wake
    relative to its maximum."""
    o = c1_oracle
    Rs, Ps, _ = o["start"]
    Q = o["inputs"][2]
    m = mw_for_c1(fp32)
    m.init_bunch(Rs, Ps, Q)
    offs = []
    b = O.Bend(Nx=C1["Nx"], Ny=C1["Ny"], **GEOM)
    for k in range(C1["steps"]):
        m.step(1)
        offs.append(m.info()["pos"][0] - o["pos"][k + 1][0])
    # the two windows move along x at the same steps: the wake sample k of
    # the production window is the oracle's sample k + offset
    assert len(set(offs[5:])) == 1, "windows did not move in step: %s" % offs
    off = offs[-1]
    Rm, Pm = m.particles()
    err = {}
    err["R"] = np.abs(Rm - o["R"]).max() / b.hr[0]
    dP = o["P"] - Ps
    err["P"] = np.abs((Pm - Ps) - dP).max() / np.abs(dP).max()
    F = m.fields()
    inf = m.info()
    oi = o["info"]
    W, H = oi["W"], oi["H"]
    # oracle window cell (i, j) = global cell (pos + i, pos + j) = production logical cell (global - wx0)
    sx, sy = oi["pos"][0] - inf["pos"][0], oi["pos"][1] - inf["pos"][1]
    mE, mH = o["fields"]["mE"], o["fields"]["mH"]
    # profile of the field difference by distance from the oracle window's edges
    full_o = np.stack([mE[0][1:H + 1, 1:W + 1], mE[1][1:H + 1, 1:W + 1]])
    full_m = np.stack([F[0][1 + sy:H + 1 + sy, 1 + sx:W + 1 + sx], F[1][1 + sy:H + 1 + sy, 1 + sx:W + 1 + sx]])
    jj, ii = np.mgrid[0:H, 0:W]
    dist = np.minimum(np.minimum(ii, W - 1 - ii), np.minimum(jj, H - 1 - jj))
    dE = np.abs(full_m - full_o).max(0) / np.abs(full_o).max()
    prof = ["%d-%d: %.1e" % (a, b, dE[(dist >= a) & (dist < b)].max()) for a, b in ((0, 2), (2, 4), (4, 8), (8, 16), (16, 999))]
    print("  E difference by distance from the oracle window's edge (cells): " + ", ".join(prof))
    jmax, imax = np.unravel_index(np.argmax(dE), dE.shape)
    print("  largest at oracle window cell (%d, %d) of %d x %d" % (imax, jmax, W, H))
    # the comparison keeps 8 cells away from the reference window's edges
    b8 = 8
    sl_o = (slice(1 + b8, H + 1 - b8), slice(1 + b8, W + 1 - b8))
    sl_m = (slice(1 + b8 + sy, H + 1 - b8 + sy), slice(1 + b8 + sx, W + 1 - b8 + sx))
    e_o = np.stack([mE[0][sl_o], mE[1][sl_o]])
    e_m = np.stack([F[0][sl_m], F[1][sl_m]])
    h_o = (mH[0] + mH[1])[sl_o]
    h_m = (F[2] + F[3])[sl_m]
    err["E"] = max(np.abs(e_m - e_o).max() / np.abs(e_o).max(), np.abs(h_m - h_o).max() / np.abs(h_o).max())
    wm, wo = m.wake(), o["wake"]
    k = np.arange(len(wm))
    ko = k + off
    ok = (ko >= 0) & (ko < len(wo))
    err["wake"] = np.abs(wm[ok] - wo[ko[ok]]).max() / np.abs(wo).max()
    print("c1 workload, production %s vs oracle after %d steps: %s" % ("FP32" if fp32 else "FP64", C1["steps"],
          ", ".join("%s %.2e" % kv for kv in err.items())), flush=True)
    for key, v in err.items():
        assert v < tol[key], (key, v)


def arc_bunch(n, into_arc, tilt_deg, seed):
    b = O.Bend(Nx=1100, Ny=120, **GEOM)
    phi0 = into_arc / b.R
    c = b.center
    centre = (c[0] + b.R * np.sin(phi0), c[1] - b.R * np.cos(phi0))
    th = phi0 + np.radians(tilt_deg)
    return gaussian_bunch(n, centre, 0.3e-3, 0.3e-3, 100.0, 1.0, seed=seed, direction=(np.cos(th), np.sin(th)),
                          energy_spread=1e-3)


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("comp", [0, 1])
@pytest.mark.parametrize("tilt", [0.0, 25.0])
def test_production_push_vs_oracle(gpu, fp32, comp, tilt):
    """csrst_mw_push on uploaded fields against orc_push_patch on the same
    (stored) fields and particles, 5 mm into the arc where Bz acts; at a 25
    deg tilt the particles also cross cell rows.  The production push keeps
    the IntS2 gather without the reference's index slip (s2bug off)."""
    b = O.Bend(Nx=1100, Ny=120, **GEOM)
    hx, hy = b.hr
    R, P, Q = arc_bunch(3000, 5e-3, tilt, seed=7 + comp)
    m = MovingWindowStep(256, 64, hx, hy, order=2, fp32=fp32, sortEvery=0, statModulus=0, **GEOM)
    m.init_bunch(R, P, Q)
    inf = m.info()
    (wx0, wy0), dt = inf["pos"], inf["dt"]
    Wn, Hn = m.W, m.H
    rng = np.random.default_rng(3 + comp)
    jj, ii = np.mgrid[0:Hn + 2, 0:Wn + 2].astype(float)
    F = np.zeros((6, Hn + 2, Wn + 2))
    for c, amp in ((0, 3e7), (1, 3e7), (2, 3e7 / Z0)):
        f = np.zeros_like(ii)
        for _ in range(4):
            kx, ky = rng.uniform(0.05, 0.4, 2)
            f += rng.normal() * np.sin(kx * ii + rng.uniform(0, 6)) * np.cos(ky * jj + rng.uniform(0, 6))
        F[c] = amp * f
    F[3] = F[2]                     # Hzx = Hzy = Hz / 2
    m.set_fields(F)
    Fd = m.fields()                 # as stored (FP32 rounded in the FP32 form)
    R0, P0 = m.particles()
    origin = (inf["origin"][0] - hx, inf["origin"][1] - hy)     # node (0, 0) of the patch: logical (-1, -1)
    Ro, Po, _, Rmid = O.push_patch(2, b, dt, Fd[0], Fd[1], Fd[2] + Fd[3], origin, (hx, hy), R0, P0, s2bug=False)
    assert (b.external_field(R0) != 0).all()
    m.push(comp)
    R1, P1 = m.particles()
    J = m.fields()[4 + comp][1:-1, 1:-1]
    Jo = O.deposit(2, comp, (wx0, wy0, Wn, Hn), tuple(b.origin), (hx, hy), dt, R0, Rmid, np.full(len(R0), Q[0]))
    dPo = Po - P0
    err_p = np.abs((P1 - P0) - dPo).max() / np.abs(dPo).max()
    err_r = np.abs(R1 - Ro).max() / hx
    err_j = np.abs(J - Jo).max() / np.abs(Jo).max()
    rows = (np.floor((Rmid[:, 1] - b.origin[1]) / hy) != np.floor((R0[:, 1] - b.origin[1]) / hy)).sum()
    print("push %s comp %d tilt %g: dP %.2e, R %.2e cells, J %.2e (%d row crossings)"
          % ("FP32" if fp32 else "FP64", comp, tilt, err_p, err_r, err_j, rows), flush=True)
    if tilt:
        assert rows > 100
    tol = (2e-5, 2e-6, 2e-5) if fp32 else (1e-11, 1e-12, 1e-12)
    assert err_p < tol[0] and err_r < tol[1] and err_j < tol[2]


def test_line_fields_vs_oracle(gpu):
    """SaveEMFieldsCommand (SaveEMFields.cpp:34-160) through csrst_line_fields
    after 40 steps into the bend and the copy of the window's E and H that
    precedes it (main.cpp:468-476), against the oracle.  The gathers follow
    ScatterGatherImpl; the reference's own IPPL gather is parity unpinned."""
    Nx, Ny = 1100, 120
    R, P, Q = gaussian_bunch(2000, (0.046, 0.0064), 1e-3, 0.5e-3, 100.0, 1.0, seed=6)
    g, o = twogrid_pair(2, Nx, Ny, R, P, Q, stat=1)
    for _ in range(40):
        g.step(1)
        o.step(1)
    g.copyFields(False)
    o.copy_fields(False)
    lg, lo = g.lineFields(), o.line_fields()
    assert lg.shape == lo.shape and len(lg) > 50
    for col in range(7):
        scale = np.abs(lo[:, col]).max()
        if scale > 0:
            assert np.abs(lg[:, col] - lo[:, col]).max() <= 1e-12 * scale, col
    assert np.abs(lo[:, 1]).max() > 0 and np.abs(lo[:, 3]).max() > 0


def test_gauss_residual(gpu):
    """csrst_gauss_residual: div E - rho/eps0 on the window's vertices.  At
    the -I start (E = 0) it is -rho/eps0 with the oracle's charge density;
    after steps into the bend it equals the same formula on the downloaded
    fields and particles.  Prints how far Gauss's law drifts from the start
    (the split step updates Ex and Ey with H of different half steps, and
    the pushes deposit Jx and Jy half a step apart, so the reference's scheme
    does not hold it at round-off)."""
    eps0 = 8.854187817e-12
    Nx, Ny = 1100, 120
    b = O.Bend(Nx=Nx, Ny=Ny, **GEOM)
    hx, hy = b.hr
    R, P, Q = gaussian_bunch(3000, (0.046, 0.0064), 1e-3, 0.5e-3, 100.0, 1.0, seed=8)
    g, _ = twogrid_pair(1, Nx, Ny, R, P, Q, stat=1)

    def expected():
        inf = g.info()
        Rg, _, _ = g.particles()
        rho = O.charge_density(1, (0, 0, inf["W"], inf["H"]), inf["origin"], (hx, hy), Rg, g.charges())
        E = g.field("mE")
        div = (E[0][1:-1, 1:-1] - E[0][1:-1, :-2]) / hx + (E[1][1:-1, 1:-1] - E[1][:-2, 1:-1]) / hy
        return div - rho / eps0, np.abs(rho).max() / eps0

    G0 = g.gaussResidual()
    want, scale = expected()
    assert np.abs(G0 - want).max() <= 1e-12 * scale
    pos0 = g.info()["pos"]
    drift = []
    for k in range(40):
        g.step(1)
        if k % 10 == 9:
            G = g.gaussResidual()
            want, scale = expected()
            assert np.abs(G - want).max() <= 1e-11 * scale
            (px, py), H, W = g.info()["pos"], *G.shape
            sx, sy = px - pos0[0], py - pos0[1]
            a = G[max(0, -sy):H - max(0, sy), max(0, -sx):W - max(0, sx)]
            b0 = G0[max(0, sy):H - max(0, -sy), max(0, sx):W - max(0, -sx)]
            n0, n1 = min(a.shape[0], b0.shape[0]), min(a.shape[1], b0.shape[1])
            drift.append(np.abs(a[:n0, :n1] - b0[:n0, :n1]).max() / scale)
    print("Gauss residual drift from the start, relative to max rho/eps0, after 10/20/30/40 steps: %s"
          % ", ".join("%.2e" % d for d in drift), flush=True)
