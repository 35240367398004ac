"""The CPU oracle against the reference's known answers and against what the
mathematics fixes (partition of unity, linear reproduction, continuity,
brute-force crossings, analytic wall intersections, fast-rsqrt error bound).
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "scattergather.json")))
ORDERS = {"CIC": 0, "S1": 1, "S2": 2}


@pytest.mark.parametrize("name", ["CIC", "S1", "S2"])
def test_scatter_golden_vectors(name):
    """test/scattergather.cpp:90-342: particle (3.3, 3.2), unit mesh, tol 1e-8."""
    order = ORDERS[name]
    R = np.array([GOLD["particle"]])
    shape = tuple(GOLD["patch"])
    ex = O.scatter_edge(order, 0, shape, (0, 0), (1, 1), R, [1.0])
    ey = O.scatter_edge(order, 1, shape, (0, 0), (1, 1), R, [1.0])
    bz = O.scatter_cell(order, shape, (0, 0), (1, 1), R, [1.0])
    for i, j, vx, vy in GOLD[name]["edge"]:
        assert abs(ex[j, i] - vx) <= GOLD["tolerance"]
        assert abs(ey[j, i] - vy) <= GOLD["tolerance"]
    for i, j, v in GOLD[name]["cell"]:
        assert abs(bz[j, i] - v) <= GOLD["tolerance"]
    # everything not listed is zero and the weights sum to one
    listed = {(i, j) for i, j, _, _ in GOLD[name]["edge"]}
    for (j, i), v in np.ndenumerate(ex):
        if (i, j) not in listed:
            assert v == 0.0
    assert abs(ex.sum() - 1) < 1e-14 and abs(ey.sum() - 1) < 1e-14 and abs(bz.sum() - 1) < 1e-14


@pytest.mark.parametrize("name", ["CIC", "S1", "S2"])
def test_gather_partition_of_unity(name):
    """test/scattergather.cpp:130-166, 218-269, 344-416: gather of 1 gives 1."""
    order = ORDERS[name]
    ones = np.ones((21, 21))
    R = np.array([GOLD["particle"]])
    for bug in (False, True):
        assert abs(O.gather_edge(order, 0, ones, (0, 0), (1, 1), R, bug)[0] - 1) < 1e-14
        assert abs(O.gather_edge(order, 1, ones, (0, 0), (1, 1), R, bug)[0] - 1) < 1e-14
        assert abs(O.gather_cell(order, ones, (0, 0), (1, 1), R, bug)[0] - 1) < 1e-14


@pytest.mark.parametrize("order", [0, 1, 2])
def test_gather_reproduces_linear_fields(order):
    """Edge field Ex(i,j) at (i+1/2, j), Ey at (i, j+1/2), Hz at (i+1/2, j+1/2):
    a linear field is interpolated exactly (pins the staggering and indices).
    The reference's IntS2 index at ScatterGatherImpl.hh:467 breaks this for a
    field varying in x -- the bug switch shows it."""
    rng = np.random.default_rng(0)
    jj, ii = np.mgrid[0:24, 0:24].astype(float)
    a, b, c = 0.3, 1.7, -0.9
    fex = a + b * (ii + 0.5) + c * jj
    fey = a + b * ii + c * (jj + 0.5)
    fhz = a + b * (ii + 0.5) + c * (jj + 0.5)
    R = rng.uniform(6, 16, (50, 2))
    lin = a + b * R[:, 0] + c * R[:, 1]
    gx = O.gather_edge(order, 0, fex, (0, 0), (1, 1), R, s2bug=False)
    gy = O.gather_edge(order, 1, fey, (0, 0), (1, 1), R, s2bug=False)
    gh = O.gather_cell(order, fhz, (0, 0), (1, 1), R, s2bug=False)
    assert np.abs(gx - lin).max() < 1e-12
    assert np.abs(gy - lin).max() < 1e-12
    assert np.abs(gh - lin).max() < 1e-12
    if order == 2:
        gb = O.gather_edge(order, 0, fex, (0, 0), (1, 1), R, s2bug=True)
        assert np.abs(gb - lin).max() > 1e-6


def test_q_rsqrt_error_bound():
    """utils.hh:491-504: two Newton steps from the 0x5f3759df guess; the
    Kadlec second step keeps the relative error below ~1e-6."""
    xs = np.concatenate([np.linspace(1.0, 10.0, 500), np.logspace(0, 12, 2000)])
    err = np.array([abs(O.q_rsqrt(float(x)) * np.sqrt(np.float32(x)) - 1.0) for x in xs])
    assert err.max() < 2e-6
    assert err.max() > 1e-9      # it is an approximation, not 1/sqrt


def brute_crossings(o, e, dual):
    """Fractions where floor(x - s) or floor(y - s) changes along o -> e."""
    sh = 0.5 if dual else 0.0
    out = []
    for d in range(2):
        a, b = o[d] - sh, e[d] - sh
        lo, hi = sorted((a, b))
        for line in range(int(np.floor(lo)) + 1, int(np.floor(hi)) + 1):
            out.append((line - a) / (b - a))
    return np.sort(np.array(out))


@pytest.mark.parametrize("dual", [False, True])
def test_crossings_brute_force(dual):
    rng = np.random.default_rng(7)
    for _ in range(400):
        o = rng.uniform(2, 8, 2)
        e = o + rng.uniform(-2.5, 2.5, 2)
        frac, sub = O.crossings(o, e, (1.0, 1.0), (0.0, 0.0), dual)
        ref = brute_crossings(o, e, dual)
        assert frac[0] == 0.0 and frac[-1] == 1.0
        assert len(frac) == len(ref) + 2
        assert np.abs(frac[1:-1] - ref).max(initial=0) < 1e-12
        # sub-steps lie on the segment
        t = frac
        pts = o[None, :] + t[:, None] * (e - o)[None, :]
        assert np.abs(sub - pts).max() < 1e-12


def continuity(order, oldR, R, q, dx, dt):
    """rho(R) - rho(oldR) + dt (dJx/dx + dJy/dy) on vertices."""
    box = (-6, -6, 40, 40)
    org = (0.0, 0.0)
    Jx = O.deposit(order, 0, box, org, dx, dt, oldR, R, q)
    Jy = O.deposit(order, 1, box, org, dx, dt, oldR, R, q)
    r1 = O.charge_density(order, box, org, dx, R, q)
    r0 = O.charge_density(order, box, org, dx, oldR, q)
    div = np.zeros_like(r1)
    div[:, 1:] += (Jx[:, 1:] - Jx[:, :-1]) / dx[0]
    div[1:, :] += (Jy[1:, :] - Jy[:-1, :]) / dx[1]
    return (r1 - r0) + dt * div, np.abs(r1).max()


@pytest.mark.parametrize("order", [0, 1, 2])
def test_deposit_continuity_reference_trajectories(order):
    """Trajectories of test/calcCurrentS1.cpp:63-66 and test/calcCurrent.cpp:66-73;
    the deposit must satisfy the discrete continuity equation to round-off."""
    cases = [((0.164, 0.36), (0.128, 0.315), (0.04, 0.05)),
             ((3.3, 3.2), (4.2, 3.3), (1.0, 0.5))]
    for o, e, dx in cases:
        oldR = np.array([o])
        R = np.array([e])
        res, scale = continuity(order, oldR, R, np.array([1.0]), dx, 1.0)
        assert np.abs(res).max() < 1e-12 * scale


@pytest.mark.parametrize("order", [0, 1, 2])
def test_deposit_continuity_random(order):
    rng = np.random.default_rng(order + 11)
    n = 300
    oldR = rng.uniform(8, 20, (n, 2)) * np.array([0.7, 1.3])
    step = rng.uniform(-0.99, 0.99, (n, 2)) * np.array([0.7, 1.3])
    R = oldR + step
    q = rng.uniform(-1, 1, n)
    res, scale = continuity(order, oldR, R, q, (0.7, 1.3), 0.37)
    assert np.abs(res).max() < 1e-12 * max(scale, 1.0)


def test_bend_geometry_analytic():
    """Cut-edge fractions on the arc agree with circle/line intersections;
    masks and areas are consistent (Bend.cpp:126-353, 427-729)."""
    geo = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=0.1704)
    b = O.Bend(Nx=1100, Ny=120, **geo)
    m = b.mask()
    ij, data = b.bcells()
    assert m.sum() > 0 and len(ij) > 0
    lam0, lam1, area, ln0, ln1 = data.T
    assert (area > 0).all() and (area <= 1 + 1e-10).all()
    assert (lam0 >= 0).all() and (lam0 <= 1).all() and (lam1 >= 0).all() and (lam1 <= 1).all()
    # limitArea: area >= every lambda / 2
    assert (area >= np.maximum.reduce([lam0, lam1, ln0, ln1]) / 2 - 1e-15).all()
    # no cut cell is also a bulk cell
    assert not m[ij[:, 1], ij[:, 0]].any()
    # arc cells: an x-edge crossing the outer circle |X - C| = R + w/2
    hx, hy = b.hr
    Rout = b.R + 0.5 * geo["width"]
    checked = 0
    for (i, j), row in zip(ij, data):
        x0 = b.origin[0] + i * hx
        y0 = b.origin[1] + j * hy
        inside0 = b.is_inside(x0, y0)
        inside1 = b.is_inside(x0 + hx, y0)
        if inside0 and not inside1 and x0 > geo["lenBefore"] + 1e-3:
            dx0, dy0 = x0 - b.center[0], y0 - b.center[1]
            if b.phi != 0 and dy0 * np.sin(b.phi) <= dx0 * np.cos(b.phi):
                continue   # beyond the arc
            # solve |(dx0 + t hx, dy0)| = Rout for t in [0, 1]
            disc = Rout ** 2 - dy0 ** 2
            if disc <= 0:
                continue
            t = (np.sqrt(disc) - dx0) / hx
            if 0 <= t <= 1:
                assert abs(row[0] - t) < 1e-9
                checked += 1
    assert checked > 0


def test_external_field_strip():
    """Bend.cpp:865-907: Bz inside the chord strip of the arc, zero outside."""
    geo = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=0.1704)
    b = O.Bend(Nx=1100, Ny=120, **geo)
    R = np.array([[0.04, 0.0064], [0.06, 0.0064], [0.26, 0.0064]])   # chord 2R sin(phi/2) = 0.174 m
    bf = b.external_field(R)
    assert bf[0] == 0.0 and bf[1] == geo["bz"] and bf[2] == 0.0
    # radius from Ekin and Bz (Bend.cpp:76-78, electron mass 0.511 MeV)
    ek = geo["ekinMeV"] * 1e6
    assert abs(b.R - np.sqrt(ek * ek + 2 * ek * 0.511e6) / (299792458.0 * geo["bz"])) < 1e-12 * b.R


def test_two_grid_step_window_follows_bunch():
    """The reference loop on the CPU: the window moves one column per step at
    dt = hx/c and stays around the bunch; fields stay finite."""
    geo = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=0.1704)
    from csrst_amd.bunch import gaussian_bunch
    s = O.Sim(1100, 120, order=1, statModulus=1, **geo)
    R, P, Q = gaussian_bunch(500, (0.02, 0.0064), 1e-3, 0.5e-3, 100.0, 1.0, seed=2)
    s.init(R, P, Q)
    p0 = s.info()["pos"]
    s.step(10)
    p1 = s.info()["pos"]
    assert 8 <= p1[0] - p0[0] <= 10
    assert np.isfinite(s.field("mE")).all() and np.abs(s.field("mJ")).max() > 0


def test_split_field_step_moves_div_e():
    """The reference's split step (TETMSolver.hh:110-266) updates Ex with H
    after the first half step and Ey with H after the second: the divergence
    of the curl terms does not cancel, so with J = 0 div E changes by far more
    than round-off in bulk cells.  A standard Yee step (both E components from
    one H) would leave it unchanged; this pins the reference's scheme."""
    from scipy.ndimage import binary_erosion
    geo = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=0.1704)
    nx, ny = 1100, 120
    s = O.Sim(Nx=nx, Ny=ny, order=0, statModulus=0, **geo)
    b = O.Bend(Nx=nx, Ny=ny, **geo)
    hx, hy = b.hr
    rng = np.random.default_rng(5)
    mask = b.mask().astype(bool)
    yy, xx = np.mgrid[0:ny, 0:nx]
    E = np.zeros((2, ny + 2, nx + 2))
    H = np.zeros((2, ny + 2, nx + 2))
    for arr, k in ((E, 0), (E, 1), (H, 0)):
        f = np.zeros((ny, nx))
        for _ in range(4):
            kx, ky = rng.uniform(0.02, 0.2, 2)
            f += rng.normal() * np.sin(kx * xx + rng.uniform(0, 6)) * np.cos(ky * yy + rng.uniform(0, 6))
        arr[k, 1:-1, 1:-1] = np.where(mask, f * 1e5, 0.0)
    H[0, 1:-1, 1:-1] *= 1e-3
    H[1, 1:-1, 1:-1] = H[0, 1:-1, 1:-1]
    s.set_field("E", E)
    s.set_field("H", H)

    def div(F):
        d = np.zeros((ny, nx))
        d[:, 1:] += (F[0, 1:-1, 2:-1] - F[0, 1:-1, 1:-2]) / hx
        d[1:, :] += (F[1, 2:-1, 1:-1] - F[1, 1:-2, 1:-1]) / hy
        return d

    inner = binary_erosion(mask, iterations=4)
    d0 = div(s.field("E"))
    s.field_steps(1)
    E1 = s.field("E")
    change = np.abs(div(E1) - d0)[inner].max() / (np.abs(E1[0]).max() / hx)
    print("div E change in one field step: %.2e of max|E|/hx" % change)
    assert change > 1e-5


def test_line_fields_projection_and_geometry():
    """SaveEMFieldsCommand (SaveEMFields.cpp:34-160) on uniform fields: every
    sample is the projection of the field onto the mean momentum and its
    normal (the gathers reproduce constants), the samples sit on the global
    grid's columns on the line through the centroid along the mean momentum,
    s is the signed distance along it, and the columns run from
    floor(1.5 gmin - 0.5 gmax) to 2 gmax - gmin of the bunch's cell range."""
    geo = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=0.1704)
    s = O.Sim(1100, 120, order=1, statModulus=1, **geo)
    th = np.radians(5.0)
    rng = np.random.default_rng(3)
    n = 500
    R = np.stack([0.03 + rng.normal(0, 1e-3, n), 0.006 + rng.normal(0, 0.5e-3, n)], 1)
    P = np.stack([np.full(n, 99.9 * np.cos(th)), np.full(n, 99.9 * np.sin(th))], 1) + rng.normal(0, 0.05, (n, 2))
    s.init(R, P, np.full(n, -1e-12))
    shape = (2, 122, 1102)
    for name, vals in (("E", (3.0, -2.0)), ("H", (0.25, 0.5)), ("J", (7.0, 1.0))):
        f = np.zeros(shape)
        f[0], f[1] = vals
        s.set_field(name, f)
    rows = s.line_fields()
    Rs, Ps, _ = s.particles()
    Rm, Pm = Rs.mean(0), Ps.mean(0)
    c, sn = Pm / np.hypot(*Pm)
    b = O.Bend(Nx=1100, Ny=120, **geo)
    assert np.allclose(rows[:, 1], c * 3.0 + sn * -2.0, rtol=1e-14, atol=0)
    assert np.allclose(rows[:, 2], -sn * 3.0 + c * -2.0, rtol=1e-14, atol=0)
    assert np.allclose(rows[:, 3], 0.75, rtol=1e-14, atol=0)
    assert np.allclose(rows[:, 4], c * 7.0 + sn * 1.0, rtol=1e-14, atol=0)
    x, y = rows[:, 5], rows[:, 6]
    dx, dy = x - Rm[0], y - Rm[1]
    assert np.abs(dx * Pm[1] - dy * Pm[0]).max() < 1e-12 * np.hypot(*Pm) * np.abs(dx).max()
    assert np.allclose(rows[:, 0], np.sign(dx) * np.hypot(dx, dy), rtol=1e-12, atol=1e-15)
    cols = (x - b.origin[0]) / b.hr[0]
    assert np.abs(cols - np.round(cols)).max() < 1e-9
    gmin = np.floor((Rs[:, 0].min() - b.origin[0]) / b.hr[0])
    gmax = np.floor((Rs[:, 0].max() - b.origin[0]) / b.hr[0])
    assert round(cols[0]) == np.floor(1.5 * gmin - 0.5 * gmax) and round(cols[-1]) == 2 * gmax - gmin
    assert np.all(np.diff(np.round(cols)) == 1)
