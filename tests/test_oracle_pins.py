"""Mathematical pins of the oracle's field sweeps, conformal wall cells, push
and two-grid bookkeeping -- checks that do not retype the oracle's stencils:

* Yee sweeps (TETMSolver.hh:142-249): a discrete Fourier mode of the split
  step H/2, Ex, H/2, Ey is an eigenvector of its 4x4 amplification matrix,
  assembled here from the difference symbols 2i sin(k h / 2) of the staggered
  grid.  The oracle's fields must follow lambda^n of that eigenvector in the
  interior; along an axis at c dt = h the phase advance per step is exactly
  k h (the dispersion-free magic time step of the 1-D Yee scheme).
* Conformal cells (TETMSolver.cpp:87-127, SCBoundaryCondition.hh:392-431):
  walls that pass through mesh vertices must reduce to the perfectly
  conducting masked update, so the same exact modes stay exact beside them.
* Ampere's law with a uniform current: E changes by -J dt / eps0.
* Push (PartBunch.cpp:366-383): with no self-field the rotation of one push
  is the scaled rotation u' = r(a) R(-theta(a)) u with cos(theta) r =
  1/(1+a^2), sin(theta) r = a (2+a^2)/(1+a^2), a from the halved external
  field; a particle in the bend moves on the circle of radius
  |u| / (c |q/m| Bz) around the orbit centre of its initial state.
* Exchange and moveMesh (Communicator.hh:575-876, main.cpp:562-652): with no
  charge, the window and the global grid integrate the same equations, so the
  window must equal the global grid under it after every step while it moves
  (this pins the global-to-window ghost layers and the refill of the layers
  a move exposes; the window-to-global write-back changes nothing without
  charge).

The check_* functions are also run against deliberately broken builds of the
oracle (tests/test_oracle_mutants.py): each must fail there.
"""
import numpy as np
import pytest

from oracle import oracle as O

C_LIGHT, MU_0, EPS_0 = 299792458.0, 1.256637061e-06, 8.854187817e-12   # Physics.cpp:41-51
M_E_GEV = 0.51099906e-03
Z0 = np.sqrt(MU_0 / EPS_0)

# a straight chamber (no bend field): walls on mesh vertices x = 0, L and y = 0, width
STRAIGHT = dict(lenBefore=0.01, lenAfter=0.01, width=0.02, phiDeg=0.0, ekinMeV=50.0, bz=0.0)
NSTRAIGHT = 202                       # 200 cells of 100 um each way


def amplification(kx, ky, hx, hy, dt):
    """One split step (H/2, Ex, H/2, Ey) acting on the complex amplitudes
    (Ex, Ey, Hzx, Hzy) of the mode exp(i k.r), with Ex at ((i+1/2)hx, j hy),
    Ey at (i hx, (j+1/2) hy) and Hz at the cell centre."""
    ex, ey = C_LIGHT * dt / hx, C_LIGHT * dt / hy
    gx, gy = 2j * np.sin(kx * hx / 2), 2j * np.sin(ky * hy / 2)
    Hh = np.eye(4, dtype=complex)
    Hh[2, 1] = -0.5 * ex / Z0 * gx       # dHzx/dt = -dEy/dx / mu0
    Hh[3, 0] = +0.5 * ey / Z0 * gy       # dHzy/dt = +dEx/dy / mu0
    Ex = np.eye(4, dtype=complex)
    Ex[0, 2] = Ex[0, 3] = ey * Z0 * gy   # dEx/dt = +dHz/dy / eps0
    Ey = np.eye(4, dtype=complex)
    Ey[1, 2] = Ey[1, 3] = -ex * Z0 * gx  # dEy/dt = -dHz/dx / eps0
    return Ey @ Hh @ Ex @ Hh


def mode_fields(sim, kx, ky, v):
    """The complex mode with amplitudes v on the global grid (guards incl.),
    positions relative to the mesh origin."""
    hx, hy = O.Bend(Nx=sim.p.Nx, Ny=sim.p.Ny, **STRAIGHT).hr
    n, m = sim.p.Nx, sim.p.Ny
    jj, ii = np.mgrid[-1:m + 1, -1:n + 1].astype(float)
    ph_ex = np.exp(1j * (kx * (ii + 0.5) * hx + ky * jj * hy))
    ph_ey = np.exp(1j * (kx * ii * hx + ky * (jj + 0.5) * hy))
    ph_hz = np.exp(1j * (kx * (ii + 0.5) * hx + ky * (jj + 0.5) * hy))
    E = np.stack([v[0] * ph_ex, v[1] * ph_ey])
    H = np.stack([v[2] * ph_hz, v[3] * ph_hz])
    return E, H


def straight_sim():
    return O.Sim(NSTRAIGHT, NSTRAIGHT, order=0, statModulus=1, **STRAIGHT)


def run_mode(kx, ky, nsteps, region):
    """Evolve the nontrivial eigenmode of (kx, ky) in the oracle and return
    (max deviation from lambda^n v in `region`, max amplitude, lambda)."""
    s = straight_sim()
    hx, hy = O.Bend(Nx=NSTRAIGHT, Ny=NSTRAIGHT, **STRAIGHT).hr
    dt = hx / C_LIGHT                                    # main.cpp:187
    M = amplification(kx, ky, hx, hy, dt)
    w, V = np.linalg.eig(M)
    k = int(np.argmax(np.abs(np.angle(w))))
    lam, v = w[k], V[:, k] / np.abs(V[:, k]).max()
    E, H = mode_fields(s, kx, ky, v)
    s.set_field("E", E.real)
    s.set_field("H", H.real)
    s.field_steps(nsteps)
    E1, H1 = mode_fields(s, kx, ky, v * lam ** nsteps)
    (i0, i1), (j0, j1) = region
    sl = (slice(None), slice(j0 + 1, j1 + 2), slice(i0 + 1, i1 + 2))
    got_e, got_h = s.field("E")[sl], s.field("H")[sl]
    scale = max(np.abs(E.real).max(), np.abs(H.real).max() * Z0)
    dev = max(np.abs(got_e - E1.real[sl]).max(), np.abs(got_h - H1.real[sl]).max() * Z0)
    return dev, scale, lam


INTERIOR = ((70, 130), (70, 130))    # at least 69 cells from every wall


def check_free_space_modes():
    h = 1e-4
    for kx, ky, n in ((2 * np.pi / (40 * h), 0.0, 20), (0.0, 2 * np.pi / (25 * h), 20),
                      (2 * np.pi / (30 * h), 2 * np.pi / (45 * h), 15), (2 * np.pi / (7 * h), -2 * np.pi / (9 * h), 15)):
        dev, scale, lam = run_mode(kx, ky, n, INTERIOR)
        assert dev < 1e-11 * scale, (kx, ky, dev / scale)
        assert abs(abs(lam) - 1.0) < 1e-12
        if ky == 0.0 or kx == 0.0:
            # c dt = h along an axis: the phase advance is exactly k h
            assert abs(abs(np.angle(lam)) - (kx + abs(ky)) * h) < 1e-8


def check_walls_reduce_to_pec():
    """Walls through vertices: x = 0 (cells i = 1) and y = 0 (cells j = 1)
    are conformal cells with lambda = 1 or 0 and area 1.  The modes that meet
    the perfect-conductor conditions there (Ey = 0 on x = 0, Ex = 0 on y = 0)
    stay exact up to the wall, away from the corners."""
    h = 1e-4
    kx = np.pi / (20 * h)          # Ey ~ sin(kx x): zero on the wall x = 0 (vertex i = 1)
    ky = np.pi / (16 * h)
    for (a, b), region in (((kx, 0.0), ((1, 30), (80, 120))), ((0.0, ky), ((80, 120), (1, 30)))):
        s = straight_sim()
        hx, hy = O.Bend(Nx=NSTRAIGHT, Ny=NSTRAIGHT, **STRAIGHT).hr
        M = amplification(a, b, hx, hy, hx / C_LIGHT)
        w, V = np.linalg.eig(M)
        k = int(np.argmax(np.abs(np.angle(w))))
        lam, v = w[k], V[:, k] / np.abs(V[:, k]).max()
        # standing wave: sin along the wall normal, with the wall vertex (index 1) at x or y = 0
        n, m = NSTRAIGHT, NSTRAIGHT
        jj, ii = np.mgrid[-1:m + 1, -1:n + 1].astype(float)

        def standing(vv):
            if b == 0.0:     # Ey(i) at x = (i - 1) h, Hz(i) at (i - 1/2) h
                ey = vv[1] * np.sin(a * (ii - 1) * hx)
                hz = -1j * np.cos(a * (ii - 0.5) * hx)
                E = np.stack([0 * ii, ey])
                H = np.stack([vv[2] * hz, vv[3] * hz])
            else:            # Ex(j) at y = (j - 1) h, Hz(j) at (j - 1/2) h
                ex = vv[0] * np.sin(b * (jj - 1) * hy)
                hz = -1j * np.cos(b * (jj - 0.5) * hy)
                E = np.stack([ex, 0 * jj])
                H = np.stack([vv[2] * hz, vv[3] * hz])
            return E, H
        # the standing wave is the +k mode minus the -k mode, whose eigenvector is v with
        # the H amplitudes negated: (2i v_E sin, 2 v_H cos), here scaled by -i/2
        E0, H0 = standing(v)
        s.set_field("E", E0.real)
        s.set_field("H", H0.real)
        nsteps = 12
        s.field_steps(nsteps)
        E1, H1 = standing(v * lam ** nsteps)
        (i0, i1), (j0, j1) = region
        sl = (slice(None), slice(j0 + 1, j1 + 2), slice(i0 + 1, i1 + 2))
        scale = np.abs(E0.real).max()
        dev = max(np.abs(s.field("E")[sl] - E1.real[sl]).max(), np.abs(s.field("H")[sl] - H1.real[sl]).max() * Z0)
        assert dev < 1e-9 * scale, ((a, b), dev / scale)


# the bend turning the other way (Bz < 0): its first straight section's lower
# wall y = 0 falls inside a cell row, so its cut cells have edge fractions
# lambda = f and area f (case 12 of Bend::areaCases)
BEND_DOWN = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=-0.1704)


def check_fractional_wall():
    """A wave travelling along x, uniform in y (Ey, Hz), is an exact solution
    beside any wall parallel to x.  Beside the fractional lower wall the cut
    cells must reproduce it: their increments are lambda-weighted differences
    divided by the area, hx and hy kept apart."""
    Nx, Ny = 1100, 117            # f = 0.67: limitArea's floor of lambdaN/2 = 1/2 leaves area = f
    b = O.Bend(Nx=Nx, Ny=Ny, **BEND_DOWN)
    hx, hy = b.hr
    ij, data = b.bcells()
    jw = int(np.floor(-b.origin[1] / hy))        # cell row that holds y = 0
    row = (ij[:, 1] == jw) & (ij[:, 0] > 40) & (ij[:, 0] < 200)
    f = data[row, 1]
    assert row.sum() > 100 and np.all(np.abs(f - f[0]) < 1e-12) and 0.55 < f[0] < 0.95
    assert np.all(np.abs(data[row, 2] - f[0]) < 1e-12)               # area = f
    assert abs(hx / hy - 1) > 0.1
    s = O.Sim(Nx, Ny, order=0, statModulus=1, **BEND_DOWN)
    kx = 2 * np.pi / (37 * hx)
    M = amplification(kx, 0.0, hx, hy, hx / C_LIGHT)
    w, V = np.linalg.eig(M)
    k = int(np.argmax(np.abs(np.angle(w))))
    lam, v = w[k], V[:, k] / np.abs(V[:, k]).max()
    jj, ii = np.mgrid[-1:Ny + 1, -1:Nx + 1].astype(float)

    def wave(vv):
        ey = vv[1] * np.exp(1j * kx * ii * hx)
        hz = np.exp(1j * kx * (ii + 0.5) * hx)
        return np.stack([0 * ii, ey]), np.stack([vv[2] * hz, vv[3] * hz])
    E0, H0 = wave(v)
    s.set_field("E", E0.real)
    s.set_field("H", H0.real)
    nsteps = 20
    s.field_steps(nsteps)
    E1, H1 = wave(v * lam ** nsteps)
    sl = (slice(None), slice(jw + 1, jw + 21), slice(60 + 1, 150 + 1))
    scale = np.abs(E0.real).max()
    dev = max(np.abs(s.field("E")[sl] - E1.real[sl]).max(), np.abs(s.field("H")[sl] - H1.real[sl]).max() * Z0)
    assert dev < 1e-9 * scale, dev / scale
    # the cut cells' Hz is in the compared rows, their Ey edges too
    assert np.abs(E1.real[1, jw + 1, 61:151]).max() > 0.1 * scale


def check_uniform_current():
    """Ampere's law: with E = H = 0 and a uniform Jx, one step gives
    Ex = -Jx dt / eps0 in the bulk (c dt Z0 = dt / eps0) and leaves Ey, Hz."""
    s = straight_sim()
    hx = O.Bend(Nx=NSTRAIGHT, Ny=NSTRAIGHT, **STRAIGHT).hr[0]
    dt = hx / C_LIGHT
    J = np.zeros((2, NSTRAIGHT + 2, NSTRAIGHT + 2))
    J[0] = 3.0e4
    J[1] = -1.5e4
    s.set_field("J", J)
    s.field_steps(1)
    E, H = s.field("E"), s.field("H")
    inner = (slice(11, NSTRAIGHT - 10), slice(11, NSTRAIGHT - 10))
    # Ey sees the current after the H half steps, which a uniform Ex leaves at zero
    assert np.allclose(E[0][inner], -3.0e4 * dt / EPS_0, rtol=1e-9, atol=0)
    assert np.allclose(E[1][inner], +1.5e4 * dt / EPS_0, rtol=1e-9, atol=0)
    assert np.abs(H[:, inner[0], inner[1]]).max() < 1e-12 * 3.0e4 * dt / EPS_0 / Z0


BEND = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=0.1704)


def bend_sim(**kw):
    return O.Sim(1100, 120, order=kw.pop("order", 1), statModulus=1, **BEND, **kw)


def uncharged_bunch(center, direction_deg, gamma, n=24, spread=2e-4, seed=3):
    rng = np.random.default_rng(seed)
    th = np.radians(direction_deg)
    d = np.array([np.cos(th), np.sin(th)])
    R = np.asarray(center) + rng.uniform(-spread, spread, (n, 2))
    P = np.outer(np.full(n, np.sqrt(gamma * gamma - 1.0)), d)
    return R, P, np.zeros(n)


def scaled_rotation(a):
    """(r, theta) of the push's momentum rotation for rotation parameter a."""
    c = 1.0 / (1.0 + a * a)
    s = a * (2.0 + a * a) / (1.0 + a * a)
    return np.hypot(c, s), np.arctan2(s, c)


def check_push_rotation():
    """One push (op 0) of uncharged particles inside the bend's field strip:
    u' = r(a) R(-theta) u with a = (dt/2) c (q/m) c (Bz/2) / gamma(float)."""
    s = bend_sim()
    b = O.Bend(Nx=1100, Ny=120, **BEND)
    g = 100.0
    # a point on the centre line 5 mm into the arc, moving along its tangent
    phi0 = 0.005 / b.R
    cen = b.center
    R0 = (cen[0] + b.R * np.sin(phi0), cen[1] - b.R * np.cos(phi0))
    R, P, Q = uncharged_bunch(R0, np.degrees(phi0), g)
    s.init(R, P, Q)
    dt = s.info()["dt"]
    qm = -1e-9 / M_E_GEV                                   # main.cpp:225
    for _ in range(3):
        _, P0, _ = s.particles()
        s.substep(0)
        R1, P1, O1 = s.particles()
        bext = b.external_field(O1 + 0.5 * (R1 - O1))      # inside the strip all along the push
        assert np.all(bext == BEND["bz"])
        for k in range(len(P0)):
            u = P0[k]
            gi = O.q_rsqrt(float(1.0 + u @ u))
            a = (0.5 * dt * C_LIGHT * qm) * C_LIGHT * (0.5 * bext[k]) * gi
            r, th = scaled_rotation(a)
            rot = np.array([[np.cos(th), np.sin(th)], [-np.sin(th), np.cos(th)]])
            want = r * rot @ u
            assert np.abs(P1[k] - want).max() < 1e-13 * np.abs(u).max(), k
        for op in (2, 5, 8, 3, 6, 9, 2, 5, 8, 11, 1, 4, 7, 10, 11, 13, 14):
            s.substep(op)


def check_bend_orbit():
    """Uncharged particles entering the bend follow the uniform-field orbit:
    |u| grows by r(a) and turns by theta(a) at every push (the halved field,
    two pushes per step, a from the float 1/gamma of the current |u|), and
    the positions stay on the circle of radius |u| / (c |q/m| Bz) about the
    centre of curvature of the initial state."""
    s = bend_sim()
    b = O.Bend(Nx=1100, Ny=120, **BEND)
    g = 100.0
    R, P, Q = uncharged_bunch((BEND["lenBefore"] + 1e-3, 0.5 * BEND["width"]), 0.0, g, n=8, spread=1e-4)
    s.init(R, P, Q)
    dt = s.info()["dt"]
    qm = -1e-9 / M_E_GEV
    R0, P0, _ = s.particles()
    u0 = np.hypot(P0[:, 0], P0[:, 1])
    rg = u0 / (C_LIGHT * abs(qm) * BEND["bz"])
    # electrons in +Bz turn towards +y: the centre lies to the left of the momentum
    centre = R0 + rg[:, None] * np.stack([-P0[:, 1], P0[:, 0]], 1) / u0[:, None]
    nsteps = 300
    s.step(nsteps)
    R1, P1, _ = s.particles()
    for k in range(len(u0)):
        u, turn = u0[k], 0.0
        for _ in range(2 * nsteps):
            a = (0.5 * dt * C_LIGHT * qm) * C_LIGHT * (0.5 * BEND["bz"]) * O.q_rsqrt(float(1.0 + u * u))
            r, th = scaled_rotation(a)
            u, turn = u * r, turn - th
        assert abs(np.hypot(*P1[k]) / u - 1) < 1e-12
        ang = np.arctan2(P1[k, 1], P1[k, 0]) - np.arctan2(P0[k, 1], P0[k, 0])
        assert abs(ang - turn) < 1e-9, (ang, turn)
    assert np.degrees(turn) > 3.0                        # well inside the arc
    dist = np.hypot(*(R1 - centre).T)
    assert np.abs(dist / rg - 1).max() < 5e-6
    assert np.abs(rg / b.R - 1).max() < 1e-4            # the bend is built for this momentum (masses of A11)


def check_two_grid_invariant():
    """No charge: after every step (window moving in x and y), the window's E
    and Hz equal the global grid's under the window."""
    s = bend_sim(order=0)
    b = O.Bend(Nx=1100, Ny=120, **BEND)
    rng = np.random.default_rng(11)
    n, m = 1100, 120
    jj, ii = np.mgrid[-1:m + 1, -1:n + 1].astype(float)
    E = np.zeros((2, m + 2, n + 2))
    H = np.zeros((2, m + 2, n + 2))
    for _ in range(6):
        kx, ky = rng.uniform(-0.5, 0.5, 2)
        ph = rng.uniform(0, 2 * np.pi, 4)
        amp = rng.normal(size=4)
        E[0] += amp[0] * np.cos(kx * ii + ky * jj + ph[0])
        E[1] += amp[1] * np.cos(kx * ii + ky * jj + ph[1])
        H[0] += amp[2] * np.cos(kx * ii + ky * jj + ph[2]) / Z0
        H[1] += amp[3] * np.cos(kx * ii + ky * jj + ph[3]) / Z0
    mask = b.mask()
    inside = np.zeros((m + 2, n + 2), bool)
    inside[1:-1, 1:-1] = mask.astype(bool)
    E *= inside
    H *= inside
    s.set_field("E", E)
    s.set_field("H", H)
    R, P, Q = uncharged_bunch((0.02, 0.004), 20.0, 100.0, n=16, spread=3e-4)
    s.init(R, P, Q)
    moves = set()
    pos_prev = s.info()["pos"]
    for _ in range(40):
        s.step(1)
        inf = s.info()
        moves.add((inf["pos"][0] - pos_prev[0], inf["pos"][1] - pos_prev[1]))
        pos_prev = inf["pos"]
        (p0, p1), W, Hh = inf["pos"], inf["W"], inf["H"]
        gE, gH = s.field("E"), s.field("H")
        wE, wH = s.field("mE"), s.field("mH")
        ge = gE[:, p1 + 1:p1 + Hh + 1, p0 + 1:p0 + W + 1]
        gh = (gH[0] + gH[1])[p1 + 1:p1 + Hh + 1, p0 + 1:p0 + W + 1]
        we = wE[:, 1:Hh + 1, 1:W + 1]
        wh = (wH[0] + wH[1])[1:Hh + 1, 1:W + 1]
        scale = np.abs(ge).max()
        assert np.abs(we - ge).max() < 1e-12 * scale
        assert np.abs(wh - gh).max() * Z0 < 1e-12 * scale
    assert (1, 1) in moves and (1, 0) in moves     # the window moved along x alone and diagonally


def test_free_space_modes():
    check_free_space_modes()


def test_walls_reduce_to_pec():
    check_walls_reduce_to_pec()


def test_fractional_wall():
    check_fractional_wall()


def test_uniform_current():
    check_uniform_current()


def test_push_rotation():
    check_push_rotation()


def test_bend_orbit():
    check_bend_orbit()


def test_two_grid_invariant():
    check_two_grid_invariant()
