"""Production moving-window step on the GPU.

- field sweeps (fused H/2-Ex-H/2 row march on the ring buffer, conformal
  cut cells, merged Hz) against the CPU oracle's global-grid sweeps;
- FP32 split precision against FP64 on a moving bunch (fields and wake);
- window motion keeps the bunch inside and positions advance at ~c.
"""
import numpy as np
import pytest

from csrst_amd.mw import MovingWindowStep
from csrst_amd import gaussian_bunch
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GEOM = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=0.1704)
NX, NY = 1100, 120


def seeded_fields(o, seed=5):
    """Smooth random E/H inside the chamber, zero outside."""
    rng = np.random.default_rng(seed)
    b = O.Bend(Nx=NX, Ny=NY, **GEOM)
    mask = b.mask().astype(bool)
    E = np.zeros((2, NY + 2, NX + 2))
    H = np.zeros((2, NY + 2, NX + 2))
    yy, xx = np.mgrid[0:NY, 0:NX]
    for arr, k in ((E, 0), (E, 1), (H, 0)):
        f = np.zeros((NY, NX))
        for _ in range(4):
            kx, ky = rng.uniform(0.02, 0.2, 2)
            f += rng.normal() * np.sin(kx * xx + rng.uniform(0, 6)) * np.cos(ky * yy + rng.uniform(0, 6))
        arr[k, 1:-1, 1:-1] = np.where(mask, f * 1e5, 0.0)
    H[0, 1:-1, 1:-1] *= 1e-3
    H[1, 1:-1, 1:-1] = H[0, 1:-1, 1:-1]
    return E, H


@pytest.mark.parametrize("fp32,tol", [(False, 5e-12), (True, 3e-5)])
def test_field_sweeps_match_oracle(gpu, fp32, tol):
    o = O.Sim(NX, NY, order=0, statModulus=0, **GEOM)
    b = O.Bend(Nx=NX, Ny=NY, **GEOM)
    E, H = seeded_fields(o)
    o.set_field("E", E)
    o.set_field("H", H)
    m = MovingWindowStep(NX, NY, b.hr[0], b.hr[1], order=0, fp32=fp32, sortEvery=0, statModulus=0, **GEOM)
    m.place(0, 0)
    F = np.zeros((6, NY + 2, NX + 2))
    F[0], F[1], F[2], F[3] = E[0], E[1], H[0], H[1]
    m.set_fields(F)
    for k in range(1, 21):
        o.field_steps(1)
        m.step(1)
        if k in (1, 5, 20):
            g = m.fields()
            Eo, Ho = o.field("E"), o.field("H")
            scaleE = np.abs(Eo).max()
            scaleH = np.abs(Ho[0] + Ho[1]).max()
            dE = max(np.abs(g[0] - Eo[0]).max(), np.abs(g[1] - Eo[1]).max()) / scaleE
            dH = np.abs((g[2] + g[3]) - (Ho[0] + Ho[1])).max() / scaleH
            assert dE < tol * k and dH < tol * k, "step %d: rel err E %.3e H %.3e" % (k, dE, dH)


def run_bunch(fp32, steps, n=200000, order=2, seed=11):
    R, P, Q = gaussian_bunch(n, (0.03, 0.0064), 0.4e-3, 0.3e-3, 100.0, 1.0, seed=seed, energy_spread=1e-3)
    b = O.Bend(Nx=NX, Ny=NY, **GEOM)
    m = MovingWindowStep(512, 100, b.hr[0], b.hr[1], order=order, fp32=fp32, sortEvery=10, statModulus=1, **GEOM)
    m.init_bunch(R, P, Q)
    m.pos0 = m.info()["pos"]
    m.step(steps)
    return m


def test_fp32_split_tracks_fp64(gpu):
    steps = 60
    m64 = run_bunch(False, steps)
    m32 = run_bunch(True, steps)
    i64, i32 = m64.info(), m32.info()
    assert i64["pos"] == i32["pos"], "window moves diverged"
    assert i64["pos"][0] - m64.pos0[0] > steps // 2, "window did not follow the bunch"
    w64, w32 = m64.wake(), m32.wake()
    err = np.abs(w32 - w64).max() / np.abs(w64).max()
    assert err < 1e-4, "FP32 wake rel. error %.3e" % err
    f64, f32 = m64.fields(), m32.fields()
    scale = np.abs(f64[0:2]).max()
    assert np.abs(f32[0:2] - f64[0:2]).max() / scale < 1e-4
    # particle order follows the cell sort of each run, so compare the
    # distributions: centroid, rms size and mean momentum
    R64, P64 = m64.particles()
    R32, P32 = m32.particles()
    h = b_h()
    assert np.abs(R32.mean(0) - R64.mean(0)).max() < 1e-4 * h
    assert np.abs(R32.std(0) - R64.std(0)).max() < 1e-3 * h
    assert np.abs(P32.mean(0) - P64.mean(0)).max() < 1e-6 * np.abs(P64.mean(0)).max()
    xs32, xs64 = np.sort(R32[:, 0]), np.sort(R64[:, 0])
    assert np.abs(xs32 - xs64).max() < 1e-2 * h


def b_h():
    return O.Bend(Nx=NX, Ny=NY, **GEOM).hr[0]


def test_bunch_advances_at_c(gpu):
    m = run_bunch(True, 30, n=20000)
    R, P = m.particles()
    info = m.info()
    dt = info["dt"]
    # the bunch centroid moved ~ 30 c dt along the beam direction
    R0, _, _ = gaussian_bunch(20000, (0.03, 0.0064), 0.4e-3, 0.3e-3, 100.0, 1.0, seed=11, energy_spread=1e-3)
    adv = np.mean(R[:, 0] - R0[:, 0])
    assert abs(adv / (30 * 299792458.0 * dt) - 1.0) < 1e-3


@pytest.mark.parametrize("order,fp32,tol,angle", [(0, False, 1e-12, 0.0), (1, False, 1e-12, 0.0),
                                                  (2, False, 1e-12, 0.0), (1, True, 2e-6, 0.0), (2, True, 2e-6, 0.0),
                                                  (1, False, 1e-12, 25.0), (2, False, 1e-12, 25.0),
                                                  (2, True, 2e-6, 25.0)])
def test_first_step_ex_is_deposit(gpu, order, fp32, tol, angle):
    """From zero fields, after one step Ex = -c dt Z0 Jx exactly (the first
    half step of H is zero, the second does not feed back into Ex), with Jx
    the oracle deposit of the first push R0 -> R0 + 2 (0.5 c dt) P/gamma.
    At 25 degrees the particles also cross cell rows: three segments."""
    b = O.Bend(Nx=NX, Ny=NY, **GEOM)
    hx, hy = b.hr
    n = 3000
    rng = np.random.default_rng(5)
    R = np.stack([0.02 + rng.normal(0, 0.3e-3, n), 0.0064 + rng.normal(0, 0.4e-3, n)], 1)
    th = np.radians(angle)
    P = np.stack([np.full(n, 99.995 * np.cos(th)), np.full(n, 99.995 * np.sin(th))], 1)
    Q = np.full(n, -1e-22)
    m = MovingWindowStep(256, 64, hx, hy, order=order, fp32=fp32, sortEvery=0, statModulus=0, **GEOM)
    m.init_bunch(R, P, Q)
    inf = m.info()
    wx0, wy0 = inf["pos"]
    dt = inf["dt"]
    R0, P0 = m.particles()
    g = np.array([O.q_rsqrt(float(1.0 + p @ p)) for p in P0])
    half = 0.5 * dt * 299792458.0
    R2 = R0 + 2 * half * P0 * g[:, None]
    Jx = O.deposit(order, 0, (wx0, wy0, m.W, m.H), tuple(b.origin), (hx, hy), dt, R0, R2, Q)
    m.step(1)
    sx = m.info()["pos"][0] - wx0
    assert m.info()["pos"][1] == wy0 and sx in (0, 1)
    F = m.fields()
    Ex = F[0][1:-1, 1:-1][:, :m.W - sx]       # new window, columns shared with the old one
    # bulk cells use dt/eps0 in the form c dt Z0 (TETMSolver.hh:125); the
    # two differ by sqrt(c^2 mu0 eps0) = 1 - 2.1e-10 with Physics.cpp's constants
    mu0, eps0 = 1.256637061e-06, 8.854187817e-12
    dteps = 299792458.0 * dt * np.sqrt(mu0 / eps0)
    expect = -dteps * Jx[:, sx:]
    scale = np.abs(expect).max()
    d = np.abs(Ex - expect)
    j, i = np.unravel_index(np.argmax(d), d.shape)
    print("order %d fp32 %s angle %g: max |Ex - expect| / max |expect| = %.2e" % (order, fp32, angle, d.max() / scale))
    assert d.max() < tol * scale, "max |Ex - expect| = %.3e of %.3e at (%d, %d): %r vs %r" % (
        d.max(), scale, i, j, Ex[j, i], expect[j, i])


@pytest.mark.parametrize("fp32", [False, True])
def test_pml_absorbs_outgoing_pulse(gpu, fp32):
    """A TEM pulse (Ey = Z0 Hz, travelling +x) between the walls of the
    straight chamber section runs into the window's leading end.  Without a
    PML the end reflects it (the guard column is a conductor); with 16 PML
    columns the wave coming back is at the percent level or below.  Measured
    as the peak |Ey| in the band the reflected pulse then crosses."""
    hy = GEOM["width"] / 50                    # walls on grid lines: the TEM mode is exact
    hx = 0.8 * hy
    W, H = 200, 56
    Z0 = np.sqrt(1.256637061e-06 / 8.854187817e-12)
    back = {}
    for pml in (0, 16):
        m = MovingWindowStep(W, H, hx, hy, order=0, fp32=fp32, sortEvery=0, statModulus=0, pml=pml, **GEOM)
        m.place(10, -2)                       # rows -2..53 hold the chamber (cell rows 1..50)
        F = np.zeros((6, H + 2, W + 2))
        i = np.arange(W)
        g = lambda x: np.exp(-0.5 * ((x - 60.0) / 6.0) ** 2)
        inside = slice(1 + 3, 1 + 3 + 50)      # logical rows 3..52 = global rows 1..50
        F[1][inside, 1:-1] = g(i)[None, :]
        F[2][inside, 1:-1] = 0.5 * g(i + 0.5)[None, :] / Z0
        F[3][inside, 1:-1] = F[2][inside, 1:-1]
        m.set_fields(F)
        m.step(200)                            # reflected pulse (if any) centred near column 138
        Ey = m.fields()[1][inside, 1:-1]
        back[pml] = np.abs(Ey[:, 110:181]).max()
        print("pml %d: peak |Ey| coming back %.3e" % (pml, back[pml]))
        m.close()
    assert back[0] > 0.5
    assert back[16] < 5e-3
