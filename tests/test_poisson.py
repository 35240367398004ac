"""Poisson initialisation: the oracle's restatement of PoissonSolver
(PoissonSolver.cpp:85-201, 441-507, 678-766) pinned on the CPU.

- the direct solve returns a manufactured potential exactly, with the operator
  assembled here independently as a sparse matrix from cutoffStencil's rules;
- the three-solve field start reproduces the closed-form field of a line
  charge between two grounded plates (the straight chamber section) in the
  boosted frame, which fixes the gamma scaling of hx, the 1/(eps0 gamma)
  coupling, the charge normalisation and the fillFields factors.
No reference test covers the Poisson start; these checks are the pin.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O

GEOM = dict(lenBefore=0.05, lenAfter=0.01, width=0.0128, phiDeg=10.0, ekinMeV=50.59, bz=0.1704)
EPS0, MU0, C = 8.854187817e-12, 1.256637061e-06, 299792458.0


def assemble(hx, hy, flag, active):
    """cutoffStencil rows (PoissonSolver.cpp:678-766) as a sparse matrix over all
    nodes: stencil rows for active nodes, identity rows elsewhere."""
    ny, nx = flag.shape
    EW, NS = -1 / hx ** 2, -1 / hy ** 2
    rows, cols, vals = [], [], []
    for j in range(ny):
        for i in range(nx):
            g = j * nx + i
            if not active[j, i]:
                rows.append(g); cols.append(g); vals.append(1.0)
                continue
            for (ii, jj, w) in ((i, j - 1, NS), (i - 1, j, EW), (i + 1, j, EW), (i, j + 1, NS)):
                if 0 <= ii < nx and 0 <= jj < ny and flag[jj, ii]:
                    rows.append(g); cols.append(jj * nx + ii); vals.append(w)
            rows.append(g); cols.append(g); vals.append(-2 * (EW + NS))
    return sp.csr_matrix((vals, (rows, cols)), shape=(nx * ny, nx * ny))


@pytest.mark.parametrize("ratio", [60.0, 1.0, 0.2])
def test_direct_solve_manufactured(ratio):
    rng = np.random.default_rng(7)
    ny, nx = 37, 53
    flag = np.zeros((ny, nx), np.uint8)
    flag[2:34, 1:50] = 1
    flag[12:17, 20:31] = 0                   # an interior hole
    active = flag.copy()
    active[:, 41:] = 0                       # identity rows past maxX, still flagged
    hx, hy = ratio * 1e-4, 1e-4
    A = assemble(hx, hy, flag, active)
    phi = rng.normal(size=(ny, nx))
    b = (A @ phi.ravel()).reshape(ny, nx)
    x = O.poisson_solve(hx, hy, flag, active, b)
    assert np.abs(x - phi).max() < 1e-11 * np.abs(phi).max()
    # identity rows return b exactly
    assert np.array_equal(x[active == 0], b[active == 0])


def test_cutoff_classes():
    b = O.Bend(Nx=1100, Ny=120, **GEOM)
    flag, active, maxX = b.poisson_classes()
    assert maxX == int(GEOM["lenBefore"] / b.hr[0])
    assert not (active & ~flag).any()
    assert not active[:, maxX:].any()
    assert flag[:, maxX:].any()              # the chamber continues past the straight section
    # the inside mask is part of the flag; grid edges are outside
    assert not (b.mask() & ~flag).any()
    assert not flag[:, 0].any() and not flag[0].any()


def line_charge_phi(Q, d, xp, y, x0p, y0):
    """Potential of a line charge Q (per unit length) at (x0p, y0) between
    grounded plates y = 0 and y = d (method of images, summed in closed form)."""
    ch = np.cosh(np.pi * (xp - x0p) / d)
    return Q / (4 * np.pi * EPS0) * np.log((ch - np.cos(np.pi * (y + y0) / d)) / (ch - np.cos(np.pi * (y - y0) / d)))


def test_field_start_line_charge_closed_form():
    Nx, Ny = 1100, 120
    b = O.Bend(Nx=Nx, Ny=Ny, **GEOM)
    hx, hy = b.hr
    ox, oy = b.origin
    flag, active, _ = b.poisson_classes()
    inside = b.mask().astype(bool)
    I0, J0, q, g = 117, 26, 1e-12, 1.5
    R = np.array([[ox + I0 * hx, oy + J0 * hy]])       # on a node: CIC puts all charge there
    s = O.Sim(Nx, Ny, order=0, statModulus=1000, **GEOM)
    s.init(R, np.zeros((1, 2)), np.array([q]))
    s.poisson(g)
    E, H = s.field("E"), s.field("H")
    Ex, Ey, Hz = E[0][1:-1, 1:-1], E[1][1:-1, 1:-1], H[0][1:-1, 1:-1]
    assert np.array_equal(H[0], H[1])
    col = np.nonzero(flag[:, I0])[0]
    ya = oy + (col.min() - 1) * hy                      # first Dirichlet rows on either side
    d = oy + (col.max() + 1) * hy - ya
    X = (ox + np.arange(Nx) * hx)[None, :] * g
    Y = (oy + np.arange(Ny) * hy)[:, None] - ya
    y0 = oy + J0 * hy - ya
    with np.errstate(all="ignore"):
        pc = line_charge_phi(q, d, X, Y, X[0, I0], y0)
        pH = 0.5 * line_charge_phi(q, d, X, Y, X[0, I0 - 1], y0) + 0.5 * line_charge_phi(q, d, X, Y, X[0, I0], y0)
    pc[~flag.astype(bool)] = 0
    pH[~flag.astype(bool)] = 0
    up = lambda a: np.vstack([a[1:], np.zeros((1, Nx))])
    right = lambda a: np.hstack([a[:, 1:], np.zeros((Ny, 1))])
    hr0 = hx * g
    expect = {"Ey": g * (pc - up(pc)) / hy, "Ex": (pc - right(pc)) / hr0,
              "Hz": np.sqrt(g * g - 1) * (pH - up(pH)) / (2 * C * MU0 * hy)}
    got = {"Ey": Ey, "Ex": Ex, "Hz": Hz}
    for lo, hi, tol in ((5, 25, 3e-2), (25, 60, 3e-3)):
        sel = np.zeros((Ny, Nx), bool)
        sel[:, I0 - hi:I0 - lo + 1] = True
        sel[:, I0 + lo:I0 + hi + 1] = True
        sel &= inside & np.roll(inside, -1, 0) & np.roll(inside, -1, 1)
        for k in expect:
            scale = np.abs(expect[k][sel]).max()
            err = np.abs(got[k][sel] - expect[k][sel]).max() / scale
            assert err < tol, "%s at %d..%d cells: relative error %.2e" % (k, lo, hi, err)
