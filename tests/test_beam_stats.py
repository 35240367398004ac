"""Beam parameters of PartBunch::calcMoments / calcBeamParameters
(PartBunch.cpp:584-756): the oracle against closed forms, textbook
covariances in the bunch frame and rotation invariance."""
import numpy as np

from csrst_amd import stats
from oracle import oracle as O

M0 = 0.51099906   # MeV, Physics.cpp:51 (m_e in GeV) * 1000


def test_field_names_agree():
    assert tuple(stats.FIELDS) == tuple(O.BEAM_FIELDS)


def test_four_particle_box():
    """(x0 +- a, y0 +- b), all with momentum (p, 0): rms sizes a, b, no momentum
    spread, no emittance; Ekin from gamma = sqrt(1 + p^2)."""
    a, b, x0, y0, p = 1e-3, 2e-3, 0.5, 0.2, 100.0
    R = np.array([[x0 + a, y0 + b], [x0 + a, y0 - b], [x0 - a, y0 + b], [x0 - a, y0 - b]])
    P = np.tile([p, 0.0], (4, 1))
    s = O.beam_parameters(R, P)
    assert abs(s["rrms_x"] - a) < 1e-15 and abs(s["rrms_y"] - b) < 1e-15
    assert s["prms_x"] == 0 and s["prms_y"] == 0 and s["emit_x"] == 0 and s["dE"] == 0
    assert abs(s["Ekin"] - (np.sqrt(1 + p * p) - 1) * M0) < 1e-12
    assert s["rmin_x"] == x0 - a and s["rmax_y"] == y0 + b


def bunch(n, seed, angle=0.0):
    """Correlated Gaussian phase space along x, then rotated by angle."""
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1e-3, n)
    y = rng.normal(0, 2e-4, n)
    px = 200.0 + 0.3 * x / 1e-3 + rng.normal(0, 0.05, n)
    py = 0.02 * y / 2e-4 + rng.normal(0, 0.01, n) + 0.004 * (px - 200.0)
    c, s = np.cos(angle), np.sin(angle)
    R = np.stack([0.3 + c * x - s * y, 0.1 + s * x + c * y], 1)
    P = np.stack([c * px - s * py, s * px + c * py], 1)
    return R, P


def test_textbook_covariances_in_the_bunch_frame():
    R, P = bunch(20000, 3, angle=0.3)
    s = O.beam_parameters(R, P)
    # bunch frame: x along the mean momentum, centred on the centroid
    pm = P.mean(0)
    c, sn = pm / np.linalg.norm(pm)
    d = R - R.mean(0)
    xl = c * d[:, 0] + sn * d[:, 1]
    xt = -sn * d[:, 0] + c * d[:, 1]
    pl = c * P[:, 0] + sn * P[:, 1]
    pt = -sn * P[:, 0] + c * P[:, 1]
    cov_l = np.cov(np.stack([xl, pl]), bias=True)
    cov_t = np.cov(np.stack([xt, pt]), bias=True)
    assert abs(s["rrms_x"] - np.sqrt(cov_l[0, 0])) < 1e-12 * s["rrms_x"]
    assert abs(s["prms_x"] - np.sqrt(cov_l[1, 1])) < 1e-9 * s["prms_x"]
    assert abs(s["emit_x"] - np.sqrt(np.linalg.det(cov_l))) < 1e-8 * s["emit_x"]
    assert abs(s["rrms_y"] - np.sqrt(cov_t[0, 0])) < 1e-12 * s["rrms_y"]
    assert abs(s["prms_y"] - np.sqrt(np.mean(pt * pt))) < 1e-9 * s["prms_y"]
    assert abs(s["Dy"] - np.mean(xt * pl)) < 1e-9 * s["rrms_y"] * np.linalg.norm(pm)
    assert abs(s["DDy"] - np.mean(pt * pl)) < 1e-9 * s["prms_y"] * np.linalg.norm(pm)
    g = np.sqrt(1 + (P * P).sum(1)).mean()
    ek = (g - 1) * M0
    assert abs(s["Ekin"] - ek) < 1e-12 * ek
    de = s["prms_x"] * np.sqrt(ek * (ek + 2 * M0) / (1 + ek * (ek + 2 * M0) / M0 ** 2))
    assert abs(s["dE"] - de) < 1e-12 * de


def test_rotation_invariance():
    R0, P0 = bunch(5000, 4, angle=0.0)
    R1, P1 = bunch(5000, 4, angle=0.7)
    a, b = O.beam_parameters(R0, P0), O.beam_parameters(R1, P1)
    # the longitudinal momentum spread is formed as sum p^2 - N pmean^2
    # (PartBunch.cpp:625), which loses (|p| / prms)^2 ~ 4e5 in relative
    # precision: the reference's own rounding, not the rotation, sets these
    for k, tol in (("rrms_x", 1e-13), ("rrms_y", 1e-13), ("prms_y", 1e-12), ("emit_y", 1e-12), ("Ekin", 1e-14),
                   ("prms_x", 1e-9), ("dE", 1e-9), ("emit_x", 1e-7)):
        assert abs(a[k] - b[k]) < tol * abs(a[k]), k
    assert abs(a["Dy"] - b["Dy"]) < 1e-8 * abs(a["Dy"])
