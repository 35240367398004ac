"""The five workloads of the benchmark (synthetic seeded bunches).

Bend fields follow from the bend radius with the electron mass of Bend.cpp:66
(0.511 MeV): Bz = p / (c R).  Grids are the co-moving window of the
moving-window step; chamber width slightly below the window height.
"""
import numpy as np

M_BEND = 0.511      # MeV
C = 299792458.0


def ekin_mev(gamma):
    return (gamma - 1.0) * M_BEND


def bz_for(gamma, R):
    ek = ekin_mev(gamma) * 1e6
    p = np.sqrt(ek * ek + 2.0 * ek * M_BEND * 1e6)
    return p / (C * R)


def config(name):
    if name == "c1":
        g, R = 100.0, 1.0
        return dict(name=name, W=512, H=256, dx=50e-6, dy=50e-6, width=12.0e-3, phiDeg=10.0, gamma=g,
                    ekinMeV=ekin_mev(g), bz=bz_for(g, R), R=R, lenBefore=0.05, lenAfter=0.02, n=100000,
                    charge_nC=1.0, sigma_s=1e-3, sigma_x=0.5e-3, espread=0.0, chirp=0.0, order=1, steps=2000,
                    sortEvery=20, cfl=0.5, pml=16, poisson=True)
    if name == "c2":
        g, R = 500.0, 5.0
        return dict(name=name, W=8192, H=2048, dx=10e-6, dy=10e-6, width=20.0e-3, phiDeg=30.0, gamma=g,
                    ekinMeV=ekin_mev(g), bz=bz_for(g, R), R=R, lenBefore=0.10, lenAfter=0.10, n=10_000_000,
                    charge_nC=0.5, sigma_s=200e-6, sigma_x=100e-6, espread=1e-3, chirp=0.0, order=2,
                    steps=20000, sortEvery=20, cfl=1.0, pml=32)
    if name == "c3":
        g, R = 1000.0, 10.0
        return dict(name=name, W=32768, H=4096, dx=5e-6, dy=5e-6, width=20.0e-3, phiDeg=20.0, gamma=g,
                    ekinMeV=ekin_mev(g), bz=bz_for(g, R), R=R, lenBefore=0.20, lenAfter=0.10, n=100_000_000,
                    charge_nC=1.0, sigma_s=50e-6, sigma_x=50e-6, espread=0.0, chirp=-10.0, order=2,
                    steps=10000, sortEvery=20, cfl=1.0, pml=0)
    if name == "c5":
        g, R = 300.0, 2.0
        return dict(name=name, W=16384, H=8192, dx=2e-6, dy=2e-6, width=16.0e-3, phiDeg=15.0, gamma=g,
                    ekinMeV=ekin_mev(g), bz=bz_for(g, R), R=R, lenBefore=0.10, lenAfter=0.05, n=50_000_000,
                    charge_nC=0.2, sigma_s=300e-6, sigma_x=50e-6, espread=0.0, chirp=0.0, order=2,
                    steps=15000, sortEvery=20, modulation=(0.1, 20e-6), cfl=1.0, pml=32, poisson=True)
    raise KeyError(name)


def arc_point(cfg, into_arc):
    """Centre line point `into_arc` metres along the bend's arc and the
    direction of the arc there (the bend of Bend.cpp:60-124: centre of
    curvature at (lenBefore, width/2 + R) for Bz > 0)."""
    R = cfg["R"]
    phi0 = into_arc / R
    sgn = 1.0 if cfg["bz"] > 0 else -1.0
    cx, cy = cfg["lenBefore"], 0.5 * cfg["width"] + sgn * R
    pos = (cx + R * np.sin(phi0), cy - sgn * R * np.cos(phi0))
    return pos, sgn * phi0


def bunch_in_bend(cfg, n=None, seed=1, arc_fraction=0.5):
    """Seeded bunch of the configuration centred on the bend's centre line at
    `arc_fraction` of the arc, moving along the arc's tangent there."""
    from .bunch import gaussian_bunch
    n = cfg["n"] if n is None else n
    arc = np.radians(cfg["phiDeg"]) * cfg["R"]
    centre, th = arc_point(cfg, arc_fraction * arc)
    return gaussian_bunch(n, centre, cfg["sigma_s"], cfg["sigma_x"], cfg["gamma"], cfg["charge_nC"], seed=seed,
                          direction=(float(np.cos(th)), float(np.sin(th))), energy_spread=cfg["espread"],
                          chirp=cfg["chirp"], modulation=cfg.get("modulation"))


def bunch_for(cfg, n=None, seed=1, s_offset=-5e-3, angle_deg=0.0):
    """Seeded bunch of the configuration, centred s_offset before the bend entry, mid-chamber,
    moving at angle_deg to the beam axis (0: along it)."""
    from .bunch import gaussian_bunch
    n = cfg["n"] if n is None else n
    center = (cfg["lenBefore"] + s_offset, 0.5 * cfg["width"])
    th = np.radians(angle_deg)
    return gaussian_bunch(n, center, cfg["sigma_s"], cfg["sigma_x"], cfg["gamma"], cfg["charge_nC"], seed=seed,
                          direction=(float(np.cos(th)), float(np.sin(th))), energy_spread=cfg["espread"],
                          chirp=cfg["chirp"], modulation=cfg.get("modulation"))
