"""The C++ driver (csrst_amd/csrst) on the GPU:
- the statistics file of DataSinkCommand (Data/maxwadis.stat) has the
  reference's SDDS header and one row per --statmodulus steps whose values
  equal what the library returns for the same run at the same point of the
  step, before the window moves (PartBunch::calcBeamParameters,
  CalcEnergyCommand);
- the field plots (PlotBinaryVTKCommand) are byte for byte the oracle
  writer's files of the library's fields for the same run, with copyFields
  before every plot after the start."""
import os
import subprocess

import numpy as np
import pytest

from csrst_amd import Commands, TwoGridStep, gaussian_bunch
from oracle import vtk as V

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "csrst_amd", "csrst")
GEOM = dict(lengthbefore=0.05, length=0.01, width=0.0128, phi=10.0, Ekin=50.59, Bz=0.1704)
EPS0, MU0 = 8.854187817e-12, 1.256637061e-06   # Physics.cpp:41-51


def test_driver_statistics_file(gpu, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("driver not built")
    R, P, Q = gaussian_bunch(3000, (0.02, 0.0064), 1e-3, 0.5e-3, 100.0, 1.0, seed=12, energy_spread=1e-3)
    np.column_stack([R, P, Q]).astype(np.float64).tofile(str(tmp_path / "bunch.bin"))
    steps, stat = 6, 3
    args = [EXE, "--numgridpx", "1100", "--numgridpy", "120", "--order", "1", "-I", "--numsteps", str(steps),
            "--statmodulus", str(stat), "--bunch", "bunch.bin"]
    for k, v in GEOM.items():
        args += ["--" + k, repr(v)]
    print("driver:", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = open(tmp_path / "Data" / "maxwadis.stat").read().splitlines()
    header = open(os.path.join(ROOT, "tests", "golden", "maxwadis_header.txt")).read().splitlines()
    assert lines[:len(header)] == header
    assert lines[len(header)].startswith("CSRST ")
    rows = [np.array([float(x) for x in l.split()]) for l in lines[len(header) + 1:]]
    assert len(rows) == steps // stat and all(len(row) == 16 for row in rows)
    # the same run through the library
    cmd = Commands(numgridpx=1100, numgridpy=120, lengthbefore=GEOM["lengthbefore"], length=GEOM["length"],
                   width=GEOM["width"], phi=GEOM["phi"], Ekin=GEOM["Ekin"], Bz=GEOM["Bz"], order=1, statmodulus=stat)
    g = TwoGridStep(cmd)
    g.init_bunch(R, P, Q)
    dt = g.info()["dt"]
    e_old = 0.0
    for k, row in enumerate(rows):
        g.step(stat - 1)
        g.updateField_firstStep()
        g.updateField_secondStep()
        s = g.calcBeamParameters()
        sums = g.fieldEnergySums()
        energy = (0.25 * EPS0 * (sums[0] + e_old) + 0.5 * MU0 * sums[1]) * 1e9
        e_old = sums[0]
        want = np.array([(k + 1) * stat * dt, s["Ekin"], s["rrms_x"], s["rrms_y"], s["prms_x"], s["prms_y"],
                         s["emit_x"], s["emit_y"], s["rmean_x"], s["rmean_y"], s["pmean_x"], s["pmean_y"], s["Dy"],
                         s["DDy"], s["dE"], energy * 1e9])
        print("row %d: %s" % (k, " ".join("%.6e" % v for v in row)))
        assert energy > 0
        np.testing.assert_allclose(row, want, rtol=2e-14, atol=0)
        g.moveMesh()


def run_driver(tmp_path, R, P, Q, extra):
    np.column_stack([R, P, Q]).astype(np.float64).tofile(str(tmp_path / "bunch.bin"))
    args = [EXE, "--numgridpx", "1100", "--numgridpy", "120", "--order", "1", "-I", "--bunch", "bunch.bin"] + extra
    for k, v in GEOM.items():
        args += ["--" + k, repr(v)]
    print("driver:", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


def test_driver_vtk_frames(gpu, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("driver not built")
    R, P, Q = gaussian_bunch(2000, (0.02, 0.0064), 1e-3, 0.5e-3, 100.0, 1.0, seed=13, energy_spread=1e-3)
    steps, pm = 7, 3
    run_driver(tmp_path, R, P, Q, ["--numsteps", str(steps), "--plotmodulus", str(pm)])
    cmd = Commands(numgridpx=1100, numgridpy=120, lengthbefore=GEOM["lengthbefore"], length=GEOM["length"],
                   width=GEOM["width"], phi=GEOM["phi"], Ekin=GEOM["Ekin"], Bz=GEOM["Bz"], order=1)
    g = TwoGridStep(cmd)
    g.init_bunch(R, P, Q)
    dt = g.info()["dt"]
    want, times = {}, []

    def snap(t):
        k = len(times)
        times.append(t)
        want.update(V.frame_files("c_%04d" % k, g.field("E"), g.field("H"), g.field("J"), g.hr, g.origin))
        want.update(V.frame_files("mm_%04d" % k, g.field("mE"), g.field("mH"), g.field("mJ"), g.hr,
                                  g.info()["origin"]))

    snap(0.0)
    for it in range(steps):
        g.updateField_firstStep()
        g.updateField_secondStep()
        if it % pm == pm - 1:
            g.copyFields(True)
            snap((it + 1) * dt)
        g.moveMesh()
    want["c_collection.pvd"] = V.collection_bytes("c", times)
    want["mm_collection.pvd"] = V.collection_bytes("mm", times)
    data = tmp_path / "Data"
    got = sorted(n for n in os.listdir(data) if n.endswith((".vtr", ".pvtr", ".pvd")))
    assert got == sorted(want)
    for name, b in want.items():
        assert (data / name).read_bytes() == b, name
    assert len(times) == 3


def read_h5part_last(path, out):
    """The last step of an H5Part file as the driver's record (csrst_h5part read)."""
    from tests.test_restart_files import unpack
    tool = os.path.join(ROOT, "csrst_amd", "csrst_h5part")
    subprocess.run([tool, "read", str(path), str(out)], check=True)
    (rec,) = unpack(open(out, "rb").read())
    return rec


def test_driver_save_and_restart(gpu, tmp_path):
    """-M 3 --h5file run.h5part saves the start and steps 3 and 6 (SaveH5Part
    after copyFields); -r run.h5part then continues (LoadH5Part: the file's
    arguments, startstep ((0 + 6) / 3) * 3 = 6, numsteps 6) and appends steps
    9 and 12.  The last record equals the library restored from step 6 and
    stepped the same way."""
    if not os.path.exists(EXE) or not os.path.exists(os.path.join(ROOT, "csrst_amd", "csrst_h5part")):
        pytest.skip("driver or csrst_h5part not built")
    if not os.path.exists("/opt/conda/lib/libhdf5.so"):
        pytest.skip("the HDF5 library csrst_h5part was built against is not installed")
    R, P, Q = gaussian_bunch(2000, (0.02, 0.0064), 1e-3, 0.5e-3, 100.0, 1.0, seed=17, energy_spread=1e-3)
    run_driver(tmp_path, R, P, Q, ["--numsteps", "6", "--savemodulus", "3", "--h5file", "run.h5part"])
    rec6 = read_h5part_last(tmp_path / "run.h5part", tmp_path / "a.rs")
    attrs = dict(rec6["attrs"])
    assert rec6["step"] == 2 and (attrs["numsteps"], attrs["savemodulus"], attrs["numgridpx"]) == (6, 3, 1100)
    print("restart:", EXE, "-r run.h5part", flush=True)
    r = subprocess.run([EXE, "-r", "run.h5part"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "restart from run.h5part, step 6" in r.stdout
    rec12 = read_h5part_last(tmp_path / "run.h5part", tmp_path / "b.rs")
    assert rec12["step"] == 4
    assert rec12["time"] == pytest.approx(12 * rec6["time"] / 6, rel=1e-12)
    # the library, restored from the step-6 record
    cmd = Commands(numgridpx=1100, numgridpy=120, lengthbefore=GEOM["lengthbefore"], length=GEOM["length"],
                   width=GEOM["width"], phi=GEOM["phi"], Ekin=GEOM["Ekin"], Bz=GEOM["Bz"], order=1)
    g = TwoGridStep(cmd)
    Nx, Ny = 1100, 120

    def guarded(v):
        a = np.zeros((2, Ny + 2, Nx + 2))
        a[:, 1:-1, 1:-1] = np.asarray(v).reshape(2, Ny, Nx)
        return a

    g.restore(guarded(rec6["E"]), guarded(rec6["H"]), np.stack([rec6["x"], rec6["y"]], 1),
              np.stack([rec6["oldx"], rec6["oldy"]], 1), np.stack([rec6["px"], rec6["py"]], 1), rec6["q"], step=6)
    for it in range(6, 12):
        g.updateField_firstStep()
        g.updateField_secondStep()
        if it % 3 == 2:
            g.copyFields(True)
        if it < 11:
            g.moveMesh()
    for k, f in (("E", "E"), ("H", "H"), ("J", "J")):
        assert np.array_equal(rec12[k], g.field(f)[:, 1:-1, 1:-1].ravel()), k
    Rg, Pg, Og = g.particles()
    for k, v in (("x", Rg[:, 0]), ("y", Rg[:, 1]), ("oldx", Og[:, 0]), ("oldy", Og[:, 1]), ("px", Pg[:, 0]),
                 ("py", Pg[:, 1]), ("q", g.charges())):
        assert np.array_equal(rec12[k], v), k
    # a width on the command line overrides the file's: the bend's mesh then
    # differs from the one the stored fields were written on, which is refused
    r = subprocess.run([EXE, "-r", "run.h5part", "--width", "0.0132"], cwd=str(tmp_path), capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "mesh" in (r.stdout + r.stderr)


def test_driver_wake_files(gpu, tmp_path):
    """IntegrateElongCommand::write (IntegrateElong.cpp:91-125) as main.cpp
    drives it: Data/energyChangeBefore.dat when the bunch's rmax enters the
    bend (then the accumulator restarts), Data/energyChangeAfter.dat at the
    end; each line x - rmean_x and -statModulus dt times the accumulated
    field, six digits.  Expected from the library over the same run."""
    if not os.path.exists(EXE):
        pytest.skip("driver not built")
    R, P, Q = gaussian_bunch(2000, (0.0455, 0.0064), 0.4e-3, 0.4e-3, 100.0, 1.0, seed=19, energy_spread=1e-3)
    steps, stat = 40, 2
    run_driver(tmp_path, R, P, Q, ["--numsteps", str(steps), "--statmodulus", str(stat)])
    cmd = Commands(numgridpx=1100, numgridpy=120, lengthbefore=GEOM["lengthbefore"], length=GEOM["length"],
                   width=GEOM["width"], phi=GEOM["phi"], Ekin=GEOM["Ekin"], Bz=GEOM["Bz"], order=1,
                   statmodulus=stat)
    g = TwoGridStep(cmd)
    # Bend.hh:307-323 on the entrance plane x = lengthbefore, on the beam axis
    assert not g.isInsideBend((GEOM["lengthbefore"] - 1e-4, 0.5 * GEOM["width"]))
    assert g.isInsideBend((GEOM["lengthbefore"] + 1e-4, 0.5 * GEOM["width"]))
    g.init_bunch(R, P, Q)
    dt = g.info()["dt"]

    def wake_text(rmean_x):
        w = g.wake()
        o, h = g.info()["origin"][0], g.hr[0]
        return "".join("%g\t%g\n" % (o + (k + 2) * h - rmean_x, -stat * dt * w[k]) for k in range(len(w)))

    want, entered = {}, None
    for it in range(steps):
        g.updateField_firstStep()
        g.updateField_secondStep()
        if it % stat == stat - 1:
            s = g.calcBeamParameters()
            if entered is None and g.isInsideBend((s["rmax_x"], s["rmax_y"])):
                want["energyChangeBefore.dat"] = wake_text(s["rmean_x"])
                g.wake_reset()
                entered = it
        g.moveMesh()
    s = g.calcBeamParameters()
    want["energyChangeAfter.dat"] = wake_text(s["rmean_x"])
    print("bunch entered the bend at step", entered, flush=True)
    assert entered is not None and entered > 2
    for name, text in want.items():
        got = open(tmp_path / "Data" / name).read()
        assert got == text, name
    assert not os.path.exists(tmp_path / "Data" / "energyChangeInBend.dat")


def test_driver_line_files(gpu, tmp_path):
    """--linemodulus 4: Data/FieldOnLine_0000.dat and _0001.dat after steps 4
    and 8 (main.cpp:468-476), each line the library's csrst_line_fields row
    of the same run printed as the reference's stream does (six digits)."""
    if not os.path.exists(EXE):
        pytest.skip("driver not built")
    R, P, Q = gaussian_bunch(2000, (0.02, 0.0064), 1e-3, 0.5e-3, 100.0, 1.0, seed=19, energy_spread=1e-3)
    run_driver(tmp_path, R, P, Q, ["--numsteps", "9", "--linemodulus", "4"])
    cmd = Commands(numgridpx=1100, numgridpy=120, lengthbefore=GEOM["lengthbefore"], length=GEOM["length"],
                   width=GEOM["width"], phi=GEOM["phi"], Ekin=GEOM["Ekin"], Bz=GEOM["Bz"], order=1)
    g = TwoGridStep(cmd)
    g.init_bunch(R, P, Q)
    files = []
    for it in range(9):
        g.updateField_firstStep()
        g.updateField_secondStep()
        if it % 4 == 3:
            g.copyFields(False)
            rows = g.lineFields()
            files.append("".join("\t".join("%g" % v for v in r) + "\n" for r in rows))
        g.moveMesh()
    data = tmp_path / "Data"
    assert sorted(n for n in os.listdir(data) if n.startswith("FieldOnLine")) == ["FieldOnLine_0000.dat",
                                                                                 "FieldOnLine_0001.dat"]
    for k, want in enumerate(files):
        assert (data / ("FieldOnLine_%04d.dat" % k)).read_text() == want
