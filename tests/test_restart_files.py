"""Restart files on the CPU: the driver's records (csrst_amd/host/checkpoint.hpp)
through csrst_h5part to an H5Part file and back.

- h5dump (HDF5's own tool) shows the layout SaveH5Part's H5hut calls produce
  (SaveH5Part.cpp:79-353): Step#k groups with TIME in ns, the ten particle
  arrays (z = 1.00000001, pz = 0, id int64), Block/{E,H,J}field with
  components 0 1 2 of shape (1, Ny, Nx), __Spacing__ and __Origin__ (the H
  field's shifted by half a cell), the units and the arguments as file
  attributes;
- reading takes the last step (LoadH5Part.cpp:49-58) and gives back that
  record bit for bit, with the arguments;
- writing into an existing file appends steps after the last one and updates
  the reduced argument set (SaveH5Part.cpp:355-428).

H5hut itself is not available here: the layout follows its documented file
format; parity with an H5hut-written file is unpinned.
"""
import os
import re
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "csrst_amd", "csrst_h5part")
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"

pytestmark = pytest.mark.skipif(not os.path.exists(TOOL), reason="csrst_h5part not built (no HDF5)")


def make_record(rng, Nx, Ny, n, step, t, attrs):
    return dict(Nx=Nx, Ny=Ny, step=step, time=t, hx=2.5e-5, hy=3.5e-5, ox=-0.01, oy=-0.002,
                E=rng.normal(size=2 * Nx * Ny), H=rng.normal(size=2 * Nx * Ny), J=rng.normal(size=2 * Nx * Ny),
                x=rng.normal(size=n), y=rng.normal(size=n), oldx=rng.normal(size=n), oldy=rng.normal(size=n),
                px=rng.normal(size=n), py=rng.normal(size=n), q=rng.normal(size=n),
                id=np.arange(n, dtype=np.int64), attrs=attrs)


def pack(records):
    b = bytearray(b"CSRSTRS1")
    for r in records:
        n = len(r["x"])
        b += b"STEP" + struct.pack("<4q", r["Nx"], r["Ny"], n, r["step"])
        b += struct.pack("<5d", r["time"], r["hx"], r["hy"], r["ox"], r["oy"])
        for k in ("E", "H", "J", "x", "y", "oldx", "oldy", "px", "py", "q"):
            b += np.asarray(r[k], "<f8").tobytes()
        b += np.asarray(r["id"], "<i8").tobytes()
        b += struct.pack("<q", len(r["attrs"]))
        for name, v in r["attrs"]:
            b += struct.pack("<i", len(name)) + name.encode()
            if isinstance(v, str):
                b += b"s" + struct.pack("<i", len(v)) + v.encode()
            elif isinstance(v, float):
                b += b"d" + struct.pack("<d", v)
            else:
                b += b"i" + struct.pack("<q", v)
    return bytes(b)


def unpack(data):
    assert data[:8] == b"CSRSTRS1"
    o, out = 8, []
    while o < len(data):
        assert data[o:o + 4] == b"STEP"
        Nx, Ny, n, step = struct.unpack_from("<4q", data, o + 4)
        o += 36
        r = dict(Nx=Nx, Ny=Ny, step=step)
        r["time"], r["hx"], r["hy"], r["ox"], r["oy"] = struct.unpack_from("<5d", data, o)
        o += 40
        for k, m in (("E", 2 * Nx * Ny), ("H", 2 * Nx * Ny), ("J", 2 * Nx * Ny)) + tuple(
                (k, n) for k in ("x", "y", "oldx", "oldy", "px", "py", "q")):
            r[k] = np.frombuffer(data, "<f8", m, o)
            o += 8 * m
        r["id"] = np.frombuffer(data, "<i8", n, o)
        o += 8 * n
        (na,) = struct.unpack_from("<q", data, o)
        o += 8
        attrs = []
        for _ in range(na):
            (l,) = struct.unpack_from("<i", data, o)
            name = data[o + 4:o + 4 + l].decode()
            o += 4 + l
            kind = data[o:o + 1]
            o += 1
            if kind == b"s":
                (l,) = struct.unpack_from("<i", data, o)
                v = data[o + 4:o + 4 + l].decode()
                o += 4 + l
            elif kind == b"d":
                (v,) = struct.unpack_from("<d", data, o)
                o += 8
            else:
                (v,) = struct.unpack_from("<q", data, o)
                o += 8
            attrs.append((name, v))
        r["attrs"] = attrs
        out.append(r)
    return out


ATTRS = [("restart", ""), ("startstep", 0), ("numsteps", 6), ("savemodulus", 3), ("numgridpx", 7),
         ("phi", 10.0), ("h5file", "Output.h5part"), ("no_initials", 1), ("order", 1)]


def dump(path, *args):
    r = subprocess.run([H5DUMP] + list(args) + [path], capture_output=True, text=True, check=True)
    return r.stdout


@pytest.fixture
def two_steps(tmp_path):
    rng = np.random.default_rng(4)
    recs = [make_record(rng, 7, 3, 5, 0, 0.0, ATTRS), make_record(rng, 7, 3, 5, 3, 2.5e-11, ATTRS)]
    (tmp_path / "a.rs").write_bytes(pack(recs))
    subprocess.run([TOOL, "write", str(tmp_path / "a.rs"), str(tmp_path / "a.h5part")], check=True)
    return recs, tmp_path / "a.h5part"


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not available")
def test_h5part_layout(two_steps):
    recs, h5 = two_steps
    hdr = dump(str(h5), "-H")
    for k in (0, 1):
        assert 'GROUP "Step#%d"' % k in hdr
    assert 'GROUP "Step#2"' not in hdr
    for name in ("x", "y", "z", "oldx", "oldy", "px", "py", "pz", "q", "id"):
        assert 'DATASET "%s"' % name in hdr
    assert hdr.count("DATASPACE  SIMPLE { ( 5 ) / ( 5 ) }") == 2 * 10
    assert hdr.count("DATASPACE  SIMPLE { ( 1, 3, 7 ) / ( 1, 3, 7 ) }") == 2 * 9
    for a in ("TIME", "__Spacing__", "__Origin__", "tUnit", "TIMEUnit", "pxUnit", "numgridpx", "phi", "h5file"):
        assert 'ATTRIBUTE "%s"' % a in hdr
    vals = lambda path: [float(v) for v in dump(str(h5), "-d", path, "-y", "-w", "0", "-m", "%.17g")
                         .split("DATA {")[1].split("}")[0].replace(",", " ").split()]
    assert vals("/Step#1/z") == [1.00000001] * 5 and vals("/Step#0/pz") == [0.0] * 5
    np.testing.assert_array_equal(vals("/Step#1/Block/Efield/1"), recs[1]["E"][21:])
    np.testing.assert_array_equal(vals("/Step#1/Block/Hfield/2"), recs[1]["H"][:21] + recs[1]["H"][21:])
    assert not any(vals("/Step#0/Block/Jfield/2"))
    att = lambda path, name: [float(v) for v in re.sub(r"\(\d+\):", " ", dump(
        str(h5), "-a", path + "/" + name, "-m", "%.17g").split("DATA {")[1].split("}")[0]).replace(",", " ").split()]
    assert att("/Step#1", "TIME") == [2.5e-11 * 1e9]
    assert att("/Step#0/Block/Hfield", "__Origin__") == [-0.01 + 0.5 * 2.5e-5, -0.002 + 0.5 * 3.5e-5, 0.0]
    assert att("/Step#0/Block/Efield", "__Spacing__") == [2.5e-5, 3.5e-5, 1.0]


def test_read_gives_last_step(two_steps, tmp_path):
    recs, h5 = two_steps
    subprocess.run([TOOL, "read", str(h5), str(tmp_path / "b.rs")], check=True)
    (got,) = unpack((tmp_path / "b.rs").read_bytes())
    want = recs[1]
    assert (got["Nx"], got["Ny"], got["step"]) == (7, 3, 1)
    assert got["time"] == pytest.approx(want["time"], rel=1e-15)
    assert (got["hx"], got["hy"], got["ox"], got["oy"]) == (want["hx"], want["hy"], want["ox"], want["oy"])
    for k in ("E", "H", "J", "x", "y", "oldx", "oldy", "px", "py", "q", "id"):
        assert np.array_equal(got[k], want[k]), k
    attrs = dict(got["attrs"])
    for name, v in ATTRS:
        assert attrs[name] == v, name
    assert attrs["TIMEUnit"] == "ns" and attrs["qUnit"] == "Cb"


def test_write_appends_after_last_step(two_steps, tmp_path):
    recs, h5 = two_steps
    rng = np.random.default_rng(9)
    more = [("startstep", 6), ("numsteps", 9), ("savemodulus", 3), ("order", 1), ("phi", 99.0)]
    rec = make_record(rng, 7, 3, 5, 9, 5e-11, more)
    (tmp_path / "c.rs").write_bytes(pack([rec]))
    subprocess.run([TOOL, "write", str(tmp_path / "c.rs"), str(h5)], check=True)
    subprocess.run([TOOL, "read", str(h5), str(tmp_path / "d.rs")], check=True)
    (got,) = unpack((tmp_path / "d.rs").read_bytes())
    assert got["step"] == 2 and np.array_equal(got["x"], rec["x"])
    attrs = dict(got["attrs"])
    # the reduced set changes, the rest (phi) keeps the first write's value
    assert (attrs["startstep"], attrs["numsteps"], attrs["phi"]) == (6, 9, 10.0)


def test_rejects_bad_input(tmp_path):
    (tmp_path / "bad.rs").write_bytes(b"NOTAFILE")
    r = subprocess.run([TOOL, "write", str(tmp_path / "bad.rs"), str(tmp_path / "x.h5part")],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "not a restart file" in r.stderr
    r = subprocess.run([TOOL, "read", str(tmp_path / "missing.h5part"), str(tmp_path / "y.rs")],
                       capture_output=True, text=True)
    assert r.returncode != 0


@pytest.fixture(scope="module")
def args_tool(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("rsargs") / "restart_args_main")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "restart_args_main.cpp"), "-o", exe], check=True)
    return exe


def restart_args(exe, rs, *args):
    r = subprocess.run([exe, str(rs)] + list(args), capture_output=True, text=True, check=True)
    return dict(line.split(" ", 1) for line in r.stdout.splitlines())


def test_restart_argument_rules(args_tool, tmp_path):
    """LoadH5Part::readH5FileAttributes (LoadH5Part.cpp:60-159): arguments on
    the command line win, except numgridpx, numgridpy, Np and phi, which always
    come from the file; startstep = ((startstep + numsteps) / savemodulus) *
    savemodulus of the file's values; switches are int64 0/1."""
    attrs = [("restart", ""), ("startstep", 4), ("numsteps", 11), ("savemodulus", 5), ("numgridpx", 700),
             ("numgridpy", 90), ("Np", 1234), ("phi", 12.5), ("width", 0.02), ("Ekin", 80.0),
             ("no_initials", 1), ("h5file", "old.h5part"), ("order", 2)]
    rec = make_record(np.random.default_rng(1), 3, 2, 2, 0, 1e-10, attrs)
    (tmp_path / "r.rs").write_bytes(pack([rec]))
    a = restart_args(args_tool, tmp_path / "r.rs", "-r", "r.rs", "--numgridpx", "64", "-w", "0.05", "-P", "3",
                     "--numsteps", "7")
    assert a["restart"] == "r.rs"                      # set: kept
    assert a["numgridpx"] == "700" and a["phi"] == "12.5" and a["numgridpy"] == "90" and a["Np"] == "1234"
    assert a["width"] == "0.050000000000000003"         # set: kept
    assert a["numsteps"] == "7"                         # set: kept
    assert a["startstep"] == "15"                       # ((4 + 11) / 5) * 5
    assert a["Ekin"] == "80" and a["no_initials"] == "1" and a["h5file"] == "old.h5part" and a["order"] == "2"
    assert a["Bz"] == "0.10000000000000001"             # absent from the file: default
