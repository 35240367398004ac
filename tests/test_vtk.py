"""Binary VTK field output (PlotBinaryVTKCommand / BinaryVtkFile, one rank).

- the oracle's files decode, by the VTK XML file format's own rules (raw
  appended data: each array a UInt32 byte count and its values, offsets
  counted from the byte after '_'), to the fields it was given: coordinates
  i*h + origin, (x, y, 0) vectors, Hzx + Hzy, extents, names and sources;
- the driver's C++ writer (csrst_amd/host/vtk.hpp) writes byte-identical
  files, collection included, on ragged and one-cell domains;
- copy_window (Communicator::copy) places the window at its offset and clips
  it to the global domain.
"""
import os
import shutil
import struct
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from oracle import oracle as O
from oracle import vtk as V

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def fields(rng, nx, ny):
    return [rng.normal(size=(2, ny + 2, nx + 2)) * s for s in (1e5, 1e-2, 3e2)]   # E, J, H


def decode_piece(b):
    """Split a .vtr into its XML (data removed) and the raw appended bytes."""
    start = b.index(b'<AppendedData encoding="raw">')
    us = b.index(b"_", start)
    end = b.rindex(b"\n  </AppendedData>")
    raw = b[us + 1:end]
    xml = b[:us] + b[end:]
    return ET.fromstring(xml.decode()), raw


def read_array(raw, off):
    (nbytes,) = struct.unpack_from("<I", raw, off)
    return np.frombuffer(raw, "<f4", nbytes // 4, off + 4), off + 4 + nbytes


@pytest.mark.parametrize("nx,ny", [(7, 5), (1, 1), (33, 2)])
def test_oracle_piece_decodes_to_fields(nx, ny):
    rng = np.random.default_rng(nx * 100 + ny)
    E, J, H = fields(rng, nx, ny)
    hr, origin = (2.5e-5, 3.1e-5), (-0.0123, 0.0047)
    files = V.frame_files("c_0003", E, H, J, hr, origin)
    assert sorted(files) == ["c_0003.pvtr", "c_0003_node00000.vtr"]
    root, raw = decode_piece(files["c_0003_node00000.vtr"])
    assert root.tag == "VTKFile" and root.get("type") == "RectilinearGrid" and root.get("byte_order") == "LittleEndian"
    grid = root.find("RectilinearGrid")
    ext = "0 %d 0 %d 1 1" % (nx - 1, ny - 1)
    assert grid.get("WholeExtent") == ext and grid.find("Piece").get("Extent") == ext
    pd = grid.find("Piece/PointData")
    assert pd.get("Scalars") == "H-Field" and pd.get("Vectors") == "E-Field, J-Field"
    arrays = {a.get("Name") or a.get("name"): a for a in grid.iter("DataArray")}
    assert [a.get("Name") for a in pd] == ["E-Field", "H-Field", "J-Field"]
    # walk the appended data in offset order: each block starts where the last ended
    offs = sorted((int(a.get("offset")), n) for n, a in arrays.items())
    pos, got = 0, {}
    for off, n in offs:
        assert off == pos, (n, off, pos)
        got[n], pos = read_array(raw, off)
    assert pos == len(raw)
    assert [n for _, n in offs] == ["X_COORDINATES", "Y_COORDINATES", "Z_COORDINATES", "E-Field", "J-Field",
                                    "H-Field"]
    f32 = lambda a: np.asarray(a, np.float64).astype(np.float32)
    np.testing.assert_array_equal(got["X_COORDINATES"], f32(np.arange(nx) * hr[0] + origin[0]))
    np.testing.assert_array_equal(got["Y_COORDINATES"], f32(np.arange(ny) * hr[1] + origin[1]))
    np.testing.assert_array_equal(got["Z_COORDINATES"], [0.0])
    for name, F in (("E-Field", E), ("J-Field", J)):
        assert arrays[name].get("NumberOfComponents") == "3"
        v = got[name].reshape(ny, nx, 3)
        np.testing.assert_array_equal(v[..., 0], f32(F[0, 1:-1, 1:-1]))
        np.testing.assert_array_equal(v[..., 1], f32(F[1, 1:-1, 1:-1]))
        assert not v[..., 2].any()
    h = got["H-Field"].reshape(ny, nx)
    want = f32(H[0, 1:-1, 1:-1] + H[1, 1:-1, 1:-1])
    np.testing.assert_array_equal(h[:-1, :-1], want[:-1, :-1])
    assert not h[-1].any() and not h[:, -1].any()
    # parallel header: same names, one piece pointing at the file
    ph = ET.fromstring(files["c_0003.pvtr"].decode())
    pg = ph.find("PRectilinearGrid")
    assert ph.get("type") == "PRectilinearGrid" and pg.get("WholeExtent") == ext and pg.get("GhostLevel") == "1"
    assert [(a.get("Name"), a.get("NumberOfComponents")) for a in pg.find("PPointData")] == [
        ("E-Field", "3"), ("H-Field", "1"), ("J-Field", "3")]
    assert [a.get("name") for a in pg.find("PCoordinates")] == ["X_COORDINATES", "Y_COORDINATES", "Z_COORDINATES"]
    (piece,) = pg.findall("Piece")
    assert piece.get("Extent") == ext and piece.get("Source") == "c_0003_node00000.vtr"


def test_oracle_collection_lists_frames():
    b = V.collection_bytes("mm", [0.0, 2.5e-12, 1.0 / 3.0])
    root = ET.fromstring(b.decode())
    sets = root.findall("Collection/DataSet")
    assert [(d.get("timestep"), d.get("file"), d.get("part")) for d in sets] == [
        ("0", "mm_0000.pvtr", "0"), ("2.5e-12", "mm_0001.pvtr", "0"), ("0.333333", "mm_0002.pvtr", "0")]
    assert b.endswith(b"</VTKFile>\n\n")


@pytest.fixture(scope="module")
def writer(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("vtkw") / "vtk_writer_main")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror",
                    os.path.join(ROOT, "tests", "vtk_writer_main.cpp"), "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("nx,ny", [(7, 5), (1, 1), (64, 3), (5, 40)])
def test_cpp_writer_bytes_equal_oracle(writer, tmp_path, nx, ny):
    rng = np.random.default_rng(7 * nx + ny)
    E, J, H = fields(rng, nx, ny)
    hr, origin = (1.0 / 3.0 * 1e-4, 2.7e-5), (0.0731, -0.00125)
    np.concatenate([E, J, H]).astype(np.float64).tofile(str(tmp_path / "f.bin"))
    out = tmp_path / "Data"
    out.mkdir()
    times = [0.0, 1.234567891e-11, 7e-10]
    subprocess.run([writer, str(nx), str(ny), repr(hr[0]), repr(hr[1]), repr(origin[0]), repr(origin[1]),
                    str(tmp_path / "f.bin"), str(out), "c"] + [repr(t) for t in times], check=True)
    want = {}
    for k in range(len(times)):
        s = k + 1.0
        want.update(V.frame_files("c_%04d" % k, E * s, H * s, J * s, hr, origin))
    want["c_collection.pvd"] = V.collection_bytes("c", times)
    assert sorted(os.listdir(out)) == sorted(want)
    for name, b in want.items():
        got = (out / name).read_bytes()
        assert got == b, name


def test_cpp_writer_continues_after_collection_file(writer, tmp_path):
    """The reference's constructor counts the DataSets of Data/collection.pvd
    (PlotBinaryVTK.cpp:90-119): frames continue after them."""
    E, J, H = fields(np.random.default_rng(3), 4, 3)
    np.concatenate([E, J, H]).tofile(str(tmp_path / "f.bin"))
    out = tmp_path / "Data"
    out.mkdir()
    (out / "collection.pvd").write_bytes(V.collection_bytes("x", [1e-9, 2e-9]))
    subprocess.run([writer, "4", "3", "1e-4", "1e-4", "0", "0", str(tmp_path / "f.bin"), str(out), "c", "3e-9"],
                   check=True)
    assert (out / "c_0002.pvtr").exists() and not (out / "c_0000.pvtr").exists()
    assert (out / "c_collection.pvd").read_bytes() == V.collection_bytes("c", [1e-9, 2e-9, 3e-9])


def test_copy_window_offsets_and_clipping():
    G = np.zeros((2, 8 + 2, 12 + 2))
    Wf = np.arange(2 * 5 * 6, dtype=float).reshape(2, 5, 6) + 1.0      # window 4 x 3 (+ guards)
    out = O.copy_window(G, Wf, (2, 3))
    assert (out[:, 4:7, 3:7] == Wf[:, 1:-1, 1:-1]).all()
    out[:, 4:7, 3:7] = 0
    assert not out.any()
    # hanging over the top-right corner and the left edge: clipped
    out = O.copy_window(G, Wf, (10, 6))
    assert (out[:, 7:9, 11:13] == Wf[:, 1:3, 1:3]).all()
    out[:, 7:9, 11:13] = 0
    assert not out.any()
    out = O.copy_window(G, Wf, (-3, 0))
    assert (out[:, 1:4, 1:2] == Wf[:, 1:4, 4:5]).all()
    out[:, 1:4, 1:2] = 0
    assert not out.any()

