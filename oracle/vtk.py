"""CPU restatement of the reference's binary VTK field output for one rank:
BinaryVtkFile (BinaryVTKFile.hh:25-381, BinaryVTKFile.cpp:25-217) as
PlotBinaryVTKCommand drives it (PlotBinaryVTK.cpp:30-88, PlotBinaryVTK.hh:58-80).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the reference writer for the driver's
writer.  The product never imports this module.

The restatement follows the reference's string building step by step
(fragments with embedded newlines, re-indented by replacing each newline with
newline + indent) rather than writing the final layout out directly, so it is
an independent route to the same bytes.
"""
import struct

import numpy as np

L0, L1, L2, L3 = "", "  ", "    ", "      "     # indent_l0..3, defs.hh:33-36


def _indent(text, ind):
    """boost::regex_replace(text, "\\n", "\\n" + ind)."""
    return text.replace("\n", "\n" + ind)


class BinaryVtkFile:
    """One rank's file: addVectorField / addScalarField append data in call
    order; the XML lists fields in std::map order of "name,components"."""

    def __init__(self):
        self.data = bytearray()
        self.fields = {}          # "name,ncomp" -> DataArray line (BinaryVTKFile.hh:296-330)
        self.headers = {}         # name -> PDataArray line (BinaryVTKFile.hh:158-176)
        self.extent = None
        self.coordinates = None
        self.origin = (0.0, 0.0)

    def _extent(self, nx, ny):
        # BinaryVTKFile.hh:116-134: first last first last 1 1, quoted
        return '"%d %d %d %d %d %d"' % (0, nx - 1, 0, ny - 1, 1, 1)

    def _append_array(self, values):
        """uint32 byte count, then float32 values (double -> float rounding)."""
        v = np.asarray(values, dtype=np.float64).astype("<f4")
        self.data += struct.pack("<I", v.nbytes)
        self.data += v.tobytes()

    def _grid(self, nx, ny, hx, hy):
        # addGridExtent + addCoordinates (BinaryVTKFile.hh:103-140, BinaryVTKFile.cpp:88-107)
        self.nx, self.ny = nx, ny
        self.extent = self._extent(nx, ny)
        parts = []
        for name, n, h, o in (("X", nx, hx, self.origin[0]), ("Y", ny, hy, self.origin[1])):
            off = len(self.data)
            parts.append('<DataArray type="Float32" name="%s_COORDINATES" NumberOfComponents="1" '
                         'format="appended" offset="%d"/>' % (name, off))
            self._append_array(np.arange(n, dtype=np.float64) * h + o)
        off = len(self.data)
        parts.append('<DataArray type="Float32" name="Z_COORDINATES" NumberOfComponents="1" '
                     'format="appended" offset="%d"/>' % off)
        self._append_array([0.0])
        self.coordinates = (L0 + "<Coordinates>\n" + "".join(L1 + _indent(p, L1) + "\n" for p in parts)
                            + L0 + "</Coordinates>")

    def add_vector_field(self, F, name, hx, hy, origin):
        """F: [2, ny, nx] (component, j, i) over the index domain."""
        self.origin = origin
        if self.extent is None:
            self._grid(F.shape[2], F.shape[1], hx, hy)
        off = len(self.data)
        self.fields["%s,2" % name] = ('<DataArray type="Float32" Name="%s" NumberOfComponents="3" '
                                      'format="appended" offset="%d"/>' % (name, off))
        xyz = np.stack([F[0], F[1], np.zeros_like(F[0])], axis=-1)      # point-major, i fastest
        self._append_array(xyz.ravel())
        self.headers[name] = '<PDataArray type="Float32" Name="%s" NumberOfComponents="3"/>' % name

    def add_scalar_field(self, S, name, hx, hy, origin):
        self.origin = origin
        if self.extent is None:
            self._grid(S.shape[1], S.shape[0], hx, hy)
        off = len(self.data)
        self.fields["%s,1" % name] = ('<DataArray type="Float32" Name="%s" NumberOfComponents="1" '
                                      'format="appended" offset="%d"/>' % (name, off))
        self._append_array(np.asarray(S).ravel())
        self.headers[name] = '<PDataArray type="Float32" Name="%s" NumberOfComponents="1"/>' % name

    def _names(self):
        # concatenateFieldNames (BinaryVTKFile.hh:284-290) with the scalar and
        # vector name lists (BinaryVTKFile.cpp:109-161): "..." plus a blank
        def names(k):
            sel = [key.rsplit(",", 1)[0] for key in sorted(self.fields) if key.rsplit(",", 1)[1] == str(k)]
            return '"' + ", ".join(sel) + '" ' if sel else '""'
        return "Scalars=" + names(1) + " Vectors=" + names(2)

    def piece_bytes(self):
        point = (L0 + "<PointData " + self._names() + ">\n"
                 + "".join(L1 + _indent(self.fields[k], L3) + "\n" for k in sorted(self.fields))
                 + L0 + "</PointData>")
        head = ('<?xml version="1.0"?>\n'
                + L0 + '<VTKFile type="RectilinearGrid" version="0.1" byte_order="LittleEndian">\n'
                + L1 + "<RectilinearGrid WholeExtent=" + self.extent + ">" + "\n"
                + L2 + "<Piece Extent=" + self.extent + ">\n"
                + L3 + _indent(point, L3) + "\n"
                + L3 + _indent(self.coordinates, L3) + "\n"
                + L2 + "</Piece>\n"
                + L1 + "</RectilinearGrid>\n"
                + L1 + '<AppendedData encoding="raw">\n'
                + L2 + "_")
        tail = "\n" + L1 + "</AppendedData>\n" + L0 + "</VTKFile>" + "\n"
        return head.encode() + bytes(self.data) + tail.encode()

    def header_bytes(self, source):
        hdrs = (L0 + "<PPointData " + self._names() + ">\n"
                + "".join(L1 + _indent(self.headers[k], L1) + "\n" for k in sorted(self.headers))
                + L0 + "</PPointData>")
        pieces = L0 + '<Piece Extent="%d %d %d %d %d %d" Source="%s"/>' % (0, self.nx - 1, 0, self.ny - 1, 1, 1,
                                                                           source)
        text = ('<?xml version="1.0"?>\n'
                + L0 + '<VTKFile type="PRectilinearGrid" version="0.1" byte_order="LittleEndian">\n'
                + L1 + "<PRectilinearGrid WholeExtent=" + self.extent + ' GhostLevel="1">\n'
                + L2 + _indent(hdrs, L2) + "\n"
                + L2 + "<PCoordinates>\n"
                + L3 + '<PDataArray type="Float32" name="X_COORDINATES" NumberOfComponents="1"/>\n'
                + L3 + '<PDataArray type="Float32" name="Y_COORDINATES" NumberOfComponents="1"/>\n'
                + L3 + '<PDataArray type="Float32" name="Z_COORDINATES" NumberOfComponents="1"/>\n'
                + L2 + "</PCoordinates>\n"
                + L2 + _indent(pieces, L2) + "\n"
                + L1 + "</PRectilinearGrid>\n"
                + L0 + "</VTKFile>" + "\n")
        return text.encode()


def frame_files(stem, E, H, J, hr, origin):
    """PlotBinaryVTKCommand::execute (PlotBinaryVTK.cpp:52-88) for one rank.
    E, H, J: guarded arrays [2, ny+2, nx+2] (component, j, i) over [-1, nx] x
    [-1, ny]; stem "c_0003".  Returns {file name: bytes} for the piece and
    its parallel header.  H-Field = Hzx + Hzy on [0, nx-2] x [0, ny-2]; the
    last column and row of that new field are 0 here (the reference does not
    set them)."""
    Ei, Ji, Hi = E[:, 1:-1, 1:-1], J[:, 1:-1, 1:-1], H[:, 1:-1, 1:-1]
    ny, nx = Ei.shape[1], Ei.shape[2]
    S = np.zeros((ny, nx))
    S[:ny - 1, :nx - 1] = Hi[0, :ny - 1, :nx - 1] + Hi[1, :ny - 1, :nx - 1]
    f = BinaryVtkFile()
    f.add_vector_field(Ei, "E-Field", hr[0], hr[1], origin)
    f.add_vector_field(Ji, "J-Field", hr[0], hr[1], origin)
    f.add_scalar_field(S, "H-Field", hr[0], hr[1], origin)
    piece = "%s_node%05d.vtr" % (stem, 0)
    return {piece: f.piece_bytes(), stem + ".pvtr": f.header_bytes(piece)}


def collection_bytes(base, times):
    """PlotBinaryVTKCommand::closeFile (PlotBinaryVTK.hh:57-80); times as an
    ostream prints a double (%g)."""
    s = '<?xml version="1.0"?>\n' + L0 + '<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n'
    s += L1 + "<Collection>\n"
    for i, t in enumerate(times):
        s += L2 + '<DataSet timestep="%g" group="" part="0" file="%s_%04d.pvtr"/>\n' % (t, base, i)
    s += L1 + "</Collection>\n" + L0 + "</VTKFile>\n" + "\n"
    return s.encode()
