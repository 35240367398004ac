// Binary VTK field output of the reference's PlotBinaryVTKCommand
// (PlotBinaryVTK.cpp:30-88, PlotBinaryVTK.hh:58-80) and BinaryVtkFile
// (BinaryVTKFile.hh:25-381, BinaryVTKFile.cpp:25-217), for one rank.
//
// Every execute() writes
//   Data/<base>_NNNN_node00000.vtr  rectilinear-grid piece: XML header, then
//                                   the appended raw data (each array a
//                                   uint32 byte count and float32 values);
//   Data/<base>_NNNN.pvtr           the parallel header naming that piece;
// and close() writes Data/<base>_collection.pvd, the list of (time, file).
// The arrays, in the order the data is appended: X, Y, Z coordinates of the
// index domain [0, nx) x [0, ny) (x_i = i hx + ox computed in double, then
// rounded to float), E-Field and J-Field (three floats per point: x, y, 0)
// and the scalar H-Field Hzx + Hzy.  The XML lists the fields sorted by name
// ("E-Field", "H-Field", "J-Field") with the offsets of that data order.
// The file bytes, whitespace included, equal the reference writer's on one
// rank (a piece extent equal to the whole extent).
//
// Only the standard library is used: this header is independent of the GPU
// library and is what the driver links.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <regex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace csrst {
namespace vtk {

// One plane of a field with one guard layer: element (i, j), i in [-1, nx],
// j in [-1, ny], at p[(j + 1) * pitch + i + 1].
struct Plane {
    const double *p = nullptr;
    size_t pitch = 0;
    double at(int i, int j) const { return p[(size_t)(j + 1) * pitch + (size_t)(i + 1)]; }
};

// One output frame on the index domain [0, nx) x [0, ny): mesh spacing and
// origin of the E field's mesh, E and J (x and y planes) and the two split
// H planes.
struct Frame {
    int nx = 0, ny = 0;
    double hx = 0, hy = 0, ox = 0, oy = 0;
    Plane E[2], J[2], H[2];
};

inline const char *byte_order()
{
    const uint16_t one = 1;
    unsigned char b[2];
    std::memcpy(b, &one, 2);
    return b[0] == 1 ? "LittleEndian" : "BigEndian";
}

class Appended {
public:
    size_t offset() const { return d_.size(); }
    void count(uint32_t bytes) { d_.append(reinterpret_cast<const char *>(&bytes), sizeof bytes); }
    void value(double v)
    {
        const float f = (float)v;
        d_.append(reinterpret_cast<const char *>(&f), sizeof f);
    }
    const std::string &bytes() const { return d_; }

private:
    std::string d_;
};

inline std::string extent(int nx, int ny)
{
    std::ostringstream s;
    s << "\"0 " << nx - 1 << " 0 " << ny - 1 << " 1 1\"";
    return s.str();
}

// The piece file (BinaryVtkFile::writeFile, BinaryVTKFile.cpp:25-52).
// H-Field: Hzx + Hzy over [0, nx-2] x [0, ny-2], as PlotBinaryVTK.cpp:72-80
// fills it; the reference leaves the last column and row of that scalar field
// as the new field holds them, written as 0 here.
inline std::string piece(const Frame &f)
{
    if (f.nx <= 0 || f.ny <= 0) throw std::invalid_argument("vtk: empty domain");
    Appended a;
    const size_t oX = a.offset();
    a.count((uint32_t)(sizeof(float) * f.nx));
    for (int i = 0; i < f.nx; ++i) a.value(i * f.hx + f.ox);
    const size_t oY = a.offset();
    a.count((uint32_t)(sizeof(float) * f.ny));
    for (int j = 0; j < f.ny; ++j) a.value(j * f.hy + f.oy);
    const size_t oZ = a.offset();
    a.count((uint32_t)sizeof(float));
    a.value(0.0);
    const uint32_t vbytes = (uint32_t)(sizeof(float) * 3 * (size_t)f.nx * f.ny);
    auto vector_field = [&](const Plane (&v)[2]) {
        a.count(vbytes);
        for (int j = 0; j < f.ny; ++j)
            for (int i = 0; i < f.nx; ++i) {
                a.value(v[0].at(i, j));
                a.value(v[1].at(i, j));
                a.value(0.0);
            }
    };
    const size_t oE = a.offset();
    vector_field(f.E);
    const size_t oJ = a.offset();
    vector_field(f.J);
    const size_t oH = a.offset();
    a.count((uint32_t)(sizeof(float) * (size_t)f.nx * f.ny));
    for (int j = 0; j < f.ny; ++j)
        for (int i = 0; i < f.nx; ++i)
            a.value(i < f.nx - 1 && j < f.ny - 1 ? f.H[0].at(i, j) + f.H[1].at(i, j) : 0.0);

    const std::string ext = extent(f.nx, f.ny);
    auto array = [](const char *nameAttr, const char *name, int comps, size_t off) {
        std::ostringstream s;
        s << "<DataArray type=\"Float32\" " << nameAttr << "=\"" << name << "\" NumberOfComponents=\"" << comps
          << "\" format=\"appended\" offset=\"" << off << "\"/>";
        return s.str();
    };
    std::ostringstream x;
    x << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"RectilinearGrid\" version=\"0.1\" byte_order=\"" << byte_order() << "\">\n"
      << "  <RectilinearGrid WholeExtent=" << ext << ">\n"
      << "    <Piece Extent=" << ext << ">\n"
      << "      <PointData Scalars=\"H-Field\"  Vectors=\"E-Field, J-Field\" >\n"
      << "        " << array("Name", "E-Field", 3, oE) << "\n"
      << "        " << array("Name", "H-Field", 1, oH) << "\n"
      << "        " << array("Name", "J-Field", 3, oJ) << "\n"
      << "      </PointData>\n"
      << "      <Coordinates>\n"
      << "        " << array("name", "X_COORDINATES", 1, oX) << "\n"
      << "        " << array("name", "Y_COORDINATES", 1, oY) << "\n"
      << "        " << array("name", "Z_COORDINATES", 1, oZ) << "\n"
      << "      </Coordinates>\n"
      << "    </Piece>\n"
      << "  </RectilinearGrid>\n"
      << "  <AppendedData encoding=\"raw\">\n"
      << "    _";
    std::string out = x.str();
    out += a.bytes();
    out += "\n  </AppendedData>\n</VTKFile>\n";
    return out;
}

// The parallel header (BinaryVtkFile::writeHeaderFile, BinaryVTKFile.cpp:54-78);
// source: the piece's file name without its directory.
inline std::string header(int nx, int ny, const std::string &source)
{
    const std::string ext = extent(nx, ny);
    std::ostringstream x;
    x << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"PRectilinearGrid\" version=\"0.1\" byte_order=\"" << byte_order() << "\">\n"
      << "  <PRectilinearGrid WholeExtent=" << ext << " GhostLevel=\"1\">\n"
      << "    <PPointData Scalars=\"H-Field\"  Vectors=\"E-Field, J-Field\" >\n"
      << "      <PDataArray type=\"Float32\" Name=\"E-Field\" NumberOfComponents=\"3\"/>\n"
      << "      <PDataArray type=\"Float32\" Name=\"H-Field\" NumberOfComponents=\"1\"/>\n"
      << "      <PDataArray type=\"Float32\" Name=\"J-Field\" NumberOfComponents=\"3\"/>\n"
      << "    </PPointData>\n"
      << "    <PCoordinates>\n"
      << "      <PDataArray type=\"Float32\" name=\"X_COORDINATES\" NumberOfComponents=\"1\"/>\n"
      << "      <PDataArray type=\"Float32\" name=\"Y_COORDINATES\" NumberOfComponents=\"1\"/>\n"
      << "      <PDataArray type=\"Float32\" name=\"Z_COORDINATES\" NumberOfComponents=\"1\"/>\n"
      << "    </PCoordinates>\n"
      << "    <Piece Extent=" << ext << " Source=\"" << source << "\"/>\n"
      << "  </PRectilinearGrid>\n"
      << "</VTKFile>\n";
    return x.str();
}

// The time collection (PlotBinaryVTKCommand::closeFile, PlotBinaryVTK.hh:58-80);
// times printed as an ostream prints a double (six significant digits).
inline std::string collection(const std::string &base, const std::vector<double> &times)
{
    std::ostringstream x;
    x << "<?xml version=\"1.0\"?>\n"
      << "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"" << byte_order() << "\">\n"
      << "  <Collection>\n";
    char idx[16];
    for (size_t k = 0; k < times.size(); ++k) {
        std::snprintf(idx, sizeof idx, "%04u", (unsigned)k);
        x << "    <DataSet timestep=\"" << times[k] << "\" group=\"\" part=\"0\" file=\"" << base << "_" << idx
          << ".pvtr\"/>\n";
    }
    x << "  </Collection>\n"
      << "</VTKFile>\n\n";
    return x.str();
}

inline void write_file(const std::string &path, const std::string &bytes)
{
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    f.write(bytes.data(), (std::streamsize)bytes.size());
    if (!f) throw std::runtime_error("write failed: " + path);
}

// PlotBinaryVTKCommand for one rank: execute(t, frame) writes frame number
// k = 0, 1, ... as dir/<base>_kkkk.pvtr and its piece; close() the collection.
// The constructor counts the DataSet lines of dir/collection.pvd, as the
// reference's does (PlotBinaryVTK.cpp:90-119), so frame numbers continue
// after those; the reference never writes a file of that name itself.
class Plot {
public:
    Plot(std::string dir, std::string base) : dir_(std::move(dir)), base_(std::move(base))
    {
        std::ifstream is(dir_ + "/collection.pvd");
        const std::regex ts("<DataSet timestep=\"(.*?)\"");
        std::string line;
        std::smatch m;
        while (std::getline(is, line))
            if (std::regex_search(line, m, ts)) {
                std::istringstream v(m[1].str());
                double t = 0;
                v >> t;
                times_.push_back(t);
            }
    }
    void execute(double t, const Frame &f)
    {
        char idx[16];
        std::snprintf(idx, sizeof idx, "%04u", (unsigned)times_.size());
        times_.push_back(t);
        const std::string stem = base_ + "_" + idx;
        const std::string pieceName = stem + "_node00000.vtr";
        write_file(dir_ + "/" + pieceName, piece(f));
        write_file(dir_ + "/" + stem + ".pvtr", header(f.nx, f.ny, pieceName));
    }
    void close() const { write_file(dir_ + "/" + base_ + "_collection.pvd", collection(base_, times_)); }
    size_t frames() const { return times_.size(); }

private:
    std::string dir_, base_;
    std::vector<double> times_;
};

}  // namespace vtk
}  // namespace csrst
