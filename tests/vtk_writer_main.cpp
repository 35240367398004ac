// Test driver of csrst_amd/host/vtk.hpp: writes the VTK frames of one field
// set at the given times into OUTDIR, then the collection.
//   vtk_writer_main NX NY HX HY OX OY FIELDS OUTDIR BASE T0 [T1 ...]
// FIELDS: 6 planes of (NY+2) x (NX+2) doubles (guards included): Ex, Ey,
// Jx, Jy, Hzx, Hzy.  Frame k scales every field value by (k + 1).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "../csrst_amd/host/vtk.hpp"

int main(int argc, char **argv)
{
    if (argc < 11) {
        std::fprintf(stderr, "usage: %s NX NY HX HY OX OY FIELDS OUTDIR BASE T0 [T1 ...]\n", argv[0]);
        return 2;
    }
    const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]);
    const size_t plane = (size_t)(nx + 2) * (ny + 2);
    std::vector<double> in(6 * plane);
    std::ifstream f(argv[7], std::ios::binary);
    if (!f.read(reinterpret_cast<char *>(in.data()), (std::streamsize)(in.size() * sizeof(double)))) {
        std::fprintf(stderr, "short field file\n");
        return 2;
    }
    try {
        csrst::vtk::Plot plot(argv[8], argv[9]);
        for (int a = 10; a < argc; ++a) {
            std::vector<double> v(in);
            for (double &x : v) x *= (double)(a - 9);
            csrst::vtk::Frame fr;
            fr.nx = nx; fr.ny = ny;
            fr.hx = std::atof(argv[3]); fr.hy = std::atof(argv[4]);
            fr.ox = std::atof(argv[5]); fr.oy = std::atof(argv[6]);
            for (int k = 0; k < 2; ++k) {
                fr.E[k] = {v.data() + k * plane, (size_t)nx + 2};
                fr.J[k] = {v.data() + (2 + k) * plane, (size_t)nx + 2};
                fr.H[k] = {v.data() + (4 + k) * plane, (size_t)nx + 2};
            }
            plot.execute(std::strtod(argv[a], nullptr), fr);
        }
        plot.close();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
