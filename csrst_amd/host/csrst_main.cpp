// csrst -- driver of the GPU time step with the reference's command line and
// main loop (main.cpp:76-532): Bend and fields, bunch, Poisson start unless
// -I, then per step updateField_firstStep, updateField_secondStep, the wake
// sample and moveMesh (main.cpp:390-506).  Every --statmodulus steps prints a
// summary line and appends the beam statistics to Data/maxwadis.stat
// (DataSinkCommand); writes the CSR wake to --wake FILE.
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <fstream>
#include <memory>

#include "csrst.hpp"

namespace {

csrst::Bunch read_bunch(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw csrst::Error(CSRST_EINVAL, "cannot open bunch file " + path);
    std::vector<double> v;
    double x;
    while (f.read(reinterpret_cast<char *>(&x), sizeof(x))) v.push_back(x);
    if (v.empty() || v.size() % 5) throw csrst::Error(CSRST_EINVAL, "bunch file must hold n x 5 doubles");
    const size_t n = v.size() / 5;
    csrst::Bunch b;
    b.R.resize(2 * n); b.P.resize(2 * n); b.Q.resize(n);
    for (size_t k = 0; k < n; ++k) {
        b.R[2 * k] = v[5 * k]; b.R[2 * k + 1] = v[5 * k + 1];
        b.P[2 * k] = v[5 * k + 2]; b.P[2 * k + 1] = v[5 * k + 3];
        b.Q[k] = v[5 * k + 4];
    }
    return b;
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        const csrst::Commands cmd = csrst::Commands::parse(argc, argv);
        csrst::Engine engine(cmd);
        csrst::Bend geo(engine);
        const csrst::Bunch b = cmd.bunch.empty()
                                   ? csrst::Distribution::generate(cmd, engine.hr(), engine.origin())
                                   : read_bunch(cmd.bunch);
        csrst::PartBunch bunch(engine, b);
        std::printf("csrst: %s, grid %d x %d, dx %.6g dy %.6g dt %.6g, %zu particles, order %d\n",
                    csrst_version(), cmd.numgridpx, cmd.numgridpy, engine.hr()[0], engine.hr()[1], engine.dt(),
                    bunch.getTotalNum(), cmd.order);
        if (!cmd.no_initials) {
            csrst::PoissonSolver poisson(engine);
            poisson.computeField(cmd.poissontol, cmd.poissonmaxit);
            std::printf("csrst: Poisson start, V-cycles %d %d %d, residuals %.2e %.2e %.2e\n", poisson.cycles[0],
                        poisson.cycles[1], poisson.cycles[2], poisson.resid[0], poisson.resid[1], poisson.resid[2]);
        }
        csrst::TETMSolver solver(engine);
        // DataSinkCommand writes Data/maxwadis.stat (DataSink.cpp:22-30)
        std::unique_ptr<csrst::DataSink> sink;
        csrst::PartBunchState state;
        if (cmd.statmodulus > 0) {
            ::mkdir("Data", 0755);
            sink.reset(new csrst::DataSink(engine, "Data/maxwadis.stat"));
        }
        const auto t0 = std::chrono::steady_clock::now();
        for (int it = cmd.startstep; it < cmd.startstep + cmd.numsteps; ++it) {
            solver.updateField_firstStep();
            solver.updateField_secondStep();
            engine.moveMesh();
            if (cmd.statmodulus > 0 && it % cmd.statmodulus == cmd.statmodulus - 1) {
                double rm[2], pm[2];
                bunch.get_rmean(rm);
                bunch.get_pmean(pm);
                std::printf("step %d: rmean %.9e %.9e pmean %.9e %.9e energy %.9e\n", it + 1, rm[0], rm[1], pm[0],
                            pm[1], bunch.fieldEnergy());
                sink->execute(engine, (it - cmd.startstep + 1) * engine.dt(), state);   // main.cpp:435-444
            }
        }
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const auto inf = engine.info();
        std::printf("csrst: %d steps in %.3f s, window at (%d, %d) of %d x %d\n", cmd.numsteps, sec, (int)inf[3],
                    (int)inf[4], (int)inf[5], (int)inf[6]);
        if (!cmd.wake.empty()) {
            const std::vector<double> w = engine.wake();
            std::FILE *f = std::fopen(cmd.wake.c_str(), "w");
            if (!f) throw csrst::Error(CSRST_EINVAL, "cannot write " + cmd.wake);
            for (double v : w) std::fprintf(f, "%.17g\n", v);
            std::fclose(f);
        }
        return EXIT_SUCCESS;
    } catch (const csrst::Error &e) {
        std::fprintf(stderr, "csrst: error %d: %s\n", e.code, e.what());
        return EXIT_FAILURE;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "csrst: %s\n", e.what());
        return EXIT_FAILURE;
    }
}
