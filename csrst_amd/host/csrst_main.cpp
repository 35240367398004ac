// csrst -- driver of the GPU time step with the reference's command line and
// main loop (main.cpp:76-532): Bend and fields, bunch, Poisson start unless
// -I, then per step updateField_firstStep, updateField_secondStep, the field
// plots, the statistics, the wake sample and moveMesh (main.cpp:390-506).
// Field plots (PlotBinaryVTKCommand): the global fields as Data/c_NNNN.pvtr
// and the window as Data/mm_NNNN.pvtr at the start and every --plotmodulus
// steps, after copyFields; Data/c_collection.pvd and mm_collection.pvd at the
// end.  Every --statmodulus steps prints a summary line and appends the beam
// statistics to Data/maxwadis.stat (DataSinkCommand); the CSR wake goes to
// Data/energyChange{Before,InBend,After}.dat (IntegrateElongCommand::write).  -M/--savemodulus saves
// restart records (SaveH5Part) to --h5file, -r/--restart FILE continues from
// the last one (LoadH5Part); H5Part names go through csrst_h5part.  Writes the
// CSR wake to --wake FILE.
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <fstream>
#include <memory>

#include "csrst.hpp"
#include "restart.hpp"
#include "vtk.hpp"

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

// One VTK frame from planar [2][m+2][n+2] field downloads (guards included).
csrst::vtk::Frame frame(const std::vector<double> &E, const std::vector<double> &H, const std::vector<double> &J,
                        int n, int m, const double *hr, double ox, double oy)
{
    const size_t plane = (size_t)(n + 2) * (m + 2);
    if (E.size() != 2 * plane || H.size() != 2 * plane || J.size() != 2 * plane)
        throw csrst::Error(CSRST_EINVAL, "field download size");
    csrst::vtk::Frame f;
    f.nx = n; f.ny = m;
    f.hx = hr[0]; f.hy = hr[1]; f.ox = ox; f.oy = oy;
    for (int k = 0; k < 2; ++k) {
        f.E[k] = {E.data() + k * plane, (size_t)n + 2};
        f.H[k] = {H.data() + k * plane, (size_t)n + 2};
        f.J[k] = {J.data() + k * plane, (size_t)n + 2};
    }
    return f;
}

// pavc.execute and mmpavc.execute (main.cpp:358-366, 416-423)
void plot(const csrst::Engine &engine, csrst::vtk::Plot &global, csrst::vtk::Plot &window, double t)
{
    const csrst::Commands &cmd = engine.commands();
    const auto inf = engine.info();
    global.execute(t, frame(engine.field(0), engine.field(1), engine.field(5), cmd.numgridpx, cmd.numgridpy,
                            engine.hr(), engine.origin()[0], engine.origin()[1]));
    window.execute(t, frame(engine.field(2), engine.field(3), engine.field(4), (int)inf[5], (int)inf[6], engine.hr(),
                            inf[1], inf[2]));
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        csrst::Commands cmd = csrst::Commands::parse(argc, argv);
        // -r FILE: the last step of a restart file and its arguments
        // (main.cpp:144-153, LoadH5Part.cpp:60-159), read before the GPU is used
        const bool restarting = !cmd.restart.empty();
        csrst::ckpt::Record rs;
        if (restarting) {
            rs = csrst::restart::load(cmd.restart);
            csrst::restart::apply(cmd, rs);
        }
        csrst::Engine engine(cmd);
        csrst::Bend geo(engine);
        csrst::Bunch b;
        if (restarting)
            b = csrst::restart::restore_fields(engine, rs);   // restoreFields + restoreBunch (main.cpp:226-228)
        else
            b = cmd.bunch.empty() ? csrst::Distribution::generate(cmd, engine.hr(), engine.origin())
                                  : read_bunch(cmd.bunch);
        csrst::PartBunch bunch(engine, b, restarting ? cmd.startstep : 0);
        std::printf("csrst: %s, grid %d x %d, dx %.6g dy %.6g dt %.6g, %zu particles, order %d\n",
                    csrst_version(), cmd.numgridpx, cmd.numgridpy, engine.hr()[0], engine.hr()[1], engine.dt(),
                    bunch.getTotalNum(), cmd.order);
        if (restarting)
            std::printf("csrst: restart from %s, step %d, t %.9e\n", cmd.restart.c_str(), cmd.startstep, rs.time);
        else if (!cmd.no_initials) {
            csrst::PoissonSolver poisson(engine);
            poisson.computeField(cmd.poissontol, cmd.poissonmaxit);
            std::printf("csrst: Poisson start, V-cycles %d %d %d, residuals %.2e %.2e %.2e\n", poisson.cycles[0],
                        poisson.cycles[1], poisson.cycles[2], poisson.resid[0], poisson.resid[1], poisson.resid[2]);
        }
        double t = restarting ? rs.time : 0.0;   // currentSimulatedTime
        csrst::TETMSolver solver(engine);
        ::mkdir("Data", 0755);
        // DataSinkCommand writes Data/maxwadis.stat (DataSink.cpp:22-30)
        std::unique_ptr<csrst::DataSink> sink;
        csrst::PartBunchState state;
        if (cmd.statmodulus > 0) sink.reset(new csrst::DataSink(engine, "Data/maxwadis.stat"));
        // restart records every --savemodulus steps, as the SaveH5Part calls
        // of main.cpp:361-366,425-430 (commented out there) would write them:
        // into --h5file when -M is given, or appended to the restart file on
        // a restart; an H5Part name is written by csrst_h5part at the end
        const bool saving = (cmd.isSet("savemodulus") || restarting) && cmd.savemodulus > 0;
        const std::string saveTo = restarting ? cmd.restart : cmd.h5file;
        const bool viaTool = csrst::restart::is_h5part(saveTo);
        const std::string records = viaTool ? saveTo + ".rs.tmp" + std::to_string(::getpid()) : saveTo;
        std::unique_ptr<csrst::ckpt::Writer> saver;
        if (saving) saver.reset(new csrst::ckpt::Writer(records, !viaTool));
        // the start state is plotted (and saved) without a copy (main.cpp:358-366)
        csrst::vtk::Plot pavc("Data", "c"), mmpavc("Data", "mm");
        if (!restarting) {
            plot(engine, pavc, mmpavc, 0.0);
            if (saving) saver->write(csrst::restart::snapshot(engine, bunch, t, cmd.startstep));
        }
        csrst::SaveEMFields semfc(engine);
        bool inStraightSectionBeforeBend = true, inBend = false;   // main.cpp:91-92
        const auto t0 = std::chrono::steady_clock::now();
        for (int it = cmd.startstep; it < cmd.startstep + cmd.numsteps; ++it) {
            solver.updateField_firstStep();
            solver.updateField_secondStep();
            const double tn = t + engine.dt();   // currentSimulatedTime + dt
            const bool plotStep = cmd.plotmodulus > 0 && it % cmd.plotmodulus == cmd.plotmodulus - 1;
            if (plotStep) {   // main.cpp:416-423
                engine.copyFields(true);
                plot(engine, pavc, mmpavc, tn);
            }
            if (saving && it % cmd.savemodulus == cmd.savemodulus - 1) {   // main.cpp:425-430
                engine.copyFields(true);
                saver->write(csrst::restart::snapshot(engine, bunch, tn, it + 1));
            }
            if (cmd.statmodulus > 0 && it % cmd.statmodulus == cmd.statmodulus - 1) {   // main.cpp:435-444
                double rm[2], pm[2];
                bunch.get_rmean(rm);
                bunch.get_pmean(pm);
                std::printf("step %d: rmean %.9e %.9e pmean %.9e %.9e energy %.9e\n", it + 1, rm[0], rm[1], pm[0],
                            pm[1], bunch.fieldEnergy());
                sink->execute(engine, tn, state);
                // the wake files at the bend's entry and exit (main.cpp:445-463);
                // the wake is then sampled in moveMesh (main.cpp:465)
                const double rmax[2] = {state.get_rmax(0), state.get_rmax(1)};
                const double rmin[2] = {state.get_rmin(0), state.get_rmin(1)};
                if (inStraightSectionBeforeBend && geo.isInsideBend(rmax)) {
                    engine.writeWake("Data/energyChangeBefore.dat", state.get_rmean(0));
                    engine.wakeReset();
                    inStraightSectionBeforeBend = false;
                    inBend = true;
                }
                if (inBend && !geo.isInsideBend(rmax) && !geo.isInsideBend(rmin)) {
                    engine.writeWake("Data/energyChangeInBend.dat", state.get_rmean(0));
                    engine.wakeReset();
                    inBend = false;
                }
            }
            // main.cpp:468-476: the line dump (SaveEMFieldsCommand) after the window's
            // E and H are copied into the global fields
            if (cmd.linemodulus > 0 && it % cmd.linemodulus == cmd.linemodulus - 1) {
                if (!plotStep) engine.copyFields(false);
                semfc.execute();
            }
            engine.moveMesh();   // wake sample, J = 0, moveMesh (main.cpp:445-501)
            t += engine.dt();
        }
        csrst::calcBeamParameters(engine, state);   // main.cpp:516-519
        engine.writeWake("Data/energyChangeAfter.dat", state.get_rmean(0));
        pavc.close();   // main.cpp:521-522
        mmpavc.close();
        if (saver) {
            saver.reset();
            if (viaTool) {
                csrst::restart::run_tool("write", records, saveTo);
                std::remove(records.c_str());
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
