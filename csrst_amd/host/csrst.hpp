// csrst.hpp -- C++ front end of libcsrst (include/csrst_hip.h) with the
// reference's class interfaces for the particle-in-cell time step:
//   Commands      the command line of main.cpp:111-136
//   Distribution  the bunch generator of Distribution.cpp:55-130
//   Engine        AllocateFields + Bend::Init + the window (main.cpp:168-217);
//                 owns the csrst_ctx
//   Bend          geometry queries (Bend.hh:373-485)
//   PartBunch     particle set-up and statistics (PartBunch.hh:35-206)
//   TETMSolver    updateField_firstStep / secondStep (TETMSolver.hh:38-103)
//   PoissonSolver computeField (PoissonSolver.hh:47-66)
// Errors are thrown as csrst::Error; the reference aborts (PAssert) or
// returns EXIT_FAILURE from main instead.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <stdexcept>
#include <string>
#include <vector>

#include "csrst_hip.h"

namespace csrst {

class Error : public std::runtime_error {
public:
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const void *handle = nullptr)
{
    if (rc != CSRST_OK) {
        const char *m = csrst_last_error(handle);
        throw Error(rc, m ? m : "csrst error");
    }
}

// ---------------------------------------------------------------------------
// main.cpp:111-136 (TCLAP flags, same names, short forms and defaults), plus
// the choices the reference makes at compile time (IntOp_t, defs.hh:55-57).
struct Commands {
    std::string restart;
    int startstep = 0;
    int numsteps = 10;
    int numgridpx = 100;
    int numgridpy = 100;
    int plotmodulus = 100000;
    int statmodulus = 100000;
    int linemodulus = 100000;
    int reballance = 100000;
    int savemodulus = 100000;   // -M: restart records every M steps (main.cpp:118, commented out there)
    double phi = 0.0;
    double width = 0.10;
    double length = 0.10;
    double lengthbefore = 0.10;
    double Ekin = 250.0;
    double Bz = 0.10;
    double Qtotal = -1.0;
    double mean_Rx = 0.05;
    double mean_Ry = 0.0;
    double sigma_x = 0.01;
    double sigma_p = 1.0;
    int Np = 50000;
    std::string h5file = "Output.h5part";
    bool no_initials = false;
    bool debug = false;
    // compile-time in the reference
    int order = 0;          // --order: 0 IntCIC, 1 IntS1, 2 IntS2
    bool s2bug = true;      // --no-s2bug: gather IntS2 at (I+2, J+2) instead of ScatterGatherImpl.hh:467
    // additions of this front end
    std::string bunch;      // --bunch FILE: n x 5 doubles (x, y, px, py, q) instead of the generator
    std::string wake;       // --wake FILE: write the CSR wake samples at the end
    double poissontol = 1e-12;   // main.cpp:94 relativeConvergenceTolerance
    int poissonmaxit = 500;      // main.cpp:93 maxNumPoissonIterations

    std::vector<std::string> given;   // long names of the arguments on the command line

    // One argument: TCLAP short and long name, kind (i int, d double,
    // s string, b switch) and the member it sets.
    struct Opt {
        const char *shortName, *longName;
        char kind;
        void *dst;
    };
    std::vector<Opt> options();
    static Commands parse(int argc, char **argv);
    bool isSet(const std::string &name) const
    {
        return std::find(given.begin(), given.end(), name) != given.end();
    }
    csrst_params params() const
    {
        csrst_params p;
        p.Nx = numgridpx;
        p.Ny = numgridpy;
        p.lenBefore = lengthbefore;
        p.lenAfter = length;
        p.width = width;
        p.phiDeg = phi;
        p.ekinMeV = Ekin;
        p.bz = Bz;
        p.order = order;
        p.s2bug = s2bug ? 1 : 0;
        p.statModulus = statmodulus;
        return p;
    }
};

// ---------------------------------------------------------------------------
// Distribution.cpp:55-130: two-dimensional Halton points (bases 2 and 3, one
// stream for positions and one for momenta) through the Box-Muller map of
// gaussian2DSample (Distribution.cpp:146-157); R = mean + x, P = (p0, 0) + p
// with gamma = Ekin / (m_e 1e3) + 1; Q = -Qtotal 1e-9 / Np each; R alone is
// then sorted by Utils::XFirst (utils.hh:116-140), which re-pairs positions
// and momenta exactly as the reference does.
struct Bunch {
    std::vector<double> R, P, Q;   // R, P: 2 per particle
    std::vector<double> oldR;      // restart only (LoadH5Part::restoreBunch); empty: makeReady
    size_t size() const { return Q.size(); }
};

class Distribution {
public:
    // radical inverse of k in base b (the Halton coordinate)
    static double radical_inverse(unsigned long long k, unsigned b)
    {
        double f = 1.0, r = 0.0;
        while (k) {
            f /= b;
            r += f * (double)(k % b);
            k /= b;
        }
        return r;
    }
    static Bunch generate(const Commands &c, const double hr[2], const double origin[2]);
};

// ---------------------------------------------------------------------------
struct BoundaryCell {       // BoundaryCell.hh:28-33
    int i, j;
    double lambda[2];
    double area;
    double lambdaToNeighbours[2];
};

class Engine {
public:
    explicit Engine(const Commands &cmd) : cmd_(cmd)
    {
        const csrst_params p = cmd.params();
        check(csrst_create(&p, &ctx_));
        check(csrst_geometry(ctx_, geo_), ctx_);
    }
    ~Engine() { csrst_destroy(ctx_); }
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    csrst_ctx *ctx() const { return ctx_; }
    const Commands &commands() const { return cmd_; }
    double dt() const { return geo_[4]; }
    const double *hr() const { return geo_; }
    const double *origin() const { return geo_ + 2; }
    double radius() const { return geo_[5]; }
    double phi() const { return geo_[6]; }

    // main.cpp:494-501 (moveMesh) with the wake sample and J reset before it
    std::array<int, 2> moveMesh()
    {
        int mi = 0, mj = 0;
        check(csrst_move_window(ctx_, &mi, &mj), ctx_);
        return {mi, mj};
    }
    // copyFields (main.cpp:534-548): window E, H (and J) into the global fields
    void copyFields(bool copyCurrent) { check(csrst_copy_fields(ctx_, copyCurrent ? 1 : 0), ctx_); }
    // IntegrateElongCommand accumulator (IntegrateElong.cpp:38-125)
    std::vector<double> wake() const
    {
        int n = 0;
        check(csrst_wake(ctx_, nullptr, 0, &n), ctx_);
        std::vector<double> w(n > 0 ? n : 1);
        check(csrst_wake(ctx_, w.data(), (int)w.size(), &n), ctx_);
        w.resize(n);
        return w;
    }
    // IntegrateElongCommand::reset (IntegrateElong.hh:59-62)
    void wakeReset() { check(csrst_wake_reset(ctx_), ctx_); }
    // IntegrateElongCommand::write (IntegrateElong.cpp:91-125): per sample k
    // (window column k + 2) its x relative to the bunch centroid and
    // -statModulus dt times the accumulated field, as "x\tvalue" lines
    // printed with the stream's default six digits
    void writeWake(const std::string &path, double rmeanX) const
    {
        const std::vector<double> w = wake();
        const std::array<double, 9> inf = info();
        const double h = hr()[0], wdt = cmd_.statmodulus * dt();
        std::ofstream out(path);
        if (!out) throw Error(CSRST_EINVAL, "cannot write " + path);
        for (size_t k = 0; k < w.size(); ++k) {
            const double curPos = inf[1] + (double)(k + 2) * h - rmeanX;
            out << curPos << "\t" << -wdt * w[k] << "\n";
        }
    }
    // dt, window origin (2), window position (2), W, H, steps, wake samples
    std::array<double, 9> info() const
    {
        std::array<double, 9> a{};
        check(csrst_info(ctx_, a.data()), ctx_);
        return a;
    }
    // which: 0 E, 1 H (global); 2 E, 3 H, 4 J (window); 5 J (global)
    void setField(int which, const std::vector<double> &f)
    {
        check(csrst_upload_field(ctx_, which, f.data(), (int)f.size()), ctx_);
    }
    std::vector<double> field(int which) const
    {
        int n = 0;
        check(csrst_download_field(ctx_, which, nullptr, 0, &n), ctx_);
        std::vector<double> f(n);
        check(csrst_download_field(ctx_, which, f.data(), n, &n), ctx_);
        return f;
    }
    // The caller's geometry in place of the one Bend computes: the inside mask
    // (numgridpx x numgridpy bytes) and SCBoundaryCondition::getBoundaryCells()
    // (SCBoundaryCondition.hh:39-84).  Before the PartBunch is created.
    void setGeometry(const std::vector<uint8_t> &mask, const std::vector<BoundaryCell> &cells)
    {
        if (!mask.empty()) check(csrst_set_mask(ctx_, mask.data()), ctx_);
        std::vector<int> ij(2 * cells.size());
        std::vector<double> d(5 * cells.size());
        for (size_t k = 0; k < cells.size(); ++k) {
            const BoundaryCell &c = cells[k];
            ij[2 * k] = c.i; ij[2 * k + 1] = c.j;
            d[5 * k] = c.lambda[0]; d[5 * k + 1] = c.lambda[1]; d[5 * k + 2] = c.area;
            d[5 * k + 3] = c.lambdaToNeighbours[0]; d[5 * k + 4] = c.lambdaToNeighbours[1];
        }
        check(csrst_set_boundary_cells(ctx_, ij.data(), d.data(), (int)cells.size()), ctx_);
    }
    // The engine's GPU work on the caller's stream (a hipStream_t; nullptr:
    // the engine's own)
    void setStream(void *stream) { check(csrst_set_stream(ctx_, stream), ctx_); }

private:
    Commands cmd_;
    csrst_ctx *ctx_ = nullptr;
    double geo_[7] = {0, 0, 0, 0, 0, 0, 0};
};

class Bend {
public:
    explicit Bend(const Engine &e) : e_(e) {}
    double getRadius() const { return e_.radius(); }
    double getAngle() const { return e_.phi(); }
    std::array<double, 2> getMeshSpacing() const { return {e_.hr()[0], e_.hr()[1]}; }
    std::array<double, 2> getOrigin() const { return {e_.origin()[0], e_.origin()[1]}; }
    // Bend.hh:307-323
    bool isInsideBend(const double R[2]) const
    {
        int in = 0;
        check(csrst_is_inside_bend(e_.ctx(), R, &in), e_.ctx());
        return in != 0;
    }
    // Bend.hh:288-291
    size_t getLengthStraightSection() const { return (size_t)(e_.commands().lengthbefore / e_.hr()[0]); }
    std::vector<uint8_t> getInsideMask() const
    {
        std::vector<uint8_t> m((size_t)e_.commands().numgridpx * e_.commands().numgridpy);
        int n = 0;
        check(csrst_boundary_cells(e_.ctx(), m.data(), nullptr, nullptr, 0, &n), e_.ctx());
        return m;
    }
    std::vector<BoundaryCell> getBoundaryCells() const
    {
        int n = 0;
        check(csrst_boundary_cells(e_.ctx(), nullptr, nullptr, nullptr, 0, &n), e_.ctx());
        std::vector<int> ij(2 * (size_t)n);
        std::vector<double> d(5 * (size_t)n);
        check(csrst_boundary_cells(e_.ctx(), nullptr, ij.data(), d.data(), n, &n), e_.ctx());
        std::vector<BoundaryCell> out(n);
        for (int k = 0; k < n; ++k)
            out[k] = BoundaryCell{ij[2 * k], ij[2 * k + 1], {d[5 * k], d[5 * k + 1]}, d[5 * k + 2],
                                  {d[5 * k + 3], d[5 * k + 4]}};
        return out;
    }

private:
    const Engine &e_;
};

class PartBunch {
public:
    // PartBunch(...) + makeReady(dt/2) + the window set-up of main.cpp:199-347;
    // with b.oldR given, the restart instead (no makeReady, the window filled
    // from the global fields as uploaded, the step counter at step)
    PartBunch(Engine &e, const Bunch &b, int step = 0) : e_(e), n_(b.size())
    {
        if (b.oldR.empty())
            check(csrst_init_bunch(e.ctx(), b.R.data(), b.P.data(), b.Q.data(), (int)n_), e.ctx());
        else if (b.oldR.size() != 2 * n_)
            throw Error(CSRST_EINVAL, "oldR size");
        else
            check(csrst_restore_bunch(e.ctx(), b.R.data(), b.oldR.data(), b.P.data(), b.Q.data(), (int)n_, step),
                  e.ctx());
    }
    size_t getTotalNum() const { return n_; }
    void get_rmean(double out[2]) const { diag(); out[0] = d_[0]; out[1] = d_[1]; }
    void get_pmean(double out[2]) const { diag(); out[0] = d_[2]; out[1] = d_[3]; }
    double get_qtotal() const { diag(); return d_[4]; }
    double fieldEnergy() const { diag(); return d_[5]; }
    void getParticles(std::vector<double> &R, std::vector<double> &P, std::vector<double> &oldR) const
    {
        R.resize(2 * n_); P.resize(2 * n_); oldR.resize(2 * n_);
        check(csrst_download_particles(e_.ctx(), R.data(), P.data(), oldR.data(), (int)n_), e_.ctx());
    }
    std::vector<double> getCharges() const
    {
        std::vector<double> Q(n_);
        check(csrst_download_charges(e_.ctx(), Q.data(), (int)n_), e_.ctx());
        return Q;
    }

private:
    void diag() const { check(csrst_diag(e_.ctx(), d_), e_.ctx()); }
    Engine &e_;
    size_t n_;
    mutable double d_[6] = {0, 0, 0, 0, 0, 0};
};

// A node patch of a mesh: nodes [i0, i0 + nx) x [j0, j0 + ny), node (0, 0) of
// the mesh at `origin`, spacing dx.  Patch arrays are [ny][nx] doubles.
struct NodePatch {
    int i0 = 0, j0 = 0, nx = 0, ny = 0;
    double origin[2] = {0.0, 0.0};
    double dx[2] = {1.0, 1.0};
};

// ShapeFunction<IntOp> (ShapeFunction.hh:30-508) on device arrays: ORDER 0
// IntCIC, 1 IntS1, 2 IntS2.  R, oldR: interleaved (x, y) per particle; every
// pointer is device memory; `stream` a hipStream_t (nullptr: the null stream).
template <int ORDER>
struct ShapeFunction {
    static_assert(ORDER >= 0 && ORDER <= 2, "IntCIC, IntS1 or IntS2");
    // getCurrentDensity{X,Y} with its Base part: J over the patch (overwritten)
    static void getCurrentDensity(int comp, double *J, const NodePatch &p, double dt, const double *oldR,
                                  const double *R, const double *Q, int n, void *stream = nullptr)
    {
        check(csrst_sf_current_density(ORDER, comp, J, p.i0, p.j0, p.nx, p.ny, p.origin, p.dx, dt, oldR, R, Q, n,
                                       stream));
    }
    static void getCurrentDensityX(double *J, const NodePatch &p, double dt, const double *oldR, const double *R,
                                   const double *Q, int n, void *stream = nullptr)
    {
        getCurrentDensity(0, J, p, dt, oldR, R, Q, n, stream);
    }
    static void getCurrentDensityY(double *J, const NodePatch &p, double dt, const double *oldR, const double *R,
                                   const double *Q, int n, void *stream = nullptr)
    {
        getCurrentDensity(1, J, p, dt, oldR, R, Q, n, stream);
    }
    // getChargeDensity: rho over the patch (overwritten)
    static void getChargeDensity(double *rho, const NodePatch &p, const double *R, const double *Q, int n,
                                 void *stream = nullptr)
    {
        check(csrst_sf_charge_density(ORDER, rho, p.i0, p.j0, p.nx, p.ny, p.origin, p.dx, R, Q, n, stream));
    }
    // getChargeDensityDiff: rho(R) - rho(oldR)
    static void getChargeDensityDiff(double *rho, const NodePatch &p, const double *R, const double *oldR,
                                     const double *Q, int n, void *stream = nullptr)
    {
        check(csrst_sf_charge_density_diff(ORDER, rho, p.i0, p.j0, p.nx, p.ny, p.origin, p.dx, R, oldR, Q, n,
                                           stream));
    }
};

// ScatterGatherImpl<IntOp, double> (ScatterGatherImpl.hh:27-64) on device
// arrays: a field on the nodes of an nx x ny patch with node (0, 0) at
// `origin`.  Edge fields take comp 0 (x edges) or 1 (y edges); cell fields
// comp -1.  s2bug: the reference's (I+1, J+2) term of the IntS2 gather.
template <int ORDER>
struct ScatterGatherImpl {
    static_assert(ORDER >= 0 && ORDER <= 2, "IntCIC, IntS1 or IntS2");
    static void gather(const double *field, int nx, int ny, const double origin[2], const double dx[2], int comp,
                       const double *R, int n, double *out, bool s2bug = true, void *stream = nullptr)
    {
        check(csrst_sg_gather(ORDER, s2bug ? 1 : 0, comp < 0 ? 2 : comp, field, nx, ny, origin, dx, R, n, out,
                              stream));
    }
    // field += the particles' values times their weights
    static void scatter(double *field, int nx, int ny, const double origin[2], const double dx[2], int comp,
                        const double *R, const double *val, int n, void *stream = nullptr)
    {
        check(csrst_sg_scatter(ORDER, comp < 0 ? 2 : comp, field, nx, ny, origin, dx, R, val, n, stream));
    }
};

// PartBunchState (PartBunch.hh:290-323): what PartBunch::calcBeamParameters
// fills (csrst_beam_stats); d = 0, 1 for x, y
struct PartBunchState {
    double v[CSRST_BEAM_STATS] = {};
    double get_meanEnergy() const { return v[0]; }
    double get_rrms(int d) const { return v[1 + d]; }
    double get_prms(int d) const { return v[3 + d]; }
    double get_norm_emit(int d) const { return v[5 + d]; }
    double get_rmean(int d) const { return v[7 + d]; }
    double get_pmean(int d) const { return v[9 + d]; }
    double get_Dy() const { return v[11]; }
    double get_DDy() const { return v[12]; }
    double get_dE() const { return v[13]; }
    double get_rmin(int d) const { return v[14 + d]; }
    double get_rmax(int d) const { return v[16 + d]; }
};

inline void calcBeamParameters(const Engine &e, PartBunchState &state)
{
    check(csrst_beam_stats(e.ctx(), state.v), e.ctx());
}

// CalcEnergyCommand (CalcEnergy.cpp:60-153): total field energy in nJ, from the
// sums over the global grid and the window (csrst_field_energy); the E term
// averages this call's sum with the previous call's.
class CalcEnergy {
public:
    explicit CalcEnergy(const Engine &e) : e_(e) {}
    double execute()
    {
        const double eps0 = 8.854187817e-12, mu0 = 1.256637061e-06;   // Physics.cpp:41-51
        double sums[2];
        check(csrst_field_energy(e_.ctx(), sums), e_.ctx());
        const double total = (0.25 * eps0 * (sums[0] + eOld_) + 0.5 * mu0 * sums[1]) * 1e9;
        eOld_ = sums[0];
        return total;
    }

private:
    const Engine &e_;
    double eOld_ = 0.0;
};

// DataSinkCommand (DataSink.cpp:22-148): the statistics file in SDDS, one row
// per call, columns as the reference writes them (width 25, 15 digits,
// scientific).  The last column is the CalcEnergy value times 1e9, as
// DataSink.cpp:146 prints it.
class DataSink {
public:
    DataSink(const Engine &e, const std::string &path) : cec_(e), out_(path)
    {
        if (!out_) throw Error(CSRST_EINVAL, "cannot write " + path);
        out_.precision(15);
        out_.setf(std::ios::scientific, std::ios::floatfield);
        header();
    }
    void execute(const Engine &e, double t, PartBunchState &state)
    {
        calcBeamParameters(e, state);
        const double energy = cec_.execute();
        const int w = 25;
        out_ << std::setw(w) << t << std::setw(w) << state.get_meanEnergy();
        for (int d = 0; d < 2; ++d) out_ << std::setw(w) << state.get_rrms(d);
        for (int d = 0; d < 2; ++d) out_ << std::setw(w) << state.get_prms(d);
        for (int d = 0; d < 2; ++d) out_ << std::setw(w) << state.get_norm_emit(d);
        for (int d = 0; d < 2; ++d) out_ << std::setw(w) << state.get_rmean(d);
        for (int d = 0; d < 2; ++d) out_ << std::setw(w) << state.get_pmean(d);
        out_ << std::setw(w) << state.get_Dy() << std::setw(w) << state.get_DDy() << std::setw(w) << state.get_dE()
             << std::setw(w) << energy * 1e9 << std::endl;
    }

private:
    void header()
    {
        struct Col {
            const char *name, *units, *desc;
        };
        static const Col cols[] = {
            {"t,      ", "s", "1 Time"},
            {"energy, ", "MeV", "2 Mean Energy"},
            {"rms_x,  ", "m ", "3 RMS Beamsize in x  "},
            {"rms_y,  ", "m ", "4 RMS Beamsize in y  "},
            {"rms_px, ", "1 ", "5 RMS Momenta in x  "},
            {"rms_py, ", "1 ", "6 RMS Momenta in y  "},
            {"emit_x, ", "m ", "7 Normalized Emittance x  "},
            {"emit_y, ", "m ", "8 Normalized Emittance y  "},
            {"mean_x, ", "m ", "9 Mean Beam Position in x  "},
            {"mean_y, ", "m ", "10 Mean Beam Position in y  "},
            {"mean_px, ", "1 ", "11 Mean Beam Momentum in x  "},
            {"mean_py, ", "1 ", "12 Mean Beam Momentum in y  "},
            {"Dy,     ", "m ", "13 Dispersion in y  "},
            {"DDy,    ", "1 ", "14 Derivative of dispersion in y  "},
            {"dE,     ", "MeV ", "15 energy spread of the beam  "},
            {"Efld,     ", "nJ ", "16 total field energy   "},
        };
        out_ << "SDDS1\n"
             << "&description text=\"Statistics data \", contents=\"stat parameters\" &end\n"
             << "&parameter name=processors, type=long, description=\"Number of Processors\" &end\n"
             << "&parameter name=revision, type=string, description=\"svn revision of CSRST\" &end\n";
        for (const Col &c : cols)
            out_ << "&column name=" << c.name << "type=double, units=" << c.units << ", description=\"" << c.desc
                 << "\" &end\n";
        out_ << "&data mode=ascii &end\n"
             << "Cores used 1\n"
             << "CSRST " << csrst_version() << std::endl;
    }

    CalcEnergy cec_;
    std::ofstream out_;
};

// SaveEMFieldsCommand (SaveEMFields.cpp:34-160): Data/FieldOnLine_NNNN.dat,
// one file per call, numbered from 0; per line s, E along and across the
// mean momentum, Hz, J along it, x, y, tab-separated in the stream's default
// six digits (csrst_line_fields computes them on the GPU).
class SaveEMFields {
public:
    explicit SaveEMFields(const Engine &e) : e_(e) {}
    void execute()
    {
        int n = 0;
        check(csrst_line_fields(e_.ctx(), nullptr, 0, &n), e_.ctx());
        std::vector<double> rows(7 * (size_t)(n > 0 ? n : 1));
        check(csrst_line_fields(e_.ctx(), rows.data(), n, &n), e_.ctx());
        char name[64];
        std::snprintf(name, sizeof(name), "Data/FieldOnLine_%04u.dat", iteration_);
        std::ofstream out(name);
        if (!out) throw Error(CSRST_EINVAL, std::string("cannot write ") + name);
        for (int i = 0; i < n; ++i) {
            const double *r = &rows[7 * (size_t)i];
            out << r[0] << "\t" << r[1] << "\t" << r[2] << "\t" << r[3] << "\t" << r[4] << "\t" << r[5] << "\t"
                << r[6] << "\n";
        }
        ++iteration_;
    }

private:
    const Engine &e_;
    unsigned iteration_ = 0;
};

class TETMSolver {
public:
    explicit TETMSolver(Engine &e) : e_(e) {}
    void updateField_firstStep() { check(csrst_first_step(e_.ctx()), e_.ctx()); }
    void updateField_secondStep() { check(csrst_second_step(e_.ctx()), e_.ctx()); }

private:
    Engine &e_;
};

class PoissonSolver {
public:
    explicit PoissonSolver(Engine &e) : e_(e) {}
    // three solves, fillFields, resetOutsideDomain, window copy
    void computeField(double tol, int maxIterations)
    {
        check(csrst_poisson_compute_field(e_.ctx(), tol, maxIterations, cycles, resid), e_.ctx());
    }
    int cycles[3] = {0, 0, 0};
    double resid[3] = {0, 0, 0};

private:
    Engine &e_;
};

// ---------------------------------------------------------------------------
// inline definitions

// Flags as TCLAP accepts them: "-x 64", "--numgridpx 64", "--numgridpx=64";
// switches -I/--no_initials and -d/--debug.  An unknown flag or a bad value
// throws (the reference prints the TCLAP error and returns EXIT_FAILURE).
inline std::vector<Commands::Opt> Commands::options()
{
    Commands &c = *this;
    return {
        {"r", "restart", 's', &c.restart},         {"n", "startstep", 'i', &c.startstep},
        {"N", "numsteps", 'i', &c.numsteps},       {"x", "numgridpx", 'i', &c.numgridpx},
        {"y", "numgridpy", 'i', &c.numgridpy},     {"m", "plotmodulus", 'i', &c.plotmodulus},
        {"", "statmodulus", 'i', &c.statmodulus},  {"", "linemodulus", 'i', &c.linemodulus},
        {"", "reballance", 'i', &c.reballance},    {"M", "savemodulus", 'i', &c.savemodulus},
        {"P", "phi", 'd', &c.phi},
        {"w", "width", 'd', &c.width},             {"l", "length", 'd', &c.length},
        {"", "lengthbefore", 'd', &c.lengthbefore}, {"E", "Ekin", 'd', &c.Ekin},
        {"B", "Bz", 'd', &c.Bz},                   {"Q", "Qtotal", 'd', &c.Qtotal},
        {"", "mean_Rx", 'd', &c.mean_Rx},          {"", "mean_Ry", 'd', &c.mean_Ry},
        {"s", "sigma_x", 'd', &c.sigma_x},         {"S", "sigma_p", 'd', &c.sigma_p},
        {"", "Np", 'i', &c.Np},                    {"", "h5file", 's', &c.h5file},
        {"I", "no_initials", 'b', &c.no_initials}, {"d", "debug", 'b', &c.debug},
        {"", "order", 'i', &c.order},              {"", "bunch", 's', &c.bunch},
        {"", "wake", 's', &c.wake},                {"", "poissontol", 'd', &c.poissontol},
        {"", "poissonmaxit", 'i', &c.poissonmaxit},
    };
}

inline Commands Commands::parse(int argc, char **argv)
{
    Commands c;
    const std::vector<Opt> opts = c.options();
    for (int k = 1; k < argc; ++k) {
        std::string a = argv[k], val;
        bool hasVal = false;
        const Opt *o = nullptr;
        if (a == "--no-s2bug") { c.s2bug = false; continue; }
        if (a.size() > 2 && a[0] == '-' && a[1] == '-') {
            std::string name = a.substr(2);
            const size_t eq = name.find('=');
            if (eq != std::string::npos) { val = name.substr(eq + 1); name = name.substr(0, eq); hasVal = true; }
            for (const Opt &x : opts) if (name == x.longName) o = &x;
        } else if (a.size() == 2 && a[0] == '-') {
            for (const Opt &x : opts) if (a.substr(1) == x.shortName) o = &x;
        }
        if (!o) throw Error(CSRST_EINVAL, "unknown argument: " + a);
        c.given.push_back(o->longName);
        if (o->kind == 'b') { *(bool *)o->dst = true; continue; }
        if (!hasVal) {
            if (k + 1 >= argc) throw Error(CSRST_EINVAL, "missing value for " + a);
            val = argv[++k];
        }
        char *end = nullptr;
        if (o->kind == 'i') {
            const long v = std::strtol(val.c_str(), &end, 10);
            if (!end || *end) throw Error(CSRST_EINVAL, "not an integer for " + a + ": " + val);
            *(int *)o->dst = (int)v;
        } else if (o->kind == 'd') {
            const double v = std::strtod(val.c_str(), &end);
            if (!end || *end) throw Error(CSRST_EINVAL, "not a number for " + a + ": " + val);
            *(double *)o->dst = v;
        } else {
            *(std::string *)o->dst = val;
        }
    }
    if (c.order < 0 || c.order > 2) throw Error(CSRST_EINVAL, "--order must be 0, 1 or 2");
    return c;
}

inline Bunch Distribution::generate(const Commands &c, const double hr[2], const double origin[2])
{
    const int n = c.Np;
    if (n <= 0) throw Error(CSRST_EINVAL, "--Np must be positive");
    const double twoPi = 2.0 * 3.14159265358979323846;
    const double mass = 0.51099906e-03 * 1e3;              // Distribution.cpp:26, MeV
    const double gamma = c.Ekin / mass + 1;
    const double meanP = std::sqrt(gamma * gamma - 1);
    const double qEach = (-c.Qtotal * 1e-9) / n;
    Bunch b;
    b.R.resize(2 * (size_t)n);
    b.P.resize(2 * (size_t)n);
    b.Q.assign(n, qEach);
    auto sample = [&](unsigned long long k, double sigma, double out[2]) {
        const double v0 = radical_inverse(k, 2), v1 = radical_inverse(k, 3);
        const double dist = sigma * std::sqrt(-2 * std::log(v0));
        const double angle = twoPi * v1;
        out[0] = dist * std::cos(angle);
        out[1] = dist * std::sin(angle);
    };
    for (int i = 0; i < n; ++i) {
        double x[2], p[2];
        sample((unsigned long long)i + 1, c.sigma_x, x);   // one Halton stream each
        sample((unsigned long long)i + 1, c.sigma_p, p);
        b.R[2 * i] = c.mean_Rx + x[0];
        b.R[2 * i + 1] = c.mean_Ry + x[1];
        b.P[2 * i] = meanP + p[0];
        b.P[2 * i + 1] = 0.0 + p[1];
    }
    // Utils::XFirst (utils.hh:116-140), applied to R only: by y cell, then by
    // x value.  XFirst returns true for equal x (not a strict weak order);
    // equal x never occurs for these samples, so ">" orders the same.
    std::vector<std::array<double, 2>> r(n);
    for (int i = 0; i < n; ++i) r[i] = {b.R[2 * i], b.R[2 * i + 1]};
    std::sort(r.begin(), r.end(), [&](const std::array<double, 2> &u, const std::array<double, 2> &v) {
        const int J1 = (int)((u[1] - origin[1]) / hr[1]), J2 = (int)((v[1] - origin[1]) / hr[1]);
        if (J2 > J1) return true;
        if (J2 < J1) return false;
        return v[0] > u[0];
    });
    for (int i = 0; i < n; ++i) { b.R[2 * i] = r[i][0]; b.R[2 * i + 1] = r[i][1]; }
    return b;
}

}  // namespace csrst
