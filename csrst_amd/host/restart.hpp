// Restart files of the driver: SaveH5Part (SaveH5Part.cpp:79-428) and
// LoadH5Part (LoadH5Part.cpp:60-321) over the records of checkpoint.hpp.
// An H5Part file goes through the csrst_h5part tool next to the driver,
// started as a child process, so the driver itself never links HDF5.
#pragma once

#include <spawn.h>
#include <sys/wait.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "checkpoint.hpp"
#include "csrst.hpp"

extern char **environ;

namespace csrst {
namespace restart {

inline bool ends_with(const std::string &s, const std::string &t)
{
    return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

// H5Part files by name, as the reference's default Output.h5part
inline bool is_h5part(const std::string &path) { return ends_with(path, ".h5part") || ends_with(path, ".h5"); }

inline std::string tool_path()
{
    char buf[4096];
    const ssize_t l = ::readlink("/proc/self/exe", buf, sizeof buf - 1);
    if (l <= 0) throw Error(CSRST_EINVAL, "cannot locate the driver executable");
    buf[l] = 0;
    std::string dir(buf);
    dir = dir.substr(0, dir.find_last_of('/'));
    return dir + "/csrst_h5part";
}

// csrst_h5part MODE A B as a child process; throws unless it exits with 0
inline void run_tool(const char *mode, const std::string &a, const std::string &b)
{
    const std::string tool = tool_path();
    if (::access(tool.c_str(), X_OK) != 0) throw Error(CSRST_EINVAL, "H5Part files need " + tool);
    std::vector<std::string> args = {tool, mode, a, b};
    std::vector<char *> argv;
    for (std::string &s : args) argv.push_back(&s[0]);
    argv.push_back(nullptr);
    pid_t pid;
    if (::posix_spawn(&pid, tool.c_str(), nullptr, nullptr, argv.data(), environ) != 0)
        throw Error(CSRST_EINVAL, "cannot start " + tool);
    int st = 0;
    if (::waitpid(pid, &st, 0) < 0 || !WIFEXITED(st) || WEXITSTATUS(st) != 0)
        throw Error(CSRST_EINVAL, std::string("csrst_h5part ") + mode + " failed for " + a);
}

// LoadH5Part::openFile + the last step (LoadH5Part.cpp:48-58)
inline ckpt::Record load(const std::string &path)
{
    if (!is_h5part(path)) return ckpt::read_all(path).back();
    const std::string tmp = path + ".rs.tmp" + std::to_string(::getpid());
    run_tool("read", path, tmp);
    ckpt::Record r = ckpt::read_all(tmp).back();
    std::remove(tmp.c_str());
    return r;
}

// LoadH5Part::readH5FileAttributes (LoadH5Part.cpp:60-159): an argument set
// on the command line keeps its value unless it is numgridpx, numgridpy, Np
// or phi; the others take the file's.  startstep becomes
// ((startstep + numsteps) / savemodulus) * savemodulus of the file's values.
inline void apply(Commands &c, const ckpt::Record &r)
{
    const std::vector<std::string> always = {"numgridpx", "numgridpy", "Np", "phi"};
    for (const Commands::Opt &o : c.options()) {
        const std::string name = o.longName;
        const bool forced = std::find(always.begin(), always.end(), name) != always.end();
        if (c.isSet(name) && !forced) continue;
        if (name == "startstep") {
            const ckpt::Attr *s = r.attr("startstep"), *n = r.attr("numsteps"), *m = r.attr("savemodulus");
            if (s && n && m && s->kind == 'i' && n->kind == 'i' && m->kind == 'i' && m->i > 0)
                c.startstep = (int)(((s->i + n->i) / m->i) * m->i);
            continue;
        }
        const ckpt::Attr *a = r.attr(name);
        if (!a) continue;
        if (o.kind == 's' && a->kind == 's') *(std::string *)o.dst = a->s;
        else if (o.kind == 'd' && a->kind == 'd') *(double *)o.dst = a->d;
        else if (o.kind == 'i' && a->kind == 'i') *(int *)o.dst = (int)a->i;
        else if (o.kind == 'b' && a->kind == 'i') *(bool *)o.dst = a->i == 1;
    }
}

// the arguments as SaveH5Part::writeH5FileAttributes stores them
// (SaveH5Part.cpp:288-352): strings, float64, int64 (switches 0/1)
inline std::vector<ckpt::Attr> attributes(Commands c)
{
    std::vector<ckpt::Attr> out;
    for (const Commands::Opt &o : c.options()) {
        ckpt::Attr a;
        a.name = o.longName;
        switch (o.kind) {
        case 's': a.kind = 's'; a.s = *(const std::string *)o.dst; break;
        case 'd': a.kind = 'd'; a.d = *(const double *)o.dst; break;
        case 'i': a.kind = 'i'; a.i = *(const int *)o.dst; break;
        default: a.kind = 'i'; a.i = *(const bool *)o.dst ? 1 : 0; break;
        }
        out.push_back(a);
    }
    return out;
}

// the record SaveH5Part::write stores (SaveH5Part.cpp:79-259): the global
// E, H, J over the domain cells, the bunch with its oldR; id = the particle's
// place in the engine's order (the engine keeps no separate particle ids)
inline ckpt::Record snapshot(const Engine &e, const PartBunch &b, double t, int step)
{
    const Commands &c = e.commands();
    ckpt::Record r;
    r.Nx = c.numgridpx; r.Ny = c.numgridpy; r.step = step; r.time = t;
    r.hx = e.hr()[0]; r.hy = e.hr()[1]; r.ox = e.origin()[0]; r.oy = e.origin()[1];
    const size_t nx = (size_t)r.Nx, ny = (size_t)r.Ny, pitch = nx + 2, plane = pitch * (ny + 2);
    const int which[3] = {0, 1, 5};
    std::vector<double> *dst[3] = {&r.E, &r.H, &r.J};
    for (int f = 0; f < 3; ++f) {
        const std::vector<double> g = e.field(which[f]);
        if (g.size() != 2 * plane) throw Error(CSRST_EINVAL, "field size");
        dst[f]->resize(2 * nx * ny);
        for (size_t k = 0; k < 2; ++k)
            for (size_t j = 0; j < ny; ++j)
                for (size_t i = 0; i < nx; ++i)
                    (*dst[f])[k * nx * ny + j * nx + i] = g[k * plane + (j + 1) * pitch + i + 1];
    }
    std::vector<double> R, P, O;
    b.getParticles(R, P, O);
    const size_t n = b.getTotalNum();
    std::vector<double> Q = b.getCharges();
    for (std::vector<double> *v : {&r.x, &r.y, &r.oldx, &r.oldy, &r.px, &r.py, &r.q}) v->resize(n);
    r.id.resize(n);
    for (size_t k = 0; k < n; ++k) {
        r.x[k] = R[2 * k]; r.y[k] = R[2 * k + 1];
        r.oldx[k] = O[2 * k]; r.oldy[k] = O[2 * k + 1];
        r.px[k] = P[2 * k]; r.py[k] = P[2 * k + 1];
        r.q[k] = Q[k];
        r.id[k] = (int64_t)k;
    }
    r.attrs = attributes(c);
    return r;
}

// LoadH5Part::restoreFields (LoadH5Part.cpp:161-240) into the engine's global
// E and H (J is zeroed before the first step, main.cpp:388) and the bunch of
// restoreBunch (LoadH5Part.cpp:242-321)
inline Bunch restore_fields(Engine &e, const ckpt::Record &r)
{
    const Commands &c = e.commands();
    if (r.Nx != c.numgridpx || r.Ny != c.numgridpy) throw Error(CSRST_EINVAL, "restart file grid size");
    // restoreFields takes the file's mesh spacing and origin (LoadH5Part.cpp:223-233);
    // the engine's mesh comes from the bend arguments, which the command line may
    // have changed: the stored fields only fit the mesh they were written on
    auto near = [](double a, double b) { return std::fabs(a - b) <= 1e-12 * std::max(std::fabs(a), std::fabs(b)); };
    if (!near(r.hx, e.hr()[0]) || !near(r.hy, e.hr()[1]) || !near(r.ox, e.origin()[0]) || !near(r.oy, e.origin()[1]))
        throw Error(CSRST_EINVAL, "restart file mesh (spacing and origin) differs from the mesh of the bend "
                                  "arguments; give the arguments the file was written with");
    const size_t nx = (size_t)r.Nx, ny = (size_t)r.Ny, pitch = nx + 2, plane = pitch * (ny + 2);
    const std::vector<double> *src[2] = {&r.E, &r.H};
    for (int f = 0; f < 2; ++f) {
        std::vector<double> g(2 * plane, 0.0);
        for (size_t k = 0; k < 2; ++k)
            for (size_t j = 0; j < ny; ++j)
                for (size_t i = 0; i < nx; ++i)
                    g[k * plane + (j + 1) * pitch + i + 1] = (*src[f])[k * nx * ny + j * nx + i];
        e.setField(f, g);
    }
    Bunch b;
    const size_t n = r.n();
    b.R.resize(2 * n); b.P.resize(2 * n); b.oldR.resize(2 * n); b.Q = r.q;
    for (size_t k = 0; k < n; ++k) {
        b.R[2 * k] = r.x[k]; b.R[2 * k + 1] = r.y[k];
        b.oldR[2 * k] = r.oldx[k]; b.oldR[2 * k + 1] = r.oldy[k];
        b.P[2 * k] = r.px[k]; b.P[2 * k + 1] = r.py[k];
    }
    return b;
}

}  // namespace restart
}  // namespace csrst
