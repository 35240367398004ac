// The C++ front end's shims of the reference interfaces that sit on device
// arrays and caller geometry, against the CPU oracle:
//   ShapeFunction<ORDER>   getCurrentDensity{X,Y}, getChargeDensity(Diff)
//   ScatterGatherImpl<ORDER> gather (x edges, y edges, cells), scatter
//   Engine::setGeometry    the Bend's own mask and SCBoundaryCondition cells
//                          handed back; Engine::setStream on a caller stream
// Every comparison is bit for bit.  Prints one line per check and "all ok";
// exits 1 at the first difference.  Built by __graft_entry__.build() into
// tests/shim_check (run by tests/test_shim_gpu.py).
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "../csrst_amd/host/csrst.hpp"
#include "../oracle/csrst_oracle.h"

namespace {

int failures = 0;

void expect(bool ok, const char *what)
{
    std::printf("%s: %s\n", what, ok ? "ok" : "DIFFERS");
    if (!ok) ++failures;
}

bool same(const std::vector<double> &a, const std::vector<double> &b)
{
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(double)) == 0;
}

// 64-bit LCG, uniform in [0, 1)
struct Rng {
    unsigned long long s;
    double next()
    {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return ((s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    }
};

struct Dev {
    double *p = nullptr;
    size_t n = 0;
    explicit Dev(size_t count) : n(count)
    {
        if (hipMalloc((void **)&p, (count ? count : 1) * sizeof(double)) != hipSuccess)
            throw std::runtime_error("hipMalloc");
    }
    Dev(const std::vector<double> &h) : Dev(h.size()) { put(h); }
    ~Dev() { (void)hipFree(p); }
    void put(const std::vector<double> &h)
    {
        if (hipMemcpy(p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("hipMemcpy");
    }
    std::vector<double> get() const
    {
        std::vector<double> h(n);
        if (hipMemcpy(h.data(), p, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("hipMemcpy");
        return h;
    }
};

template <int ORDER>
void shape_function_checks()
{
    Rng rng{0x1234u + ORDER};
    const int n = 3000;
    const double dxy[2] = {0.7, 1.3}, org[2] = {0.0, 0.0}, dt = 0.37;
    std::vector<double> oldR(2 * n), R(2 * n), Q(n);
    for (int k = 0; k < n; ++k) {
        for (int d = 0; d < 2; ++d) {
            oldR[2 * k + d] = (8.0 + 12.0 * rng.next()) * dxy[d];
            R[2 * k + d] = oldR[2 * k + d] + (1.98 * rng.next() - 0.99) * dxy[d];
        }
        Q[k] = 2.0 * rng.next() - 1.0;
    }
    csrst::NodePatch p;
    p.i0 = -6; p.j0 = -6; p.nx = 40; p.ny = 40;
    p.origin[0] = org[0]; p.origin[1] = org[1]; p.dx[0] = dxy[0]; p.dx[1] = dxy[1];
    const size_t nodes = (size_t)p.nx * p.ny;
    Dev dOld(oldR), dR(R), dQ(Q), dJ(nodes);
    char what[96];
    for (int comp = 0; comp < 2; ++comp) {
        if (comp == 0) csrst::ShapeFunction<ORDER>::getCurrentDensityX(dJ.p, p, dt, dOld.p, dR.p, dQ.p, n);
        else csrst::ShapeFunction<ORDER>::getCurrentDensityY(dJ.p, p, dt, dOld.p, dR.p, dQ.p, n);
        std::vector<double> ref(nodes, 0.0);
        const int rc = orc_deposit(ORDER, comp, ref.data(), p.i0, p.j0, p.nx, p.ny, org, dxy, dt, oldR.data(),
                                   R.data(), Q.data(), n);
        std::snprintf(what, sizeof what, "ShapeFunction<%d>::getCurrentDensity%c", ORDER, comp ? 'Y' : 'X');
        expect(rc == 0 && same(dJ.get(), ref), what);
    }
    csrst::ShapeFunction<ORDER>::getChargeDensity(dJ.p, p, dR.p, dQ.p, n);
    std::vector<double> r1(nodes, 0.0), r0(nodes, 0.0);
    const int rc1 = orc_charge_density(ORDER, r1.data(), p.i0, p.j0, p.nx, p.ny, org, dxy, R.data(), Q.data(), n);
    std::snprintf(what, sizeof what, "ShapeFunction<%d>::getChargeDensity", ORDER);
    expect(rc1 == 0 && same(dJ.get(), r1), what);
    // the difference against the two densities: equal to round-off
    csrst::ShapeFunction<ORDER>::getChargeDensityDiff(dJ.p, p, dR.p, dOld.p, dQ.p, n);
    orc_charge_density(ORDER, r0.data(), p.i0, p.j0, p.nx, p.ny, org, dxy, oldR.data(), Q.data(), n);
    const std::vector<double> diff = dJ.get();
    double err = 0.0, scale = 0.0;
    for (size_t m = 0; m < nodes; ++m) {
        err = std::max(err, std::fabs(diff[m] - (r1[m] - r0[m])));
        scale = std::max(scale, std::fabs(r1[m]));
    }
    std::snprintf(what, sizeof what, "ShapeFunction<%d>::getChargeDensityDiff (to 1e-13)", ORDER);
    expect(err <= 1e-13 * scale, what);

    // ScatterGatherImpl on a 30 x 40 node field
    const int fx = 40, fy = 30;
    std::vector<double> field((size_t)fx * fy);
    for (double &v : field) v = 2.0 * rng.next() - 1.0;
    const double forg[2] = {0.013, -0.021};
    std::vector<double> G(2 * 500);
    for (int k = 0; k < 500; ++k) {
        G[2 * k] = forg[0] + (4.0 + 30.0 * rng.next()) * dxy[0];
        G[2 * k + 1] = forg[1] + (4.0 + 20.0 * rng.next()) * dxy[1];
    }
    Dev dF(field), dG(G), dOut(500);
    for (int comp = -1; comp < 2; ++comp)
        for (int bug = 0; bug < 2; ++bug) {
            csrst::ScatterGatherImpl<ORDER>::gather(dF.p, fx, fy, forg, dxy, comp, dG.p, 500, dOut.p, bug != 0);
            std::vector<double> ref(500);
            if (comp < 0) orc_gather_cell(ORDER, bug, field.data(), fx, fy, forg, dxy, G.data(), 500, ref.data());
            else orc_gather_edge(ORDER, comp, bug, field.data(), fx, fy, forg, dxy, G.data(), 500, ref.data());
            std::snprintf(what, sizeof what, "ScatterGatherImpl<%d>::gather comp %d s2bug %d", ORDER, comp, bug);
            expect(same(dOut.get(), ref), what);
        }
    std::vector<double> val(500);
    for (double &v : val) v = 2.0 * rng.next() - 1.0;
    Dev dV(val);
    for (int comp = -1; comp < 2; ++comp) {
        std::vector<double> ref(field.size(), 0.0);
        Dev dS(ref);
        csrst::ScatterGatherImpl<ORDER>::scatter(dS.p, fx, fy, forg, dxy, comp, dG.p, dV.p, 500);
        if (comp < 0) orc_scatter_cell(ORDER, ref.data(), fx, fy, forg, dxy, G.data(), val.data(), 500);
        else orc_scatter_edge(ORDER, comp, ref.data(), fx, fy, forg, dxy, G.data(), val.data(), 500);
        std::snprintf(what, sizeof what, "ScatterGatherImpl<%d>::scatter comp %d", ORDER, comp);
        expect(same(dS.get(), ref), what);
    }
}

void geometry_and_stream_checks()
{
    csrst::Commands cmd;
    cmd.numgridpx = 1100; cmd.numgridpy = 120;
    cmd.lengthbefore = 0.05; cmd.length = 0.01; cmd.width = 0.0128; cmd.phi = 10.0;
    cmd.Ekin = 50.59; cmd.Bz = 0.1704; cmd.order = 1; cmd.statmodulus = 1;
    csrst::Engine e1(cmd), e2(cmd);
    csrst::Bend b1(e1);
    e2.setGeometry(b1.getInsideMask(), b1.getBoundaryCells());
    hipStream_t s = nullptr;
    if (hipStreamCreate(&s) != hipSuccess) throw std::runtime_error("hipStreamCreate");
    e2.setStream(s);
    // a bunch 4 mm before the bend, moving along x (gamma 100)
    csrst::Bunch b;
    Rng rng{77};
    const int n = 1500;
    for (int k = 0; k < n; ++k) {
        b.R.push_back(0.046 + 1e-3 * (rng.next() - 0.5));
        b.R.push_back(0.0064 + 0.5e-3 * (rng.next() - 0.5));
        b.P.push_back(99.995 + 0.01 * rng.next());
        b.P.push_back(0.0);
        b.Q.push_back(-1e-9 / n);
    }
    csrst::PartBunch p1(e1, b), p2(e2, b);
    csrst::TETMSolver s1(e1), s2(e2);
    for (int k = 0; k < 20; ++k) {
        s1.updateField_firstStep(); s1.updateField_secondStep(); e1.moveMesh();
        s2.updateField_firstStep(); s2.updateField_secondStep(); e2.moveMesh();
    }
    e2.setStream(nullptr);
    (void)hipStreamDestroy(s);
    bool ok = true;
    for (int which = 0; which < 6; ++which) ok = ok && same(e1.field(which), e2.field(which));
    std::vector<double> R1, P1, O1, R2, P2, O2;
    p1.getParticles(R1, P1, O1);
    p2.getParticles(R2, P2, O2);
    expect(ok && same(R1, R2) && same(P1, P2), "Engine::setGeometry + setStream: 20 steps bit-identical");
}

}  // namespace

int main()
{
    try {
        shape_function_checks<0>();
        shape_function_checks<1>();
        shape_function_checks<2>();
        geometry_and_stream_checks();
    } catch (const std::exception &e) {
        std::printf("error: %s\n", e.what());
        return 2;
    }
    if (failures) return 1;
    std::printf("all ok\n");
    return 0;
}
