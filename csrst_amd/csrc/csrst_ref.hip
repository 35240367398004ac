// csrst_ref.hip -- the reference's two-grid PIC time step on gfx950, FP64.
//
// Global chamber grid (masked Yee sweeps + conformal boundary cells) and the
// bunch-following window grid (unmasked sweeps, particles), exchanged along
// the window edges, as in main.cpp:390-506 and TETMSolver.hh:110-266.
//
// Arithmetic follows the reference order exactly and the file is compiled
// with -ffp-contract=off, so every field and particle value is bit-identical
// to a plain CPU evaluation of the reference formulas.  The one place where
// summation order matters -- the current deposit, where the reference adds
// particle contributions to each grid node in particle order -- is made
// deterministic: contributions are emitted in (particle, segment, node)
// order, stably radix-sorted by node, and each node's run is summed
// sequentially from 0.0 by one thread.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "csrst_hip.h"
#include "bend_geom.h"
#include "hip_util.h"
#include "pic_math.h"
#include "poisson_mg.h"
#include "beam_stats.h"

using hu::DBuf;
using hu::Error;
using pm::vec2;

namespace {

constexpr int kMaxSeg = 4;     // segments per particle per push (dt = hx/c keeps it <= 3)
constexpr int kMaxNodes = 12;  // nodes per segment (S2)
constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// A 2-component field over [-1, n] x [-1, m]: two planes, pitch n+2.
struct FieldDesc {
    double *c0, *c1;
    int n, m;
};
__host__ __device__ inline size_t fidx(const FieldDesc &f, int i, int j)
{
    return (size_t)(j + 1) * (f.n + 2) + (i + 1);
}

enum FieldId { F_GE = 0, F_GH = 1, F_ME = 2, F_MH = 3, F_MJ = 4, F_GJ = 5, F_COUNT = 6 };
struct FieldSet {
    FieldDesc f[F_COUNT];
};

// Shape-function margins, *OrderShapeFunction.hh
struct Margins {
    int edge[2][2];
    int cell[2];
    int cur[2][2];
};
Margins margins(int order)
{
    Margins m;
    if (order == 0) {
        const int e[2][2] = {{-1, 1}, {0, 1}}, c[2][2] = {{0, 0}, {0, 1}};
        memcpy(m.edge, e, sizeof(e)); memcpy(m.cur, c, sizeof(c));
        m.cell[0] = -1; m.cell[1] = 1;
    } else if (order == 1) {
        const int e[2][2] = {{-1, 1}, {-1, 2}}, c[2][2] = {{-1, 1}, {-1, 2}};
        memcpy(m.edge, e, sizeof(e)); memcpy(m.cur, c, sizeof(c));
        m.cell[0] = -1; m.cell[1] = 1;
    } else {
        const int e[2][2] = {{-2, 2}, {-1, 2}}, c[2][2] = {{-1, 1}, {-1, 2}};
        memcpy(m.edge, e, sizeof(e)); memcpy(m.cur, c, sizeof(c));
        m.cell[0] = -2; m.cell[1] = 2;
    }
    return m;
}

// Bounding box accumulator: ord-encoded min x, min y, max x, max y.
struct BBox {
    unsigned long long v[4];
};

// Box of R (and oldR) of the workgroup's particles into *bb.  Called by every
// thread of the workgroup (blockDim.x <= kBlock, a multiple of 64).
__device__ inline void bbox_push(BBox *bb, double x, double y, double ox, double oy, bool withOld, bool live)
{
    double mnx = x, mny = y, mxx = x, mxy = y;
    if (!live) {
        mnx = mny = __longlong_as_double(0x7ff0000000000000ll);   // +inf
        mxx = mxy = -mnx;
    } else if (withOld) {
        mnx = ox < mnx ? ox : mnx; mny = oy < mny ? oy : mny;
        mxx = ox > mxx ? ox : mxx; mxy = oy > mxy ? oy : mxy;
    }
    unsigned long long a = hu::ord(mnx), b = hu::ord(mny), c = hu::ord(mxx), d = hu::ord(mxy);
    a = hu::wave_min(a); b = hu::wave_min(b); c = hu::wave_max(c); d = hu::wave_max(d);
    // one set of atomics per workgroup: atomics on one address serialise
    __shared__ unsigned long long part[4][kBlock / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[0][w] = a; part[1][w] = b; part[2][w] = c; part[3][w] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            a = min(a, part[0][k]); b = min(b, part[1][k]);
            c = max(c, part[2][k]); d = max(d, part[3][k]);
        }
        atomicMin(&bb->v[0], a); atomicMin(&bb->v[1], b);
        atomicMax(&bb->v[2], c); atomicMax(&bb->v[3], d);
    }
}

__global__ void k_bbox_reset(BBox *bb)
{
    bb->v[0] = ~0ull; bb->v[1] = ~0ull; bb->v[2] = 0ull; bb->v[3] = 0ull;
}

// ---------------------------------------------------------------------------
// Particles (SoA)
struct Parts {
    double *rx, *ry, *ox, *oy, *px, *py, *q;
    int n;
};

// push_particles, first loop: oldR = R; R += (0.5 dt c) P / gamma.  PartBunch.cpp:226-232
// Grid-stride over the particles (a capped grid: fewer workgroups, fewer
// same-address box atomics); the box of each thread's particles is kept in
// registers and pushed once.
__global__ void k_drift_half(Parts p, double half, BBox *bb)
{
    const double inf = __longlong_as_double(0x7ff0000000000000ll);
    double mnx = inf, mny = inf, mxx = -inf, mxy = -inf;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < p.n; k += gridDim.x * blockDim.x) {
        double x = p.rx[k], y = p.ry[k];
        p.ox[k] = x; p.oy[k] = y;
        const vec2<double> P = pm::mk(p.px[k], p.py[k]);
        const double g = pm::inv_gamma(P);
        x += half * P.x * g;
        y += half * P.y * g;
        p.rx[k] = x; p.ry[k] = y;
        mnx = fmin(mnx, x); mny = fmin(mny, y);
        mxx = fmax(mxx, x); mxy = fmax(mxy, y);
    }
    const bool any = mnx <= mxx;
    bbox_push(bb, any ? mnx : 0, any ? mny : 0, any ? mxx : 0, any ? mxy : 0, true, any);
}

// drift_particles, PartBunch.cpp:179-187
__global__ void k_drift(Parts p, double full)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.n) return;
    const vec2<double> P = pm::mk(p.px[k], p.py[k]);
    const double g = pm::inv_gamma(P);
    p.rx[k] += full * P.x * g;
    p.ry[k] += full * P.y * g;
}

struct PushArgs {
    double mmo[2];   // window origin
    double hr[2];
    int fEx[2], fEy[2], fHz[2];   // margin offsets of the patches
    double half, fact1, fact2;
    int order, s2bug;
};

// Gather from the window patches, external field, kick, rotation and second
// half drift of push_particles (PartBunch.cpp:237-383); patch origins follow
// getFields (PartBunch.cpp:957-1074).
__global__ void k_push_kick(Parts p, FieldSet fs, geom::Bend bend, PushArgs a, const BBox *bb, BBox *bbOld,
                            int *err)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = k < p.n;
    const FieldDesc mE = fs.f[F_ME], mH = fs.f[F_MH];
    // bunch domain of R (getLocalPDomain): floor((Rmin - origin) / dx)
    const double rminx = hu::unord(bb->v[0]), rminy = hu::unord(bb->v[1]);
    const int pd0 = (int)floor((rminx - a.mmo[0]) / a.hr[0]);
    const int pd1 = (int)floor((rminy - a.mmo[1]) / a.hr[1]);
    const double po0 = a.mmo[0] + pd0 * a.hr[0], po1 = a.mmo[1] + pd1 * a.hr[1];
    const vec2<double> dx = pm::mk(a.hr[0], a.hr[1]);
    const vec2<double> oEx = pm::mk(po0 + a.fEx[0] * dx.x, po1 + a.fEx[1] * dx.y);
    const vec2<double> oEy = pm::mk(po0 + a.fEy[0] * dx.x, po1 + a.fEy[1] * dx.y);
    const vec2<double> oHz = pm::mk(po0 + a.fHz[0] * dx.x, po1 + a.fHz[1] * dx.y);
    const int bEx0 = pd0 + a.fEx[0], bEx1 = pd1 + a.fEx[1];
    const int bEy0 = pd0 + a.fEy[0], bEy1 = pd1 + a.fEy[1];
    const int bHz0 = pd0 + a.fHz[0], bHz1 = pd1 + a.fHz[1];
    double x = 0, y = 0, ox = 0, oy = 0;
    if (live) {
        x = p.rx[k]; y = p.ry[k];
        ox = p.ox[k]; oy = p.oy[k];
        const vec2<double> r = pm::mk(x, y);
        const vec2<double> od = pm::mk(1.0 / dx.x, 1.0 / dx.y);
        bool bad = false;
        auto fEX = [&](long long I, long long J) {
            const int i = bEx0 + (int)I, j = bEx1 + (int)J;
            if (i < -1 || i > mE.n || j < -1 || j > mE.m) { bad = true; return 0.0; }
            return mE.c0[fidx(mE, i, j)];
        };
        auto fEY = [&](long long I, long long J) {
            const int i = bEy0 + (int)I, j = bEy1 + (int)J;
            if (i < -1 || i > mE.n || j < -1 || j > mE.m) { bad = true; return 0.0; }
            return mE.c1[fidx(mE, i, j)];
        };
        auto fHZ = [&](long long I, long long J) {
            const int i = bHz0 + (int)I, j = bHz1 + (int)J;
            if (i < -1 || i > mH.n || j < -1 || j > mH.m) { bad = true; return 0.0; }
            const size_t q = fidx(mH, i, j);
            return mH.c0[q] + mH.c1[q];
        };
        const vec2<double> pEx = pm::mk((r.x - oEx.x) * od.x, (r.y - oEx.y) * od.y);
        const vec2<double> pEy = pm::mk((r.x - oEy.x) * od.x, (r.y - oEy.y) * od.y);
        const vec2<double> pHz = pm::mk((r.x - oHz.x) * od.x, (r.y - oHz.y) * od.y);
        const double ex = pm::gather(pm::edge_stencil(a.order, 0, pEx), a.s2bug != 0, fEX);
        const double ey = pm::gather(pm::edge_stencil(a.order, 1, pEy), a.s2bug != 0, fEY);
        const double hf = pm::gather(pm::cell_stencil(a.order, pHz), a.s2bug != 0, fHZ);
        if (bad) atomicOr(err, 1);
        const double bext = geom::external_bz(bend, x, y);
        const vec2<double> P = pm::kick_rotate(pm::mk(p.px[k], p.py[k]), ex, ey, hf, bext, a.fact1, a.fact2);
        p.px[k] = P.x; p.py[k] = P.y;
        const double g = pm::inv_gamma(P);
        x += a.half * P.x * g;
        y += a.half * P.y * g;
        p.rx[k] = x; p.ry[k] = y;
    }
    bbox_push(bbOld, x, y, ox, oy, true, live);
}

struct DepArgs {
    double mmo[2], hr[2], dt;
    int order, comp, W, H;
    int slots;   // nodes per segment of the order's current stencils (2, 6, 12)
};

// nodes of one segment's current stencil: S0 2, S1 6, S2 12
inline int seg_nodes(int order) { return order == 0 ? 2 : order == 1 ? 6 : kMaxNodes; }

// Current of every (particle, segment, node) into fixed slots; keys are
// window-interior node indices, nodes outside the window are dropped
// (getCurrentDensityImpl writes only the field domain).
__global__ void k_emit(Parts p, DepArgs a, unsigned *keys, double *vals, int *err)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.n) return;
    unsigned *kk = keys + (size_t)k * kMaxSeg * a.slots;
    double *vv = vals + (size_t)k * kMaxSeg * a.slots;
    for (int s = 0; s < kMaxSeg * a.slots; ++s) kk[s] = 0xffffffffu;
    const vec2<double> org = pm::mk(a.mmo[0], a.mmo[1]);
    const vec2<double> dx = pm::mk(a.hr[0], a.hr[1]);
    const vec2<double> o = pm::mk(p.ox[k] - org.x, p.oy[k] - org.y);
    const vec2<double> e = pm::mk(p.rx[k] - org.x, p.ry[k] - org.y);
    const vec2<double> od = pm::mk(1.0 / dx.x, 1.0 / dx.y);
    vec2<double> odt;
    if (a.order == 0) odt = pm::mk(1.0 / (dx.y * a.dt), 1.0 / (dx.x * a.dt));
    else odt = pm::mk(od.x / a.dt, od.y / a.dt);
    const vec2<double> D = pm::mk(p.rx[k] - p.ox[k], p.ry[k] - p.oy[k]);
    const vec2<double> delta = pm::mk(D.x * od.x, D.y * od.y);
    const double covered = sqrt(D.x * D.x + D.y * D.y);
    pm::Crossings<double> c;
    if (!pm::crossings(c, o, e, dx, a.order == 1)) { atomicOr(err, 2); return; }
    if (c.n - 1 > kMaxSeg) { atomicOr(err, 2); return; }
    const double q = p.q[k];
    for (int s = 0; s < c.n - 1; ++s) {
        const vec2<double> xi = pm::mk(c.s[s].x * od.x, c.s[s].y * od.y);
        const vec2<double> xf = pm::mk(c.s[s + 1].x * od.x, c.s[s + 1].y * od.y);
        const vec2<double> seg = pm::mk(c.s[s + 1].x - c.s[s].x, c.s[s + 1].y - c.s[s].y);
        pm::SegCurrent<double> sc;
        pm::segment_current(sc, a.order, a.comp, xi, xf, seg, covered, q, odt, delta);
        if (sc.nx * sc.ny > a.slots) { atomicOr(err, 2); return; }
        int t = 0;
        for (int jj = 0; jj < sc.ny; ++jj)
            for (int ii = 0; ii < sc.nx; ++ii, ++t) {
                const int I = sc.I0 + ii, J = sc.J0 + jj;
                const int slot = s * a.slots + t;
                if (I >= 0 && I < a.W && J >= 0 && J < a.H) {
                    kk[slot] = (unsigned)(J * a.W + I);
                    vv[slot] = sc.w[t];
                }
            }
    }
}

// J(c) = J(c) - J(c) over the window interior (getCurrentDensityImpl: the
// component is cleared this way before the patch is written).
__global__ void k_clear_comp(FieldDesc f, int comp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= f.n) return;
    double *c = comp == 0 ? f.c0 : f.c1;
    const size_t q = fidx(f, i, j);
    c[q] = c[q] - c[q];
}

// One thread per run head: sequential sum from 0.0, in particle order.
__global__ void k_reduce_runs(const unsigned *keys, const double *vals, size_t cnt, FieldDesc f, int comp)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const unsigned key = keys[t];
    if (key == 0xffffffffu) return;
    if (t > 0 && keys[t - 1] == key) return;
    // the same sequential sum; the loads of the next G entries are in flight
    // while the current G are added (two register buffers)
    constexpr int G = 16;
    double s = 0.0;
    unsigned ka[G], kb[G];
    double va[G], vb[G];
    auto fetch = [&](size_t u, unsigned (&kv)[G], double (&vv)[G]) {
#pragma unroll
        for (int m = 0; m < G; ++m) {
            const bool in = u + m < cnt;
            kv[m] = in ? keys[u + m] : ~key;
            vv[m] = in ? vals[u + m] : 0.0;
        }
    };
    auto add = [&](const unsigned (&kv)[G], const double (&vv)[G]) {
        bool more = true;
#pragma unroll
        for (int m = 0; m < G; ++m) {
            more = more && kv[m] == key;
            if (more) s += vv[m];
        }
        return more;
    };
    fetch(t, ka, va);
    for (size_t u = t;; u += 2 * G) {
        fetch(u + G, kb, vb);
        if (!add(ka, va)) break;
        fetch(u + 2 * G, ka, va);
        if (!add(kb, vb)) break;
    }
    const int i = (int)(key % (unsigned)f.n), j = (int)(key / (unsigned)f.n);
    double *c = comp == 0 ? f.c0 : f.c1;
    c[fidx(f, i, j)] = s;
}

// ---------------------------------------------------------------------------
// Yee sweeps, TETMSolver.hh:142-149, 165-172, 241-249.  mask == nullptr for
// the window (unmasked); for the global grid the update is skipped where the
// cell mask is false (where(insideMask, new, old)).
__global__ void k_h_half(FieldDesc E, FieldDesc H, const uint8_t *mask, double sx, double sy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= H.n || j >= H.m) return;
    if (mask && !mask[(size_t)j * H.n + i]) return;
    const size_t q = fidx(H, i, j);
    const double eyp = E.c1[fidx(E, i + 1, j)], ey = E.c1[fidx(E, i, j)];
    const double exp_ = E.c0[fidx(E, i, j + 1)], ex = E.c0[fidx(E, i, j)];
    H.c0[q] = H.c0[q] - sx * (eyp - ey);
    H.c1[q] = H.c1[q] + sy * (exp_ - ex);
}

__global__ void k_ex(FieldDesc E, FieldDesc H, FieldDesc J, const uint8_t *mask, double s, double dteps)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= E.n || j >= E.m) return;
    if (mask && !mask[(size_t)j * E.n + i]) return;
    const size_t q = fidx(H, i, j), qm = fidx(H, i, j - 1);
    const double d = ((H.c0[q] - H.c0[qm]) + H.c1[q]) - H.c1[qm];
    const size_t qe = fidx(E, i, j);
    E.c0[qe] = E.c0[qe] + (s * d - dteps * J.c0[fidx(J, i, j)]);
}

__global__ void k_ey(FieldDesc E, FieldDesc H, FieldDesc J, const uint8_t *mask, double s, double dteps)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= E.n || j >= E.m) return;
    if (mask && !mask[(size_t)j * E.n + i]) return;
    const size_t q = fidx(H, i, j), qm = fidx(H, i - 1, j);
    const double d = ((H.c0[q] - H.c0[qm]) + H.c1[q]) - H.c1[qm];
    const size_t qe = fidx(E, i, j);
    E.c1[qe] = E.c1[qe] - (s * d + dteps * J.c1[fidx(J, i, j)]);
}

// Conformal corrections over the boundary-cell list
// (SCBoundaryCondition.hh:392-431, TETMSolver.cpp:87-127).
struct BCells {
    const int *i, *j;
    const double *lam0, *lam1, *area, *lamN0, *lamN1;
    int n;
};

__global__ void k_sc_hz(BCells b, FieldDesc E, FieldDesc H, double etax, double etay)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.n) return;
    const int i = b.i[t], j = b.j[t];
    const double u0 = -0.5 * etax * (b.lamN1[t] * E.c1[fidx(E, i + 1, j)] - b.lam1[t] * E.c1[fidx(E, i, j)]) / b.area[t];
    const double u1 = +0.5 * etay * (b.lamN0[t] * E.c0[fidx(E, i, j + 1)] - b.lam0[t] * E.c0[fidx(E, i, j)]) / b.area[t];
    const size_t q = fidx(H, i, j);
    H.c0[q] += u0;
    H.c1[q] += u1;
}

__global__ void k_sc_ex(BCells b, FieldDesc E, FieldDesc H, FieldDesc J, double dteps, double hy)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.n) return;
    if (b.area[t] < pm::kMinArea) return;
    if (!(b.lam0[t] > pm::kSpaceEps)) return;
    const int i = b.i[t], j = b.j[t];
    const size_t q = fidx(H, i, j), qm = fidx(H, i, j - 1);
    const double u0 = (H.c0[q] - H.c0[qm]) / hy;
    const double u1 = (H.c1[q] - H.c1[qm]) / hy;
    E.c0[fidx(E, i, j)] += dteps * (u0 + u1 - J.c0[fidx(J, i, j)]);
}

__global__ void k_sc_ey(BCells b, FieldDesc E, FieldDesc H, FieldDesc J, double dteps, double hx)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.n) return;
    if (b.area[t] < pm::kMinArea) return;
    if (!(b.lam1[t] > pm::kSpaceEps)) return;
    const int i = b.i[t], j = b.j[t];
    const size_t q = fidx(H, i, j), qm = fidx(H, i - 1, j);
    const double u0 = -(H.c0[q] - H.c0[qm]) / hx;
    const double u1 = -(H.c1[q] - H.c1[qm]) / hx;
    E.c1[fidx(E, i, j)] += dteps * (u0 + u1 - J.c1[fidx(J, i, j)]);
}

// ---------------------------------------------------------------------------
// Layer exchange: every layer is first packed (op.execute) into a staging
// buffer, then unpacked (op.invert), so all reads precede all writes as in
// Communicator::exchangeBoundaries.
struct Layer {
    int src, dst, op;        // field ids; op 0 H (sum, then half/half), 1 Ex, 2 Ey
    int lo0, hi0, lo1, hi1;  // source rectangle
    int sh0, sh1;            // dst index = src index + shift
    int off;                 // staging offset
};
constexpr int kMaxLayers = 16;
struct LayerSet {
    Layer L[kMaxLayers];
    int n, total;
};

__device__ inline int find_layer(const LayerSet &ls, int t, int &local)
{
    for (int k = 0; k < ls.n; ++k) {
        const Layer &l = ls.L[k];
        const int cnt = (l.hi0 - l.lo0 + 1) * (l.hi1 - l.lo1 + 1);
        if (t >= l.off && t < l.off + cnt) { local = t - l.off; return k; }
    }
    return -1;
}

__global__ void k_pack(LayerSet ls, FieldSet fs, double *stage)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ls.total) return;
    int loc;
    const int k = find_layer(ls, t, loc);
    if (k < 0) return;
    const Layer &l = ls.L[k];
    const int w = l.hi0 - l.lo0 + 1;
    const int i = l.lo0 + loc % w, j = l.lo1 + loc / w;
    const FieldDesc &f = fs.f[l.src];
    const size_t q = fidx(f, i, j);
    stage[t] = l.op == 0 ? f.c0[q] + f.c1[q] : (l.op == 1 ? f.c0[q] : f.c1[q]);
}

// All layers in one launch.  Where destinations overlap, the value of the
// last layer writing that component there is kept (the order of the
// exchanges' layers), so an element is skipped when a later layer writes the
// same component of the same field at the same node.
__global__ void k_unpack(LayerSet ls, FieldSet fs, const double *stage)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ls.total) return;
    int k = 0;
    while (k + 1 < ls.n && ls.L[k + 1].off <= t) ++k;
    const Layer &l = ls.L[k];
    const int w = l.hi0 - l.lo0 + 1, e = t - l.off;
    const int i = l.lo0 + e % w + l.sh0, j = l.lo1 + e / w + l.sh1;
    bool w0 = l.op != 2, w1 = l.op != 1;   // components this layer writes
    for (int k2 = k + 1; k2 < ls.n; ++k2) {
        const Layer &m = ls.L[k2];
        if (m.dst != l.dst) continue;
        if (i < m.lo0 + m.sh0 || i > m.hi0 + m.sh0 || j < m.lo1 + m.sh1 || j > m.hi1 + m.sh1) continue;
        if (m.op != 2) w0 = false;
        if (m.op != 1) w1 = false;
    }
    const FieldDesc &f = fs.f[l.dst];
    const size_t q = fidx(f, i, j);
    const double v = stage[t];
    if (l.op == 0) {
        if (w0) f.c0[q] = 0.5 * v;
        if (w1) f.c1[q] = 0.5 * v;
    } else if (l.op == 1) {
        if (w0) f.c0[q] = v;
    } else if (w1) {
        f.c1[q] = v;
    }
}

// moveMesh copy: dst(i, j) = snapshot(i + mi, j + mj) over the region.
__global__ void k_shift(FieldSet fs, FieldSet snap, int i0, int i1, int j0, int j1, int mi, int mj)
{
    const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = j0 + blockIdx.y;
    if (i > i1 || j > j1) return;
    for (int f = F_ME; f <= F_MJ; ++f) {
        const FieldDesc &d = fs.f[f], &s = snap.f[f];
        const size_t qd = fidx(d, i, j), qs = fidx(s, i + mi, j + mj);
        d.c0[qd] = s.c0[qs];
        d.c1[qd] = s.c1[qs];
    }
}

// ---------------------------------------------------------------------------
// Compensated reductions for the diagnostics (double-double accumulation).
__device__ inline void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
struct DD {
    double hi, lo;
};
__device__ inline DD dd_add(DD x, DD y)
{
    double s, e;
    two_sum(x.hi, y.hi, s, e);
    e += x.lo + y.lo;
    double hi, lo;
    two_sum(s, e, hi, lo);
    return DD{hi, lo};
}

// Per-block double-double partial sums of m values per item; fixed tree.
template <int M, class F>
__device__ void block_dd_sum(F get, int n, DD *partial)
{
    __shared__ DD sh[M][kBlock];
    DD acc[M];
    for (int m = 0; m < M; ++m) acc[m] = DD{0.0, 0.0};
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        double v[M];
        get(k, v);
        for (int m = 0; m < M; ++m) acc[m] = dd_add(acc[m], DD{v[m], 0.0});
    }
    for (int m = 0; m < M; ++m) sh[m][threadIdx.x] = acc[m];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int m = 0; m < M; ++m) sh[m][threadIdx.x] = dd_add(sh[m][threadIdx.x], sh[m][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int m = 0; m < M; ++m) partial[blockIdx.x * M + m] = sh[m][0];
}

constexpr int kRedBlocks = 120;

__global__ void k_moments_partial(Parts p, DD *partial)
{
    block_dd_sum<5>([&](int k, double *v) { v[0] = p.rx[k]; v[1] = p.px[k]; v[2] = p.ry[k]; v[3] = p.py[k]; v[4] = p.q[k]; },
                    p.n, partial);
}

// Final fixed-order sum; out = rmean x, pmean x, rmean y, pmean y, qtotal
__global__ void k_moments_final(const DD *partial, int nb, int n, double *out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int m = 0; m < 5; ++m) {
        DD s{0.0, 0.0};
        for (int b = 0; b < nb; ++b) s = dd_add(s, partial[b * 5 + m]);
        out[m] = m < 4 ? (s.hi + s.lo) / n : s.hi + s.lo;
    }
}

// IntegrateElongCommand::execute (IntegrateElong.cpp:38-89): sample E along
// the line through the centroid, accumulate cosP Ex + sinP Ey.
__global__ void k_wake(FieldDesc mE, const double *mom, double mmo0, double mmo1, double hx, double hy, DD *acc)
{
    const int ii = blockIdx.x * blockDim.x + threadIdx.x;
    const int W = mE.n, H = mE.m;
    if (ii >= W - 4) return;
    const int i = ii + 2;
    const double R0 = mom[0], P0 = mom[1], R1 = mom[2], P1 = mom[3];
    const double lenP = sqrt(P0 * P0 + P1 * P1);
    const double cosP = P0 / lenP, sinP = P1 / lenP;
    const double l0 = mmo0 + i * hx - R0;
    const double x = (cosP * l0 + sinP * 0.0) + R0;
    const double y = (-sinP * l0 + cosP * 0.0) + R1;
    const int ci = (int)floor((x - mmo0) / hx), cj = (int)floor((y - mmo1) / hy);
    if (ci < 0 || ci > W - 1 || cj < 0 || cj > H - 1) return;
    const vec2<double> od = pm::mk(1.0 / hx, 1.0 / hy);
    const vec2<double> pp = pm::mk((x - mmo0) * od.x, (y - mmo1) * od.y);
    auto fx = [&](long long I, long long J) { return mE.c0[fidx(mE, (int)I, (int)J)]; };
    auto fy = [&](long long I, long long J) { return mE.c1[fidx(mE, (int)I, (int)J)]; };
    const double ex = pm::gather(pm::edge_stencil(0, 0, pp), false, fx);
    const double ey = pm::gather(pm::edge_stencil(0, 1, pp), false, fy);
    acc[ii] = dd_add(acc[ii], DD{cosP * ex + sinP * ey, 0.0});
}

__global__ void k_energy_partial(FieldDesc E, FieldDesc H, DD *partial)
{
    const int n = E.n * E.m;
    block_dd_sum<1>([&](int k, double *v) {
        const int i = k % E.n, j = k / E.n;
        const size_t q = fidx(E, i, j);
        const double ex = E.c0[q], ey = E.c1[q];
        const double hz = H.c0[q] + H.c1[q];
        v[0] = 0.5 * pm::kEps0 * (ex * ex + ey * ey) + 0.5 * pm::kMu0 * hz * hz;
    }, n, partial);
}

// Sums of E.E and (Hzx + Hzy)^2 over the ni x nj nodes from (i0, j0) of one
// grid: the loops of CalcEnergyCommand::execute (CalcEnergy.cpp:77-124).
__global__ void k_energy_sums(FieldDesc E, FieldDesc H, int i0, int j0, int ni, int nj, DD *partial)
{
    block_dd_sum<2>([&](int k, double *v) {
        const int i = i0 + k % ni, j = j0 + k / ni;
        const size_t q = fidx(E, i, j);
        const double ex = E.c0[q], ey = E.c1[q];
        const double hz = H.c0[q] + H.c1[q];
        v[0] = ex * ex + ey * ey;
        v[1] = hz * hz;
    }, ni * nj, partial);
}

// ---------------------------------------------------------------------------
struct Rect {
    int lo[2], hi[2];
};

}  // namespace

// ===========================================================================
// ---------------------------------------------------------------------------
// Window copy of main.cpp:335-337 (Communicator.hh:211-272): global E and H
// into the window over the overlap of the two index ranges.
__global__ void k_copy_to_window(FieldDesc g, FieldDesc w, int s0, int s1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;   // window column
    const int j = blockIdx.y;                               // window row
    if (i >= w.n || j >= w.m) return;
    const int gi = i + s0, gj = j + s1;
    if (gi < 0 || gj < 0 || gi >= g.n || gj >= g.m) return;
    w.c0[fidx(w, i, j)] = g.c0[fidx(g, gi, gj)];
    w.c1[fidx(w, i, j)] = g.c1[fidx(g, gi, gj)];
}

// ---------------------------------------------------------------------------
// Poisson initialisation (PoissonSolver.cpp:85-201).

// PartBunch::scatterQ (ShapeFunction getChargeDensity): vertex charge of each
// particle onto the global grid, FP64 atomics (the Poisson start is a solve
// to a tolerance, so the summation order does not matter).
template <int ORDER>
__global__ void k_scatter_q(Parts p, double o0, double o1, double h0, double h1, double *rho, int nx, int ny)
{
    constexpr int N = ORDER + 2;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.n) return;
    const double q = p.q[k];
    double wx[N], wy[N];
    int I0, J0;
    pm::charge_weights<ORDER>(pm::mk((p.rx[k] - o0) / h0, (p.ry[k] - o1) / h1), q, wx, wy, I0, J0);
#pragma unroll
    for (int b = 0; b < N; ++b)
#pragma unroll
        for (int a = 0; a < N; ++a) {
            const int I = I0 + a, J = J0 + b;
            if (I < 0 || J < 0 || I >= nx || J >= ny) continue;
            unsafeAtomicAdd(&rho[(size_t)J * nx + I], pm::charge_w<ORDER>(q, wx[a], wy[b]));
        }
}

// Gauss-law residual on the window's vertices: div E - rho / eps0 with the
// divergence of the edge fields as PartBunch::calcChargeConservation forms
// the divergence of J (PartBunch.cpp:552-556: value at i minus value at i-1
// over the spacing) and rho the bunch's vertex charge density (getChargeDensity,
// in `rho` over the cell area).
__global__ void k_gauss(FieldDesc E, const double *rho, double ihx, double ihy, double area, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= E.n) return;
    const double div = (E.c0[fidx(E, i, j)] - E.c0[fidx(E, i - 1, j)]) * ihx +
                       (E.c1[fidx(E, i, j)] - E.c1[fidx(E, i, j - 1)]) * ihy;
    const size_t q = (size_t)j * E.n + i;
    out[q] = div - rho[q] / area / pm::kEps0;
}

// fillRHS (PoissonSolver.cpp:441-460): b = rho / (hx hy) / (eps0 gamma).
__global__ void k_rhs(const double *rho, double *b, size_t n, double hh, double cc)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = (0.0 + rho[k] / hh) * cc;
}

__global__ void k_move_x(Parts p, double dx)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < p.n) p.rx[k] = p.rx[k] + dx;
}

__global__ void k_save_r(Parts p, double *sx, double *sy, int dir)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.n) return;
    if (dir == 0) { sx[k] = p.rx[k]; sy[k] = p.ry[k]; }
    else { p.rx[k] = sx[k]; p.ry[k] = sy[k]; }
}

// shiftLHS (PoissonSolver.cpp:279-319): next initial guess from the last
// potential, phi_i + tau (phi_{i+1} - phi_i) for tau > 0, with phi_{i-1}
// for tau < 0; values past the grid read as 0.
__global__ void k_shift_guess(const double *phi, double *out, int nx, int ny, double tau)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)nx * ny) return;
    const int i = (int)(k % nx);
    const double v = phi[k];
    if (tau > 0) {
        const double t = tau > 1 ? 1.0 : tau;
        const double nb = i + 1 < nx ? phi[k + 1] : 0.0;
        out[k] = v + t * nb - t * v;
    } else {
        double t = -tau;
        if (t > 1) t = 1;
        const double nb = i > 0 ? phi[k - 1] : 0.0;
        out[k] = v + t * nb - t * v;
    }
}

// fillFields (PoissonSolver.cpp:462-507): forward differences of the potential.
// comp 0: Ex = phi_i/hr0 - phi_{i+1}/hr0; 1: Ey = g phi_j/hy - g phi_{j+1}/hy;
// 2: both Hz halves = bg phi_j/(2 c mu0 hy) - bg phi_{j+1}/(2 c mu0 hy).
__global__ void k_fill_fields(const double *phi, int comp, FieldDesc f, double scale, double den)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = f.n, ny = f.m;
    if (k >= (size_t)nx * ny) return;
    const int i = (int)(k % nx), j = (int)(k / nx);
    if (comp == 0) {
        const double nb = i + 1 < nx ? phi[k + 1] : 0.0;
        f.c0[fidx(f, i, j)] = phi[k] / den - nb / den;
    } else {
        const double nb = j + 1 < ny ? phi[k + nx] : 0.0;
        const double v = scale * phi[k] / den - scale * nb / den;
        if (comp == 1) {
            f.c1[fidx(f, i, j)] = v;
        } else {
            f.c0[fidx(f, i, j)] = v;
            f.c1[fidx(f, i, j)] = v;
        }
    }
}

// SCBoundaryCondition::resetOutsideDomain (SCBoundaryCondition.hh:164-196):
// per cell bit 0 clears Ex, bit 1 clears Ey.
__global__ void k_reset_outside(FieldDesc E, const uint8_t *code)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)E.n * E.m) return;
    const uint8_t c = code[k];
    if (!c) return;
    const int i = (int)(k % E.n), j = (int)(k / E.n);
    if (c & 1) E.c0[fidx(E, i, j)] = 0.0;
    if (c & 2) E.c1[fidx(E, i, j)] = 0.0;
}

namespace {
// SaveEMFieldsCommand::execute (SaveEMFields.cpp:99-133): for the grid
// columns i0 + t, the point of the line through R along P, the gathers of E,
// H (both split parts) and J on the global fields with the run's shape, and
// the projections along and across P.  Values outside the global fields'
// guard layer read as zero.
__global__ void k_line(FieldDesc E, FieldDesc H, FieldDesc J, int order, int s2bug, double o0, double o1, double h0,
                       double h1, double R0, double R1, double P0, double P1, int i0, int n, double *out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = i0 + t;
    const double x = o0 + i * h0;
    const double y = R1 + (o0 + i * h0 - R0) / P0 * P1;
    const vec2<double> od = pm::mk(1.0 / h0, 1.0 / h1);
    const vec2<double> pp = pm::mk((x - o0) * od.x, (y - o1) * od.y);
    auto at = [](const FieldDesc &f, const double *c) {
        return [&f, c](long long I, long long Jj) {
            if (I < -1 || I > f.n || Jj < -1 || Jj > f.m) return 0.0;
            return c[fidx(f, (int)I, (int)Jj)];
        };
    };
    const bool bug = s2bug != 0;
    const pm::Stencil<double> sx = pm::edge_stencil(order, 0, pp), sy = pm::edge_stencil(order, 1, pp);
    const pm::Stencil<double> sc = pm::cell_stencil(order, pp);
    const double ex = pm::gather(sx, bug, at(E, E.c0)), ey = pm::gather(sy, bug, at(E, E.c1));
    const double ha = pm::gather(sc, bug, at(H, H.c0)), hb = pm::gather(sc, bug, at(H, H.c1));
    const double jx = pm::gather(sx, bug, at(J, J.c0)), jy = pm::gather(sy, bug, at(J, J.c1));
    const double lenP = sqrt(P0 * P0 + P1 * P1);
    const double cosP = P0 / lenP, sinP = P1 / lenP;
    out[4 * t] = cosP * ex + sinP * ey;
    out[4 * t + 1] = -sinP * ex + cosP * ey;
    out[4 * t + 2] = ha + hb;
    out[4 * t + 3] = cosP * jx + sinP * jy;
}
}  // namespace

struct csrst_ctx {
    std::string err;
    csrst_params p{};
    geom::Bend bend{};
    Margins mg{};
    double dt = 0, hr[2] = {0, 0};
    hipStream_t s = nullptr;     // the stream the context's work goes to
    hipStream_t own = nullptr;   // the context's own stream (s unless the caller set one)
    // fields
    DBuf<double> gE, gH, gJ, mEb, mHb, mJb, snapb;
    FieldSet fs{}, snap{};
    DBuf<uint8_t> mask;
    std::vector<uint8_t> hmask;
    // boundary cells
    DBuf<int> bi, bj;
    DBuf<double> blam0, blam1, barea, blamN0, blamN1;
    std::vector<int> hbij;
    std::vector<double> hbdata;
    BCells bc{};
    // window
    int W = 0, H = 0;
    int ddim[2][2] = {{0, 0}, {0, 0}};
    double mmo[2] = {0, 0};
    // particles
    int n = 0;
    DBuf<double> rx, ry, ox, oy, px, py, q;
    Parts parts{};
    DBuf<BBox> bb;         // [0]: R after the half drift, [1]: R and oldR after the push
    DBuf<int> derr;
    DBuf<unsigned> keys, keys2;
    DBuf<double> vals, vals2;
    DBuf<unsigned char> cubTmp;
    size_t cubBytes = 0;
    // exchange
    DBuf<double> stage;
    // diagnostics
    DBuf<DD> partial, wacc;
    DBuf<double> mom;
    DBuf<bs::DD> bsPart;   // beam statistics partials
    DBuf<double> bsLim;
    // Poisson start: node classes and resetOutsideDomain codes
    DBuf<uint8_t> pcls, preset;
    int steps = 0;
    bool ready = false;

    void alloc_field(DBuf<double> &b, FieldDesc &f, int nn, int mm)
    {
        const size_t plane = (size_t)(nn + 2) * (mm + 2);
        b.alloc(2 * plane);
        b.zero(s);
        f.c0 = b.p;
        f.c1 = b.p + plane;
        f.n = nn;
        f.m = mm;
    }
};

namespace {

// The mask and the boundary-cell list (host copies in c) to the device.
void upload_geometry(csrst_ctx *c)
{
    c->mask.alloc(c->hmask.size());
    HU_CHECK(hipMemcpyAsync(c->mask.p, c->hmask.data(), c->hmask.size(), hipMemcpyHostToDevice, c->s));
    const int nb = (int)(c->hbij.size() / 2);
    if (nb > 0) {
        std::vector<int> vi(nb), vj(nb);
        std::vector<double> l0(nb), l1(nb), ar(nb), n0(nb), n1(nb);
        for (int k = 0; k < nb; ++k) {
            vi[k] = c->hbij[2 * k]; vj[k] = c->hbij[2 * k + 1];
            l0[k] = c->hbdata[5 * k]; l1[k] = c->hbdata[5 * k + 1]; ar[k] = c->hbdata[5 * k + 2];
            n0[k] = c->hbdata[5 * k + 3]; n1[k] = c->hbdata[5 * k + 4];
        }
        c->bi.alloc(nb); c->bj.alloc(nb);
        c->blam0.alloc(nb); c->blam1.alloc(nb); c->barea.alloc(nb); c->blamN0.alloc(nb); c->blamN1.alloc(nb);
        HU_CHECK(hipMemcpy(c->bi.p, vi.data(), nb * sizeof(int), hipMemcpyHostToDevice));
        HU_CHECK(hipMemcpy(c->bj.p, vj.data(), nb * sizeof(int), hipMemcpyHostToDevice));
        HU_CHECK(hipMemcpy(c->blam0.p, l0.data(), nb * sizeof(double), hipMemcpyHostToDevice));
        HU_CHECK(hipMemcpy(c->blam1.p, l1.data(), nb * sizeof(double), hipMemcpyHostToDevice));
        HU_CHECK(hipMemcpy(c->barea.p, ar.data(), nb * sizeof(double), hipMemcpyHostToDevice));
        HU_CHECK(hipMemcpy(c->blamN0.p, n0.data(), nb * sizeof(double), hipMemcpyHostToDevice));
        HU_CHECK(hipMemcpy(c->blamN1.p, n1.data(), nb * sizeof(double), hipMemcpyHostToDevice));
    }
    c->bc = BCells{c->bi.p, c->bj.p, c->blam0.p, c->blam1.p, c->barea.p, c->blamN0.p, c->blamN1.p, nb};
    HU_CHECK(hipStreamSynchronize(c->s));
}

int fail(csrst_ctx *c, const std::exception &e)
{
    const Error *he = dynamic_cast<const Error *>(&e);
    const std::string m = e.what();
    if (c) c->err = m;
    hu::global_error() = m;
    return he ? he->code : CSRST_EINVAL;
}

dim3 grid2(int n, int m) { return dim3((n + kBlock - 1) / kBlock, m); }
int blocks(size_t n) { return (int)((n + kBlock - 1) / kBlock); }

void check_err(csrst_ctx *c)
{
    int h = 0;
    HU_CHECK(hipMemcpyAsync(&h, c->derr.p, sizeof(int), hipMemcpyDeviceToHost, c->s));
    HU_CHECK(hipStreamSynchronize(c->s));
    if (h & 1) throw Error(CSRST_EWINDOW, "gather stencil left the window (bunch too close to the window edge)");
    if (h & 2) throw Error(CSRST_EWINDOW, "particle crossed more than one cell line per axis in one push");
}

// Communicator::exchangeBoundaries for one op (single process),
// Communicator.hh:575-876 with the side masks of Communicator.cpp:640-642.
LayerSet build_exchange(const csrst_ctx *c, int op, int &stageOff)
{
    static const unsigned sides[3] = {(1u << 4) | (1u << 5) | (1u << 2) | (1u << 3), (1u << 7) | (1u << 1),
                                      (1u << 6) | 1u};
    const int pos[2] = {c->ddim[0][0], c->ddim[1][0]};
    const int W = c->W, H = c->H;
    Rect gEmb = {{0, 0}, {W - 1, H - 1}};
    Rect incr = {{-1 + pos[0], -1 + pos[1]}, {W + pos[0], H + pos[1]}};
    Rect g[8];
    for (int d = 0; d < 2; ++d) {
        g[d] = gEmb; g[d].lo[d] = g[d].hi[d] = gEmb.lo[d];
        g[d + 2] = gEmb; g[d + 2].lo[d] = g[d + 2].hi[d] = gEmb.hi[d];
        g[d + 4] = incr; g[d + 4].lo[d] = g[d + 4].hi[d] = incr.lo[d];
        g[d + 6] = incr; g[d + 6].lo[d] = g[d + 6].hi[d] = incr.hi[d];
    }
    const Rect surr = {{0, 0}, {c->p.Nx - 1, c->p.Ny - 1}};
    const int fsrcE = op == 0 ? F_MH : F_ME, fsrcS = op == 0 ? F_GH : F_GE;
    LayerSet ls{};
    ls.n = 0;
    for (int side = 0; side < 8; ++side) {
        if (!(sides[op] & (1u << side))) continue;
        const Rect &src = side >= 4 ? surr : gEmb;
        Rect dom;
        bool has = true;
        for (int d = 0; d < 2; ++d) {
            dom.lo[d] = std::max(src.lo[d], g[side].lo[d]);
            dom.hi[d] = std::min(src.hi[d], g[side].hi[d]);
            if (dom.lo[d] > dom.hi[d]) has = false;
        }
        if (!has) continue;
        const int idx = side < 4 ? 0 : 1;
        const Rect &fld = idx == 0 ? surr : gEmb;
        Rect ov;
        bool inside = true;
        for (int d = 0; d < 2; ++d) {
            const int diff = idx == 0 ? -pos[d] : pos[d];
            ov.lo[d] = std::max(dom.lo[d], fld.lo[d] + diff - 1);
            ov.hi[d] = std::min(dom.hi[d], fld.hi[d] + diff + 1);
            if (ov.lo[d] > ov.hi[d]) inside = false;
        }
        if (!inside) continue;
        Layer &l = ls.L[ls.n++];
        l.src = side < 4 ? fsrcE : fsrcS;
        l.dst = side < 4 ? fsrcS : fsrcE;
        l.op = op;
        l.lo0 = ov.lo[0]; l.hi0 = ov.hi[0]; l.lo1 = ov.lo[1]; l.hi1 = ov.hi[1];
        l.sh0 = side < 4 ? pos[0] : -pos[0];
        l.sh1 = side < 4 ? pos[1] : -pos[1];
        l.off = stageOff;
        stageOff += (l.hi0 - l.lo0 + 1) * (l.hi1 - l.lo1 + 1);
    }
    ls.total = stageOff;
    return ls;
}

// Communicator::exchangeBoundariesMovingMesh, Communicator.cpp:123-638.
LayerSet build_moving_exchange(const csrst_ctx *c, const int move[2])
{
    const int pos[2] = {c->ddim[0][0], c->ddim[1][0]};
    const int f0 = 0, l0 = c->W - 1, f1 = 0, l1 = c->H - 1;
    std::vector<Rect> L[6];
    auto R = [](int a0, int b0, int a1, int b1) { Rect r; r.lo[0] = a0; r.hi[0] = b0; r.lo[1] = a1; r.hi[1] = b1; return r; };
    if (move[0] == 1) {
        L[0].push_back(R(f0 - 1, f0 - 1, f1 - move[1], l1 - move[1] - 1));
        L[1].push_back(R(f0 - 1, f0 - 1, f1 - move[1] + 1, l1 - move[1]));
        L[2].push_back(R(f0, f0, f1 - move[1], l1 - move[1]));
        L[3].push_back(R(l0 + pos[0], l0 + pos[0], f1 + pos[1], l1 + pos[1]));
        L[4].push_back(R(l0 + pos[0], l0 + pos[0], f1 + pos[1], l1 + pos[1] + 1));
        L[5].push_back(R(l0 + pos[0] + 1, l0 + pos[0] + 1, f1 + pos[1], l1 + pos[1]));
    }
    if (move[1] == 1) {
        L[0].push_back(R(f0 - move[0], l0 - 1 - move[0], f1 - 1, f1 - 1));
        L[1].push_back(R(f0 - move[0], l0 - move[0], f1, f1));
        L[2].push_back(R(f0 + 1 - move[0], l0 - move[0], f1 - 1, f1 - 1));
        L[3].push_back(R(f0 + pos[0], l0 + pos[0], l1 + pos[1], l1 + pos[1]));
        L[4].push_back(R(f0 + pos[0], l0 + pos[0], l1 + 1 + pos[1], l1 + 1 + pos[1]));
        L[5].push_back(R(f0 + pos[0], l0 + pos[0] + 1, l1 + pos[1], l1 + pos[1]));
    } else if (move[1] == -1) {
        L[0].push_back(R(f0 - move[0], l0 - move[0] - 1, l1, l1));
        L[1].push_back(R(f0 - move[0], l0 - move[0], l1 + 1, l1 + 1));
        L[2].push_back(R(f0 - move[0] + 1, l0 - move[0], l1 + 1, l1 + 1));
        L[3].push_back(R(f0 + pos[0], l0 + pos[0], f1 - 1 + pos[1], f1 - 1 + pos[1]));
        L[4].push_back(R(f0 + pos[0], l0 + pos[0], f1 + pos[1], f1 + pos[1]));
        L[5].push_back(R(f0 + pos[0], l0 + pos[0] + 1, f1 + pos[1], f1 + pos[1]));
    }
    const Rect embExt = R(-1, c->W, -1, c->H);
    const Rect surr = R(0, c->p.Nx - 1, 0, c->p.Ny - 1);
    static const int srcF[6] = {F_MH, F_ME, F_ME, F_GH, F_GE, F_GE};
    static const int dstF[6] = {F_GH, F_GE, F_GE, F_MH, F_ME, F_ME};
    LayerSet ls{};
    int off = 0;
    for (int t = 0; t < 6; ++t) {
        const Rect &domain = t < 3 ? embExt : surr;
        const Rect &fld = t < 3 ? surr : embExt;
        for (const Rect &layer : L[t]) {
            Rect dom, ov;
            bool ok = true;
            for (int d = 0; d < 2; ++d) {
                dom.lo[d] = std::max(domain.lo[d], layer.lo[d]);
                dom.hi[d] = std::min(domain.hi[d], layer.hi[d]);
                if (dom.lo[d] > dom.hi[d]) ok = false;
            }
            if (!ok) continue;
            for (int d = 0; d < 2; ++d) {
                const int diff = t < 3 ? -pos[d] : pos[d];
                ov.lo[d] = std::max(dom.lo[d], fld.lo[d] + diff);
                ov.hi[d] = std::min(dom.hi[d], fld.hi[d] + diff);
                if (ov.lo[d] > ov.hi[d]) ok = false;
            }
            if (!ok) continue;
            if (ls.n >= kMaxLayers) throw Error(CSRST_EINVAL, "too many exchange layers");
            Layer &l = ls.L[ls.n++];
            l.src = srcF[t]; l.dst = dstF[t]; l.op = t % 3;
            l.lo0 = ov.lo[0]; l.hi0 = ov.hi[0]; l.lo1 = ov.lo[1]; l.hi1 = ov.hi[1];
            l.sh0 = t < 3 ? pos[0] : -pos[0];
            l.sh1 = t < 3 ? pos[1] : -pos[1];
            l.off = off;
            off += (l.hi0 - l.lo0 + 1) * (l.hi1 - l.lo1 + 1);
        }
    }
    ls.total = off;
    return ls;
}

void run_layers(csrst_ctx *c, const LayerSet &ls)
{
    if (ls.n == 0) return;
    if ((size_t)ls.total > c->stage.n) c->stage.alloc((size_t)ls.total * 2);
    k_pack<<<blocks(ls.total), kBlock, 0, c->s>>>(ls, c->fs, c->stage.p);
    HU_LAUNCH_CHECK();
    k_unpack<<<blocks(ls.total), kBlock, 0, c->s>>>(ls, c->fs, c->stage.p);
    HU_LAUNCH_CHECK();
}

void exchange(csrst_ctx *c, int op)
{
    int off = 0;
    const LayerSet ls = build_exchange(c, op, off);
    run_layers(c, ls);
}

void h_half(csrst_ctx *c)
{
    const double etax = pm::kC * c->dt / c->hr[0];
    const double etay = pm::kC * c->dt / c->hr[1];
    const double Z0 = sqrt(pm::kMu0 / pm::kEps0);
    const double sx = 0.5 * etax / Z0, sy = 0.5 * etay / Z0;
    k_h_half<<<grid2(c->p.Nx, c->p.Ny), kBlock, 0, c->s>>>(c->fs.f[F_GE], c->fs.f[F_GH], c->mask.p, sx, sy);
    k_h_half<<<grid2(c->W, c->H), kBlock, 0, c->s>>>(c->fs.f[F_ME], c->fs.f[F_MH], nullptr, sx, sy);
    HU_LAUNCH_CHECK();
}

void e_x(csrst_ctx *c)
{
    const double etay = pm::kC * c->dt / c->hr[1];
    const double Z0 = sqrt(pm::kMu0 / pm::kEps0);
    const double dteps = pm::kC * c->dt * Z0;
    const double sc = etay * Z0;
    k_ex<<<grid2(c->p.Nx, c->p.Ny), kBlock, 0, c->s>>>(c->fs.f[F_GE], c->fs.f[F_GH], c->fs.f[F_GJ], c->mask.p, sc, dteps);
    k_ex<<<grid2(c->W, c->H), kBlock, 0, c->s>>>(c->fs.f[F_ME], c->fs.f[F_MH], c->fs.f[F_MJ], nullptr, sc, dteps);
    HU_LAUNCH_CHECK();
}

void e_y(csrst_ctx *c)
{
    const double etax = pm::kC * c->dt / c->hr[0];
    const double Z0 = sqrt(pm::kMu0 / pm::kEps0);
    const double dteps = pm::kC * c->dt * Z0;
    const double sc = etax * Z0;
    k_ey<<<grid2(c->p.Nx, c->p.Ny), kBlock, 0, c->s>>>(c->fs.f[F_GE], c->fs.f[F_GH], c->fs.f[F_GJ], c->mask.p, sc, dteps);
    k_ey<<<grid2(c->W, c->H), kBlock, 0, c->s>>>(c->fs.f[F_ME], c->fs.f[F_MH], c->fs.f[F_MJ], nullptr, sc, dteps);
    HU_LAUNCH_CHECK();
}

void sc_hz(csrst_ctx *c)
{
    if (c->bc.n == 0) return;
    const double etax = c->dt / (c->hr[0] * pm::kMu0);
    const double etay = c->dt / (c->hr[1] * pm::kMu0);
    k_sc_hz<<<blocks(c->bc.n), kBlock, 0, c->s>>>(c->bc, c->fs.f[F_GE], c->fs.f[F_GH], etax, etay);
    HU_LAUNCH_CHECK();
}
void sc_ex(csrst_ctx *c)
{
    if (c->bc.n == 0) return;
    k_sc_ex<<<blocks(c->bc.n), kBlock, 0, c->s>>>(c->bc, c->fs.f[F_GE], c->fs.f[F_GH], c->fs.f[F_GJ], c->dt / pm::kEps0,
                                                   c->hr[1]);
    HU_LAUNCH_CHECK();
}
void sc_ey(csrst_ctx *c)
{
    if (c->bc.n == 0) return;
    k_sc_ey<<<blocks(c->bc.n), kBlock, 0, c->s>>>(c->bc, c->fs.f[F_GE], c->fs.f[F_GH], c->fs.f[F_GJ], c->dt / pm::kEps0,
                                                   c->hr[0]);
    HU_LAUNCH_CHECK();
}

void drift(csrst_ctx *c, double ddt)
{
    if (c->n == 0) return;
    k_drift<<<blocks(c->n), kBlock, 0, c->s>>>(c->parts, ddt * pm::kC);
    HU_LAUNCH_CHECK();
}

// PartBunch::push_particles (PartBunch.cpp:215-519) + the deposit of one
// current component into the window J.
void push(csrst_ctx *c, int comp)
{
    if (c->n == 0) throw Error(CSRST_ESTATE, "no particles");
    const double half = 0.5 * c->dt * pm::kC;
    k_bbox_reset<<<1, 1, 0, c->s>>>(c->bb.p);
    k_bbox_reset<<<1, 1, 0, c->s>>>(c->bb.p + 1);
    k_drift_half<<<std::min(blocks(c->n), 128), kBlock, 0, c->s>>>(c->parts, half, c->bb.p);
    PushArgs a{};
    a.mmo[0] = c->mmo[0]; a.mmo[1] = c->mmo[1];
    a.hr[0] = c->hr[0]; a.hr[1] = c->hr[1];
    a.fHz[0] = c->mg.cell[0]; a.fHz[1] = c->mg.cell[0];
    a.fEx[0] = c->mg.edge[0][0]; a.fEx[1] = c->mg.edge[1][0];
    a.fEy[0] = c->mg.edge[1][0]; a.fEy[1] = c->mg.edge[0][0];
    a.half = half;
    a.fact1 = half * (-1e-9 / pm::kMeGeV);   // q/m of main.cpp:225
    a.fact2 = a.fact1 * pm::kC;
    a.order = c->p.order; a.s2bug = c->p.s2bug;
    k_push_kick<<<blocks(c->n), kBlock, 0, c->s>>>(c->parts, c->fs, c->bend, a, c->bb.p, c->bb.p + 1, c->derr.p);
    HU_LAUNCH_CHECK();
    // deposit
    DepArgs d{};
    d.mmo[0] = c->mmo[0]; d.mmo[1] = c->mmo[1];
    d.hr[0] = c->hr[0]; d.hr[1] = c->hr[1];
    d.dt = c->dt; d.order = c->p.order; d.comp = comp; d.W = c->W; d.H = c->H;
    d.slots = seg_nodes(c->p.order);
    k_emit<<<blocks(c->n), kBlock, 0, c->s>>>(c->parts, d, c->keys.p, c->vals.p, c->derr.p);
    HU_LAUNCH_CHECK();
    // (particle, segment, node) slots in that order: a stable sort keeps each
    // node's contributions in particle order, whatever the slots per segment
    const size_t cnt = (size_t)c->n * kMaxSeg * d.slots;
    // node keys are below W H; the empty-slot key 0xffffffff has all of the
    // low endBit bits set, so sorting on those bits alone (a stable sort)
    // gives the same order as on all 32 in fewer passes
    int endBit = 1;
    while ((1ull << endBit) < (unsigned long long)c->W * c->H + 1) ++endBit;
    HU_CHECK(hipcub::DeviceRadixSort::SortPairs(c->cubTmp.p, c->cubBytes, c->keys.p, c->keys2.p, c->vals.p, c->vals2.p,
                                                (int)cnt, 0, endBit, c->s));
    k_clear_comp<<<grid2(c->W, c->H), kBlock, 0, c->s>>>(c->fs.f[F_MJ], comp);
    k_reduce_runs<<<blocks(cnt), kBlock, 0, c->s>>>(c->keys2.p, c->vals2.p, cnt, c->fs.f[F_MJ], comp);
    HU_LAUNCH_CHECK();
}

void first_step(csrst_ctx *c)
{
    push(c, 0);
    h_half(c);
    exchange(c, 0);
    sc_hz(c);
    e_x(c);
    exchange(c, 1);
    sc_ex(c);
    h_half(c);
    exchange(c, 0);
    sc_hz(c);
    drift(c, -0.5 * c->dt);
}

void second_step(csrst_ctx *c)
{
    push(c, 1);
    e_y(c);
    exchange(c, 2);
    sc_ey(c);
    drift(c, -0.5 * c->dt);
}

void moments(csrst_ctx *c)
{
    k_moments_partial<<<kRedBlocks, kBlock, 0, c->s>>>(c->parts, c->partial.p);
    k_moments_final<<<1, 1, 0, c->s>>>(c->partial.p, kRedBlocks, c->n, c->mom.p);
    HU_LAUNCH_CHECK();
}

void wake_sample(csrst_ctx *c)
{
    moments(c);
    const int ns = c->W - 4;
    if (ns <= 0) return;
    k_wake<<<blocks(ns), kBlock, 0, c->s>>>(c->fs.f[F_ME], c->mom.p, c->mmo[0], c->mmo[1], c->hr[0], c->hr[1], c->wacc.p);
    HU_LAUNCH_CHECK();
}

// main.cpp:562-652
void move_mesh(csrst_ctx *c, int *mvI, int *mvJ)
{
    BBox h;
    HU_CHECK(hipMemcpyAsync(&h, c->bb.p + 1, sizeof(BBox), hipMemcpyDeviceToHost, c->s));
    HU_CHECK(hipStreamSynchronize(c->s));
    check_err(c);
    int gb[2][2];
    for (int d = 0; d < 2; ++d) {
        gb[d][0] = (int)floor((hu::unord(h.v[d]) - c->mmo[d]) / c->hr[d]);
        gb[d][1] = (int)floor((hu::unord(h.v[2 + d]) - c->mmo[d]) / c->hr[d]);
    }
    const int len[2] = {c->W, c->H};
    int ideal[2][2];
    for (int d = 0; d < 2; ++d) {
        const int glen = gb[d][1] - gb[d][0] + 1;
        if (glen >= len[d]) throw Error(CSRST_EWINDOW, "bunch no longer fits the window (main.cpp:584-587)");
        const unsigned diff = (unsigned)(len[d] - glen);
        const unsigned upper = (unsigned)floor(diff / 2.0);
        const unsigned lower = diff - upper;
        ideal[d][0] = (int)lower - 1;
        ideal[d][1] = len[d] - (int)upper - 2;
    }
    const int moveI = gb[0][1] - ideal[0][1] - 1;
    const int maxJ = std::max(0, gb[1][1] - ideal[1][1] - 1);
    const int minJ = std::min(0, gb[1][1] - ideal[1][1] - 1);
    if (moveI < 0 || moveI > 1 || maxJ > 1 || minJ < -1)
        throw Error(CSRST_EWINDOW, hu::fmt("window move (%d, %d) outside {0,1}x{-1,0,1}", moveI, minJ + maxJ));
    int iI[2] = {0 - moveI, c->W - 1};
    int iJ[2] = {0 - maxJ, c->H - 1 - minJ};
    if (minJ + maxJ == 0) { iJ[0] -= 1; iJ[1] += 1; }
    const int sj = minJ + maxJ;
    if (moveI != 0 || sj != 0) {
        HU_CHECK(hipMemcpyAsync(c->snapb.p, c->mEb.p, c->mEb.bytes(), hipMemcpyDeviceToDevice, c->s));
        HU_CHECK(hipMemcpyAsync(c->snapb.p + c->mEb.n, c->mHb.p, c->mHb.bytes(), hipMemcpyDeviceToDevice, c->s));
        HU_CHECK(hipMemcpyAsync(c->snapb.p + 2 * c->mEb.n, c->mJb.p, c->mJb.bytes(), hipMemcpyDeviceToDevice, c->s));
        const int wi = iI[1] - iI[0] + 1, hj = iJ[1] - iJ[0] + 1;
        k_shift<<<dim3(blocks(wi), hj), kBlock, 0, c->s>>>(c->fs, c->snap, iI[0], iI[1], iJ[0], iJ[1], moveI, sj);
        HU_LAUNCH_CHECK();
    }
    c->ddim[0][0] += moveI; c->ddim[0][1] += moveI;
    c->ddim[1][0] += sj; c->ddim[1][1] += sj;
    c->mmo[0] += moveI * c->hr[0];
    c->mmo[1] += sj * c->hr[1];
    const int mv[2] = {moveI, sj};
    const LayerSet ls = build_moving_exchange(c, mv);
    run_layers(c, ls);
    if (mvI) *mvI = moveI;
    if (mvJ) *mvJ = sj;
}

void copy_to_window(csrst_ctx *c)
{
    for (int f = 0; f < 2; ++f) {
        const FieldDesc &gd = c->fs.f[f == 0 ? F_GE : F_GH];
        const FieldDesc &wd = c->fs.f[f == 0 ? F_ME : F_MH];
        dim3 grid((unsigned)((c->W + kBlock - 1) / kBlock), (unsigned)c->H);
        k_copy_to_window<<<grid, kBlock, 0, c->s>>>(gd, wd, c->ddim[0][0], c->ddim[1][0]);
    }
    HU_LAUNCH_CHECK();
}

// Node classes of PoissonSolver::cutoffStencil (PoissonSolver.cpp:678-766)
// and the resetOutsideDomain codes, from the host geometry (once).
void poisson_geometry(csrst_ctx *c)
{
    if (!c->pcls.p) {
        const int Nx = c->p.Nx, Ny = c->p.Ny;
        const size_t nn = (size_t)Nx * Ny;
        std::vector<int> bmap(nn, -1);
        const int nb = (int)(c->hbij.size() / 2);
        for (int k = 0; k < nb; ++k) bmap[(size_t)c->hbij[2 * k + 1] * Nx + c->hbij[2 * k]] = k;
        const long maxX = (long)(size_t)(c->p.lenBefore / c->hr[0]);    // Bend.hh:288-291
        std::vector<uint8_t> cls(nn), code(nn);
        for (int j = 0; j < Ny; ++j)
            for (int i = 0; i < Nx; ++i) {
                const size_t id = (size_t)j * Nx + i;
                const int bc = bmap[id];
                const bool flag = c->hmask[id] || (bc > -1 && c->hbdata[5 * bc + 2] > 0.5);
                cls[id] = flag ? (i < maxX ? pmg::kUnknown : pmg::kKnown) : pmg::kFixed;
                uint8_t r = 0;
                if (!c->hmask[id]) {
                    if (bc == -1) {
                        r = 3;
                    } else {
                        if (c->hbdata[5 * bc] < pm::kSpaceEps) r |= 1;
                        if (c->hbdata[5 * bc + 1] < pm::kSpaceEps) r |= 2;
                        if (c->hbdata[5 * bc + 2] < pm::kMinArea) r = 3;
                    }
                }
                code[id] = r;
            }
        c->pcls.alloc(nn);
        c->preset.alloc(nn);
        HU_CHECK(hipMemcpy(c->pcls.p, cls.data(), nn, hipMemcpyHostToDevice));
        HU_CHECK(hipMemcpy(c->preset.p, code.data(), nn, hipMemcpyHostToDevice));
    }
}

// PoissonSolver::computeField (PoissonSolver.cpp:85-201), resetOutsideDomain
// and the window copy and exchanges that follow it in main.cpp:229-341.
void poisson_compute_field(csrst_ctx *c, double tol, int maxit, int *cycles, double *resid)
{
    const int Nx = c->p.Nx, Ny = c->p.Ny;
    const size_t nn = (size_t)Nx * Ny;
    poisson_geometry(c);
    // gamma from the mean momentum (PoissonSolver.cpp:64-66)
    moments(c);
    double mom[5];
    HU_CHECK(hipMemcpyAsync(mom, c->mom.p, sizeof(mom), hipMemcpyDeviceToHost, c->s));
    HU_CHECK(hipStreamSynchronize(c->s));
    const double gamma = sqrt(mom[1] * mom[1] + mom[3] * mom[3] + 1);
    const double hr0 = c->hr[0] * gamma, hr1 = c->hr[1];
    DBuf<double> rho, b, phi, guess, sx, sy;
    rho.alloc(nn); b.alloc(nn); phi.alloc(nn); guess.alloc(nn);
    sx.alloc(c->n); sy.alloc(c->n);
    pmg::Solver mg;
    mg.setup(Nx, Ny, hr0, hr1, c->pcls.p, c->s);
    const double cc = 1 / (pm::kEps0 * gamma);
    auto rhs = [&]() {
        rho.zero(c->s);
        const double o0 = c->bend.origin[0], o1 = c->bend.origin[1];
        if (c->p.order == 0)
            k_scatter_q<0><<<blocks(c->n), kBlock, 0, c->s>>>(c->parts, o0, o1, c->hr[0], c->hr[1], rho.p, Nx, Ny);
        else if (c->p.order == 1)
            k_scatter_q<1><<<blocks(c->n), kBlock, 0, c->s>>>(c->parts, o0, o1, c->hr[0], c->hr[1], rho.p, Nx, Ny);
        else
            k_scatter_q<2><<<blocks(c->n), kBlock, 0, c->s>>>(c->parts, o0, o1, c->hr[0], c->hr[1], rho.p, Nx, Ny);
        k_rhs<<<blocks(nn), kBlock, 0, c->s>>>(rho.p, b.p, nn, c->hr[0] * c->hr[1], cc);
        HU_LAUNCH_CHECK();
    };
    auto solve = [&](int k) {
        const pmg::Stats st = mg.solve(b.p, phi.p, tol, maxit);
        if (cycles) cycles[k] = st.cycles;
        if (resid) resid[k] = st.bnorm > 0 ? st.resid / st.bnorm : st.resid;
    };
    k_save_r<<<blocks(c->n), kBlock, 0, c->s>>>(c->parts, sx.p, sy.p, 0);          // saveR
    // solve 1: Ey, charge at t + dt/2 (PoissonSolver.cpp:154-166)
    drift(c, c->dt / 2);
    rhs();
    drift(c, -c->dt / 2);
    phi.zero(c->s);
    solve(0);
    FieldDesc E = c->fs.f[F_GE], H = c->fs.f[F_GH];
    k_fill_fields<<<blocks(nn), kBlock, 0, c->s>>>(phi.p, 1, E, gamma, hr1);
    // solve 2: Hz, bunch half a cell back (PoissonSolver.cpp:167-178)
    k_move_x<<<blocks(c->n), kBlock, 0, c->s>>>(c->parts, -0.5 * hr0 / gamma);
    rhs();
    k_shift_guess<<<blocks(nn), kBlock, 0, c->s>>>(phi.p, guess.p, Nx, Ny, 1.0);
    std::swap(phi.p, guess.p);
    solve(1);
    const double betaGamma = sqrt(gamma * gamma - 1);
    k_fill_fields<<<blocks(nn), kBlock, 0, c->s>>>(phi.p, 2, H, betaGamma, 2 * pm::kC * pm::kMu0 * hr1);
    // solve 3: Ex (PoissonSolver.cpp:179-195)
    k_save_r<<<blocks(c->n), kBlock, 0, c->s>>>(c->parts, sx.p, sy.p, 1);          // restoreR
    rhs();
    k_shift_guess<<<blocks(nn), kBlock, 0, c->s>>>(phi.p, guess.p, Nx, Ny, -0.5);
    std::swap(phi.p, guess.p);
    solve(2);
    k_fill_fields<<<blocks(nn), kBlock, 0, c->s>>>(phi.p, 0, E, 1.0, hr0);
    k_reset_outside<<<blocks(nn), kBlock, 0, c->s>>>(E, c->preset.p);
    HU_LAUNCH_CHECK();
    c->gJ.zero(c->s);                                                               // main.cpp:374
    copy_to_window(c);
    exchange(c, 0);
    exchange(c, 1);
    exchange(c, 2);
    HU_CHECK(hipStreamSynchronize(c->s));
}

void full_step(csrst_ctx *c)
{
    first_step(c);
    second_step(c);
    if (c->p.statModulus > 0 && c->steps % c->p.statModulus == c->p.statModulus - 1) wake_sample(c);
    c->gJ.zero(c->s);
    move_mesh(c, nullptr, nullptr);
    ++c->steps;
}

}  // namespace

// ===========================================================================
extern "C" {

const char *csrst_version(void) { return "csrst-mi355x 0.1 (gfx950)"; }

const char *csrst_last_error(const void *ctx)
{
    // both handle types start with the error string member
    if (ctx) return static_cast<const csrst_ctx *>(ctx)->err.c_str();
    return hu::global_error().c_str();
}

int csrst_device_count(int *n)
{
    int k = 0;
    if (hipGetDeviceCount(&k) != hipSuccess) k = 0;
    *n = k;
    return CSRST_OK;
}

int csrst_create(const csrst_params *p, csrst_ctx **out)
{
    std::unique_ptr<csrst_ctx> c(new csrst_ctx());
    try {
        if (!p || !out) throw Error(CSRST_EINVAL, "null argument");
        if (p->Nx < 8 || p->Ny < 8) throw Error(CSRST_EINVAL, "grid too small");
        if (p->order < 0 || p->order > 2) throw Error(CSRST_EINVAL, "order must be 0, 1 or 2");
        c->p = *p;
        HU_CHECK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
        c->own = c->s;
        geom::bend_init(c->bend, p->lenBefore, p->lenAfter, p->width, p->phiDeg, p->ekinMeV, p->bz, p->Nx, p->Ny);
        c->hr[0] = c->bend.hr[0];
        c->hr[1] = c->bend.hr[1];
        c->dt = c->hr[0] / pm::kC;
        c->mg = margins(p->order);
        c->alloc_field(c->gE, c->fs.f[F_GE], p->Nx, p->Ny);
        c->alloc_field(c->gH, c->fs.f[F_GH], p->Nx, p->Ny);
        c->alloc_field(c->gJ, c->fs.f[F_GJ], p->Nx, p->Ny);
        // geometry: mask and boundary cells (host, once)
        const int Nx = p->Nx, Ny = p->Ny;
        c->hmask.assign((size_t)Nx * Ny, 0);
        c->hbij.clear();
        c->hbdata.clear();
        for (int cy = 0; cy < Ny; ++cy)
            for (int cx = 0; cx < Nx; ++cx) {
                geom::CutCell cut;
                bool ok;
                const int k = geom::classify(c->bend, c->bend.origin, c->bend.hr, cx, cy, Nx, Ny, &cut, &ok);
                if (!ok) throw Error(CSRST_EGEOM, hu::fmt("crossBoundary out of range at cell (%d, %d)", cx, cy));
                if (k == 1) c->hmask[(size_t)cy * Nx + cx] = 1;
                if (k == 2) {
                    c->hbij.push_back(cx); c->hbij.push_back(cy);
                    c->hbdata.push_back(cut.lam[0]); c->hbdata.push_back(cut.lam[1]);
                    c->hbdata.push_back(cut.area);
                    c->hbdata.push_back(cut.lamN[0]); c->hbdata.push_back(cut.lamN[1]);
                }
            }
        upload_geometry(c.get());
        c->bb.alloc(2);
        c->derr.alloc(1);
        c->derr.zero(c->s);
        c->partial.alloc((size_t)kRedBlocks * 5);
        c->mom.alloc(8);
        HU_CHECK(hipStreamSynchronize(c->s));
        *out = c.release();
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(nullptr, e);
    }
}

int csrst_destroy(csrst_ctx *c)
{
    if (!c) return CSRST_OK;
    if (c->s) (void)hipStreamSynchronize(c->s);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
    return CSRST_OK;
}

int csrst_set_stream(csrst_ctx *c, void *stream)
{
    try {
        if (!c) throw Error(CSRST_EINVAL, "null context");
        HU_CHECK(hipStreamSynchronize(c->s));
        c->s = stream ? (hipStream_t)stream : c->own;
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_set_mask(csrst_ctx *c, const uint8_t *mask)
{
    try {
        if (!c || !mask) throw Error(CSRST_EINVAL, "null argument");
        if (c->ready) throw Error(CSRST_ESTATE, "the geometry is set before csrst_init_bunch");
        const size_t nn = (size_t)c->p.Nx * c->p.Ny;
        for (size_t k = 0; k < nn; ++k)
            if (mask[k] > 1) throw Error(CSRST_EINVAL, "mask values must be 0 or 1");
        c->hmask.assign(mask, mask + nn);
        upload_geometry(c);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_set_boundary_cells(csrst_ctx *c, const int *ij, const double *data, int count)
{
    try {
        if (!c || count < 0 || (count > 0 && (!ij || !data))) throw Error(CSRST_EINVAL, "bad boundary-cell list");
        if (c->ready) throw Error(CSRST_ESTATE, "the geometry is set before csrst_init_bunch");
        const int Nx = c->p.Nx, Ny = c->p.Ny;
        std::vector<uint8_t> seen((size_t)Nx * Ny, 0);
        for (int k = 0; k < count; ++k) {
            const int i = ij[2 * k], j = ij[2 * k + 1];
            // the updates read E(i + 1, j), E(i, j + 1), H(i - 1, j), H(i, j - 1): interior cells only
            if (i < 0 || i >= Nx || j < 0 || j >= Ny) throw Error(CSRST_EINVAL, hu::fmt("cell %d outside the grid", k));
            if (seen[(size_t)j * Nx + i]++) throw Error(CSRST_EINVAL, hu::fmt("cell (%d, %d) listed twice", i, j));
            if (!(data[5 * k + 2] > 0)) throw Error(CSRST_EINVAL, hu::fmt("cell %d: area must be positive", k));
        }
        c->hbij.assign(ij, ij + 2 * (size_t)count);
        c->hbdata.assign(data, data + 5 * (size_t)count);
        upload_geometry(c);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_geometry(csrst_ctx *c, double *o)
{
    if (!c || !o) return CSRST_EINVAL;
    o[0] = c->hr[0]; o[1] = c->hr[1]; o[2] = c->bend.origin[0]; o[3] = c->bend.origin[1];
    o[4] = c->dt; o[5] = c->bend.R; o[6] = c->bend.phi;
    return CSRST_OK;
}

int csrst_boundary_cells(csrst_ctx *c, uint8_t *mask, int *ij, double *data, int cap, int *count)
{
    if (!c) return CSRST_EINVAL;
    const int nb = (int)(c->hbij.size() / 2);
    if (count) *count = nb;
    if (mask) memcpy(mask, c->hmask.data(), c->hmask.size());
    if (ij && data) {
        if (cap < nb) { c->err = "capacity too small"; return CSRST_EINVAL; }
        memcpy(ij, c->hbij.data(), sizeof(int) * 2 * nb);
        memcpy(data, c->hbdata.data(), sizeof(double) * 5 * nb);
    }
    return CSRST_OK;
}

// PartBunch set-up (PartBunch.cpp:1817-1844 sort), then makeReady(dt/2)
// (PartBunch.cpp:164-177) when oldR is null, else oldR as given (the restart
// path of main.cpp:147-153,228, LoadH5Part::restoreBunch); the window copy
// and exchanges of main.cpp:335-341; the step counter starts at step.
static void init_bunch(csrst_ctx *c, const double *R, const double *P, const double *Q, const double *oldR, int n,
                       int step)
{
    {
        if (!R || !P || !Q || n <= 0) throw Error(CSRST_EINVAL, "bad bunch arguments");
        if (step < 0) throw Error(CSRST_EINVAL, "negative step");
        c->n = n;
        // Distribution::getDomain (Distribution.cpp:181-207) + 5 cells, main.cpp:200-217
        double mn[2] = {R[0], R[1]}, mx[2] = {R[0], R[1]};
        for (int k = 0; k < n; ++k)
            for (int d = 0; d < 2; ++d) {
                mn[d] = std::min(mn[d], R[2 * k + d]);
                mx[d] = std::max(mx[d], R[2 * k + d]);
            }
        for (int d = 0; d < 2; ++d) {
            c->ddim[d][0] = (int)floor((mn[d] - c->bend.origin[d]) / c->hr[d]) - 5;
            c->ddim[d][1] = (int)floor((mx[d] - c->bend.origin[d]) / c->hr[d]) + 5;
        }
        if (c->ddim[0][0] < 1 || c->ddim[1][0] < 1 || c->ddim[0][1] > c->p.Nx - 2 || c->ddim[1][1] > c->p.Ny - 2)
            throw Error(CSRST_EWINDOW, "bunch window does not fit the global grid");
        c->mmo[0] = c->bend.origin[0] + c->ddim[0][0] * c->hr[0];
        c->mmo[1] = c->bend.origin[1] + c->ddim[1][0] * c->hr[1];
        c->W = c->ddim[0][1] - c->ddim[0][0] + 1;
        c->H = c->ddim[1][1] - c->ddim[1][0] + 1;
        c->alloc_field(c->mEb, c->fs.f[F_ME], c->W, c->H);
        c->alloc_field(c->mHb, c->fs.f[F_MH], c->W, c->H);
        c->alloc_field(c->mJb, c->fs.f[F_MJ], c->W, c->H);
        HU_CHECK(hipStreamSynchronize(c->s));   // the blocking copies below use the null stream
        {
            const size_t plane2 = c->mEb.n;
            c->snapb.alloc(3 * plane2);
            for (int f = F_ME; f <= F_MJ; ++f) {
                double *b = c->snapb.p + (size_t)(f - F_ME) * plane2;
                c->snap.f[f] = FieldDesc{b, b + plane2 / 2, c->W, c->H};
            }
        }
        // PartBunch::sortParticles (PartBunch.cpp:1817-1844): y slices, x
        // ascending inside a slice (stable for equal x).
        double rmin1 = R[1];
        for (int k = 0; k < n; ++k) rmin1 = std::min(rmin1, R[2 * k + 1]);
        const int pd1 = (int)floor((rmin1 - c->mmo[1]) / c->hr[1]);
        std::vector<int> idx(n), slice(n);
        for (int k = 0; k < n; ++k) {
            idx[k] = k;
            slice[k] = (int)floor((R[2 * k + 1] - c->mmo[1]) / c->hr[1]) - pd1;
        }
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
            if (slice[a] != slice[b]) return slice[a] < slice[b];
            return R[2 * a] < R[2 * b];
        });
        std::vector<double> hx(n), hy(n), hpx(n), hpy(n), hq(n), hox(n), hoy(n);
        const double halfdt = c->dt / 2 * pm::kC;   // makeReady(dt/2), PartBunch.cpp:164-177
        for (int k = 0; k < n; ++k) {
            const int s = idx[k];
            hx[k] = R[2 * s]; hy[k] = R[2 * s + 1];
            hpx[k] = P[2 * s]; hpy[k] = P[2 * s + 1];
            hq[k] = Q[s];
            if (oldR) {
                hox[k] = oldR[2 * s]; hoy[k] = oldR[2 * s + 1];
            } else {
                const double g = (double)pm::q_rsqrt((float)(1 + (hpx[k] * hpx[k] + hpy[k] * hpy[k])));
                hox[k] = hx[k] - halfdt * hpx[k] * g;
                hoy[k] = hy[k] - halfdt * hpy[k] * g;
            }
        }
        DBuf<double> *bufs[7] = {&c->rx, &c->ry, &c->ox, &c->oy, &c->px, &c->py, &c->q};
        std::vector<double> *hv[7] = {&hx, &hy, &hox, &hoy, &hpx, &hpy, &hq};
        for (int k = 0; k < 7; ++k) {
            bufs[k]->alloc(n);
            HU_CHECK(hipMemcpy(bufs[k]->p, hv[k]->data(), n * sizeof(double), hipMemcpyHostToDevice));
        }
        c->parts = Parts{c->rx.p, c->ry.p, c->ox.p, c->oy.p, c->px.p, c->py.p, c->q.p, n};
        const size_t cnt = (size_t)n * kMaxSeg * kMaxNodes;
        c->keys.alloc(cnt); c->keys2.alloc(cnt); c->vals.alloc(cnt); c->vals2.alloc(cnt);
        c->cubBytes = 0;
        HU_CHECK(hipcub::DeviceRadixSort::SortPairs((void *)nullptr, c->cubBytes, c->keys.p, c->keys2.p, c->vals.p,
                                                    c->vals2.p, (int)cnt, 0, 32, c->s));
        c->cubTmp.alloc(c->cubBytes + 16);
        copy_to_window(c);
        exchange(c, 0);
        exchange(c, 1);
        exchange(c, 2);
        const int ns = std::max(1, c->W - 4);
        c->wacc.alloc(ns);
        c->wacc.zero(c->s);
        c->gJ.zero(c->s);
        c->steps = step;
        c->ready = true;
        HU_CHECK(hipStreamSynchronize(c->s));
    }
}

int csrst_init_bunch(csrst_ctx *c, const double *R, const double *P, const double *Q, int n)
{
    try {
        if (!c) throw Error(CSRST_EINVAL, "null context");
        init_bunch(c, R, P, Q, nullptr, n, 0);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_restore_bunch(csrst_ctx *c, const double *R, const double *oldR, const double *P, const double *Q, int n,
                        int step)
{
    try {
        if (!c || !oldR) throw Error(CSRST_EINVAL, "null argument");
        init_bunch(c, R, P, Q, oldR, n, step);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_poisson_compute_field(csrst_ctx *c, double tol, int maxit, int *cycles, double *resid)
{
    try {
        if (!c || !c->ready) throw Error(CSRST_ESTATE, "csrst_poisson_compute_field before csrst_init_bunch");
        if (c->steps != 0) throw Error(CSRST_ESTATE, "the Poisson start must precede the first step");
        if (maxit < 0 || !(tol >= 0)) throw Error(CSRST_EINVAL, "bad tolerance or iteration limit");
        poisson_compute_field(c, tol, maxit, cycles, resid);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_substep(csrst_ctx *c, int op)
{
    try {
        if (!c || !c->ready) throw Error(CSRST_ESTATE, "not initialised");
        switch (op) {
        case 0: push(c, 0); break;
        case 1: push(c, 1); break;
        case 2: h_half(c); break;
        case 3: e_x(c); break;
        case 4: e_y(c); break;
        case 5: exchange(c, 0); break;
        case 6: exchange(c, 1); break;
        case 7: exchange(c, 2); break;
        case 8: sc_hz(c); break;
        case 9: sc_ex(c); break;
        case 10: sc_ey(c); break;
        case 11: drift(c, -0.5 * c->dt); break;
        case 12: wake_sample(c); break;
        case 13: c->gJ.zero(c->s); break;
        case 14: move_mesh(c, nullptr, nullptr); ++c->steps; break;
        default: throw Error(CSRST_EINVAL, hu::fmt("unknown substep %d", op));
        }
        HU_CHECK(hipStreamSynchronize(c->s));
        check_err(c);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_step(csrst_ctx *c, int nsteps)
{
    try {
        if (!c || !c->ready) throw Error(CSRST_ESTATE, "csrst_step before csrst_init_bunch");
        for (int k = 0; k < nsteps; ++k) full_step(c);
        HU_CHECK(hipStreamSynchronize(c->s));
        check_err(c);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_first_step(csrst_ctx *c)
{
    try {
        if (!c || !c->ready) throw Error(CSRST_ESTATE, "not initialised");
        first_step(c);
        HU_CHECK(hipStreamSynchronize(c->s));
        check_err(c);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_second_step(csrst_ctx *c)
{
    try {
        if (!c || !c->ready) throw Error(CSRST_ESTATE, "not initialised");
        second_step(c);
        HU_CHECK(hipStreamSynchronize(c->s));
        check_err(c);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_move_window(csrst_ctx *c, int *moveI, int *moveJ)
{
    try {
        if (!c || !c->ready) throw Error(CSRST_ESTATE, "not initialised");
        if (c->p.statModulus > 0 && c->steps % c->p.statModulus == c->p.statModulus - 1) wake_sample(c);
        c->gJ.zero(c->s);
        move_mesh(c, moveI, moveJ);
        ++c->steps;
        HU_CHECK(hipStreamSynchronize(c->s));
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

// copyFields (main.cpp:534-548) with Communicator::copy (Communicator.hh:212-268):
// the window cells (i, j) whose global index (ddim0 + i, ddim1 + j) lies in the
// global domain overwrite the global E and both H planes, and J when
// copyCurrent is set; guard layers are not copied.  Strided device copies.
int csrst_copy_fields(csrst_ctx *c, int copyCurrent)
{
    try {
        if (!c || !c->ready) throw Error(CSRST_ESTATE, "not initialised");
        const int Nx = c->p.Nx, Ny = c->p.Ny, W = c->W, H = c->H;
        const int ox = c->ddim[0][0], oy = c->ddim[1][0];
        const int i0 = std::max(0, -ox), i1 = std::min(W - 1, Nx - 1 - ox);
        const int j0 = std::max(0, -oy), j1 = std::min(H - 1, Ny - 1 - oy);
        if (i0 <= i1 && j0 <= j1) {
            const size_t gp = (size_t)Nx + 2, wp = (size_t)W + 2;
            const size_t gplane = gp * ((size_t)Ny + 2), wplane = wp * ((size_t)H + 2);
            auto copy = [&](DBuf<double> &g, const DBuf<double> &w) {
                if (g.n != 2 * gplane || w.n != 2 * wplane) throw Error(CSRST_ESTATE, "field sizes");
                for (int k = 0; k < 2; ++k) {
                    double *dst = g.p + k * gplane + (size_t)(oy + j0 + 1) * gp + (size_t)(ox + i0 + 1);
                    const double *src = w.p + k * wplane + (size_t)(j0 + 1) * wp + (size_t)(i0 + 1);
                    HU_CHECK(hipMemcpy2DAsync(dst, gp * sizeof(double), src, wp * sizeof(double),
                                              (size_t)(i1 - i0 + 1) * sizeof(double), (size_t)(j1 - j0 + 1),
                                              hipMemcpyDeviceToDevice, c->s));
                }
            };
            copy(c->gE, c->mEb);
            copy(c->gH, c->mHb);
            if (copyCurrent) copy(c->gJ, c->mJb);
        }
        HU_CHECK(hipStreamSynchronize(c->s));
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

static DBuf<double> *field_buf(csrst_ctx *c, int which, int &nn, int &mm)
{
    switch (which) {
    case 0: nn = c->p.Nx; mm = c->p.Ny; return &c->gE;
    case 1: nn = c->p.Nx; mm = c->p.Ny; return &c->gH;
    case 2: nn = c->W; mm = c->H; return &c->mEb;
    case 3: nn = c->W; mm = c->H; return &c->mHb;
    case 4: nn = c->W; mm = c->H; return &c->mJb;
    case 5: nn = c->p.Nx; mm = c->p.Ny; return &c->gJ;
    default: return nullptr;
    }
}

int csrst_download_field(csrst_ctx *c, int which, double *out, int cap, int *count)
{
    try {
        if (!c) throw Error(CSRST_EINVAL, "null context");
        int nn, mm;
        DBuf<double> *b = field_buf(c, which, nn, mm);
        if (!b || !b->p) throw Error(CSRST_EINVAL, "no such field");
        if (count) *count = (int)b->n;
        if (out) {
            if ((size_t)cap < b->n) throw Error(CSRST_EINVAL, "capacity too small");
            HU_CHECK(hipStreamSynchronize(c->s));
            HU_CHECK(hipMemcpy(out, b->p, b->bytes(), hipMemcpyDeviceToHost));
        }
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_upload_field(csrst_ctx *c, int which, const double *in, int count)
{
    try {
        if (!c || !in) throw Error(CSRST_EINVAL, "null argument");
        int nn, mm;
        DBuf<double> *b = field_buf(c, which, nn, mm);
        if (!b || !b->p) throw Error(CSRST_EINVAL, "no such field");
        if ((size_t)count != b->n) throw Error(CSRST_EINVAL, "size mismatch");
        HU_CHECK(hipStreamSynchronize(c->s));
        HU_CHECK(hipMemcpy(b->p, in, b->bytes(), hipMemcpyHostToDevice));
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_download_particles(csrst_ctx *c, double *R, double *P, double *oldR, int n)
{
    try {
        if (!c || n != c->n) throw Error(CSRST_EINVAL, "particle count mismatch");
        HU_CHECK(hipStreamSynchronize(c->s));
        std::vector<double> a(n), b(n);
        auto pair = [&](double *dst, DBuf<double> &x, DBuf<double> &y) {
            if (!dst) return;
            HU_CHECK(hipMemcpy(a.data(), x.p, n * sizeof(double), hipMemcpyDeviceToHost));
            HU_CHECK(hipMemcpy(b.data(), y.p, n * sizeof(double), hipMemcpyDeviceToHost));
            for (int k = 0; k < n; ++k) { dst[2 * k] = a[k]; dst[2 * k + 1] = b[k]; }
        };
        pair(R, c->rx, c->ry);
        pair(P, c->px, c->py);
        pair(oldR, c->ox, c->oy);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_download_charges(csrst_ctx *c, double *Q, int n)
{
    try {
        if (!c || !Q || n != c->n) throw Error(CSRST_EINVAL, "particle count mismatch");
        HU_CHECK(hipStreamSynchronize(c->s));
        HU_CHECK(hipMemcpy(Q, c->q.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_info(csrst_ctx *c, double *info)
{
    if (!c || !info) return CSRST_EINVAL;
    info[0] = c->dt; info[1] = c->mmo[0]; info[2] = c->mmo[1];
    info[3] = c->ddim[0][0]; info[4] = c->ddim[1][0]; info[5] = c->W; info[6] = c->H;
    info[7] = c->steps; info[8] = std::max(0, c->W - 4);
    return CSRST_OK;
}

int csrst_wake(csrst_ctx *c, double *out, int cap, int *count)
{
    try {
        if (!c || !c->ready) throw Error(CSRST_ESTATE, "not initialised");
        const int ns = std::max(0, c->W - 4);
        if (count) *count = ns;
        if (out) {
            if (cap < ns) throw Error(CSRST_EINVAL, "capacity too small");
            std::vector<DD> h(ns);
            HU_CHECK(hipStreamSynchronize(c->s));
            HU_CHECK(hipMemcpy(h.data(), c->wacc.p, ns * sizeof(DD), hipMemcpyDeviceToHost));
            for (int k = 0; k < ns; ++k) out[k] = h[k].hi + h[k].lo;
        }
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_wake_reset(csrst_ctx *c)
{
    try {
        if (!c || !c->ready) throw Error(CSRST_ESTATE, "not initialised");
        c->wacc.zero(c->s);
        HU_CHECK(hipStreamSynchronize(c->s));
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

// Bend::isInsideBend (Bend.hh:307-323): rotate R - center by phi; inside when
// x of R - center is positive and x of the rotated vector negative
int csrst_is_inside_bend(csrst_ctx *c, const double *R, int *inside)
{
    if (!c || !R || !inside) return CSRST_EINVAL;
    const geom::Bend &b = c->bend;
    const double dx0 = R[0] - b.center[0], dx1 = R[1] - b.center[1];
    const double rot0 = cos(b.phi) * dx0 - sin(b.phi) * dx1;
    *inside = (dx0 < 0.0 || rot0 > 0.0) ? 0 : (dx0 > 0.0 && rot0 < 0.0) ? 1 : 0;
    return CSRST_OK;
}

// particle source of the beam statistics: the FP64 SoA bunch
struct BeamSrc {
    const double *rx, *ry, *px, *py;
    __device__ void get(int k, double &x, double &y, double &a, double &b) const
    {
        x = rx[k]; y = ry[k]; a = px[k]; b = py[k];
    }
};

int csrst_beam_stats(csrst_ctx *c, double *out)
{
    try {
        if (!c || !c->ready || !out) throw Error(CSRST_ESTATE, "not initialised");
        if (!c->bsPart.p) {
            c->bsPart.alloc((size_t)bs::kBlocks * 11);
            c->bsLim.alloc((size_t)bs::kBlocks * 4);
        }
        double raw[bs::kRaw];
        bs::raw_stats(BeamSrc{c->parts.rx, c->parts.ry, c->parts.px, c->parts.py}, c->n, c->s, c->bsPart.p,
                      c->bsLim.p, raw);
        bs::parameters(raw, out);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_line_fields(csrst_ctx *c, double *out, int cap, int *count)
{
    try {
        if (!c || !c->ready) throw Error(CSRST_ESTATE, "not initialised");
        double st[CSRST_BEAM_STATS];
        const int rc = csrst_beam_stats(c, st);
        if (rc != CSRST_OK) return rc;
        const double R0 = st[7], R1 = st[8], P0 = st[9], P1 = st[10];
        const double h0 = c->hr[0], h1 = c->hr[1], o0 = c->bend.origin[0], o1 = c->bend.origin[1];
        const int Nx = c->p.Nx, Ny = c->p.Ny;
        // the line's column range from the bunch's cells (SaveEMFields.cpp:55-71)
        const int gb0 = (int)floor((st[14] - o0) / h0), gb1 = (int)floor((st[16] - o0) / h0);
        double l0 = floor(1.5 * gb0 - 0.5 * gb1), u0 = 2 * gb1 - gb0;
        double l1 = (R1 + (o0 + l0 * h0 - R0) / P0 * P1 - o1) / h1;
        double u1 = (R1 + (o0 + u0 * h0 - R0) / P0 * P1 - o1) / h1;
        const long startIdx = (long)l0;
        const int ns = (int)floor(u0 - l0 + 0.5) + 1;
        if (count) *count = ns;
        if (!out) return CSRST_OK;
        if (cap < ns) throw Error(CSRST_EINVAL, "capacity too small");
        // getPartInside (SaveEMFields.cpp:162-189) on the one-rank field domain
        {
            const int lower = (int)floor(l0 + 0.5), upper = (int)floor(u0 + 0.5);
            int minL = upper, maxL = upper;
            const double p0 = u0 - l0, p1 = u1 - l1;
            for (int i = lower; i <= upper; ++i) {
                const int yi = (int)floor(l1 + (i - lower) / p0 * p1);
                if (i >= 0 && i < Nx && yi >= 0 && yi < Ny) {
                    if (i < minL) minL = i;
                    if (i > minL) maxL = i;   // as written at SaveEMFields.cpp:180
                }
            }
            const double t1 = l1 + (minL - lower) / p0 * p1;
            u1 = l1 + (maxL - lower) / p0 * p1;
            u0 = maxL;
            l0 = minL;
            l1 = t1;
        }
        int dom[2][2] = {{(int)floor(l0 + 0.5), (int)floor(u0 + 0.5)}, {(int)floor(l1), (int)floor(u1)}};
        const int fdom[2][2] = {{0, Nx - 1}, {0, Ny - 1}};
        bool inside = true;
        for (int d = 0; d < 2; ++d) {
            const int lo = std::max(fdom[d][0], dom[d][0]), hi = std::min(fdom[d][1], dom[d][1]);
            if (lo > hi) { inside = false; break; }
            dom[d][0] = lo; dom[d][1] = hi;
        }
        std::vector<double> line(4 * (size_t)ns, 0.0);
        if (inside) {
            const int n = dom[0][1] - dom[0][0] + 1;
            DBuf<double> dev;
            dev.alloc(4 * (size_t)n);
            k_line<<<blocks(n), kBlock, 0, c->s>>>(c->fs.f[F_GE], c->fs.f[F_GH], c->fs.f[F_GJ], c->p.order,
                                                   c->p.s2bug, o0, o1, h0, h1, R0, R1, P0, P1, dom[0][0], n, dev.p);
            HU_LAUNCH_CHECK();
            HU_CHECK(hipMemcpyAsync(line.data() + 4 * (size_t)(dom[0][0] - startIdx), dev.p, 4 * (size_t)n * 8,
                                    hipMemcpyDeviceToHost, c->s));
            HU_CHECK(hipStreamSynchronize(c->s));
        }
        long ii = startIdx;
        for (int i = 0; i < ns; ++i, ++ii) {
            const double c0 = o0 + ii * h0 - R0, c1 = c0 / P0 * P1;
            double *row = out + 7 * (size_t)i;
            row[0] = sqrt(c0 * c0 + c1 * c1) * (c0 > 0 ? 1 : -1);
            for (int k = 0; k < 4; ++k) row[1 + k] = line[4 * (size_t)i + k];
            row[5] = c0 + R0;
            row[6] = c1 + R1;
        }
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_gauss_residual(csrst_ctx *c, double *out, int cap, int *count)
{
    try {
        if (!c || !c->ready) throw Error(CSRST_ESTATE, "not initialised");
        const int W = c->W, H = c->H;
        const size_t nn = (size_t)W * H;
        if (count) *count = (int)nn;
        if (!out) return CSRST_OK;
        if ((size_t)cap < nn) throw Error(CSRST_EINVAL, "capacity too small");
        DBuf<double> rho, res;
        rho.alloc(nn);
        res.alloc(nn);
        rho.zero(c->s);
        const double o0 = c->mmo[0], o1 = c->mmo[1], h0 = c->hr[0], h1 = c->hr[1];
        if (c->p.order == 0)
            k_scatter_q<0><<<blocks(c->n), kBlock, 0, c->s>>>(c->parts, o0, o1, h0, h1, rho.p, W, H);
        else if (c->p.order == 1)
            k_scatter_q<1><<<blocks(c->n), kBlock, 0, c->s>>>(c->parts, o0, o1, h0, h1, rho.p, W, H);
        else
            k_scatter_q<2><<<blocks(c->n), kBlock, 0, c->s>>>(c->parts, o0, o1, h0, h1, rho.p, W, H);
        k_gauss<<<grid2(W, H), kBlock, 0, c->s>>>(c->fs.f[F_ME], rho.p, 1.0 / h0, 1.0 / h1, h0 * h1, res.p);
        HU_LAUNCH_CHECK();
        HU_CHECK(hipMemcpyAsync(out, res.p, nn * sizeof(double), hipMemcpyDeviceToHost, c->s));
        HU_CHECK(hipStreamSynchronize(c->s));
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_field_energy(csrst_ctx *c, double *out)
{
    try {
        if (!c || !c->ready || !out) throw Error(CSRST_ESTATE, "not initialised");
        // global grid + window - the global nodes under the window
        const int Nx = c->p.Nx, Ny = c->p.Ny;
        const int s0 = c->ddim[0][0], s1 = c->ddim[1][0];
        const int lo0 = std::max(0, s0), lo1 = std::max(0, s1);
        const int hi0 = std::min(Nx - 1, s0 + c->W - 1), hi1 = std::min(Ny - 1, s1 + c->H - 1);
        long double acc[2] = {0, 0};
        std::vector<DD> part(kRedBlocks * 2);
        auto sum = [&](FieldDesc E, FieldDesc H, int i0, int j0, int ni, int nj, long double sign) {
            if (ni <= 0 || nj <= 0) return;
            k_energy_sums<<<kRedBlocks, kBlock, 0, c->s>>>(E, H, i0, j0, ni, nj, c->partial.p);
            HU_LAUNCH_CHECK();
            HU_CHECK(hipMemcpyAsync(part.data(), c->partial.p, kRedBlocks * 2 * sizeof(DD), hipMemcpyDeviceToHost,
                                    c->s));
            HU_CHECK(hipStreamSynchronize(c->s));
            for (int b = 0; b < kRedBlocks; ++b)
                for (int m = 0; m < 2; ++m)
                    acc[m] += sign * ((long double)part[b * 2 + m].hi + (long double)part[b * 2 + m].lo);
        };
        sum(c->fs.f[F_GE], c->fs.f[F_GH], 0, 0, Nx, Ny, 1);
        sum(c->fs.f[F_ME], c->fs.f[F_MH], 0, 0, c->W, c->H, 1);
        sum(c->fs.f[F_GE], c->fs.f[F_GH], lo0, lo1, hi0 - lo0 + 1, hi1 - lo1 + 1, -1);
        const long double area = (long double)c->hr[0] * c->hr[1];
        out[0] = (double)(acc[0] * area);
        out[1] = (double)(acc[1] * area);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int csrst_diag(csrst_ctx *c, double *out)
{
    try {
        if (!c || !c->ready || !out) throw Error(CSRST_ESTATE, "not initialised");
        moments(c);
        double m[5];
        HU_CHECK(hipMemcpyAsync(m, c->mom.p, 5 * sizeof(double), hipMemcpyDeviceToHost, c->s));
        k_energy_partial<<<kRedBlocks, kBlock, 0, c->s>>>(c->fs.f[F_GE], c->fs.f[F_GH], c->partial.p);
        HU_LAUNCH_CHECK();
        std::vector<DD> part(kRedBlocks * 5);
        HU_CHECK(hipMemcpyAsync(part.data(), c->partial.p, kRedBlocks * sizeof(DD), hipMemcpyDeviceToHost, c->s));
        HU_CHECK(hipStreamSynchronize(c->s));
        long double en = 0;
        for (int b = 0; b < kRedBlocks; ++b) en += (long double)part[b].hi + (long double)part[b].lo;
        out[0] = m[0]; out[1] = m[2]; out[2] = m[1]; out[3] = m[3]; out[4] = m[4];
        out[5] = (double)(en * c->hr[0] * c->hr[1]);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

}  // extern "C"
