// csrst_sf.hip -- the reference's ShapeFunction and ScatterGatherImpl
// interfaces as stateless kernels on caller memory, FP64.
//
//   csrst_sg_gather / csrst_sg_scatter       ScatterGatherImpl<IntOp,double>
//                                            ::gather / ::scatter
//                                            (ScatterGatherImpl.hh:27-641)
//   csrst_sf_current_density                 ShapeFunction<IntOp>::
//                                            getCurrentDensity{X,Y}
//   csrst_sf_charge_density                  ShapeFunction<IntOp>::getChargeDensity
//   csrst_sf_charge_density_diff             ShapeFunction<IntOp>::getChargeDensityDiff
//                                            (ShapeFunction.hh:30-508)
//
// Fields are dense patches of nodes [ny][nx] (x fastest) of a mesh with the
// given origin and spacing; particle positions are (x, y) pairs, the layout
// of the reference's Vektor<double,2> attributes.  Every array argument is a
// device pointer; work is queued on the caller's stream and the call
// returns when it has finished.
//
// Arithmetic follows the reference order and this file is compiled without
// FMA contraction.  Sums over particles are made deterministic the way the
// reference orders them: every contribution is emitted into a fixed slot in
// (particle, statement) order, the slots are stably radix-sorted by node, and
// each node's run is summed sequentially from 0.0 -- the per-node sums equal
// a CPU loop over the particles bit for bit.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <mutex>
#include <string>

#include "csrst_hip.h"
#include "hip_util.h"
#include "pic_math.h"

using hu::DBuf;
using hu::Error;
using pm::vec2;

namespace {

constexpr int kB = 256;
constexpr unsigned kEmpty = 0xffffffffu;
constexpr int kMaxSeg = 4;       // segments of one trajectory (c dt <= h: at most 3)

struct Patch {
    int i0, j0, nx, ny;          // absolute index of node (0, 0) and extent
    double o[2], h[2];           // mesh origin (absolute node 0) and spacing
};

int blocks(size_t n) { return (int)((n + kB - 1) / kB); }

// slots per particle of each operation
int slots_current(int order) { return kMaxSeg * (order == 0 ? 2 : order == 1 ? 6 : 12); }
int slots_charge(int order) { return (order + 2) * (order + 2); }
int slots_scatter(int order) { return (order + 2) * (order + 2); }

// ---------------------------------------------------------------------------
// Emitters.  Each writes its particle's slots [k * S, (k + 1) * S): key = the
// node's index in the patch, kEmpty for an unused slot.  A node outside the
// patch sets err bit 1 and is dropped.
struct Slots {
    unsigned *key;
    double *val;
    int per;                     // slots per particle
    int *err;
};

__device__ inline void put(const Slots &s, const Patch &p, size_t slot, long long I, long long J, double v)
{
    const long long i = I - p.i0, j = J - p.j0;
    if (i < 0 || j < 0 || i >= p.nx || j >= p.ny) {
        atomicOr(s.err, 1);
        return;
    }
    s.key[slot] = (unsigned)(j * p.nx + i);
    s.val[slot] = v;
}

__device__ inline void clear_slots(const Slots &s, int k)
{
    for (int t = 0; t < s.per; ++t) s.key[(size_t)k * s.per + t] = kEmpty;
}

// getCurrentDensityBase{X,Y} + getCurrentDensity{X,Y} for one particle:
// the trajectory oldR -> R split at the (dual) cell lines, each segment's
// current on its stencil (ZerothOrderShapeFunction.cpp:202-294,
// FirstOrderShapeFunction.cpp:226-329, 540-683,
// SecondOrderShapeFunction.cpp:222-345, 583-826).
__global__ void k_emit_current(Patch p, int order, int comp, double dt, const double *oldR, const double *R,
                               const double *Q, int n, Slots s)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    clear_slots(s, k);
    const vec2<double> org = pm::mk(p.o[0], p.o[1]);
    const vec2<double> dx = pm::mk(p.h[0], p.h[1]);
    const double ox = oldR[2 * k], oy = oldR[2 * k + 1], rx = R[2 * k], ry = R[2 * k + 1];
    const vec2<double> o = pm::mk(ox - org.x, oy - org.y);
    const vec2<double> e = pm::mk(rx - org.x, ry - org.y);
    const vec2<double> od = pm::mk(1.0 / dx.x, 1.0 / dx.y);
    vec2<double> odt;
    if (order == 0) odt = pm::mk(1.0 / (dx.y * dt), 1.0 / (dx.x * dt));
    else odt = pm::mk(od.x / dt, od.y / dt);
    const vec2<double> D = pm::mk(rx - ox, ry - oy);
    const vec2<double> delta = pm::mk(D.x * od.x, D.y * od.y);
    const double covered = sqrt(D.x * D.x + D.y * D.y);
    pm::Crossings<double> c;
    if (!pm::crossings(c, o, e, dx, order == 1) || c.n - 1 > kMaxSeg) {
        atomicOr(s.err, 2);
        return;
    }
    const double q = Q[k];
    const int per = s.per / kMaxSeg;
    for (int g = 0; g < c.n - 1; ++g) {
        const vec2<double> xi = pm::mk(c.s[g].x * od.x, c.s[g].y * od.y);
        const vec2<double> xf = pm::mk(c.s[g + 1].x * od.x, c.s[g + 1].y * od.y);
        const vec2<double> seg = pm::mk(c.s[g + 1].x - c.s[g].x, c.s[g + 1].y - c.s[g].y);
        pm::SegCurrent<double> sc;
        pm::segment_current(sc, order, comp, xi, xf, seg, covered, q, odt, delta);
        int t = 0;
        for (int jj = 0; jj < sc.ny; ++jj)
            for (int ii = 0; ii < sc.nx; ++ii, ++t)
                put(s, p, (size_t)k * s.per + g * per + t, sc.I0 + ii, sc.J0 + jj, sc.w[t]);
    }
}

// getChargeDensityImpl (ZerothOrderShapeFunction.cpp:296-383,
// FirstOrderShapeFunction.cpp:331-430, SecondOrderShapeFunction.cpp:347-450):
// sign = +1; with sign = -1 at oldR the second half of getChargeDensityDiffImpl
template <int ORDER>
__device__ void emit_charge(const Slots &s, const Patch &p, size_t base, double x, double y, double q, double sign)
{
    constexpr int N = ORDER + 2;
    double wx[N], wy[N];
    int I0, J0;
    pm::charge_weights<ORDER>(pm::mk((x - p.o[0]) / p.h[0], (y - p.o[1]) / p.h[1]), q, wx, wy, I0, J0);
#pragma unroll
    for (int b = 0; b < N; ++b)
#pragma unroll
        for (int a = 0; a < N; ++a) {
            const double w = pm::charge_w<ORDER>(q, wx[a], wy[b]);
            put(s, p, base + b * N + a, I0 + a, J0 + b, sign > 0 ? w : -w);
        }
}

template <int ORDER>
__global__ void k_emit_charge(Patch p, const double *R, const double *oldR, const double *Q, int n, Slots s)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    clear_slots(s, k);
    constexpr int N2 = (ORDER + 2) * (ORDER + 2);
    emit_charge<ORDER>(s, p, (size_t)k * s.per, R[2 * k], R[2 * k + 1], Q[k], 1.0);
    if (oldR) emit_charge<ORDER>(s, p, (size_t)k * s.per + N2, oldR[2 * k], oldR[2 * k + 1], Q[k], -1.0);
}

// ScatterGatherImpl::scatter of one value (ScatterGatherImpl.hh:136-194,
// 303-392, 529-641): edge component lc (0 Ex, 1 Ey) or the cell field (lc 2).
__global__ void k_emit_scatter(Patch p, int order, int lc, const double *R, const double *val, int n, Slots s)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    clear_slots(s, k);
    const vec2<double> od = pm::mk(1.0 / p.h[0], 1.0 / p.h[1]);
    const vec2<double> pp = pm::mk((R[2 * k] - p.o[0]) * od.x, (R[2 * k + 1] - p.o[1]) * od.y);
    const double v = val[k];
    pm::Stencil<double> st = lc < 2 ? pm::edge_stencil(order, lc, pp) : pm::cell_stencil(order, pp);
    if (lc == 2 && order == 1) {
        // the S1 cell scatter forms its outer weights directly (ScatterGatherImpl.hh:372-380)
        const double dx = pp.x - ((double)(st.I0 + 1) + 0.5), dy = pp.y - ((double)(st.J0 + 1) + 0.5);
        st.wx[0] = 0.5 * (0.5 - dx) * (0.5 - dx);
        st.wy[0] = 0.5 * (0.5 - dy) * (0.5 - dy);
    }
    // node (0, 0) of the stencil is absolute node I0 + i0 ... the stencil is
    // relative to the patch origin, which is absolute node i0
    size_t t = (size_t)k * s.per;
    for (int b = 0; b < st.n; ++b)
        for (int a = 0; a < st.n; ++a, ++t)
            put(s, p, t, (long long)p.i0 + st.I0 + a, (long long)p.j0 + st.J0 + b, pm::scatter_w(st, a, b, v));
}

// ScatterGatherImpl::gather (ScatterGatherImpl.hh:70-134, 202-301, 407-527)
__global__ void k_gather(Patch p, int order, int lc, int s2bug, const double *f, const double *R, int n, double *out,
                         int *err)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const vec2<double> od = pm::mk(1.0 / p.h[0], 1.0 / p.h[1]);
    const vec2<double> pp = pm::mk((R[2 * k] - p.o[0]) * od.x, (R[2 * k + 1] - p.o[1]) * od.y);
    bool bad = false;
    auto fetch = [&](long long I, long long J) {
        if (I < 0 || J < 0 || I >= p.nx || J >= p.ny) { bad = true; return 0.0; }
        return f[(size_t)J * p.nx + (size_t)I];
    };
    const pm::Stencil<double> st = lc < 2 ? pm::edge_stencil(order, lc, pp) : pm::cell_stencil(order, pp);
    out[k] = pm::gather(st, s2bug != 0, fetch);
    if (bad) atomicOr(err, 1);
}

// Sum of each node's run, in slot order, from 0.0.  mode 0: f[node] = 0.0 +
// sum / div (the ShapeFunction results, written over a cleared patch; div is
// hx hy for a density, 1 for a current); mode 1: f[node] += sum (scatter
// adds into the caller's field).
__global__ void k_runs(const unsigned *key, const double *val, size_t cnt, double *f, double div, int mode)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const unsigned kk = key[t];
    if (kk == kEmpty || (t > 0 && key[t - 1] == kk)) return;
    double s = 0.0;
    for (size_t u = t; u < cnt && key[u] == kk; ++u) s += val[u];
    if (mode == 0) f[kk] = 0.0 + s / div;
    else f[kk] += s;
}

// the patch cleared (fill kernel: nothing here is a memset node)
__global__ void k_fill(double *f, size_t n, double v)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) f[t] = v;
}
__global__ void k_zero_int(int *p) { *p = 0; }

// ---------------------------------------------------------------------------
// Scratch: keys and values twice (sort input and output), the radix sort's
// temporary storage and an error word, carved out of one block.
struct Scratch {
    unsigned *k0, *k1;
    double *v0, *v1;
    void *tmp;
    size_t tmpBytes;
    int *err;
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int key_bits(size_t nodes)
{
    int b = 1;
    while ((1ull << b) < (unsigned long long)nodes + 1) ++b;
    return b;
}

size_t sort_tmp_bytes(size_t cnt)
{
    size_t bytes = 0;
    HU_CHECK(hipcub::DeviceRadixSort::SortPairs((void *)nullptr, bytes, (unsigned *)nullptr, (unsigned *)nullptr,
                                                (double *)nullptr, (double *)nullptr, (int)cnt, 0, 32));
    return bytes;
}

size_t scratch_bytes(size_t cnt, size_t tmp)
{
    return 2 * align256(cnt * sizeof(unsigned)) + 2 * align256(cnt * sizeof(double)) + align256(tmp) + 256;
}

Scratch carve(unsigned char *base, size_t cnt, size_t tmp)
{
    Scratch s;
    size_t off = 0;
    auto take = [&](size_t b) { unsigned char *p = base + off; off += align256(b); return p; };
    s.k0 = (unsigned *)take(cnt * sizeof(unsigned));
    s.k1 = (unsigned *)take(cnt * sizeof(unsigned));
    s.v0 = (double *)take(cnt * sizeof(double));
    s.v1 = (double *)take(cnt * sizeof(double));
    s.tmp = take(tmp);
    s.tmpBytes = tmp;
    s.err = (int *)take(sizeof(int));
    return s;
}

// Stable sort of the slots by node, then the run sums into f.
void reduce(const Scratch &s, size_t cnt, size_t nodes, double *f, double div, int mode, hipStream_t st)
{
    size_t tb = s.tmpBytes;
    HU_CHECK(hipcub::DeviceRadixSort::SortPairs(s.tmp, tb, s.k0, s.k1, s.v0, s.v1, (int)cnt, 0, key_bits(nodes), st));
    k_runs<<<blocks(cnt), kB, 0, st>>>(s.k1, s.v1, cnt, f, div, mode);
    HU_LAUNCH_CHECK();
}

int read_err(const Scratch &s, hipStream_t st)
{
    int e = 0;
    HU_CHECK(hipMemcpyAsync(&e, s.err, sizeof(int), hipMemcpyDeviceToHost, st));
    HU_CHECK(hipStreamSynchronize(st));
    return e;
}

void check_patch(const Patch &p)
{
    if (p.nx <= 0 || p.ny <= 0) throw Error(CSRST_EINVAL, "empty patch");
    if ((unsigned long long)p.nx * (unsigned long long)p.ny >= 0xffffffffull)
        throw Error(CSRST_EINVAL, "patch has more than 2^32 - 1 nodes");
    if (!(p.h[0] > 0) || !(p.h[1] > 0)) throw Error(CSRST_EINVAL, "spacing must be positive");
}

Patch make_patch(int i0, int j0, int nx, int ny, const double origin[2], const double dx[2])
{
    if (!origin || !dx) throw Error(CSRST_EINVAL, "null origin or spacing");
    Patch p{i0, j0, nx, ny, {origin[0], origin[1]}, {dx[0], dx[1]}};
    check_patch(p);
    return p;
}

int sf_fail(const std::exception &e)
{
    const Error *he = dynamic_cast<const Error *>(&e);
    hu::global_error() = e.what();
    return he ? he->code : CSRST_EINVAL;
}

// The current deposit on a given scratch block.
void current_density(const Patch &p, int order, int comp, double *J, double dt, const double *oldR, const double *R,
                     const double *Q, int n, const Scratch &s, size_t cnt, hipStream_t st)
{
    const size_t nodes = (size_t)p.nx * p.ny;
    k_zero_int<<<1, 1, 0, st>>>(s.err);
    k_fill<<<blocks(nodes), kB, 0, st>>>(J, nodes, 0.0);
    if (n > 0) {
        const Slots sl{s.k0, s.v0, slots_current(order), s.err};
        k_emit_current<<<blocks(n), kB, 0, st>>>(p, order, comp, dt, oldR, R, Q, n, sl);
        HU_LAUNCH_CHECK();
        reduce(s, cnt, nodes, J, 1.0, 0, st);
    }
    const int e = read_err(s, st);
    if (e & 2) throw Error(CSRST_EINVAL, "a trajectory crossed more cell lines than one push can");
    if (e & 1) throw Error(CSRST_EWINDOW, "a current stencil left the patch (its contribution was dropped)");
}

void check_current_args(int order, int comp, double dt, const double *J, const double *oldR, const double *R,
                        const double *Q, int n)
{
    if (order < 0 || order > 2) throw Error(CSRST_EINVAL, "order must be 0..2");
    if (comp != 0 && comp != 1) throw Error(CSRST_EINVAL, "comp must be 0 or 1");
    if (!(dt > 0)) throw Error(CSRST_EINVAL, "dt must be positive");
    if (n < 0 || !J || (n > 0 && (!oldR || !R || !Q))) throw Error(CSRST_EINVAL, "bad particle arguments");
}

// ---------------------------------------------------------------------------
// Scratch of csrst_sf_current_density_noalloc: one block per device in
// static device memory, held until the process ends.  In host code `scratch`
// only names the symbol; its device address comes from hipGetSymbolAddress,
// once per device with that device current, and is kept per device.
constexpr size_t kScratchBytes = (size_t)32 << 20;   // 32 MiB
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};

unsigned char *device_scratch()
{
    int dev = 0;
    HU_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(CSRST_EINVAL, hu::fmt("device %d beyond the scratch table", dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        HU_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = (unsigned char *)p;
    }
    return g_scratch[dev];
}

}  // namespace

extern "C" {

int csrst_sg_gather(int order, int s2bug, int kind, const double *field, int nx, int ny, const double origin[2],
                    const double dx[2], const double *R, int n, double *out, void *stream)
{
    try {
        if (order < 0 || order > 2 || kind < 0 || kind > 2) throw Error(CSRST_EINVAL, "order 0..2, kind 0..2");
        if (n < 0 || (n > 0 && (!field || !R || !out))) throw Error(CSRST_EINVAL, "bad arguments");
        const Patch p = make_patch(0, 0, nx, ny, origin, dx);
        const hipStream_t st = (hipStream_t)stream;
        DBuf<int> err;
        err.alloc(1);
        k_zero_int<<<1, 1, 0, st>>>(err.p);
        if (n > 0) k_gather<<<blocks(n), kB, 0, st>>>(p, order, kind, s2bug, field, R, n, out, err.p);
        HU_LAUNCH_CHECK();
        int e = 0;
        HU_CHECK(hipMemcpyAsync(&e, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HU_CHECK(hipStreamSynchronize(st));
        if (e) throw Error(CSRST_EWINDOW, "a gather stencil left the patch (its nodes read as 0)");
        return CSRST_OK;
    } catch (const std::exception &e) {
        return sf_fail(e);
    }
}

int csrst_sg_scatter(int order, int kind, double *field, int nx, int ny, const double origin[2], const double dx[2],
                     const double *R, const double *val, int n, void *stream)
{
    try {
        if (order < 0 || order > 2 || kind < 0 || kind > 2) throw Error(CSRST_EINVAL, "order 0..2, kind 0..2");
        if (n < 0 || !field || (n > 0 && (!R || !val))) throw Error(CSRST_EINVAL, "bad arguments");
        const Patch p = make_patch(0, 0, nx, ny, origin, dx);
        if (n == 0) return CSRST_OK;
        const hipStream_t st = (hipStream_t)stream;
        const size_t cnt = (size_t)n * slots_scatter(order);
        const size_t tmp = sort_tmp_bytes(cnt);
        DBuf<unsigned char> buf;
        buf.alloc(scratch_bytes(cnt, tmp));
        const Scratch s = carve(buf.p, cnt, tmp);
        k_zero_int<<<1, 1, 0, st>>>(s.err);
        const Slots sl{s.k0, s.v0, slots_scatter(order), s.err};
        k_emit_scatter<<<blocks(n), kB, 0, st>>>(p, order, kind, R, val, n, sl);
        HU_LAUNCH_CHECK();
        reduce(s, cnt, (size_t)nx * ny, field, 1.0, 1, st);
        if (read_err(s, st) & 1) throw Error(CSRST_EWINDOW, "a scatter stencil left the patch (dropped)");
        return CSRST_OK;
    } catch (const std::exception &e) {
        return sf_fail(e);
    }
}

int csrst_sf_current_density(int order, int comp, double *J, int i0, int j0, int nx, int ny, const double origin[2],
                             const double dx[2], double dt, const double *oldR, const double *R, const double *Q,
                             int n, void *stream)
{
    try {
        check_current_args(order, comp, dt, J, oldR, R, Q, n);
        const Patch p = make_patch(i0, j0, nx, ny, origin, dx);
        const size_t cnt = (size_t)n * slots_current(order);
        const size_t tmp = sort_tmp_bytes(cnt);
        DBuf<unsigned char> buf;
        buf.alloc(scratch_bytes(cnt, tmp));
        current_density(p, order, comp, J, dt, oldR, R, Q, n, carve(buf.p, cnt, tmp), cnt, (hipStream_t)stream);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return sf_fail(e);
    }
}

int csrst_sf_current_density_noalloc(int order, int comp, double *J, int i0, int j0, int nx, int ny,
                                     const double origin[2], const double dx[2], double dt, const double *oldR,
                                     const double *R, const double *Q, int n, void *stream)
{
    try {
        check_current_args(order, comp, dt, J, oldR, R, Q, n);
        const Patch p = make_patch(i0, j0, nx, ny, origin, dx);
        const size_t cnt = (size_t)n * slots_current(order);
        const size_t tmp = sort_tmp_bytes(cnt);      // a size query: nothing is allocated
        const size_t need = scratch_bytes(cnt, tmp);
        if (need > kScratchBytes)
            throw Error(CSRST_EINVAL, hu::fmt("%d particles of order %d need %zu bytes of scratch; the no-allocation "
                                              "entry holds %zu (use csrst_sf_current_density)",
                                              n, order, need, kScratchBytes));
        current_density(p, order, comp, J, dt, oldR, R, Q, n, carve(device_scratch(), cnt, tmp), cnt,
                        (hipStream_t)stream);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return sf_fail(e);
    }
}

static int charge(int order, double *rho, int i0, int j0, int nx, int ny, const double origin[2], const double dx[2],
                  const double *R, const double *oldR, const double *Q, int n, void *stream)
{
    try {
        if (order < 0 || order > 2) throw Error(CSRST_EINVAL, "order must be 0..2");
        if (n < 0 || !rho || (n > 0 && (!R || !Q))) throw Error(CSRST_EINVAL, "bad arguments");
        const Patch p = make_patch(i0, j0, nx, ny, origin, dx);
        const hipStream_t st = (hipStream_t)stream;
        const size_t nodes = (size_t)nx * ny;
        const int per = slots_charge(order) * (oldR ? 2 : 1);
        const size_t cnt = (size_t)n * per;
        const size_t tmp = n > 0 ? sort_tmp_bytes(cnt) : 0;
        DBuf<unsigned char> buf;
        buf.alloc(scratch_bytes(cnt, tmp));
        const Scratch s = carve(buf.p, cnt, tmp);
        k_zero_int<<<1, 1, 0, st>>>(s.err);
        k_fill<<<blocks(nodes), kB, 0, st>>>(rho, nodes, 0.0);
        if (n > 0) {
            const Slots sl{s.k0, s.v0, per, s.err};
            if (order == 0) k_emit_charge<0><<<blocks(n), kB, 0, st>>>(p, R, oldR, Q, n, sl);
            else if (order == 1) k_emit_charge<1><<<blocks(n), kB, 0, st>>>(p, R, oldR, Q, n, sl);
            else k_emit_charge<2><<<blocks(n), kB, 0, st>>>(p, R, oldR, Q, n, sl);
            HU_LAUNCH_CHECK();
            reduce(s, cnt, nodes, rho, dx[0] * dx[1], 0, st);
        }
        if (read_err(s, st) & 1) throw Error(CSRST_EWINDOW, "a charge stencil left the patch (dropped)");
        return CSRST_OK;
    } catch (const std::exception &e) {
        return sf_fail(e);
    }
}

int csrst_sf_charge_density(int order, double *rho, int i0, int j0, int nx, int ny, const double origin[2],
                            const double dx[2], const double *R, const double *Q, int n, void *stream)
{
    return charge(order, rho, i0, j0, nx, ny, origin, dx, R, nullptr, Q, n, stream);
}

int csrst_sf_charge_density_diff(int order, double *rho, int i0, int j0, int nx, int ny, const double origin[2],
                                 const double dx[2], const double *R, const double *oldR, const double *Q, int n,
                                 void *stream)
{
    if (n > 0 && !oldR) {
        hu::global_error() = "null oldR";
        return CSRST_EINVAL;
    }
    return charge(order, rho, i0, j0, nx, ny, origin, dx, R, oldR, Q, n, stream);
}

}  // extern "C"
