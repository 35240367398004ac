// csrst_mw.hip -- production PIC step on gfx950: one W x H window grid that
// follows the bunch and carries the conformal chamber walls.
//
// Layout in HBM
//   fields   Ex, Ey, Hz (Hz = Hzx + Hzy merged: the split only matters with
//            PML, which the reference never applies), Jx, Jy as real planes
//            of (W+2) x (H+2) cells (one guard layer), used as a 2-D ring
//            buffer: moving the window by one column/row only clears the
//            recycled guard column/row -- no field copy.
//   geometry one byte per cell (0 no update, 1 bulk, 2 cut cell) in the same
//            ring, plus per-cell cut data (lambda0, lambda1, area, lambdaN0,
//            lambdaN1) read only at cut cells.  New columns/rows get their
//            geometry from the analytic bend on the device (FP64).
//   particles int32 global cell index + real offset in [0, 1) per axis
//            (split precision), momentum (beta*gamma) as the deviation from a
//            reference momentum p0 held in FP64; uniform charge.  Arrays are
//            padded to 64 elements and cell-sorted every few steps.
//
// Kernels per step
//   k_push<comp=0>  drift, gather, kick+rotation, drift, S0/S1/S2 deposit of
//                   Jx (register slots per stencil anchor, wavefront sums
//                   into an LDS tile, one global atomic per node), drift back
//   k_heh           H half-step, Ex, H half-step in one row-marching pass
//   k_push<comp=1>  same, depositing Jy
//   k_ey            Ey
//   k_move          window move decision on the device, ring-offset update
//   k_fresh         clear + geometry for the recycled guard column/row
// plus a periodic cell sort and the wake sampling.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <memory>
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

namespace mw {

constexpr int kB = 256;
constexpr int G = 1;  // guard layer

// Device-resident window state (changes every step without host round trips).
constexpr int kBBSlots = 64;

struct WinState {
    int wx0, wy0;        // global cell index of logical cell (0, 0)
    int rx, ry;          // ring offsets
    int mi, mj;          // last move
    int bb[4];           // particle bounding box in global cells: min x, min y, -max x, -max y
    // per-slot boxes (atomicMin), one 128-byte line each: workgroups spread
    // their updates over kBBSlots addresses, k_move reduces and clears them
    int bbs[kBBSlots][32];
    int jbox[2][4];      // current footprint per component (global nodes): lo x, lo y, -hi x, -hi y
    int err;             // 1: stencil left the window, 2: too many crossings, 4: move out of range
    int steps;
    double p0[2];        // reference momentum: particles store P - p0 (re-centred at every sort)
    int sbox[4];         // cell box of the particles at the last sort: min x, min y, -max x, -max y
};

template <class T>
struct Fields {
    T *ex, *ey, *hz, *jx, *jy;
    T *ex2, *hz2;        // ping-pong partners of ex and hz (written by k_heh)
    uint8_t *cls;
    T *cut;              // 5 planes: lam0, lam1, area, lamN0, lamN1
    int W, H, Wp, Hp;    // logical interior size, physical (padded) size
    size_t plane;
    // convolutional PML along x at both window ends (npml columns each):
    // psi of the x-derivative in the Hz update (ping-pong with k_heh) and in
    // the Ey update; per logical column: decay b and weight a = b - 1 for
    // the Hz half step (at i + 1/2) and the Ey step (at i)
    int npml;
    T *psih, *psih2, *psie;
    const T *pml;        // [4][W]: bh, ah, be, ae
};

template <class T>
struct Parts {
    int *gx, *gy;
    T *fx, *fy, *px, *py;
    int n;
};

// Buffer resource over one field plane (bytes, below 4 GiB: the host refuses
// larger windows), for loads at 32-bit byte offsets.  A run of loads at
// constant distances from one offset takes immediate offsets and merges into
// wide loads; an offset past the plane reads 0.  Word 3 is the gfx9 family's
// raw 32-bit data format.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void *plane, size_t elems, size_t esize)
{
    const size_t bytes = elems * esize;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(plane), (short)0,
                                             (int)(uint32_t)(bytes < 0xffffffffu ? bytes : 0xffffffffu), 0x00020000);
}
template <class T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const T *plane, size_t elems)
{
    return plane_rsrc(plane, elems, sizeof(T));
}
template <class T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

struct Geo {
    geom::Bend bend;
    double org[2], h[2];  // global virtual mesh: vertex (gx, gy) at org + (gx*h0, gy*h1)
};

// logical -> physical ring index
__device__ __forceinline__ int ringx(int i, int rx, int Wp)
{
    int p = i + G + rx;
    p = p >= Wp ? p - Wp : p;
    return p < 0 ? p + Wp : p;
}
__device__ __forceinline__ int ringy(int j, int ry, int Hp)
{
    int p = j + G + ry;
    p = p >= Hp ? p - Hp : p;
    return p < 0 ? p + Hp : p;
}

// ---------------------------------------------------------------------------
// geometry of one logical cell (global index gx, gy)
template <class T>
__device__ void cell_geometry(const Geo &g, const Fields<T> &f, int gx, int gy, size_t q)
{
    geom::CutCell cut;
    bool ok;
    const long long big = (long long)1 << 40;
    const int k = geom::classify(g.bend, g.org, g.h, gx, gy, big, big, &cut, &ok);
    f.cls[q] = (uint8_t)(k == 1 ? 1 : (k == 2 ? 2 : 0));
    if (k == 2) {
        f.cut[q] = (T)cut.lam[0];
        f.cut[f.plane + q] = (T)cut.lam[1];
        f.cut[2 * f.plane + q] = (T)cut.area;
        f.cut[3 * f.plane + q] = (T)cut.lamN[0];
        f.cut[4 * f.plane + q] = (T)cut.lamN[1];
    }
}

// full-window geometry + clear (set-up)
template <class T>
__global__ void k_geom_all(Fields<T> f, Geo g, const WinState *ws)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x - G;
    const int j = blockIdx.y - G;
    if (i > f.W - 1 + G) return;
    const size_t q = (size_t)ringy(j, ws->ry, f.Hp) * f.Wp + ringx(i, ws->rx, f.Wp);
    f.ex[q] = 0; f.ey[q] = 0; f.hz[q] = 0; f.jx[q] = 0; f.jy[q] = 0; f.ex2[q] = 0; f.hz2[q] = 0;
    if (f.npml) { f.psih[q] = 0; f.psih2[q] = 0; f.psie[q] = 0; }
    cell_geometry(g, f, ws->wx0 + i, ws->wy0 + j, q);
}

// ---------------------------------------------------------------------------
// particle push + deposit
template <class T>
struct PushConst {
    T hx, hy;            // half drift in cells: 0.5 c dt / h
    T fact1, fact2;      // kick factors (PartBunch.cpp:364-365)
    T qodt_x, qodt_y;    // q / (h dt) per axis
    T odt_x, odt_y;      // 1 / (h dt) per axis
    T q;
};

// Wave-wide (64 lanes) integer reductions with DPP row shifts and row
// broadcasts (no LDS traffic): lane 63 ends with the result, which readlane
// returns to every lane.  A lane without a DPP source reads the identity.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_i(int identity, int v)
{
    return __builtin_amdgcn_update_dpp(identity, v, CTRL, ROWMASK, 0xf, false);
}
template <class Op>
__device__ __forceinline__ int wave_reduce_i(int v, int identity, Op op)
{
    v = op(v, dpp_i<0x111, 0xf>(identity, v));   // row_shr:1
    v = op(v, dpp_i<0x112, 0xf>(identity, v));   // row_shr:2
    v = op(v, dpp_i<0x114, 0xf>(identity, v));   // row_shr:4
    v = op(v, dpp_i<0x118, 0xf>(identity, v));   // row_shr:8
    v = op(v, dpp_i<0x142, 0xa>(identity, v));   // row_bcast:15
    v = op(v, dpp_i<0x143, 0xc>(identity, v));   // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_min_i(int v)
{
    return wave_reduce_i(v, INT_MAX, [](int a, int b) { return a < b ? a : b; });
}
__device__ __forceinline__ int wave_max_i(int v)
{
    return wave_reduce_i(v, INT_MIN, [](int a, int b) { return a > b ? a : b; });
}
__device__ __forceinline__ int wave_or(int v)
{
    return wave_reduce_i(v, 0, [](int a, int b) { return a | b; });
}

// Kick and rotation of PartBunch.cpp:366-383 on a momentum stored as
// P = p0 + dP: the increments are formed directly so that FP32 keeps the
// resolution of dP (|dP| << |P|).  Algebraically identical to
// P' = ((1+a^2) s + a s_perp)/(1+a^2) + kick, s = um + a (um_y, -um_x).
// S2 (cubic B-spline) weights of the four nodes for offset d in [0, 1) from
// the second node, in the arithmetic of ScatterGatherImpl.hh:436-446.
template <class T>
__device__ __forceinline__ void s2_weights(T d, T (&w)[4])
{
    const T d2 = d * d, tt = T(2.0 / 3.0), os = T(1.0 / 6.0);
    w[1] = tt - d2 * (1 - T(0.5) * d);
    w[2] = tt - T(0.5) * (1 - d) * (1 - d2);
    w[3] = os * d2 * d;
    w[0] = T(1.0) - (w[1] + w[2] + w[3]);
}

// Quotient of the split-precision form: FP32 uses the hardware reciprocal
// (1 ulp) instead of the IEEE division sequence; FP64 divides exactly.
__device__ __forceinline__ float quot(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double quot(double a, double b) { return a / b; }

template <class T>
__device__ __forceinline__ vec2<T> kick_rotate_split(vec2<T> p0, vec2<T> dP, T ex, T ey, T hf, T bext, T fact1, T fact2)
{
    const T B = T(pm::kMu0) * hf + T(0.5) * bext;
    const vec2<T> kick = pm::mk(fact1 * ex, fact1 * ey);
    const vec2<T> um = pm::mk(p0.x + (dP.x + kick.x), p0.y + (dP.y + kick.y));
    const T a = fact2 * B * (T)pm::q_rsqrt((float)(T(1.0) + (um.x * um.x + um.y * um.y)));
    const T a2 = a * a;
    const T r = quot(T(1.0), T(1.0) + a2);
    const T c1 = a * (T(1.0) + r);
    const T dx = c1 * um.y - a2 * r * um.x;
    const T dy = -c1 * um.x - a2 * r * um.y;
    return pm::mk(dP.x + ((kick.x + kick.x) + dx), dP.y + ((kick.y + kick.y) + dy));
}

// f stays in [0, 1): f - floor(f) rounds to 1 for a tiny negative f, which
// is renormalised to the next cell (the same position)
template <class T>
__device__ __forceinline__ void split_add(int &g, T &f, T d)
{
    f = f + d;
    const T fl = pm::floor_(f);
    g += (int)fl;
    f = f - fl;
    const bool one = f >= T(1);
    g += one ? 1 : 0;
    f = one ? f - T(1) : f;
}

// Wave-level (64 lanes) sum with DPP row shifts and row broadcasts: lane 63
// ends with the total, which readlane returns to every lane.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_f(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, ROWMASK, 0xf, false));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_d(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(u & 0xffffffffu), CTRL, ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, ROWMASK, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
template <class T>
__device__ __forceinline__ T dpp(T v);

__device__ __forceinline__ float wave_sum(float v)
{
    v += dpp_f<0x111, 0xf>(v);   // row_shr:1
    v += dpp_f<0x112, 0xf>(v);   // row_shr:2
    v += dpp_f<0x114, 0xf>(v);   // row_shr:4
    v += dpp_f<0x118, 0xf>(v);   // row_shr:8  -> lane 15 of each row holds the row sum
    v += dpp_f<0x142, 0xa>(v);   // row_bcast:15 into rows 1, 3
    v += dpp_f<0x143, 0xc>(v);   // row_bcast:31 into rows 2, 3 -> lane 63 holds the total
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_d<0x111, 0xf>(v);
    v += dpp_d<0x112, 0xf>(v);
    v += dpp_d<0x114, 0xf>(v);
    v += dpp_d<0x118, 0xf>(v);
    v += dpp_d<0x142, 0xa>(v);
    v += dpp_d<0x143, 0xc>(v);
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(u & 0xffffffffu), 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(u >> 32), 63);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// sum over the 4 lanes of each quad (DPP quad permutations [1,0,3,2], [2,3,0,1])
__device__ __forceinline__ float quad_sum(float v)
{
    v += dpp_f<0xb1, 0xf>(v);
    v += dpp_f<0x4e, 0xf>(v);
    return v;
}
__device__ __forceinline__ double quad_sum(double v)
{
    v += dpp_d<0xb1, 0xf>(v);
    v += dpp_d<0x4e, 0xf>(v);
    return v;
}

template <class T>
__device__ __forceinline__ void atomic_add_real(T *p, T v);
template <>
__device__ __forceinline__ void atomic_add_real<float>(float *p, float v) { unsafeAtomicAdd(p, v); }
template <>
__device__ __forceinline__ void atomic_add_real<double>(double *p, double v) { unsafeAtomicAdd(p, v); }

// Trajectory split for one push: with c dt <= hx <= hy a particle crosses at
// most one cell (or dual-cell) line per axis, so the list has 2..4 entries
// and the loops below have fixed trip counts (registers, no scratch).
template <class T>
struct Split {
    T f[4];
    vec2<T> s[4];
    int n;   // entries (segments = n - 1)
};

template <class T>
__device__ __forceinline__ T absv(T v) { return v < T(0) ? -v : v; }

// component-wise select: a select of whole structs can be lowered to a select
// of stack addresses, which puts the operands in scratch memory
template <class T>
__device__ __forceinline__ vec2<T> sel(bool c, vec2<T> a, vec2<T> b)
{
    return pm::mk(c ? a.x : b.x, c ? a.y : b.y);
}

template <class T>
__device__ __forceinline__ bool split_path(Split<T> &c, vec2<T> o, vec2<T> e, bool dual)
{
    const T sh = dual ? T(0.5) : T(0);
    const int Ii = (int)pm::floor_(o.x - sh), Ji = (int)pm::floor_(o.y - sh);
    const int If = (int)pm::floor_(e.x - sh), Jf = (int)pm::floor_(e.y - sh);
    const int dI = If - Ii, dJ = Jf - Ji;
    const vec2<T> D = pm::mk(e.x - o.x, e.y - o.y);
    const T eps = T(1e-6);
    // the (at most two) interior crossings
    T fa = T(2), fb = T(2);
    T sax = o.x, say = o.y, sbx = o.x, sby = o.y;
    if (dI != 0) {
        const T xb = (T)(dI > 0 ? If : Ii) + sh;
        fa = quot(xb - o.x, D.x);
        sax = xb;
        say = o.y + fa * D.y;
    }
    if (dJ != 0) {
        const T yb = (T)(dJ > 0 ? Jf : Ji) + sh;
        fb = quot(yb - o.y, D.y);
        sbx = o.x + fb * D.x;
        sby = yb;
    }
    const vec2<T> sa = pm::mk(sax, say), sb = pm::mk(sbx, sby);
    const bool swp = fb < fa;
    const T f1 = swp ? fb : fa, f2 = swp ? fa : fb;
    const vec2<T> s1 = sel(swp, sb, sa), s2 = sel(swp, sa, sb);
    // duplicates within 1e-6 in fraction merge into the earlier entry
    bool v1 = f1 < T(2), v2 = f2 < T(2);
    v1 = v1 && !(absv(f1) < eps);
    v2 = v2 && !(absv(f2 - (v1 ? f1 : T(0))) < eps);
    // the end point: merges into the last kept crossing if within 1e-6
    const T flast = v2 ? f2 : (v1 ? f1 : T(0));
    const bool endMerged = (v1 || v2) && absv(T(1) - flast) < eps;
    // fixed slots: 0 start, then kept crossings, then the end
    c.f[0] = T(0); c.s[0] = o;
    c.f[1] = v1 ? f1 : (v2 ? f2 : T(1));
    c.s[1] = sel(v1, s1, sel(v2, s2, e));
    c.f[2] = (v1 && v2) ? f2 : T(1);
    c.s[2] = sel(v1 && v2, s2, e);
    c.f[3] = T(1); c.s[3] = e;
    const int n = 2 + (v1 ? 1 : 0) + (v2 ? 1 : 0) - (endMerged ? 1 : 0);
    // a merged end: the last kept crossing takes the end position (its fraction stays)
    c.s[1] = sel(endMerged && n == 2, e, c.s[1]);
    c.s[2] = sel(endMerged && n == 3, e, c.s[2]);
    c.n = n;
    return (dI >= -1 && dI <= 1 && dJ >= -1 && dJ <= 1);
}

// Segment currents of the S1 and S2 deposits (FirstOrderShapeFunction.cpp:
// 556-587, SecondOrderShapeFunction.cpp:583-826, the same time integrals as
// pm::seg_s1 / pm::seg_s2) with the node weights regrouped as separable
// sums: per node w = X[i] S[j] + C[i] P[j] - B[i] Q[j], the per-axis factors
// formed once per segment.  Equal in exact arithmetic; fewer operations.
template <class T>
__device__ __forceinline__ void seg_s1_sep(pm::SegCurrent<T> &o, int comp, T tau, vec2<T> a, vec2<T> b, vec2<T> qodt,
                                           vec2<T> d)
{
    const T h = T(0.5), o3 = T(1.0 / 3.0), o4 = h * h;
    const T tp2 = tau * tau * h;
    const T tp3 = tp2 * tau * 2 * o3;
    // along the current: A + B (linear); across: quadratic Ay + By + Cy
    const T al = comp == 0 ? a.x : a.y, dl = comp == 0 ? d.x : d.y;
    const T bt = comp == 0 ? b.y : b.x, dt = comp == 0 ? d.y : d.x;
    const T cd = (comp == 0 ? qodt.y : qodt.x) * dl;
    const T A[2] = {al, T(1.0) - al};
    const T B[2] = {-dl * tau, dl * tau};
    const T Ay[3] = {h * (h + bt) * (h + bt) * tau, (T(0.75) - bt * bt) * tau, h * (h - bt) * (h - bt) * tau};
    const T By[3] = {-(h + bt) * dt * tp2, 2 * bt * dt * tp2, (h - bt) * dt * tp2};
    const T Cy[3] = {h * dt * dt * tp3, -dt * dt * tp3, h * dt * dt * tp3};
    T S[3], Q[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        S[m] = cd * (Ay[m] + By[m] + Cy[m]);
        Q[m] = cd * (h * Ay[m] + o3 * By[m] + o4 * Cy[m]);
    }
    if (comp == 0) {
        o.nx = 2; o.ny = 3;
#pragma unroll
        for (int jj = 0; jj < 3; ++jj)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) o.w[jj * 2 + ii] = (A[ii] + B[ii]) * S[jj] - B[ii] * Q[jj];
    } else {
        o.nx = 3; o.ny = 2;
#pragma unroll
        for (int jj = 0; jj < 2; ++jj)
#pragma unroll
            for (int ii = 0; ii < 3; ++ii) o.w[jj * 3 + ii] = (A[jj] + B[jj]) * S[ii] - B[jj] * Q[ii];
    }
}

// S2 segment current (SecondOrderShapeFunction.cpp:583-826, pm::seg_s2):
// the reference's expansion in A2..C2 and A3..D3 is the exact time integral
//   w[along i][across j] = cd * int_0^tau W2_i(al - dl s) W3_j(bt - dt s) ds
// of the quadratic B-spline weights along the current and the cubic ones
// across it, positions moving linearly over the segment.  The integrand has
// degree 5 in s, so 3-point Gauss-Legendre quadrature gives the same integral
// exactly: the centre tau/2 (weight 4/9) and the pair tau (1/2 -+ r), r =
// sqrt(15)/10 (5/18 each).  The weights are polynomials, so at the pair they
// are their values at the centre's offsets am -+ da, bm -+ db: along Ab -+ D,
// across S -+ U (Taylor expansions, exact), and the pair contributes
// 2 (5/18) (Ab (x) S + D (x) U).  One evaluation at the centre and its
// offsets replace three evaluations; the three outer products remain.
template <class T>
__device__ __forceinline__ void seg_s2_gauss(pm::SegCurrent<T> &o, int comp, T tau, vec2<T> a, vec2<T> b,
                                             vec2<T> qodt, vec2<T> d)
{
    const T al = comp == 0 ? a.x : a.y, dl = comp == 0 ? d.x : d.y;
    const T bt = comp == 0 ? b.y : b.x, dt = comp == 0 ? d.y : d.x;
    const T k = (comp == 0 ? qodt.y : qodt.x) * dl * tau;
    const T o6 = T(1.0 / 6.0);
    const T ht = T(0.5) * tau, rt = T(0.38729833462074168852) * tau;
    const T am = al - dl * ht, bm = bt - dt * ht;
    const T da = dl * rt, db = dt * rt;
    // along: W2 = (1/2 (1/2 + a)^2, 3/4 - a^2, 1/2 (1/2 - a)^2) at am (A) and
    // the pair's mean (Ab) and half difference (D), scaled by the quadrature
    // weights and k
    const T P = T(0.5) + am, M = T(0.5) - am;
    const T P2 = P * P, M2 = M * M, c0 = T(0.75) - am * am, d2 = da * da;
    const T kc = k * T(4.0 / 9.0), ks = k * T(5.0 / 9.0);
    const T hc = T(0.5) * kc, hs = T(0.5) * ks, dks = ks * da;
    const T A[3] = {hc * P2, kc * c0, hc * M2};
    const T Ab[3] = {hs * (P2 + d2), ks * (c0 - d2), hs * (M2 + d2)};
    const T D[3] = {dks * P, T(-2) * (dks * am), -(dks * M)};
    // across: W3 = ((1+b)^3/6, (4 - 6b^2 - 3b^3)/6, 1 - the others, -b^3/6) at
    // bm (B), the pair's mean (S: B + W3'' db^2 / 2) and half difference
    // (U: W3' db + W3''' db^3 / 6)
    const T c1 = T(1) + bm, c2 = c1 * c1, b2 = bm * bm;
    const T B0 = o6 * (c2 * c1), B3 = -o6 * (b2 * bm), B1 = o6 * (T(4) - T(3) * b2 * (T(2) + bm));
    const T B[4] = {B0, B1, T(1) - (B0 + B1 + B3), B3};
    const T h = T(0.5) * (db * db);
    const T S0 = B0 + h * c1, S3 = B3 - h * bm, S1 = B1 - h * (T(2) + T(3) * bm);
    const T S[4] = {S0, S1, T(1) - (S0 + S1 + S3), S3};
    const T d3 = o6 * (db * db * db);
    const T U0 = T(0.5) * c2 * db + d3, U3 = -(T(0.5) * b2 * db) - d3;
    const T U1 = -((T(2) * bm + T(1.5) * b2) * db) - T(3) * d3;
    const T U[4] = {U0, U1, -(U0 + U1 + U3), U3};
    T w[12];
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
#pragma unroll
        for (int ii = 0; ii < 3; ++ii) w[jj * 3 + ii] = A[ii] * B[jj] + Ab[ii] * S[jj] + D[ii] * U[jj];
    if (comp == 0) {
        o.nx = 3; o.ny = 4;
#pragma unroll
        for (int t = 0; t < 12; ++t) o.w[t] = w[t];
    } else {
        o.nx = 4; o.ny = 3;
#pragma unroll
        for (int jj = 0; jj < 3; ++jj)
#pragma unroll
            for (int ii = 0; ii < 4; ++ii) o.w[jj * 4 + ii] = w[ii * 3 + jj];
    }
}

constexpr int kRun = 4;              // consecutive particles per vector load (16 B of FP32, 32 B of FP64)
constexpr int kPPT = 4;              // particles per thread: one run of kRun consecutive sorted particles
constexpr int kChunk = kB * kPPT;    // 1024 particles per workgroup
constexpr int kTW = 64, kTH = 32;    // LDS accumulation tile (nodes)
constexpr int kSlots = 2;            // register accumulation slots per lane (one stencil anchor each)
constexpr int kMaxSub = 8;           // sort key: at most 2^8 x 2^8 bins per cell

// Particle arrays start at multiples of 64 elements (Parts stride), so a run
// of kRun values at a multiple of kRun is one aligned vector access.
template <class T>
struct alignas(16) Run {
    T v[kRun];
};
template <class T>
__device__ __forceinline__ void load_run(const T *a, int k0, int n, T (&o)[kRun])
{
    if (k0 + kRun <= n) {
        const Run<T> r = *reinterpret_cast<const Run<T> *>(a + k0);
#pragma unroll
        for (int u = 0; u < kRun; ++u) o[u] = r.v[u];
    } else {
#pragma unroll
        for (int u = 0; u < kRun; ++u) o[u] = k0 + u < n ? a[k0 + u] : T(0);
    }
}
template <class T>
__device__ __forceinline__ void store_run(T *a, int k0, int n, const T (&o)[kRun])
{
    if (k0 + kRun <= n) {
        Run<T> r;
#pragma unroll
        for (int u = 0; u < kRun; ++u) r.v[u] = o[u];
        *reinterpret_cast<Run<T> *>(a + k0) = r;
    } else {
#pragma unroll
        for (int u = 0; u < kRun; ++u)
            if (k0 + u < n) a[k0 + u] = o[u];
    }
}
// shift a run down by one register and append v (keeps every index static)
template <class T>
__device__ __forceinline__ void rotate_in(T (&o)[kRun], T v)
{
#pragma unroll
    for (int u = 0; u < kRun - 1; ++u) o[u] = o[u + 1];
    o[kRun - 1] = v;
}

// Push of one workgroup chunk (kChunk sorted particles): thread t owns
// particles base + kPPT t .. + kPPT - 1, so its consecutive particles share a
// cell and their segment currents share stencil anchors.  Currents collect in
// kSlots register slots per lane (one anchor each).  A segment whose anchor
// matches neither (a run crossing a cell boundary, or a bunch moving at an
// angle, whose runs have up to four anchors) goes into the chunk's LDS tile;
// at the end the slots are reduced across the wavefront once per distinct
// anchor into the tile.  The tile is flushed with one global atomic per
// touched node.  Loops stay rolled where unrolling would
// not pay: the kernel has to fit the instruction cache.
// At most three waves per SIMD: at four (the register count allows it for
// FP32 S2) the push runs 10 % slower, at two 27 % slower.
// MISSCOOP: slot misses through the wavefront instead of per-lane LDS
// atomics.  Which form is faster depends on the bunch's angle and density
// (c3 at 1-10 deg: the cooperative form 13-25 % faster; c2 at 1-8 deg: 2-8 %
// slower; along x the cooperative path in the loop costs 20 %), so the step
// times the two against each other and keeps the faster (csrst_mw_step).
template <class T, int ORDER, int COMP, bool MISSCOOP>
__global__ void __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(1, 3))) k_push(Parts<T> p, Fields<T> f, Geo g, PushConst<T> c, WinState *ws)
{
    __shared__ T acc[kTH * kTW];
    __shared__ int pbox[4];          // particle cell box of the chunk (min x, min y, -max x, -max y)
    const int tid = threadIdx.x, lane = tid & 63;
    const int base = blockIdx.x * kChunk;
    const int kt = base + tid * kPPT;           // this thread's first particle
    const int wx0 = ws->wx0, wy0 = ws->wy0, rx = ws->rx, ry = ws->ry;
    const vec2<T> p0 = pm::mk((T)ws->p0[0], (T)ws->p0[1]);
    const int W = f.W, H = f.H;
    T *J = COMP == 0 ? f.jx : f.jy;
    // ---- tile anchor from the chunk's first particle: in the tile-major sort
    // order the chunk continues right and down from it, so the tile covers
    // 20 nodes left and 6 up of that cell (stencil reach included); nodes of a
    // sparse chunk that fall outside go to global atomics
    const int ax = p.gx[base] - 20, ay = p.gy[base] - 6;   // LDS node (0,0) = global node (ax, ay)
    if (tid == 0) {
        pbox[0] = INT_MAX; pbox[1] = INT_MAX; pbox[2] = INT_MAX; pbox[3] = INT_MAX;
    }
    for (int q = tid; q < kTH * kTW; q += kB) acc[q] = T(0);
    __syncthreads();
    int err = 0;
    int bminx = INT_MAX, bminy = INT_MAX, bmaxx = INT_MIN, bmaxy = INT_MIN;
    constexpr int NX = ORDER == 0 ? (COMP == 0 ? 1 : 2) : (ORDER == 1 ? (COMP == 0 ? 2 : 3) : (COMP == 0 ? 3 : 4));
    constexpr int NY = ORDER == 0 ? (COMP == 0 ? 2 : 1) : (ORDER == 1 ? (COMP == 0 ? 3 : 2) : (COMP == 0 ? 4 : 3));
    const vec2<T> odt0 = pm::mk(c.odt_y, c.odt_x);   // S0: (1/(hy dt), 1/(hx dt)), ZerothOrderShapeFunction.cpp:224
    const vec2<T> qodt = pm::mk(c.qodt_x, c.qodt_y);
    constexpr int NN = NX * NY;
    // one node value from one lane: LDS tile, else a global atomic
    auto node_add = [&](int gxn, int gyn, T v) {
        const int li = gxn - ax, lj = gyn - ay;
        if (li >= 0 && li < kTW && lj >= 0 && lj < kTH) {
            atomicAdd(&acc[lj * kTW + li], v);
        } else {
            const int i = gxn - wx0, j = gyn - wy0;
            if (i < 0 || i > W - 1 || j < 0 || j > H - 1) err |= 1;
            else atomic_add_real(&J[(size_t)ringy(j, ry, f.Hp) * f.Wp + ringx(i, rx, f.Wp)], v);
        }
    };
    // per-lane add of one stencil (LDS tile, global atomics off the tile)
    auto lane_add = [&](int I0, int J0, const T *w) {
#pragma unroll
        for (int t = 0; t < NN; ++t) node_add(I0 + (t % NX), J0 + (t / NX), w[t]);
    };
    // Wavefront-cooperative add of the lanes' stencils (lanes with `mine`
    // set), one pass per distinct anchor.  The NN x 64 values go through the
    // wave's LDS scratch (node-major, conflict-free writes); lane l < 4 NN
    // sums 16 of node l / 4's values and a quad sum completes it, then lanes
    // 4t add node t into the tile in one step.  Called in wave-uniform control
    // flow (every lane active).
    __shared__ T red[(kB / 64) * NN * 64];
    T *rw = red + (tid >> 6) * (NN * 64);
    auto coop_add = [&](bool mine0, int sI, int sJ, const T *w) {
        unsigned long long rem = __ballot(mine0);
        while (rem) {
            const int lead = __ffsll((long long)rem) - 1;
            const int lI = __builtin_amdgcn_readlane(sI, lead), lJ = __builtin_amdgcn_readlane(sJ, lead);
            const bool mine = mine0 && sI == lI && sJ == lJ;
            rem &= ~__ballot(mine);
#pragma unroll
            for (int t = 0; t < NN; ++t) rw[t * 64 + lane] = mine ? w[t] : T(0);
            __builtin_amdgcn_wave_barrier();
            T v = T(0);
            if (lane < 4 * NN) {
                const T *src = rw + (lane >> 2) * 64 + (lane & 3) * 16;
#pragma unroll
                for (int k = 0; k < 16; ++k) v += src[k];
            }
            v = quad_sum(v);
            __builtin_amdgcn_wave_barrier();
            if (lane < 4 * NN && (lane & 3) == 0) {
                const int t = lane >> 2;
                node_add(lI + (t % NX), lJ + (t / NX), v);
            }
        }
    };
    // current of one segment xi -> xf (cell units relative to the corner of
    // cell (cgx, cgy)), tau its fraction of the push, dlt the whole push's
    // displacement
    auto seg_current = [&](pm::SegCurrent<T> &sc, vec2<T> xi, vec2<T> xf, T tau, vec2<T> dlt, int cgx, int cgy) {
        if (ORDER == 0) {
            pm::seg_s0(sc, COMP, xi, xf, c.q, odt0);
        } else {
            const vec2<T> cc = pm::mk(T(0.5) * (xi.x + xf.x), T(0.5) * (xi.y + xf.y));
            const int I = ORDER == 1 ? (int)pm::floor_(cc.x + T(0.5)) : (int)pm::floor_(cc.x);
            const int Jn = ORDER == 1 ? (int)pm::floor_(cc.y + T(0.5)) : (int)pm::floor_(cc.y);
            const vec2<T> a = pm::mk((T)I + T(0.5) - xi.x, (T)Jn + T(0.5) - xi.y);
            const vec2<T> b = pm::mk((T)I - xi.x, (T)Jn - xi.y);
            if (ORDER == 1) seg_s1_sep(sc, COMP, tau, a, b, qodt, dlt);
            else seg_s2_gauss(sc, COMP, tau, a, b, qodt, dlt);
            sc.I0 = I - 1; sc.J0 = Jn - 1;
        }
        sc.I0 += cgx; sc.J0 += cgy;   // global node of the first stencil entry
    };
    // per-wave queue of third segments: xi, xf, tau, displacement; old cell
    __shared__ T qvAll[kB / 64][7][64];
    __shared__ int qgAll[kB / 64][2][64];
    T (*qv)[64] = qvAll[tid >> 6];
    int (*qg)[64] = qgAll[tid >> 6];
    int qn = 0;   // wave-uniform queue length
    auto drain_queue = [&]() {
        __builtin_amdgcn_wave_barrier();
        const bool mine = lane < qn;
        pm::SegCurrent<T> sc;
        const int k = mine ? lane : 0;
        seg_current(sc, pm::mk(qv[0][k], qv[1][k]), pm::mk(qv[2][k], qv[3][k]), qv[4][k], pm::mk(qv[5][k], qv[6][k]),
                    qg[0][k], qg[1][k]);
        __builtin_amdgcn_wave_barrier();
        coop_add(mine, sc.I0, sc.J0, sc.w);
        qn = 0;
    };
    // kSlots register accumulation slots per lane
    bool has[kSlots];
    int sI[kSlots], sJ[kSlots];
    T sacc[kSlots][NN];
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
        has[q] = false; sI[q] = 0; sJ[q] = 0;
#pragma unroll
        for (int t = 0; t < NN; ++t) sacc[q][t] = T(0);
    }
#pragma unroll 1
    for (int r = 0; r < kPPT / kRun; ++r) {
        const int k0 = kt + r * kRun;
        int GX[kRun], GY[kRun];
        T FX[kRun], FY[kRun], PX[kRun], PY[kRun];
        load_run(p.gx, k0, p.n, GX);
        load_run(p.gy, k0, p.n, GY);
        load_run(p.fx, k0, p.n, FX);
        load_run(p.fy, k0, p.n, FY);
        load_run(p.px, k0, p.n, PX);
        load_run(p.py, k0, p.n, PY);
#pragma unroll 1
        for (int u = 0; u < kRun; ++u) {
            const bool live = k0 + u < p.n;
            const int gx0 = GX[0], gy0 = GY[0];
            const T fx0 = FX[0], fy0 = FY[0];
            vec2<T> dP = pm::mk(PX[0], PY[0]), P = dP;
            int gx = gx0, gy = gy0;
            T fx = fx0, fy = fy0;
            T g2 = 0;
            T x1 = fx0, y1 = fy0, x2 = fx0, y2 = fy0;
            if (live) {
                P = pm::mk(p0.x + dP.x, p0.y + dP.y);
                // half drift (PartBunch.cpp:228-232)
                const T gi = pm::inv_gamma(P);
                x1 = fx0 + c.hx * P.x * gi;
                y1 = fy0 + c.hy * P.y * gi;
                {
                    const T flx = pm::floor_(x1), fly = pm::floor_(y1);
                    gx = gx0 + (int)flx; gy = gy0 + (int)fly;
                    fx = x1 - flx; fy = y1 - fly;
                }
                // gather at R1 (positions relative to the particle's cell corner).
                // With fx, fy in [0, 1) every stencil of the three fields starts at
                // node offset L or L + 1 from the cell and spans N nodes, so the
                // ring-wrapped byte offsets of the U = N + 1 union columns and rows
                // are formed once and each load is plane + row + column (32-bit
                // offsets from the uniform plane base).
                const vec2<T> pp = pm::mk(fx, fy);
                constexpr int N = ORDER + 2, L = ORDER == 2 ? -2 : -1, U = N + 1;
                // The union's first column and row, clamped so that all U
                // stay inside the guarded window (a particle outside it has
                // set the error bit; its loads stay in bounds), then the ring
                // offsets: rows one by one (each may wrap), columns as one
                // base when the U columns do not cross the ring's seam, so a
                // row's loads take immediate offsets from one address.
                uint32_t R[U], C0;
                bool cwrap;
                {
                    const int i0 = gx + L - wx0, j0 = gy + L - wy0;
                    if (i0 < -G || i0 + U - 1 > W - 1 + G || j0 < -G || j0 + U - 1 > H - 1 + G) err |= 1;
                    int px = min(max(i0, -G), W - U + G) + G + rx;
                    px = px >= f.Wp ? px - f.Wp : px;
                    cwrap = px > f.Wp - U;
                    C0 = (uint32_t)px * (uint32_t)sizeof(T);
                    int py = min(max(j0, -G), H - U + G) + G + ry;
                    py = py >= f.Hp ? py - f.Hp : py;
                    const uint32_t rowb = (uint32_t)(f.Wp * sizeof(T));
                    R[0] = (uint32_t)py * rowb;
#pragma unroll
                    for (int a = 1; a < U; ++a) {
                        const bool wrap = ++py == f.Hp;
                        py = wrap ? 0 : py;
                        R[a] = wrap ? 0u : R[a - 1] + rowb;
                    }
                }
                const bool anyWrap = __any(cwrap);
                auto gat = [&](const T *plane, const pm::Stencil<T> &st) -> T {
                    const bool sx = st.I0 != L, sy = st.J0 != L;
                    uint32_t rr[N];
#pragma unroll
                    for (int a = 0; a < N; ++a) rr[a] = sy ? R[a + 1] : R[a];
                    // The empty asm makes the plane pointer opaque here, so the
                    // resource (four scalar registers) is built at each gather
                    // instead of three being hoisted out of the particle loop,
                    // where they push other scalars into spills (each reload
                    // is a VALU readlane).
                    const T *pl = plane;
                    asm volatile("" : "+s"(pl));
                    const auto rs = plane_rsrc(pl, f.plane);
                    if (!anyWrap) {
                        // one offset per row; its N columns are consecutive
                        // (immediate offsets, merged into wide loads)
                        const uint32_t cb = C0 + (sx ? (uint32_t)sizeof(T) : 0u);
                        return pm::gather_n<N>(st, false, [&](int a, int b) {
                            return buf_load<T>(rs, rr[b] + cb + (uint32_t)(a * sizeof(T)));
                        });
                    }
                    // a stencil of this wavefront crosses the seam: per-column offsets
                    uint32_t cc[N];
                    {
                        int px = (int)(C0 / (uint32_t)sizeof(T)) + (sx ? 1 : 0);
#pragma unroll
                        for (int a = 0; a < N; ++a) {
                            px = px >= f.Wp ? px - f.Wp : px;
                            cc[a] = (uint32_t)px * (uint32_t)sizeof(T);
                            ++px;
                        }
                    }
                    return pm::gather_n<N>(st, false, [&](int a, int b) { return buf_load<T>(rs, rr[b] + cc[a]); });
                };
                T ex, ey, hf;
                if constexpr (ORDER == 2) {
                    // The three S2 stencils use two weight sets per axis: at the
                    // offset from the cell corner (Ex's y, Ey's x) and from the
                    // nearest cell centre (Ex's x, Ey's y, both of Hz's)
                    // (ScatterGatherImpl.hh:424-447, 486-505 with fx, fy in [0, 1)).
                    pm::Stencil<T> sx, sy;
                    const int lx = (int)pm::floor_(pp.x - T(0.5)), ly = (int)pm::floor_(pp.y - T(0.5));
                    T wxu[4], wyu[4];
                    s2_weights(pp.x - ((T)lx + T(0.5)), sx.wx);
                    s2_weights(pp.y - ((T)ly + T(0.5)), sy.wy);
                    s2_weights(pp.x, wxu);
                    s2_weights(pp.y, wyu);
                    pm::Stencil<T> st;
                    st.n = 4;
                    st.I0 = lx - 1; st.J0 = -1;
#pragma unroll
                    for (int a = 0; a < 4; ++a) { st.wx[a] = sx.wx[a]; st.wy[a] = wyu[a]; }
                    ex = gat(f.ex, st);
                    st.I0 = -1; st.J0 = ly - 1;
#pragma unroll
                    for (int a = 0; a < 4; ++a) { st.wx[a] = wxu[a]; st.wy[a] = sy.wy[a]; }
                    ey = gat(f.ey, st);
                    st.I0 = lx - 1;
#pragma unroll
                    for (int a = 0; a < 4; ++a) st.wx[a] = sx.wx[a];
                    hf = gat(f.hz, st);
                } else {
                    ex = gat(f.ex, pm::edge_stencil(ORDER, 0, pp));
                    ey = gat(f.ey, pm::edge_stencil(ORDER, 1, pp));
                    hf = gat(f.hz, pm::cell_stencil(ORDER, pp));
                }
                const double X = g.org[0] + ((double)gx + (double)fx) * g.h[0];
                const double Y = g.org[1] + ((double)gy + (double)fy) * g.h[1];
                const T bext = (T)geom::external_bz(g.bend, X, Y);
                dP = kick_rotate_split(p0, dP, ex, ey, hf, bext, c.fact1, c.fact2);
                P = pm::mk(p0.x + dP.x, p0.y + dP.y);
                g2 = pm::inv_gamma(P);
                x2 = x1 + c.hx * P.x * g2;
                y2 = y1 + c.hy * P.y * g2;
            }
            // ---- deposit of R0 -> R2, cell units relative to the old cell corner
            const vec2<T> o = pm::mk(fx0, fy0);
            const vec2<T> e = pm::mk(x2, y2);
            Split<T> xs;
            const bool oneCross = split_path(xs, o, e, ORDER == 1);   // dead lanes: o = e = 0
            if (live && !oneCross) err |= 2;
            const vec2<T> delta = pm::mk(e.x - o.x, e.y - o.y);
            const int nseg = live ? xs.n - 1 : 0;
            // third segments (a push crossing a cell column and a cell row) are
            // queued and computed 64 at a time: in-line, one lane with three
            // segments makes its whole wavefront run a third segment pass
            const bool has3 = nseg == 3;
            const unsigned long long m3 = __ballot(has3);
            if (m3) {
                const int cnt = __popcll(m3);
                if (qn + cnt > 64) drain_queue();
                if (has3) {
                    const int at = qn + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m3 >> 32),
                                                                       __builtin_amdgcn_mbcnt_lo((unsigned)m3, 0u));
                    qv[0][at] = xs.s[2].x; qv[1][at] = xs.s[2].y;
                    qv[2][at] = xs.s[3].x; qv[3][at] = xs.s[3].y;
                    qv[4][at] = xs.f[3] - xs.f[2];
                    qv[5][at] = delta.x; qv[6][at] = delta.y;
                    qg[0][at] = gx0; qg[1][at] = gy0;
                }
                qn += cnt;
            }
#pragma unroll 1
            for (int s = 0; s < 2; ++s) {
                const bool act = s < nseg;
                if (!__any(act)) break;
                pm::SegCurrent<T> sc;
                {
                    // slot selects instead of xs.s[s] keep the list in registers
                    const vec2<T> xi = sel(s == 0, xs.s[0], xs.s[1]);
                    const vec2<T> xf = sel(s == 0, xs.s[1], xs.s[2]);
                    const T tau = s == 0 ? xs.f[1] - xs.f[0] : xs.f[2] - xs.f[1];
                    seg_current(sc, xi, xf, tau, delta, gx0, gy0);
                }
                // register accumulation by anchor: the first kSlots anchors a
                // lane meets own its slots, in order (their sums start at
                // zero); a segment matching none goes to the LDS tile
                // (MISSCOOP: through the wavefront).  Independent conditional
                // adds: a wavefront whose lanes take different slots pays one
                // pass per slot taken.
                bool taken = !act;
#pragma unroll
                for (int q = 0; q < kSlots; ++q) {
                    const bool to = !taken && (!has[q] || (sc.I0 == sI[q] && sc.J0 == sJ[q]));
                    sI[q] = to ? sc.I0 : sI[q];
                    sJ[q] = to ? sc.J0 : sJ[q];
                    has[q] = has[q] || to;
                    if (to) {
#pragma unroll
                        for (int t = 0; t < NN; ++t) sacc[q][t] += sc.w[t];
                    }
                    taken = taken || to;
                }
                const bool miss = !taken;
                if (MISSCOOP) {
                    if (__any(miss)) coop_add(miss, sc.I0, sc.J0, sc.w);
                } else if (miss) {
                    lane_add(sc.I0, sc.J0, sc.w);
                }
            }
            // ---- drift back by half a step (TETMSolver.hh:210, 265)
            if (live) {
                const int ex2 = gx0 + (int)pm::floor_(x2), ey2 = gy0 + (int)pm::floor_(y2);
                bminx = min(bminx, min(gx0, ex2)); bminy = min(bminy, min(gy0, ey2));
                bmaxx = max(bmaxx, max(gx0, ex2)); bmaxy = max(bmaxy, max(gy0, ey2));
                gx = gx0; gy = gy0; fx = T(0); fy = T(0);
                split_add(gx, fx, x2 - c.hx * P.x * g2);
                split_add(gy, fy, y2 - c.hy * P.y * g2);
            }
            rotate_in(GX, gx); rotate_in(GY, gy);
            rotate_in(FX, fx); rotate_in(FY, fy);
            rotate_in(PX, dP.x); rotate_in(PY, dP.y);
        }
        store_run(p.gx, k0, p.n, GX);
        store_run(p.gy, k0, p.n, GY);
        store_run(p.fx, k0, p.n, FX);
        store_run(p.fy, k0, p.n, FY);
        store_run(p.px, k0, p.n, PX);
        store_run(p.py, k0, p.n, PY);
    }
    // ---- queued third segments, then the register slots
    if (qn) drain_queue();
#pragma unroll 1
    for (int slot = 0; slot < kSlots; ++slot) {
        T w[NN];
        bool h = has[0];
        int I = sI[0], Jn = sJ[0];
#pragma unroll
        for (int t = 0; t < NN; ++t) w[t] = sacc[0][t];
#pragma unroll
        for (int q = 1; q < kSlots; ++q) {
            h = slot == q ? has[q] : h;
            I = slot == q ? sI[q] : I;
            Jn = slot == q ? sJ[q] : Jn;
#pragma unroll
            for (int t = 0; t < NN; ++t) w[t] = slot == q ? sacc[q][t] : w[t];
        }
        coop_add(h, I, Jn, w);
    }
    // ---- chunk's particle box and error bits, then flush the tile with one
    // global atomic per nonzero node
    bminx = wave_min_i(bminx); bminy = wave_min_i(bminy);
    bmaxx = wave_max_i(bmaxx); bmaxy = wave_max_i(bmaxy);
    err = wave_or(err);
    if (lane == 0) {
        if (bminx != INT_MAX) {
            atomicMin(&pbox[0], bminx); atomicMin(&pbox[1], bminy);
            atomicMin(&pbox[2], -bmaxx); atomicMin(&pbox[3], -bmaxy);
        }
        if (err) atomicOr(&ws->err, err);
    }
    __syncthreads();
    // one update of the window's particle box per workgroup, into one of
    // kBBSlots slots: same-address atomics serialise
    if (tid == 0 && pbox[0] != INT_MAX) {
        int *b = ws->bbs[blockIdx.x % kBBSlots];
        atomicMin(&b[0], pbox[0]); atomicMin(&b[1], pbox[1]);
        atomicMin(&b[2], pbox[2]); atomicMin(&b[3], pbox[3]);
    }
    for (int q = tid; q < kTH * kTW; q += kB) {
        const T v = acc[q];
        if (v == T(0)) continue;
        const int i = ax + (q % kTW) - wx0, j = ay + (q / kTW) - wy0;
        if (i < 0 || i > W - 1 || j < 0 || j > H - 1) { atomicOr(&ws->err, 1); continue; }
        atomic_add_real(&J[(size_t)ringy(j, ry, f.Hp) * f.Wp + ringx(i, rx, f.Wp)], v);
    }
}

// ---------------------------------------------------------------------------
// Field sweeps.  Thread = column, marching over a strip of rows.
//
// k_heh: H half step, Ex, H half step (TETMSolver.hh:142-206 without the
// exchanges, with the conformal cell updates of TETMSolver.cpp:87-115 in
// place of the bulk update at cut cells).  For row j it computes H1(j)
// (needs E(j+1) old), Ex(j) (needs H1(j), H1(j-1)), then H2(j-1) (needs
// Ex(j), Ex(j-1) new).  Ey is read-only in this kernel.
template <class T>
struct SweepConst {
    T sx, sy;            // 0.5 etax / Z0, 0.5 etay / Z0 (bulk H)
    T sex, dteps;        // etay Z0 (bulk Ex), c dt Z0
    T sey;               // etax Z0 (bulk Ey)
    T cx, cy;            // 0.5 dt / (h mu0) (conformal H)
    T ce;                // dt / eps0 (conformal E)
    T ihx, ihy;          // 1/hx, 1/hy
};

// Rows per workgroup of the sweeps.  Short strips give many workgroups to
// hide HBM latency, which outweighs k_heh's re-read of the rows above and
// below its strip: at c3 the two sweeps took 1.01 ms per step with 32-row
// strips, 0.91 ms with 8 and 0.89 ms with 4 (c2: 0.198, 0.160 ms with 4).
// k_ey loads its rows in one group of four.
constexpr int kStrip = 4;

template <class T>
__device__ __forceinline__ T h_update(const Fields<T> &f, size_t q, uint8_t cls, T h, T ey, T eyp, T ex, T exn,
                                      const SweepConst<T> &c)
{
    if (cls == 1) return (h - c.sx * (eyp - ey)) + c.sy * (exn - ex);
    if (cls == 2) {
        const T l0 = f.cut[q], l1 = f.cut[f.plane + q], ar = f.cut[2 * f.plane + q];
        const T n0 = f.cut[3 * f.plane + q], n1 = f.cut[4 * f.plane + q];
        const T u0 = -c.cx * (n1 * eyp - l1 * ey) / ar;
        const T u1 = +c.cy * (n0 * exn - l0 * ex) / ar;
        return h + (u0 + u1);
    }
    return h;
}

template <class T>
__device__ __forceinline__ T ex_update(const Fields<T> &f, size_t q, uint8_t cls, T ex, T h, T hm, T jv,
                                       const SweepConst<T> &c)
{
    if (cls == 1) return ex + (c.sex * (h - hm) - c.dteps * jv);
    if (cls == 2 && f.cut[q] > T(pm::kSpaceEps) && f.cut[2 * f.plane + q] >= T(pm::kMinArea))
        return ex + c.ce * ((h - hm) * c.ihy - jv);
    return ex;
}

// Row march for one column i over the strip [j0, j1]: H1 for rows
// j0-1 .. j1+1, Ex for rows j0 .. j1+1, H2 for rows j0 .. j1; only H2 and Ex
// of the strip rows are stored, into the partner buffers (ex2, hz2), so the
// halo rows other strips read are never overwritten during the pass.  Guard
// rows (-1, H) are never updated.
// Columns: PML = false sweeps the ncol columns from cbase; PML = true the two
// end layers (t < npml: column t, else column W - 2 npml + t).
template <class T, bool PML, int STRIP>
__global__ void __launch_bounds__(kB) k_heh(Fields<T> f, SweepConst<T> c, const WinState *ws, int jbase, int cbase,
                                            int ncol)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncol) return;
    const int i = PML ? (t < f.npml ? t : f.W - 2 * f.npml + t) : cbase + t;   // logical column
    const int rx = ws->rx, ry = ws->ry;
    const int j0 = jbase + blockIdx.y * STRIP;
    if (j0 > f.H - 1) return;
    const int j1 = min(j0 + STRIP - 1, f.H - 1);
    const int px = ringx(i, rx, f.Wp), pxp = ringx(i + 1, rx, f.Wp);
    const int jlo0 = ws->jbox[0][0] - ws->wx0, jlo1 = ws->jbox[0][1] - ws->wy0;
    const int jhi0 = -ws->jbox[0][2] - ws->wx0, jhi1 = -ws->jbox[0][3] - ws->wy0;
    const bool colInJ = i >= jlo0 && i <= jhi0;
    auto row = [&](int j) { return (size_t)ringy(j, ry, f.Hp) * f.Wp; };
    // PML column: Hz's x term (eyp - ey) becomes (eyp - ey) + psi with
    // psi <- b psi + a (eyp - ey)
    const bool inPml = PML && (i < f.npml || i >= f.W - f.npml);
    const T bh = inPml ? f.pml[i] : T(0), ah = inPml ? f.pml[f.W + i] : T(0);
    // x difference of the Hz update: Ey(i+1) - Ey(i) in bulk cells, the
    // edge-length weighted n1 Ey(i+1) - l1 Ey(i) in cut cells
    // (TETMSolver.cpp:87-104); psi_in -> psi_out is its PML convolution.
    // The x difference is the same in both half steps (Ey is not updated in
    // between), so psi advances once per step (update = true in the first
    // half step) and both halves use it: the H change over the step is then
    // that of the one-step convolutional PML.
    auto hup = [&](size_t q, uint8_t cl, T h, T ey, T eyp, T ex, T exn, T psi_in, T &psi_out, bool update) -> T {
        if (!(PML && inPml)) return h_update(f, q, cl, h, ey, eyp, ex, exn, c);
        if (cl == 1) {
            const T X = eyp - ey;
            psi_out = update ? bh * psi_in + ah * X : psi_in;
            return (h - c.sx * (X + psi_out)) + c.sy * (exn - ex);
        }
        if (cl == 2) {
            const T l0 = f.cut[q], l1 = f.cut[f.plane + q], ar = f.cut[2 * f.plane + q];
            const T n0 = f.cut[3 * f.plane + q], n1 = f.cut[4 * f.plane + q];
            const T X = n1 * eyp - l1 * ey;
            psi_out = update ? bh * psi_in + ah * X : psi_in;
            const T u0 = -c.cx * (X + psi_out) / ar;
            const T u1 = +c.cy * (n0 * exn - l0 * ex) / ar;
            return h + (u0 + u1);
        }
        return h;
    };
    // prologue: H1(j0-1)
    T h1p;
    {
        const int j = j0 - 1;
        const size_t rb = row(j);
        const T hold = f.hz[rb + px];
        if (j >= 0) {
            T psi1 = 0;
            h1p = hup(rb + px, f.cls[rb + px], hold, f.ey[rb + px], f.ey[rb + pxp], f.ex[rb + px],
                      f.ex[row(j0) + px], inPml ? f.psih[rb + px] : T(0), psi1, true);
        } else {
            h1p = hold;
        }
    }
    T psi1p = 0;
    T exo = f.ex[row(j0) + px];    // Ex old at the current row
    T exnp = 0, eyp_prev = 0, ey_prev = 0, h1pp = 0;
    uint8_t clp = 0;
    size_t qprev = 0;
    // the march's loads run one row ahead of its arithmetic
    struct RowIn {
        T ey, eyp, hz, jv, psi, exnext;
        uint8_t cl;
    };
    auto fetch = [&](int j, RowIn &r) {
        const size_t rb = row(j), q = rb + px;
        r.ey = f.ey[q];
        r.eyp = f.ey[rb + pxp];
        r.hz = f.hz[q];
        r.cl = f.cls[q];
        r.exnext = f.ex[row(j + 1) + px];
        r.jv = (colInJ && j >= jlo1 && j <= jhi1 && j <= f.H - 1) ? f.jx[q] : T(0);
        r.psi = inPml ? f.psih[q] : T(0);
    };
    RowIn cur, nxt;
    fetch(j0, cur);
    for (int j = j0; j <= j1 + 1; ++j) {
        if (j + 1 <= j1 + 1) fetch(j + 1, nxt);
        const size_t q = row(j) + px;
        const T exo_next = cur.exnext;
        const T ey = cur.ey, eyp = cur.eyp;
        const uint8_t cl = cur.cl;
        T h1, exn, psi1 = 0;
        if (j <= f.H - 1) {
            h1 = hup(q, cl, cur.hz, ey, eyp, exo, exo_next, cur.psi, psi1, true);
            exn = ex_update(f, q, cl, exo, h1, h1p, cur.jv, c);
        } else {
            h1 = cur.hz;
            exn = exo;
        }
        if (j > j0) {
            // H2(j-1) from H1(j-1), Ex new at j-1 and j
            T psi2 = 0;
            f.hz2[qprev] = hup(qprev, clp, h1pp, ey_prev, eyp_prev, exnp, exn, psi1p, psi2, false);
            if (inPml) f.psih2[qprev] = psi2;
            f.ex2[qprev] = exnp;
        }
        psi1p = psi1;
        h1pp = h1;
        h1p = h1;
        exnp = exn;
        exo = exo_next;
        ey_prev = ey; eyp_prev = eyp; clp = cl; qprev = q;
        cur = nxt;
    }
}

// Ey (TETMSolver.hh:241-249, conformal TETMSolver.cpp:117-127); clears Jy in
// its footprint after use and Jx in the Jx footprint, leaving both current
// planes zero for the next deposits.
template <class T, bool PML, int STRIP>
__global__ void __launch_bounds__(kB) k_ey(Fields<T> f, SweepConst<T> c, const WinState *ws, int cbase, int ncol)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncol) return;
    const int i = PML ? (t < f.npml ? t : f.W - 2 * f.npml + t) : cbase + t;
    const int rx = ws->rx, ry = ws->ry;
    const int px = ringx(i, rx, f.Wp), pxm = ringx(i - 1, rx, f.Wp);
    const int jlo0 = ws->jbox[1][0] - ws->wx0, jlo1 = ws->jbox[1][1] - ws->wy0;
    const int jhi0 = -ws->jbox[1][2] - ws->wx0, jhi1 = -ws->jbox[1][3] - ws->wy0;
    const bool colInJ = i >= jlo0 && i <= jhi0;
    const int xlo0 = ws->jbox[0][0] - ws->wx0, xlo1 = ws->jbox[0][1] - ws->wy0;
    const int xhi0 = -ws->jbox[0][2] - ws->wx0, xhi1 = -ws->jbox[0][3] - ws->wy0;
    const bool colInJx = i >= xlo0 && i <= xhi0;
    const bool inPml = PML && (i < f.npml || i >= f.W - f.npml);
    const T be = inPml ? f.pml[2 * f.W + i] : T(0), ae = inPml ? f.pml[3 * f.W + i] : T(0);
    const int j0 = blockIdx.y * STRIP, j1 = min(j0 + STRIP - 1, f.H - 1);
    // rows are independent: each group of four issues all its loads before
    // any arithmetic or store
    constexpr int G4 = 4;
    for (int jb = j0; jb <= j1; jb += G4) {
        size_t qv[G4];
        uint8_t clv[G4];
        T dv[G4], eyv[G4], jvv[G4], psv[G4];
#pragma unroll
        for (int u = 0; u < G4; ++u) {
            const int j = min(jb + u, j1);
            const size_t rb = (size_t)ringy(j, ry, f.Hp) * f.Wp;
            const size_t q = rb + px;
            qv[u] = q;
            clv[u] = f.cls[q];
            dv[u] = f.hz[q] - f.hz[rb + pxm];
            eyv[u] = f.ey[q];
            jvv[u] = (colInJ && j >= jlo1 && j <= jhi1) ? f.jy[q] : T(0);
            psv[u] = (PML && inPml) ? f.psie[q] : T(0);
        }
#pragma unroll
        for (int u = 0; u < G4; ++u) {
            const int j = jb + u;
            if (j > j1) break;
            const size_t q = qv[u];
            const uint8_t cl = clv[u];
            const T d = dv[u], jv = jvv[u];
            if (colInJ && j >= jlo1 && j <= jhi1) f.jy[q] = 0;
            if (colInJx && j >= xlo1 && j <= xhi1) f.jx[q] = 0;
            if (cl == 1) {
                T dd = d;
                if (PML && inPml) {
                    const T psi = be * psv[u] + ae * d;
                    f.psie[q] = psi;
                    dd = d + psi;
                }
                f.ey[q] = eyv[u] - (c.sey * dd + c.dteps * jv);
            } else if (cl == 2 && f.cut[f.plane + q] > T(pm::kSpaceEps) && f.cut[2 * f.plane + q] >= T(pm::kMinArea)) {
                T dd = d;
                if (PML && inPml) {
                    const T psi = be * psv[u] + ae * d;
                    f.psie[q] = psi;
                    dd = d + psi;
                }
                f.ey[q] = eyv[u] + c.ce * (-dd * c.ihx - jv);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// window move decision: keep the bunch's leading edge `lead` cells from the
// window front and its centre near the window centre in y (at most one cell
// per axis per step, as moveMesh, main.cpp:562-652).
// one wavefront: reduce and clear the kBBSlots particle-box slots, then
// lane 0 decides the move
__global__ void k_move(WinState *ws, int W, int H, int Wp, int Hp, int lead)
{
    const int t = threadIdx.x;
    int b[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        b[c] = t < kBBSlots ? ws->bbs[t][c] : INT_MAX;
        if (t < kBBSlots) ws->bbs[t][c] = INT_MAX;
        b[c] = hu::wave_min(b[c]);
    }
    if (t != 0) return;
    for (int c = 0; c < 4; ++c) ws->bb[c] = b[c];
    const int bmaxx = -ws->bb[2] - ws->wx0;
    const int bminy = ws->bb[1] - ws->wy0, bmaxy = -ws->bb[3] - ws->wy0;
    int mi = bmaxx > W - 1 - lead ? 1 : 0;
    const int cy2 = bminy + bmaxy, mid2 = H - 1;
    int mj = cy2 > mid2 + 2 ? 1 : (cy2 < mid2 - 2 ? -1 : 0);
    if (lead < 0) { mi = 0; mj = 0; }   // footprint update only
    ws->mi = mi; ws->mj = mj;
    ws->wx0 += mi; ws->wy0 += mj;
    ws->rx = (ws->rx + mi) % Wp;
    ws->ry = (ws->ry + mj + Hp) % Hp;
    // the footprint of the next deposits: bunch box + stencil margins + one cell of motion
    for (int c = 0; c < 2; ++c) {
        ws->jbox[c][0] = ws->bb[0] - 6; ws->jbox[c][1] = ws->bb[1] - 6;
        ws->jbox[c][2] = ws->bb[2] - 6; ws->jbox[c][3] = ws->bb[3] - 6;
    }
    ws->bb[0] = INT_MAX; ws->bb[1] = INT_MAX; ws->bb[2] = INT_MAX; ws->bb[3] = INT_MAX;
    ws->steps += 1;
}

// clear + geometry of the recycled guard column (mi = 1) / row (mj = +-1)
template <class T>
__global__ void k_fresh(Fields<T> f, Geo g, const WinState *ws)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int mi = ws->mi, mj = ws->mj;
    const int ncol = f.H + 2 * G, nrow = f.W + 2 * G;
    int i, j;
    if (t < ncol) {
        if (mi == 0) return;
        i = f.W - 1 + G; j = t - G;
    } else if (t < ncol + nrow) {
        if (mj == 0) return;
        i = t - ncol - G; j = mj > 0 ? f.H - 1 + G : -G;
    } else if (t < 2 * ncol + nrow) {
        // the column that moved out of the leading PML into the interior
        if (mi == 0 || f.npml == 0) return;
        const size_t q = (size_t)ringy(t - ncol - nrow - G, ws->ry, f.Hp) * f.Wp + ringx(f.W - f.npml - 1, ws->rx, f.Wp);
        f.psih[q] = 0; f.psih2[q] = 0; f.psie[q] = 0;
        return;
    } else {
        return;
    }
    const size_t q = (size_t)ringy(j, ws->ry, f.Hp) * f.Wp + ringx(i, ws->rx, f.Wp);
    f.ex[q] = 0; f.ey[q] = 0; f.hz[q] = 0; f.jx[q] = 0; f.jy[q] = 0; f.ex2[q] = 0; f.hz2[q] = 0;
    if (f.npml) { f.psih[q] = 0; f.psih2[q] = 0; f.psie[q] = 0; }
    cell_geometry(g, f, ws->wx0 + i, ws->wy0 + j, q);
}

// ---------------------------------------------------------------------------
// Poisson start on the window grid (PoissonSolver.cpp:85-201 with the
// window in place of the global grid; window edges are Dirichlet).

// Vertex charge (getChargeDensityImpl) at the particles' positions moved by
// drift * P / gamma (metres) and by shiftx (metres, x only), in window node
// units; FP64 atomics into rho (W x H nodes, logical order).
template <class T, int ORDER>
__global__ void k_pois_scatter(Parts<T> p, Geo g, const WinState *ws, double q, double drift, double shiftx,
                               double *rho, int W, int H)
{
    constexpr int N = ORDER + 2;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.n) return;
    const double Px = ws->p0[0] + (double)p.px[k], Py = ws->p0[1] + (double)p.py[k];
    const double gi = (double)pm::q_rsqrt((float)(1 + (Px * Px + Py * Py)));
    const double u = (double)(p.gx[k] - ws->wx0) + (double)p.fx[k] + (drift * Px * gi + shiftx) / g.h[0];
    const double v = (double)(p.gy[k] - ws->wy0) + (double)p.fy[k] + (drift * Py * gi) / g.h[1];
    double wx[N], wy[N];
    int I0, J0;
    pm::charge_weights<ORDER>(pm::mk(u, v), q, wx, wy, I0, J0);
#pragma unroll
    for (int b = 0; b < N; ++b)
#pragma unroll
        for (int a = 0; a < N; ++a) {
            const int I = I0 + a, J = J0 + b;
            if (I < 0 || J < 0 || I >= W || J >= H) continue;
            unsafeAtomicAdd(&rho[(size_t)J * W + I], pm::charge_w<ORDER>(q, wx[a], wy[b]));
        }
}

// cutoffStencil node classes (PoissonSolver.cpp:678-766) from the window's
// cell classes: flagged = bulk, or cut with area > 0.5; a stencil row if also
// left of the straight-section end; window edge nodes are fixed at zero.
template <class T>
__global__ void k_pois_classes(Fields<T> f, const WinState *ws, long long maxX, uint8_t *cls)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= f.W || j >= f.H) return;
    const size_t q = (size_t)ringy(j, ws->ry, f.Hp) * f.Wp + ringx(i, ws->rx, f.Wp);
    const uint8_t c = f.cls[q];
    const bool flag = c == 1 || (c == 2 && f.cut[2 * f.plane + q] > T(0.5));
    const bool edge = i == 0 || j == 0 || i == f.W - 1 || j == f.H - 1;
    cls[(size_t)j * f.W + i] =
        edge || !flag ? pmg::kFixed : ((long long)ws->wx0 + i < maxX ? pmg::kUnknown : pmg::kKnown);
}

// fillFields (PoissonSolver.cpp:462-507) into the window planes; comp 0 Ex,
// 1 Ey, 2 Hz (the window keeps Hz = Hzx + Hzy, each half the difference).
template <class T>
__global__ void k_pois_fill(Fields<T> f, const WinState *ws, const double *phi, int comp, double scale, double den)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= f.W || j >= f.H) return;
    const size_t q = (size_t)ringy(j, ws->ry, f.Hp) * f.Wp + ringx(i, ws->rx, f.Wp);
    const size_t k = (size_t)j * f.W + i;
    if (comp == 0) {
        const double nb = i + 1 < f.W ? phi[k + 1] : 0.0;
        f.ex[q] = (T)(phi[k] / den - nb / den);
    } else {
        const double nb = j + 1 < f.H ? phi[k + f.W] : 0.0;
        const double v = scale * phi[k] / den - scale * nb / den;
        if (comp == 1) f.ey[q] = (T)v;
        else f.hz[q] = (T)(v + v);
    }
}

// resetOutsideDomain (SCBoundaryCondition.hh:164-196) on the window.
template <class T>
__global__ void k_pois_reset(Fields<T> f, const WinState *ws)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= f.W || j >= f.H) return;
    const size_t q = (size_t)ringy(j, ws->ry, f.Hp) * f.Wp + ringx(i, ws->rx, f.Wp);
    const uint8_t c = f.cls[q];
    if (c == 1) return;
    if (c == 0) { f.ex[q] = T(0); f.ey[q] = T(0); return; }
    if (f.cut[q] < T(pm::kSpaceEps)) f.ex[q] = T(0);
    if (f.cut[f.plane + q] < T(pm::kSpaceEps)) f.ey[q] = T(0);
    if (f.cut[2 * f.plane + q] < T(pm::kMinArea)) { f.ex[q] = T(0); f.ey[q] = T(0); }
}

__global__ void k_pois_rhs(const double *rho, double *b, size_t n, double hh, double cc)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = (0.0 + rho[k] / hh) * cc;
}

__global__ void k_pois_shift_guess(const double *phi, double *out, int nx, int ny, double tau)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)nx * ny) return;
    const int i = (int)(k % nx);
    const double v = phi[k];
    const double t = tau > 0 ? (tau > 1 ? 1.0 : tau) : (-tau > 1 ? 1.0 : -tau);
    const double nb = tau > 0 ? (i + 1 < nx ? phi[k + 1] : 0.0) : (i > 0 ? phi[k - 1] : 0.0);
    out[k] = v + t * nb - t * v;
}

// ---------------------------------------------------------------------------
// cell sort keys and permutation
// The low 8 bits of v moved to the even bit positions 0, 2, .., 14.
__device__ __forceinline__ unsigned spread_bits(unsigned v)
{
    v &= 0xffu;
    v = (v | (v << 4)) & 0x0f0fu;
    v = (v | (v << 2)) & 0x3333u;
    v = (v | (v << 1)) & 0x5555u;
    return v;
}

template <class T>
__global__ void k_sort_keys(Parts<T> p, const WinState *ws, unsigned *keys, int *idx)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.n) return;
    // 16x4-cell tiles over the particles' cell box, row-major; the bits the
    // tile index leaves of 32 (after 6 of cell in the tile) go to the
    // particle's bin in its cell, 2^sb x 2^sb bins (sb <= kMaxSub).
    // Particles move nearly together, so consecutive particles of one bin
    // stay in one cell, or two, for many pushes after the sort: the finer
    // the bins, the fewer runs of a push straddle a cell line.
    const int x0 = ws->sbox[0], y0 = ws->sbox[1];
    const unsigned tw = (unsigned)(-ws->sbox[2] - x0 + 16) >> 4;      // tiles per row
    const unsigned tiles = tw * ((unsigned)(-ws->sbox[3] - y0 + 4) >> 2);
    const int tb = 32 - __clz((int)(tiles > 1u ? tiles - 1u : 1u));  // bits of the largest tile index
    const int sb = max(1, min(kMaxSub, (32 - 6 - tb) / 2));
    const int i = p.gx[k] - x0, j = p.gy[k] - y0;
    const unsigned tile = (unsigned)(j >> 2) * tw + (unsigned)(i >> 4);
    const T sc = T(1 << sb);
    const unsigned bx = (unsigned)(p.fx[k] * sc), by = (unsigned)(p.fy[k] * sc);   // fx, fy in [0, 1)
    // Bins in Z order (y and x bits interleaved, y above x): the quadrant of
    // the cell ranks first (the second push of a step starts half a step on,
    // where a cell's particles have split by halves), and at any density a
    // run of consecutive particles fills a small square of its quadrant
    // rather than a thin row across it.
    const unsigned bins = (spread_bits(by) << 1) | spread_bits(bx);
    keys[k] = (tile << (6 + 2 * sb)) | ((unsigned)(j & 3) << (4 + 2 * sb)) | ((unsigned)(i & 15) << (2 * sb)) | bins;
    idx[k] = k;
}

// Cell box of the particles (min x, min y, -max x, -max y) for the sort
// keys: per-block minima, then one block reduces them into ws->sbox.
template <class T>
__global__ void __launch_bounds__(kB) k_box_partial(Parts<T> p, int *partial)
{
    __shared__ int sh[kB / 64][4];
    int b[4] = {INT_MAX, INT_MAX, INT_MAX, INT_MAX};
    for (int k = blockIdx.x * kB + threadIdx.x; k < p.n; k += gridDim.x * kB) {
        const int x = p.gx[k], y = p.gy[k];
        b[0] = min(b[0], x); b[1] = min(b[1], y); b[2] = min(b[2], -x); b[3] = min(b[3], -y);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) b[c] = hu::wave_min(b[c]);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 4; ++c) sh[threadIdx.x >> 6][c] = b[c];
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        int v = INT_MAX;
        for (int w = 0; w < kB / 64; ++w) v = min(v, sh[w][threadIdx.x]);
        partial[blockIdx.x * 4 + threadIdx.x] = v;
    }
}

__global__ void __launch_bounds__(kB) k_box_final(const int *partial, int nb, WinState *ws)
{
    __shared__ int sh[kB / 64][4];
    int b[4] = {INT_MAX, INT_MAX, INT_MAX, INT_MAX};
    for (int q = threadIdx.x; q < nb; q += kB) {
#pragma unroll
        for (int c = 0; c < 4; ++c) b[c] = min(b[c], partial[q * 4 + c]);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) b[c] = hu::wave_min(b[c]);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 4; ++c) sh[threadIdx.x >> 6][c] = b[c];
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        int v = INT_MAX;
        for (int w = 0; w < kB / 64; ++w) v = min(v, sh[w][threadIdx.x]);
        ws->sbox[threadIdx.x] = v;
    }
}

template <class T>
__global__ void k_permute(Parts<T> src, Parts<T> dst, const int *idx)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= src.n) return;
    const int s = idx[k];
    dst.gx[k] = src.gx[s]; dst.gy[k] = src.gy[s];
    dst.fx[k] = src.fx[s]; dst.fy[k] = src.fy[s];
    dst.px[k] = src.px[s]; dst.py[k] = src.py[s];
}

// ---------------------------------------------------------------------------
// diagnostics: moments (double-double) and the wake along the centroid line
struct DD {
    double hi, lo;
};
__device__ inline DD dd_add(DD x, DD y)
{
    double s = x.hi + y.hi;
    double bb = s - x.hi;
    double e = (x.hi - (s - bb)) + (y.hi - bb);
    e += x.lo + y.lo;
    const double hi = s + e;
    return DD{hi, e - (hi - s)};
}
constexpr int kRedBlocks = 1024;   // blocks of the particle reductions (partials of 4 DD each)

template <class T>
__global__ void k_moments(Parts<T> p, Geo g, const WinState *ws, DD *partial)
{
    __shared__ DD sh[4][kB];
    DD acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < p.n; k += gridDim.x * blockDim.x) {
        const double X = g.org[0] + ((double)p.gx[k] + (double)p.fx[k]) * g.h[0];
        const double Y = g.org[1] + ((double)p.gy[k] + (double)p.fy[k]) * g.h[1];
        acc[0] = dd_add(acc[0], DD{X, 0});
        acc[1] = dd_add(acc[1], DD{ws->p0[0] + (double)p.px[k], 0});
        acc[2] = dd_add(acc[2], DD{Y, 0});
        acc[3] = dd_add(acc[3], DD{ws->p0[1] + (double)p.py[k], 0});
    }
    for (int m = 0; m < 4; ++m) sh[m][threadIdx.x] = acc[m];
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int m = 0; m < 4; ++m) sh[m][threadIdx.x] = dd_add(sh[m][threadIdx.x], sh[m][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int m = 0; m < 4; ++m) partial[blockIdx.x * 4 + m] = sh[m][0];
}

// M double-double partials per block, kRedBlocks blocks: one workgroup of kB
// threads sums them in a fixed order (strided per thread, then a tree)
template <int M>
__device__ void reduce_partials(const DD *partial, DD (&out)[M])
{
    __shared__ DD sh[M][kB];
    for (int m = 0; m < M; ++m) {
        DD s{0, 0};
        for (int b = threadIdx.x; b < kRedBlocks; b += kB) s = dd_add(s, partial[b * M + m]);
        sh[m][threadIdx.x] = s;
    }
    __syncthreads();
    for (int w = kB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int m = 0; m < M; ++m) sh[m][threadIdx.x] = dd_add(sh[m][threadIdx.x], sh[m][threadIdx.x + w]);
        __syncthreads();
    }
    for (int m = 0; m < M; ++m) out[m] = sh[m][0];
}

__global__ void __launch_bounds__(kB) k_moments_final(const DD *partial, int n, double *mom)
{
    DD s[4];
    reduce_partials<4>(partial, s);
    if (threadIdx.x == 0)
        for (int m = 0; m < 4; ++m) mom[m] = (s[m].hi + s[m].lo) / n;
}

// Wake: E along the line through the centroid in the direction of the mean
// momentum, with the rotation convention of IntegrateElong.cpp:49-61;
// accumulated with compensation over the sampling steps.  Sample k sits at
//  - grid = 0: the distance (k - ns/2) h0 from the centroid (a fixed set of
//    offsets along the bunch);
//  - grid = 1: the distance X - x_centroid, X the x of the window's grid
//    column k + 2, as IntegrateElong.cpp:59-61 takes the window's grid
//    points (the offsets shift with the window).
template <class T>
__global__ void k_wake(Fields<T> f, Geo g, const WinState *ws, const double *mom, int ns, DD *acc, int grid)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const double R0 = mom[0], P0 = mom[1], R1 = mom[2], P1 = mom[3];
    const double lenP = sqrt(P0 * P0 + P1 * P1);
    const double cosP = P0 / lenP, sinP = P1 / lenP;
    const double l0 = grid ? (g.org[0] + (double)(ws->wx0 + k + 2) * g.h[0]) - R0 : (double)(k - ns / 2) * g.h[0];
    const double x = (cosP * l0) + R0;
    const double y = (-sinP * l0) + R1;
    // position in global cells
    const double ux = (x - g.org[0]) / g.h[0], uy = (y - g.org[1]) / g.h[1];
    const int gx = (int)floor(ux), gy = (int)floor(uy);
    const int i = gx - ws->wx0, j = gy - ws->wy0;
    if (i < 1 || i > f.W - 2 || j < 1 || j > f.H - 2) return;
    const vec2<T> pp = pm::mk((T)(ux - gx), (T)(uy - gy));
    auto fx = [&](long long I, long long J) {
        return f.ex[(size_t)ringy(j + (int)J, ws->ry, f.Hp) * f.Wp + ringx(i + (int)I, ws->rx, f.Wp)];
    };
    auto fy = [&](long long I, long long J) {
        return f.ey[(size_t)ringy(j + (int)J, ws->ry, f.Hp) * f.Wp + ringx(i + (int)I, ws->rx, f.Wp)];
    };
    const double ex = (double)pm::gather(pm::edge_stencil(0, 0, pp), false, fx);
    const double ey = (double)pm::gather(pm::edge_stencil(0, 1, pp), false, fy);
    acc[k] = dd_add(acc[k], DD{cosP * ex + sinP * ey, 0.0});
}

// Re-centring of the split momentum: p0 moves to the mean momentum (kept a
// value of T, so p0 + dP is unchanged up to one rounding of dP).
template <class T>
__global__ void k_dp_partial(Parts<T> p, DD *partial)
{
    __shared__ DD sh[2][kB];
    DD acc[2] = {{0, 0}, {0, 0}};
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < p.n; k += gridDim.x * blockDim.x) {
        acc[0] = dd_add(acc[0], DD{(double)p.px[k], 0});
        acc[1] = dd_add(acc[1], DD{(double)p.py[k], 0});
    }
    sh[0][threadIdx.x] = acc[0];
    sh[1][threadIdx.x] = acc[1];
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int m = 0; m < 2; ++m) sh[m][threadIdx.x] = dd_add(sh[m][threadIdx.x], sh[m][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) { partial[blockIdx.x * 2] = sh[0][0]; partial[blockIdx.x * 2 + 1] = sh[1][0]; }
}

template <class T>
__global__ void __launch_bounds__(kB) k_dp_final(const DD *partial, int n, WinState *ws, double *shift)
{
    DD s[2];
    reduce_partials<2>(partial, s);
    if (threadIdx.x) return;
    for (int m = 0; m < 2; ++m) {
        const double mean = (s[m].hi + s[m].lo) / n;
        const T pn = (T)(ws->p0[m] + mean);
        shift[m] = (double)pn - ws->p0[m];
        ws->p0[m] = (double)pn;
    }
}

template <class T>
__global__ void k_dp_apply(Parts<T> p, const double *shift)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.n) return;
    p.px[k] = p.px[k] - (T)shift[0];
    p.py[k] = p.py[k] - (T)shift[1];
}

__global__ void k_state_init(WinState *ws, int wx0, int wy0)
{
    ws->wx0 = wx0; ws->wy0 = wy0; ws->rx = 0; ws->ry = 0; ws->mi = 0; ws->mj = 0;
    for (int k = 0; k < 4; ++k) ws->bb[k] = INT_MAX;
    for (int q = 0; q < kBBSlots; ++q)
        for (int k = 0; k < 4; ++k) ws->bbs[q][k] = INT_MAX;
    for (int c = 0; c < 2; ++c)
        for (int k = 0; k < 4; ++k) ws->jbox[c][k] = INT_MAX;
    ws->err = 0;
    ws->steps = 0;
}

__global__ void k_set_p0(WinState *ws, double p0x, double p0y)
{
    ws->p0[0] = p0x;
    ws->p0[1] = p0y;
}

__global__ void k_bbox_init(Parts<float> pf, Parts<double> pd, bool isf, WinState *ws)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = isf ? pf.n : pd.n;
    int b[4] = {INT_MAX, INT_MAX, INT_MAX, INT_MAX};
    if (k < n) {
        const int gx = isf ? pf.gx[k] : pd.gx[k], gy = isf ? pf.gy[k] : pd.gy[k];
        b[0] = gx; b[1] = gy; b[2] = -gx; b[3] = -gy;
    }
    // wavefront minimum first: every thread targets the same four words
#pragma unroll
    for (int c = 0; c < 4; ++c) b[c] = hu::wave_min(b[c]);
    if ((threadIdx.x & 63) == 0 && b[0] != INT_MAX) {
        int *s = ws->bbs[(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % kBBSlots];
#pragma unroll
        for (int c = 0; c < 4; ++c) atomicMin(&s[c], b[c]);
    }
}

}  // namespace mw

// ===========================================================================
struct csrst_mw {
    std::string err;
    csrst_mw_params p{};
    mw::Geo geo{};
    double dt = 0;
    hipStream_t s = nullptr;     // the stream the step's work goes to
    hipStream_t own = nullptr;   // the step's own stream (s unless the caller set one)
    int W = 0, H = 0, Wp = 0, Hp = 0;
    size_t plane = 0;
    bool f32 = true;
    // storage
    DBuf<unsigned char> fieldsBuf, partBuf, partBuf2;
    DBuf<uint8_t> cls;
    DBuf<mw::WinState> ws;
    DBuf<unsigned> keys, keys2;
    DBuf<int> idx, idx2;
    DBuf<unsigned char> cubTmp;
    size_t cubBytes = 0;
    DBuf<mw::DD> partial, wacc;
    DBuf<double> mom, shift;
    DBuf<bs::DD> bsPart;   // beam statistics partials
    DBuf<double> bsLim;
    int n = 0;
    double q = 0;
    int nsWake = 0;
    int steps = 0;
    int lead = 32;
    hipEvent_t ev[16];          // two slots of a step's eight timing events (csrst_mw_step alternates)
    double lastMs[5] = {0, 0, 0, 0, 0};
    bool ready = false;

    int cur = 0;   // which of the Ex/Hz (and PML psi) ping-pong buffers is current
    int npml = 0;  // PML columns at each window end
    // the push sends slot misses through the wavefront (MISSCOOP).  Chosen by
    // timing: every kRetune steps one step runs each form and the faster stays
    bool tilted = false;
    int tunePhase = 0;            // 0 idle, 1 timing the cooperative form, 2 the per-lane form
    int nextTune = 0;             // step at which the next comparison starts
    float tuneMs = 0;             // push time of the cooperative form
    DBuf<unsigned char> pmlBuf;   // psih, psih', psie planes + [4][W] coefficients

    template <class T>
    mw::Fields<T> fields()
    {
        mw::Fields<T> f;
        T *b = reinterpret_cast<T *>(fieldsBuf.p);
        T *exa = b, *exb = b + 5 * plane, *hza = b + 2 * plane, *hzb = b + 6 * plane;
        f.ex = cur ? exb : exa; f.ex2 = cur ? exa : exb;
        f.hz = cur ? hzb : hza; f.hz2 = cur ? hza : hzb;
        f.ey = b + plane; f.jx = b + 3 * plane; f.jy = b + 4 * plane;
        f.cut = b + 7 * plane;
        f.cls = cls.p;
        f.W = W; f.H = H; f.Wp = Wp; f.Hp = Hp; f.plane = plane;
        f.npml = npml;
        T *pb = reinterpret_cast<T *>(pmlBuf.p);
        f.psih = npml ? (cur ? pb + plane : pb) : nullptr;
        f.psih2 = npml ? (cur ? pb : pb + plane) : nullptr;
        f.psie = npml ? pb + 2 * plane : nullptr;
        f.pml = npml ? pb + 3 * plane : nullptr;
        return f;
    }
    // elements per particle array: a multiple of 64, so every array starts
    // 256-byte aligned and runs of consecutive particles load as vectors
    size_t stride() const { return ((size_t)n + 63) & ~(size_t)63; }
    size_t part_bytes() const { return stride() * (8 + 4 * (f32 ? 4 : 8)); }
    template <class T>
    mw::Parts<T> parts(DBuf<unsigned char> &buf)
    {
        mw::Parts<T> pp;
        unsigned char *b = buf.p;
        const size_t nn = stride();
        pp.gx = reinterpret_cast<int *>(b);
        pp.gy = reinterpret_cast<int *>(b + 4 * nn);
        T *r = reinterpret_cast<T *>(b + 8 * nn);
        pp.fx = r; pp.fy = r + nn; pp.px = r + 2 * nn; pp.py = r + 3 * nn;
        pp.n = n;
        return pp;
    }
};

namespace {


int mw_fail(csrst_mw *m, const std::exception &e)
{
    const Error *he = dynamic_cast<const Error *>(&e);
    if (m) m->err = e.what();
    hu::global_error() = e.what();
    return he ? he->code : CSRST_EINVAL;
}

template <class T>
void sweep_consts(const csrst_mw *m, mw::SweepConst<T> &c)
{
    const double etax = pm::kC * m->dt / m->geo.h[0];
    const double etay = pm::kC * m->dt / m->geo.h[1];
    const double Z0 = sqrt(pm::kMu0 / pm::kEps0);
    c.sx = (T)(0.5 * etax / Z0);
    c.sy = (T)(0.5 * etay / Z0);
    c.sex = (T)(etay * Z0);
    c.sey = (T)(etax * Z0);
    c.dteps = (T)(pm::kC * m->dt * Z0);
    c.cx = (T)(0.5 * (m->dt / (m->geo.h[0] * pm::kMu0)));
    c.cy = (T)(0.5 * (m->dt / (m->geo.h[1] * pm::kMu0)));
    c.ce = (T)(m->dt / pm::kEps0);
    c.ihx = (T)(1.0 / m->geo.h[0]);
    c.ihy = (T)(1.0 / m->geo.h[1]);
}

template <class T>
void push_consts(const csrst_mw *m, mw::PushConst<T> &c)
{
    const double half = 0.5 * m->dt * pm::kC;
    c.hx = (T)(half / m->geo.h[0]);
    c.hy = (T)(half / m->geo.h[1]);
    const double f1 = half * (-1e-9 / pm::kMeGeV);
    c.fact1 = (T)f1;
    c.fact2 = (T)(f1 * pm::kC);
    c.q = (T)m->q;
    c.qodt_x = (T)(m->q / (m->geo.h[0] * m->dt));
    c.qodt_y = (T)(m->q / (m->geo.h[1] * m->dt));
    c.odt_x = (T)(1.0 / (m->geo.h[0] * m->dt));
    c.odt_y = (T)(1.0 / (m->geo.h[1] * m->dt));
}

template <class T, int ORDER, int COMP>
void launch_push(csrst_mw *m)
{
    mw::PushConst<T> c;
    push_consts(m, c);
    const dim3 grid((m->n + mw::kChunk - 1) / mw::kChunk);
    if (m->tilted)
        hipLaunchKernelGGL((mw::k_push<T, ORDER, COMP, true>), grid, dim3(mw::kB), 0, m->s, m->parts<T>(m->partBuf),
                           m->fields<T>(), m->geo, c, m->ws.p);
    else
        hipLaunchKernelGGL((mw::k_push<T, ORDER, COMP, false>), grid, dim3(mw::kB), 0, m->s, m->parts<T>(m->partBuf),
                           m->fields<T>(), m->geo, c, m->ws.p);
}

template <class T, int COMP>
void push_dispatch(csrst_mw *m)
{
    if (m->p.order == 0) launch_push<T, 0, COMP>(m);
    else if (m->p.order == 1) launch_push<T, 1, COMP>(m);
    else launch_push<T, 2, COMP>(m);
    HU_LAUNCH_CHECK();
}

template <class T>
void sort_particles(csrst_mw *m)
{
    const int nb = (m->n + mw::kB - 1) / mw::kB;
    int *boxPart = reinterpret_cast<int *>(m->partial.p);   // kRedBlocks x 4 ints
    mw::k_box_partial<T><<<mw::kRedBlocks, mw::kB, 0, m->s>>>(m->parts<T>(m->partBuf), boxPart);
    mw::k_box_final<<<1, mw::kB, 0, m->s>>>(boxPart, mw::kRedBlocks, m->ws.p);
    mw::k_sort_keys<T><<<nb, mw::kB, 0, m->s>>>(m->parts<T>(m->partBuf), m->ws.p, m->keys.p, m->idx.p);
    HU_LAUNCH_CHECK();
    HU_CHECK(hipcub::DeviceRadixSort::SortPairs(m->cubTmp.p, m->cubBytes, m->keys.p, m->keys2.p, m->idx.p, m->idx2.p,
                                                m->n, 0, 32, m->s));
    mw::k_permute<T><<<nb, mw::kB, 0, m->s>>>(m->parts<T>(m->partBuf), m->parts<T>(m->partBuf2), m->idx2.p);
    HU_LAUNCH_CHECK();
    std::swap(m->partBuf.p, m->partBuf2.p);
    std::swap(m->partBuf.n, m->partBuf2.n);
    mw::k_dp_partial<T><<<mw::kRedBlocks, mw::kB, 0, m->s>>>(m->parts<T>(m->partBuf), m->partial.p);
    mw::k_dp_final<T><<<1, mw::kB, 0, m->s>>>(m->partial.p, m->n, m->ws.p, m->shift.p);
    mw::k_dp_apply<T><<<nb, mw::kB, 0, m->s>>>(m->parts<T>(m->partBuf), m->shift.p);
    HU_LAUNCH_CHECK();
}

template <class T>
void wake(csrst_mw *m)
{
    mw::k_moments<T><<<mw::kRedBlocks, mw::kB, 0, m->s>>>(m->parts<T>(m->partBuf), m->geo, m->ws.p, m->partial.p);
    mw::k_moments_final<<<1, mw::kB, 0, m->s>>>(m->partial.p, m->n, m->mom.p);
    mw::k_wake<T><<<(m->nsWake + mw::kB - 1) / mw::kB, mw::kB, 0, m->s>>>(m->fields<T>(), m->geo, m->ws.p, m->mom.p,
                                                                          m->nsWake, m->wacc.p, m->p.wakeGrid);
    HU_LAUNCH_CHECK();
}

template <class T>
void one_step(csrst_mw *m, bool timed, int slot)
{
    hipEvent_t *ev = m->ev + 8 * slot;
    mw::SweepConst<T> sc;
    sweep_consts(m, sc);
    const mw::Fields<T> f = m->fields<T>();
    // interior columns without PML terms; the two PML layers in a launch of their own
    const int np = m->npml, nin = m->W - 2 * np;
    const dim3 gs((nin + mw::kB - 1) / mw::kB, (m->H + mw::kStrip - 1) / mw::kStrip);
    // the PML columns are few: one wavefront of columns per workgroup and
    // short strips, so that the launch still spreads over the chip
    constexpr int kPmlStrip = 8;
    const dim3 gp((2 * np + 63) / 64, (m->H + kPmlStrip - 1) / kPmlStrip);
    if (m->n > 0 && m->p.sortEvery > 0 && m->steps % m->p.sortEvery == 0) {
        if (timed) HU_CHECK(hipEventRecord(ev[6], m->s));
        sort_particles<T>(m);
        if (timed) HU_CHECK(hipEventRecord(ev[7], m->s));
    }
    if (timed) HU_CHECK(hipEventRecord(ev[0], m->s));
    if (m->n > 0) push_dispatch<T, 0>(m);
    if (timed) HU_CHECK(hipEventRecord(ev[1], m->s));
    hipLaunchKernelGGL((mw::k_heh<T, false, mw::kStrip>), gs, dim3(mw::kB), 0, m->s, f, sc, m->ws.p, 0, np, nin);
    if (np) hipLaunchKernelGGL((mw::k_heh<T, true, kPmlStrip>), gp, dim3(64), 0, m->s, f, sc, m->ws.p, 0, 0, 2 * np);
    HU_LAUNCH_CHECK();
    m->cur ^= 1;
    const mw::Fields<T> f2 = m->fields<T>();
    if (timed) HU_CHECK(hipEventRecord(ev[2], m->s));
    if (m->n > 0) push_dispatch<T, 1>(m);
    if (timed) HU_CHECK(hipEventRecord(ev[3], m->s));
    hipLaunchKernelGGL((mw::k_ey<T, false, mw::kStrip>), gs, dim3(mw::kB), 0, m->s, f2, sc, m->ws.p, np, nin);
    if (np) hipLaunchKernelGGL((mw::k_ey<T, true, kPmlStrip>), gp, dim3(64), 0, m->s, f2, sc, m->ws.p, 0, 2 * np);
    HU_LAUNCH_CHECK();
    if (m->n > 0 && m->p.statModulus > 0 && m->steps % m->p.statModulus == m->p.statModulus - 1) wake<T>(m);
    if (timed) HU_CHECK(hipEventRecord(ev[4], m->s));
    if (m->n > 0) {
        mw::k_move<<<1, 64, 0, m->s>>>(m->ws.p, m->W, m->H, m->Wp, m->Hp, m->lead);
        const int nf = 2 * (m->H + 2 * mw::G) + m->W + 2 * mw::G;
        hipLaunchKernelGGL((mw::k_fresh<T>), dim3((nf + mw::kB - 1) / mw::kB), dim3(mw::kB), 0, m->s, f2, m->geo, m->ws.p);
        HU_LAUNCH_CHECK();
    }
    if (timed) HU_CHECK(hipEventRecord(ev[5], m->s));
    ++m->steps;
}

// the two push forms are compared again after this many steps (the bunch's
// angle and density change slowly; one comparison costs at most one step in
// the slower form)
constexpr int kRetune = 200;

void check_state(csrst_mw *m)
{
    mw::WinState h;
    HU_CHECK(hipMemcpyAsync(&h, m->ws.p, sizeof(h), hipMemcpyDeviceToHost, m->s));
    HU_CHECK(hipStreamSynchronize(m->s));
    if (h.err & 1) throw Error(CSRST_EWINDOW, "particle stencil left the window");
    if (h.err & 2) throw Error(CSRST_EWINDOW, "particle crossed too many cell lines in one push");
}

}  // namespace

template <class T, int ORDER>
void mw_scatter(csrst_mw *m, double drift, double shiftx, double *rho)
{
    mw::k_pois_scatter<T, ORDER><<<(m->n + mw::kB - 1) / mw::kB, mw::kB, 0, m->s>>>(
        m->parts<T>(m->partBuf), m->geo, m->ws.p, m->q, drift, shiftx, rho, m->W, m->H);
}

// PoissonSolver::computeField (PoissonSolver.cpp:85-201) on the window, then
// resetOutsideDomain.  gamma from the mean momentum; hx scaled by gamma.
template <class T>
void mw_poisson(csrst_mw *m, double tol, int maxit, int *cycles, double *resid)
{
    const int W = m->W, H = m->H;
    const size_t nn = (size_t)W * H;
    mw::k_moments<T><<<mw::kRedBlocks, mw::kB, 0, m->s>>>(m->parts<T>(m->partBuf), m->geo, m->ws.p, m->partial.p);
    mw::k_moments_final<<<1, mw::kB, 0, m->s>>>(m->partial.p, m->n, m->mom.p);
    double mom[4];
    HU_CHECK(hipMemcpyAsync(mom, m->mom.p, sizeof(mom), hipMemcpyDeviceToHost, m->s));
    HU_CHECK(hipStreamSynchronize(m->s));
    const double gamma = sqrt(mom[1] * mom[1] + mom[3] * mom[3] + 1);
    const double hr0 = m->geo.h[0] * gamma, hr1 = m->geo.h[1];
    DBuf<uint8_t> cls;
    DBuf<double> rho, b, phi, guess;
    cls.alloc(nn); rho.alloc(nn); b.alloc(nn); phi.alloc(nn); guess.alloc(nn);
    const mw::Fields<T> f = m->fields<T>();
    const dim3 gg((W + mw::kB - 1) / mw::kB, H);
    const long long maxX = (long long)(size_t)(m->p.lenBefore / m->geo.h[0]);   // Bend.hh:288-291
    mw::k_pois_classes<T><<<gg, mw::kB, 0, m->s>>>(f, m->ws.p, maxX, cls.p);
    HU_LAUNCH_CHECK();
    pmg::Solver mg;
    mg.setup(W, H, hr0, hr1, cls.p, m->s);
    const double cc = 1 / (pm::kEps0 * gamma);
    const int nb = (int)((nn + mw::kB - 1) / mw::kB);
    auto rhs = [&](double drift, double shiftx) {
        HU_CHECK(hipMemsetAsync(rho.p, 0, nn * sizeof(double), m->s));
        if (m->p.order == 0) mw_scatter<T, 0>(m, drift, shiftx, rho.p);
        else if (m->p.order == 1) mw_scatter<T, 1>(m, drift, shiftx, rho.p);
        else mw_scatter<T, 2>(m, drift, shiftx, rho.p);
        mw::k_pois_rhs<<<nb, mw::kB, 0, m->s>>>(rho.p, b.p, nn, m->geo.h[0] * m->geo.h[1], cc);
        HU_LAUNCH_CHECK();
    };
    auto solve = [&](int k) {
        const pmg::Stats st = mg.solve(b.p, phi.p, tol, maxit);
        if (cycles) cycles[k] = st.cycles;
        if (resid) resid[k] = st.bnorm > 0 ? st.resid / st.bnorm : st.resid;
    };
    // solve 1: Ey from the charge at t + dt/2
    rhs(m->dt / 2 * pm::kC, 0.0);
    HU_CHECK(hipMemsetAsync(phi.p, 0, nn * sizeof(double), m->s));
    solve(0);
    mw::k_pois_fill<T><<<gg, mw::kB, 0, m->s>>>(f, m->ws.p, phi.p, 1, gamma, hr1);
    // solve 2: Hz with the bunch half a cell back
    rhs(0.0, -0.5 * hr0 / gamma);
    mw::k_pois_shift_guess<<<nb, mw::kB, 0, m->s>>>(phi.p, guess.p, W, H, 1.0);
    std::swap(phi.p, guess.p);
    solve(1);
    mw::k_pois_fill<T><<<gg, mw::kB, 0, m->s>>>(f, m->ws.p, phi.p, 2, sqrt(gamma * gamma - 1),
                                                 2 * pm::kC * pm::kMu0 * hr1);
    // solve 3: Ex
    rhs(0.0, 0.0);
    mw::k_pois_shift_guess<<<nb, mw::kB, 0, m->s>>>(phi.p, guess.p, W, H, -0.5);
    std::swap(phi.p, guess.p);
    solve(2);
    mw::k_pois_fill<T><<<gg, mw::kB, 0, m->s>>>(f, m->ws.p, phi.p, 0, 1.0, hr0);
    mw::k_pois_reset<T><<<gg, mw::kB, 0, m->s>>>(f, m->ws.p);
    HU_LAUNCH_CHECK();
    HU_CHECK(hipStreamSynchronize(m->s));
}

namespace {
// particle source of the beam statistics: split positions and momentum
// deviations, as csrst_mw_download_particles forms R and P
template <class T>
struct MwBeamSrc {
    mw::Parts<T> p;
    double o0, o1, h0, h1, p00, p01;
    __device__ void get(int k, double &x, double &y, double &a, double &b) const
    {
        x = o0 + ((double)p.gx[k] + (double)p.fx[k]) * h0;
        y = o1 + ((double)p.gy[k] + (double)p.fy[k]) * h1;
        a = p00 + (double)p.px[k];
        b = p01 + (double)p.py[k];
    }
};
}  // namespace

extern "C" {

int csrst_mw_create(const csrst_mw_params *p, csrst_mw **out)
{
    std::unique_ptr<csrst_mw> m(new csrst_mw());
    try {
        if (!p || !out) throw Error(CSRST_EINVAL, "null argument");
        if (p->W < 16 || p->H < 16) throw Error(CSRST_EINVAL, "window too small");
        if (p->order < 0 || p->order > 2) throw Error(CSRST_EINVAL, "order must be 0..2");
        if (p->precision != 0 && p->precision != 1) throw Error(CSRST_EINVAL, "precision must be 0 or 1");
        if (!(p->dx > 0) || !(p->dy > 0)) throw Error(CSRST_EINVAL, "spacing must be positive");
        if (p->dx > p->dy * (1 + 1e-12)) throw Error(CSRST_EINVAL, "TE/TM scheme needs dx <= dy");
        m->p = *p;
        m->f32 = p->precision == 1;
        HU_CHECK(hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking));
        m->own = m->s;
        for (auto &e : m->ev) HU_CHECK(hipEventCreate(&e));
        // geometry: the bend of Bend.cpp with the configured spacing; the
        // virtual global mesh origin as in Bend::Init (vertex 1 on x = 0)
        geom::Bend &b = m->geo.bend;
        geom::bend_init(b, p->lenBefore, p->lenAfter, p->width, p->phiDeg, p->ekinMeV, p->bz, 1000, 1000);
        const double dwid0 = b.dwid - b.hr[1];
        m->geo.h[0] = p->dx;
        m->geo.h[1] = p->dy;
        m->geo.org[0] = -p->dx;
        m->geo.org[1] = b.bz <= 0.0 ? (p->width - dwid0) - p->dy : -p->dy;
        b.hr[0] = p->dx; b.hr[1] = p->dy;
        b.origin[0] = m->geo.org[0]; b.origin[1] = m->geo.org[1];
        const double cfl = p->cfl > 0 ? p->cfl : 1.0;
        m->dt = cfl * p->dx / pm::kC;
        m->W = p->W; m->H = p->H;
        m->Wp = p->W + 2 * mw::G; m->Hp = p->H + 2 * mw::G;
        m->plane = (size_t)m->Wp * m->Hp;
        const size_t rs = m->f32 ? 4 : 8;
        // the push gathers address a field plane with 32-bit byte offsets
        if (m->plane * rs > 0xffffffffull) throw Error(CSRST_EINVAL, "window plane exceeds 4 GiB");
        m->fieldsBuf.alloc(m->plane * 12 * rs);   // ex ey hz jx jy ex' hz' + 5 cut planes
        m->fieldsBuf.zero(m->s);
        if (p->pml < 0 || 2 * p->pml > p->W - 8) throw Error(CSRST_EINVAL, "pml must be in [0, (W - 8) / 2]");
        // sort keys: a 16x4-cell tile index over the particles' box (at most
        // the window plus its guards) above 6 bits of cell in the tile and at
        // least one bit per axis of position in the cell
        if ((long long)((p->H + 16 + 3) / 4) * (long long)((p->W + 16 + 15) / 16) > (1ll << 24))
            throw Error(CSRST_EINVAL, "window too large for the 32-bit particle sort key");
        m->npml = p->pml;
        if (m->npml) {
            // Convolutional PML (kappa 1, alpha 0): psi <- b psi + (b - 1) D once
            // per step, b = exp(-sigma dt / eps0), sigma = sigma_max d^3 over the
            // layer with sigma_max eta0 hx = 1.6 (the lowest reflection of a 1-D
            // scan of the grading for this scheme at CFL 1 and 0.5).
            const int W = p->W, np = m->npml;
            std::vector<double> co(4 * (size_t)W, 0.0);
            auto depth = [&](double x) {   // x in cells (logical), depth into the layer in [0, 1]
                if (x < np) return (np - x) / np;
                if (x > W - np) return (x - (W - np)) / np;
                return 0.0;
            };
            for (int i = 0; i < W; ++i) {
                const double dh = depth(i + 0.5), de = depth((double)i);
                const double bh = exp(-1.6 * dh * dh * dh * cfl), be = exp(-1.6 * de * de * de * cfl);
                co[i] = bh; co[W + i] = bh - 1; co[2 * (size_t)W + i] = be; co[3 * (size_t)W + i] = be - 1;
            }
            m->pmlBuf.alloc((3 * m->plane + 4 * (size_t)W) * rs);
            m->pmlBuf.zero(m->s);
            HU_CHECK(hipStreamSynchronize(m->s));
            if (m->f32) {
                std::vector<float> cf(co.begin(), co.end());
                HU_CHECK(hipMemcpy(m->pmlBuf.p + 3 * m->plane * rs, cf.data(), cf.size() * 4, hipMemcpyHostToDevice));
            } else {
                HU_CHECK(hipMemcpy(m->pmlBuf.p + 3 * m->plane * rs, co.data(), co.size() * 8, hipMemcpyHostToDevice));
            }
        }
        m->cls.alloc(m->plane);
        m->cls.zero(m->s);
        m->ws.alloc(1);
        m->partial.alloc((size_t)mw::kRedBlocks * 4);
        m->mom.alloc(4);
        m->shift.alloc(2);
        HU_CHECK(hipStreamSynchronize(m->s));
        *out = m.release();
        return CSRST_OK;
    } catch (const std::exception &e) {
        return mw_fail(nullptr, e);
    }
}

int csrst_mw_destroy(csrst_mw *m)
{
    if (!m) return CSRST_OK;
    if (m->s) (void)hipStreamSynchronize(m->s);
    if (m->own) (void)hipStreamDestroy(m->own);
    for (auto &e : m->ev) if (e) (void)hipEventDestroy(e);
    delete m;
    return CSRST_OK;
}

int csrst_mw_init_bunch(csrst_mw *m, const double *R, const double *P, const double *Q, int n)
{
    try {
        if (!m || !R || !P || !Q || n <= 0) throw Error(CSRST_EINVAL, "bad bunch arguments");
        m->n = n;
        m->q = Q[0];
        for (int k = 1; k < n; ++k)
            if (Q[k] != m->q) throw Error(CSRST_EINVAL, "the moving-window step takes a uniform macro-particle charge");
        // split positions; window: bunch front `lead` cells behind the window
        // front, bunch centred in y
        std::vector<int> gx(n), gy(n);
        const bool f32 = m->f32;
        const size_t rs = f32 ? 4 : 8;
        const size_t ns = m->stride();
        std::vector<unsigned char> host(m->part_bytes(), 0);
        int mnx = INT_MAX, mxx = INT_MIN, mny = INT_MAX, mxy = INT_MIN;
        long double sp[2] = {0, 0};
        for (int k = 0; k < n; ++k) { sp[0] += P[2 * k]; sp[1] += P[2 * k + 1]; }
        double p0[2] = {(double)(sp[0] / n), (double)(sp[1] / n)};
        if (f32) { p0[0] = (double)(float)p0[0]; p0[1] = (double)(float)p0[1]; }
        for (int k = 0; k < n; ++k) {
            const double ux = (R[2 * k] - m->geo.org[0]) / m->geo.h[0];
            const double uy = (R[2 * k + 1] - m->geo.org[1]) / m->geo.h[1];
            const double fx = floor(ux), fy = floor(uy);
            gx[k] = (int)fx; gy[k] = (int)fy;
            mnx = std::min(mnx, gx[k]); mxx = std::max(mxx, gx[k]);
            mny = std::min(mny, gy[k]); mxy = std::max(mxy, gy[k]);
            memcpy(&host[4 * (size_t)k], &gx[k], 4);
            memcpy(&host[4 * ns + 4 * (size_t)k], &gy[k], 4);
            const double v[4] = {ux - fx, uy - fy, P[2 * k] - p0[0], P[2 * k + 1] - p0[1]};
            for (int c = 0; c < 4; ++c) {
                unsigned char *dst = &host[8 * ns + ((size_t)c * ns + k) * rs];
                if (f32) {
                    float fv = (float)v[c];
                    if (c < 2 && fv >= 1.0f) fv = nextafterf(1.0f, 0.0f);
                    memcpy(dst, &fv, 4);
                } else {
                    memcpy(dst, &v[c], 8);
                }
            }
        }
        if (mxx - mnx + 2 * m->lead >= m->W || mxy - mny + 16 >= m->H)
            throw Error(CSRST_EWINDOW, hu::fmt("bunch (%d x %d cells) does not fit the window", mxx - mnx + 1, mxy - mny + 1));
        const int wx0 = mxx - (m->W - 1 - m->lead);
        const int wy0 = (mny + mxy) / 2 - (m->H - 1) / 2;
        m->partBuf.alloc(host.size());
        m->partBuf2.alloc(host.size());
        HU_CHECK(hipMemcpy(m->partBuf.p, host.data(), host.size(), hipMemcpyHostToDevice));
        mw::k_state_init<<<1, 1, 0, m->s>>>(m->ws.p, wx0, wy0);
        mw::k_set_p0<<<1, 1, 0, m->s>>>(m->ws.p, p0[0], p0[1]);
        m->tilted = false;
        m->tunePhase = 0;
        m->nextTune = 0;
        const dim3 gg((m->W + 2 * mw::G + mw::kB - 1) / mw::kB, m->H + 2 * mw::G);
        if (f32) hipLaunchKernelGGL((mw::k_geom_all<float>), gg, dim3(mw::kB), 0, m->s, m->fields<float>(), m->geo, m->ws.p);
        else hipLaunchKernelGGL((mw::k_geom_all<double>), gg, dim3(mw::kB), 0, m->s, m->fields<double>(), m->geo, m->ws.p);
        HU_LAUNCH_CHECK();
        mw::Parts<float> pf{}; mw::Parts<double> pd{};
        if (f32) pf = m->parts<float>(m->partBuf); else pd = m->parts<double>(m->partBuf);
        mw::k_bbox_init<<<(n + mw::kB - 1) / mw::kB, mw::kB, 0, m->s>>>(pf, pd, f32, m->ws.p);
        // initial footprint = particle box +- 4 (k_move convention)
        mw::k_move<<<1, 64, 0, m->s>>>(m->ws.p, m->W, m->H, m->Wp, m->Hp, -1);
        HU_LAUNCH_CHECK();
        // k_move with a negative lead makes no move; reset the counters
        mw::WinState h;
        HU_CHECK(hipMemcpyAsync(&h, m->ws.p, sizeof(h), hipMemcpyDeviceToHost, m->s));
        HU_CHECK(hipStreamSynchronize(m->s));
        h.steps = 0; h.mi = 0; h.mj = 0;
        for (int c = 0; c < 4; ++c) h.bb[c] = INT_MAX;
        HU_CHECK(hipMemcpy(m->ws.p, &h, sizeof(h), hipMemcpyHostToDevice));
        // sort scratch
        m->keys.alloc(n); m->keys2.alloc(n); m->idx.alloc(n); m->idx2.alloc(n);
        m->cubBytes = 0;
        HU_CHECK(hipcub::DeviceRadixSort::SortPairs((void *)nullptr, m->cubBytes, m->keys.p, m->keys2.p, m->idx.p,
                                                    m->idx2.p, n, 0, 32, m->s));
        m->cubTmp.alloc(m->cubBytes + 16);
        m->nsWake = m->W - 4;
        m->wacc.alloc(m->nsWake);
        m->wacc.zero(m->s);
        m->steps = 0;
        m->ready = true;
        if (m->p.sortEvery > 0) {
            if (f32) sort_particles<float>(m); else sort_particles<double>(m);
        }
        HU_CHECK(hipStreamSynchronize(m->s));
        return CSRST_OK;
    } catch (const std::exception &e) {
        return mw_fail(m, e);
    }
}

int csrst_mw_poisson(csrst_mw *m, double tol, int maxit, int *cycles, double *resid)
{
    try {
        if (!m || !m->ready) throw Error(CSRST_ESTATE, "csrst_mw_poisson before csrst_mw_init_bunch");
        if (m->steps != 0) throw Error(CSRST_ESTATE, "the Poisson start must precede the first step");
        if (maxit < 0 || !(tol >= 0)) throw Error(CSRST_EINVAL, "bad tolerance or iteration limit");
        if (m->f32) mw_poisson<float>(m, tol, maxit, cycles, resid);
        else mw_poisson<double>(m, tol, maxit, cycles, resid);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return mw_fail(m, e);
    }
}

int csrst_mw_step(csrst_mw *m, int nsteps)
{
    try {
        if (!m || !m->ready) throw Error(CSRST_ESTATE, "csrst_mw_step before csrst_mw_init_bunch");
        double acc[5] = {0, 0, 0, 0, 0};
        // Step k records its timing events in slot k % 2, and the host reads
        // them after it has queued step k + 1, so the GPU never idles while
        // the host waits.  The comparison of the two slot-miss forms (a
        // cooperative trial step, then a per-lane one) therefore decides one
        // step after the per-lane trial ends; that step keeps the per-lane form.
        int role[2] = {0, 0};              // per slot: 0 none, 1 cooperative trial, 2 per-lane trial
        bool sorted[2] = {false, false};
        auto read = [&](int b) {
            hipEvent_t *e = m->ev + 8 * b;
            HU_CHECK(hipEventSynchronize(e[5]));
            float t01, t12, t23, t34, t45, t67 = 0;
            HU_CHECK(hipEventElapsedTime(&t01, e[0], e[1]));
            HU_CHECK(hipEventElapsedTime(&t12, e[1], e[2]));
            HU_CHECK(hipEventElapsedTime(&t23, e[2], e[3]));
            HU_CHECK(hipEventElapsedTime(&t34, e[3], e[4]));
            HU_CHECK(hipEventElapsedTime(&t45, e[4], e[5]));
            if (role[b] == 1) {
                m->tuneMs = t01 + t23;
            } else if (role[b] == 2) {
                m->tilted = m->tuneMs < t01 + t23;
                m->tunePhase = 0;
                m->nextTune = m->steps + kRetune;
            }
            if (sorted[b]) HU_CHECK(hipEventElapsedTime(&t67, e[6], e[7]));
            acc[0] += t01 + t23;
            acc[1] += t12 + t34;
            acc[2] += t45;
            acc[3] += t67;
        };
        for (int k = 0; k < nsteps; ++k) {
            const int b = k & 1;
            sorted[b] = m->n > 0 && m->p.sortEvery > 0 && m->steps % m->p.sortEvery == 0;
            role[b] = 0;
            if (m->n > 0 && m->tunePhase == 0 && m->steps >= m->nextTune) {
                m->tunePhase = 1;           // this step runs the cooperative form
                m->tilted = true;
                role[b] = 1;
            } else if (m->tunePhase == 1) {
                m->tunePhase = 2;           // and this one the per-lane form
                m->tilted = false;
                role[b] = 2;
            }
            if (m->f32) one_step<float>(m, true, b); else one_step<double>(m, true, b);
            if (k > 0) read(b ^ 1);
        }
        if (nsteps > 0) read((nsteps - 1) & 1);
        for (int k = 0; k < 4; ++k) m->lastMs[k] = nsteps ? acc[k] / nsteps : 0;
        check_state(m);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return mw_fail(m, e);
    }
}

int csrst_mw_set_stream(csrst_mw *m, void *stream)
{
    try {
        if (!m) throw Error(CSRST_EINVAL, "null");
        HU_CHECK(hipStreamSynchronize(m->s));
        m->s = stream ? (hipStream_t)stream : m->own;
        return CSRST_OK;
    } catch (const std::exception &e) {
        return mw_fail(m, e);
    }
}

int csrst_mw_push(csrst_mw *m, int comp)
{
    try {
        if (!m || !m->ready || m->n <= 0) throw Error(CSRST_ESTATE, "csrst_mw_push before csrst_mw_init_bunch");
        if (comp != 0 && comp != 1) throw Error(CSRST_EINVAL, "comp must be 0 (Jx) or 1 (Jy)");
        if (m->f32) { if (comp == 0) push_dispatch<float, 0>(m); else push_dispatch<float, 1>(m); }
        else { if (comp == 0) push_dispatch<double, 0>(m); else push_dispatch<double, 1>(m); }
        check_state(m);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return mw_fail(m, e);
    }
}

int csrst_mw_place(csrst_mw *m, int wx0, int wy0)
{
    try {
        if (!m) throw Error(CSRST_EINVAL, "null");
        m->n = 0;
        m->cur = 0;
        mw::k_state_init<<<1, 1, 0, m->s>>>(m->ws.p, wx0, wy0);
        mw::k_set_p0<<<1, 1, 0, m->s>>>(m->ws.p, 0.0, 0.0);
        const dim3 gg((m->W + 2 * mw::G + mw::kB - 1) / mw::kB, m->H + 2 * mw::G);
        if (m->f32) hipLaunchKernelGGL((mw::k_geom_all<float>), gg, dim3(mw::kB), 0, m->s, m->fields<float>(), m->geo, m->ws.p);
        else hipLaunchKernelGGL((mw::k_geom_all<double>), gg, dim3(mw::kB), 0, m->s, m->fields<double>(), m->geo, m->ws.p);
        HU_LAUNCH_CHECK();
        m->nsWake = m->W - 4;
        m->wacc.alloc(m->nsWake);
        m->wacc.zero(m->s);
        m->steps = 0;
        m->ready = true;
        HU_CHECK(hipStreamSynchronize(m->s));
        return CSRST_OK;
    } catch (const std::exception &e) {
        return mw_fail(m, e);
    }
}

int csrst_mw_upload_fields(csrst_mw *m, const double *in, int count)
{
    try {
        if (!m || !in || !m->ready) throw Error(CSRST_ESTATE, "not initialised");
        if ((size_t)count != 6 * m->plane) throw Error(CSRST_EINVAL, "size mismatch");
        HU_CHECK(hipStreamSynchronize(m->s));
        mw::WinState h;
        HU_CHECK(hipMemcpy(&h, m->ws.p, sizeof(h), hipMemcpyDeviceToHost));
        const size_t rs = m->f32 ? 4 : 8;
        std::vector<unsigned char> raw(7 * m->plane * rs);
        HU_CHECK(hipMemcpy(raw.data(), m->fieldsBuf.p, raw.size(), hipMemcpyDeviceToHost));
        const int exPlane = m->cur ? 5 : 0, hzPlane = m->cur ? 6 : 2;
        const int dst[5] = {exPlane, 1, hzPlane, 3, 4};
        for (int pl = 0; pl < 5; ++pl)
            for (int j = -mw::G; j < m->H + mw::G; ++j)
                for (int i = -mw::G; i < m->W + mw::G; ++i) {
                    int px = (i + mw::G + h.rx) % m->Wp, py = (j + mw::G + h.ry) % m->Hp;
                    if (px < 0) px += m->Wp;
                    if (py < 0) py += m->Hp;
                    const size_t q = (size_t)py * m->Wp + px;
                    const size_t lq = (size_t)(j + mw::G) * m->Wp + (i + mw::G);
                    double v = pl < 2 ? in[(size_t)pl * m->plane + lq]
                             : pl == 2 ? in[2 * m->plane + lq] + in[3 * m->plane + lq]
                                       : in[(size_t)(pl + 1) * m->plane + lq];
                    unsigned char *d = &raw[((size_t)dst[pl] * m->plane + q) * rs];
                    if (m->f32) { float fv = (float)v; memcpy(d, &fv, 4); }
                    else memcpy(d, &v, 8);
                }
        HU_CHECK(hipMemcpy(m->fieldsBuf.p, raw.data(), raw.size(), hipMemcpyHostToDevice));
        return CSRST_OK;
    } catch (const std::exception &e) {
        return mw_fail(m, e);
    }
}

int csrst_mw_timing(csrst_mw *m, double *ms, int cap)
{
    if (!m || !ms) return CSRST_EINVAL;
    for (int k = 0; k < cap && k < 5; ++k) ms[k] = m->lastMs[k];
    return CSRST_OK;
}

int csrst_mw_download_fields(csrst_mw *m, double *out, int cap, int *count)
{
    try {
        if (!m) throw Error(CSRST_EINVAL, "null");
        const size_t cnt = 6 * m->plane;
        if (count) *count = (int)cnt;
        if (!out) return CSRST_OK;
        if ((size_t)cap < cnt) throw Error(CSRST_EINVAL, "capacity too small");
        HU_CHECK(hipStreamSynchronize(m->s));
        mw::WinState h;
        HU_CHECK(hipMemcpy(&h, m->ws.p, sizeof(h), hipMemcpyDeviceToHost));
        const size_t rs = m->f32 ? 4 : 8;
        std::vector<unsigned char> raw(7 * m->plane * rs);
        HU_CHECK(hipMemcpy(raw.data(), m->fieldsBuf.p, raw.size(), hipMemcpyDeviceToHost));
        const int exPlane = m->cur ? 5 : 0, hzPlane = m->cur ? 6 : 2;
        // unroll the ring into logical order; planes ex ey hz/2 hz/2 jx jy
        for (int pl = 0; pl < 6; ++pl) {
            const int src = pl == 0 ? exPlane : pl == 1 ? 1 : (pl < 4 ? hzPlane : pl - 1);
            for (int j = -mw::G; j < m->H + mw::G; ++j)
                for (int i = -mw::G; i < m->W + mw::G; ++i) {
                    int px = (i + mw::G + h.rx) % m->Wp, py = (j + mw::G + h.ry) % m->Hp;
                    if (px < 0) px += m->Wp;
                    if (py < 0) py += m->Hp;
                    const size_t q = (size_t)py * m->Wp + px;
                    double v;
                    if (m->f32) { float fv; memcpy(&fv, &raw[((size_t)src * m->plane + q) * 4], 4); v = fv; }
                    else memcpy(&v, &raw[((size_t)src * m->plane + q) * 8], 8);
                    if (pl == 2 || pl == 3) v *= 0.5;
                    out[(size_t)pl * m->plane + (size_t)(j + mw::G) * m->Wp + (i + mw::G)] = v;
                }
        }
        return CSRST_OK;
    } catch (const std::exception &e) {
        return mw_fail(m, e);
    }
}


int csrst_mw_beam_stats(csrst_mw *m, double *out)
{
    try {
        if (!m || !m->ready || !out) throw Error(CSRST_ESTATE, "not initialised");
        if (!m->bsPart.p) {
            m->bsPart.alloc((size_t)bs::kBlocks * 11);
            m->bsLim.alloc((size_t)bs::kBlocks * 4);
        }
        mw::WinState h;
        HU_CHECK(hipMemcpyAsync(&h, m->ws.p, sizeof(h), hipMemcpyDeviceToHost, m->s));
        HU_CHECK(hipStreamSynchronize(m->s));
        double raw[bs::kRaw];
        const double *o = m->geo.org, *hh = m->geo.h;
        if (m->f32)
            bs::raw_stats(MwBeamSrc<float>{m->parts<float>(m->partBuf), o[0], o[1], hh[0], hh[1], h.p0[0], h.p0[1]},
                          m->n, m->s, m->bsPart.p, m->bsLim.p, raw);
        else
            bs::raw_stats(MwBeamSrc<double>{m->parts<double>(m->partBuf), o[0], o[1], hh[0], hh[1], h.p0[0], h.p0[1]},
                          m->n, m->s, m->bsPart.p, m->bsLim.p, raw);
        bs::parameters(raw, out);
        return CSRST_OK;
    } catch (const std::exception &e) {
        return mw_fail(m, e);
    }
}

int csrst_mw_download_particles(csrst_mw *m, double *R, double *P, int n)
{
    try {
        if (!m || n != m->n) throw Error(CSRST_EINVAL, "particle count mismatch");
        HU_CHECK(hipStreamSynchronize(m->s));
        const size_t rs = m->f32 ? 4 : 8, ns = m->stride();
        std::vector<unsigned char> host(m->part_bytes());
        HU_CHECK(hipMemcpy(host.data(), m->partBuf.p, host.size(), hipMemcpyDeviceToHost));
        mw::WinState hs;
        HU_CHECK(hipMemcpy(&hs, m->ws.p, sizeof(hs), hipMemcpyDeviceToHost));
        for (int k = 0; k < n; ++k) {
            int gx, gy;
            memcpy(&gx, &host[4 * (size_t)k], 4);
            memcpy(&gy, &host[4 * ns + 4 * (size_t)k], 4);
            double v[4];
            for (int c = 0; c < 4; ++c) {
                const unsigned char *src = &host[8 * ns + ((size_t)c * ns + k) * rs];
                if (m->f32) { float fv; memcpy(&fv, src, 4); v[c] = fv; }
                else memcpy(&v[c], src, 8);
            }
            if (R) {
                R[2 * k] = m->geo.org[0] + ((double)gx + v[0]) * m->geo.h[0];
                R[2 * k + 1] = m->geo.org[1] + ((double)gy + v[1]) * m->geo.h[1];
            }
            if (P) { P[2 * k] = hs.p0[0] + v[2]; P[2 * k + 1] = hs.p0[1] + v[3]; }
        }
        return CSRST_OK;
    } catch (const std::exception &e) {
        return mw_fail(m, e);
    }
}

int csrst_mw_info(csrst_mw *m, double *info)
{
    try {
        if (!m || !info) throw Error(CSRST_EINVAL, "null");
        mw::WinState h;
        HU_CHECK(hipStreamSynchronize(m->s));
        HU_CHECK(hipMemcpy(&h, m->ws.p, sizeof(h), hipMemcpyDeviceToHost));
        info[0] = m->dt;
        info[1] = m->geo.org[0] + h.wx0 * m->geo.h[0];
        info[2] = m->geo.org[1] + h.wy0 * m->geo.h[1];
        info[3] = h.wx0; info[4] = h.wy0; info[5] = m->W; info[6] = m->H;
        info[7] = m->steps; info[8] = m->nsWake; info[9] = m->n;
        return CSRST_OK;
    } catch (const std::exception &e) {
        return mw_fail(m, e);
    }
}

int csrst_mw_wake(csrst_mw *m, double *out, int cap, int *count)
{
    try {
        if (!m || !m->ready) throw Error(CSRST_ESTATE, "not initialised");
        if (count) *count = m->nsWake;
        if (!out) return CSRST_OK;
        if (cap < m->nsWake) throw Error(CSRST_EINVAL, "capacity too small");
        std::vector<mw::DD> h(m->nsWake);
        HU_CHECK(hipStreamSynchronize(m->s));
        HU_CHECK(hipMemcpy(h.data(), m->wacc.p, h.size() * sizeof(mw::DD), hipMemcpyDeviceToHost));
        for (int k = 0; k < m->nsWake; ++k) out[k] = h[k].hi + h[k].lo;
        return CSRST_OK;
    } catch (const std::exception &e) {
        return mw_fail(m, e);
    }
}

}  // extern "C"
