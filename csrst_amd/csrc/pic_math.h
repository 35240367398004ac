// pic_math.h -- per-particle arithmetic of the CSRST PIC step, written for
// gfx950 device code (and usable on the host for set-up).
//
// Every routine keeps the reference's evaluation order so that the FP64
// instantiation is bit-identical to the reference arithmetic; the translation
// unit must be compiled with -ffp-contract=off.  The same templates run in
// FP32 for the split-precision production step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PM_HD __host__ __device__ __forceinline__

namespace pm {

// Physics.cpp:41-51
constexpr double kC = 299792458.0;
constexpr double kMu0 = 1.256637061e-06;
constexpr double kEps0 = 8.854187817e-12;
constexpr double kMeGeV = 0.51099906e-03;
constexpr double kPi = 3.14159265358979323846;
// defs.hh:30-33
constexpr double kSpaceEps = 1e-10;
constexpr double kMinArea = 1e-20;

template <class T>
struct vec2 {
    T x, y;
};
template <class T>
PM_HD vec2<T> mk(T a, T b) { vec2<T> r; r.x = a; r.y = b; return r; }

// utils.hh:491-504 -- fast inverse square root, float, two Newton steps
// (1.5 and the Kadlec pair 1.5010936 / 0.5010936).
PM_HD float q_rsqrt(float x)
{
#if defined(__clang__)
    // as the reference's ISO C++ build: no fused multiply-adds, so the result
    // is the reference's float bit for bit in every translation unit
#pragma clang fp contract(off)
#endif
    const float xhalf = 0.5f * x;
    const float xqhalf = 0.5010936f * x;
    uint32_t i = __builtin_bit_cast(uint32_t, x);
    i = 0x5f3759dfu - (i >> 1);
    float y = __builtin_bit_cast(float, i);
    y = y * (1.5f - xhalf * y * y);
    y = y * (1.5010936f - xqhalf * y * y);
    return y;
}

template <class T>
PM_HD T floor_(T v);
template <>
PM_HD double floor_<double>(double v) { return floor(v); }
template <>
PM_HD float floor_<float>(float v) { return floorf(v); }
template <class T>
PM_HD T sqrt_(T v);
template <>
PM_HD double sqrt_<double>(double v) { return sqrt(v); }
template <>
PM_HD float sqrt_<float>(float v) { return sqrtf(v); }

// ---------------------------------------------------------------------------
// Trajectory splitting at cell (or dual-cell) borders, utils.hh:364-489.
// Positions are relative to the mesh origin.  Up to 1 + |dI| + |dJ| segments;
// `kMaxX` bounds the list (the step's displacement stays below one cell).
// Sorting is a stable insertion sort of the interior crossings; duplicates
// within 1e-6 in fraction are merged keeping the first (libstdc++ semantics).
constexpr int kMaxX = 8;

template <class T>
struct Crossings {
    int n;
    T f[kMaxX];
    vec2<T> s[kMaxX];
};

template <class T>
PM_HD void crossings_finish(Crossings<T> &c, int next, vec2<T> o, vec2<T> e)
{
    for (int a = 2; a < next - 1; ++a) {
        const T vf = c.f[a];
        const vec2<T> vs = c.s[a];
        int b = a;
        while (b > 1 && vf < c.f[b - 1]) {
            c.f[b] = c.f[b - 1];
            c.s[b] = c.s[b - 1];
            --b;
        }
        c.f[b] = vf;
        c.s[b] = vs;
    }
    int dest = 0;
    for (int k = 1; k < next; ++k) {
        const T d = c.f[dest] - c.f[k];
        if (!((d < 0 ? -d : d) < T(1e-6))) {
            ++dest;
            c.f[dest] = c.f[k];
            c.s[dest] = c.s[k];
        }
    }
    c.n = dest + 1;
    c.s[0] = o;
    c.s[c.n - 1] = e;
}

// shift = 0 for cell borders (x = I*h), 0.5 for dual-cell borders (x = (I+0.5)*h).
// Returns false if more than kMaxX entries would be needed.
template <class T>
PM_HD bool crossings(Crossings<T> &c, vec2<T> o, vec2<T> e, vec2<T> sp, bool dual)
{
    int Ii, Ji, If, Jf;
    if (dual) {
        Ii = (int)floor_(o.x / sp.x - T(0.5)); Ji = (int)floor_(o.y / sp.y - T(0.5));
        If = (int)floor_(e.x / sp.x - T(0.5)); Jf = (int)floor_(e.y / sp.y - T(0.5));
    } else {
        Ii = (int)floor_(o.x / sp.x); Ji = (int)floor_(o.y / sp.y);
        If = (int)floor_(e.x / sp.x); Jf = (int)floor_(e.y / sp.y);
    }
    const int need = 2 + (If > Ii ? If - Ii : Ii - If) + (Jf > Ji ? Jf - Ji : Ji - Jf);
    if (need > kMaxX) return false;
    const vec2<T> D = mk(e.x - o.x, e.y - o.y);
    int I = Ii, J = Ji, next = 0;
    c.f[next++] = T(0);
    const T up = dual ? T(1.5) : T(1), dn = dual ? T(0.5) : T(0);
    while (I < If) {
        const T xb = sp.x * (I + up);
        c.f[next] = (xb - o.x) / D.x;
        c.s[next] = mk(xb, o.y + c.f[next] * D.y);
        ++I; ++next;
    }
    while (I > If) {
        const T xb = sp.x * (I + dn);
        c.f[next] = (xb - o.x) / D.x;
        c.s[next] = mk(xb, o.y + c.f[next] * D.y);
        --I; ++next;
    }
    while (J < Jf) {
        const T yb = sp.y * (J + up);
        c.f[next] = (yb - o.y) / D.y;
        c.s[next] = mk(o.x + c.f[next] * D.x, yb);
        ++J; ++next;
    }
    while (J > Jf) {
        const T yb = sp.y * (J + dn);
        c.f[next] = (yb - o.y) / D.y;
        c.s[next] = mk(o.x + c.f[next] * D.x, yb);
        --J; ++next;
    }
    c.f[next++] = T(1);
    crossings_finish(c, next, o, e);
    return true;
}

// ---------------------------------------------------------------------------
// Interpolation weights, ScatterGatherImpl.hh.
// A stencil is described by its first relative node (I0, J0), the number of
// nodes per axis (2, 3 or 4) and the weights.  `ppos` is the particle position
// in units of the spacing, relative to the patch origin.
template <class T>
struct Stencil {
    int I0, J0;          // first node; cell indices fit 32 bits
    int n;
    T wx[4], wy[4];
};

// edge-centred component `lc` (0: Ex at (i+1/2, j); 1: Ey at (i, j+1/2))
template <class T>
PM_HD Stencil<T> edge_stencil(int order, int lc, vec2<T> ppos)
{
    Stencil<T> s;
    if (order == 0) {  // ScatterGatherImpl.hh:86-102
        T ax = lc == 0 ? T(0.5) : T(0), ay = lc == 1 ? T(0.5) : T(0);
        int lx = (int)floor_(ppos.x + ax), ly = (int)floor_(ppos.y + ay);
        if (lc == 0) lx -= 1; else ly -= 1;
        T gx = (T)lx, gy = (T)ly;
        if (lc == 0) gx += T(0.5); else gy += T(0.5);
        const T dx = ppos.x - gx, dy = ppos.y - gy;
        s.I0 = lx; s.J0 = ly; s.n = 2;
        s.wx[0] = 1 - dx; s.wx[1] = dx;
        s.wy[0] = 1 - dy; s.wy[1] = dy;
    } else if (order == 1) {  // ScatterGatherImpl.hh:218-238
        T ax = lc == 0 ? T(0) : T(0.5), ay = lc == 1 ? T(0) : T(0.5);
        int lx = (int)floor_(ppos.x + ax), ly = (int)floor_(ppos.y + ay);
        T gx = (T)lx, gy = (T)ly;
        if (lc == 0) gx += T(0.5); else gy += T(0.5);
        const T dx = ppos.x - gx, dy = ppos.y - gy;
        s.I0 = lx - 1; s.J0 = ly - 1; s.n = 3;
        s.wx[1] = T(0.75) - dx * dx;
        s.wx[2] = T(0.5) * (T(0.5) + dx) * (T(0.5) + dx);
        s.wx[0] = T(1.0) - (s.wx[1] + s.wx[2]);
        s.wy[1] = T(0.75) - dy * dy;
        s.wy[2] = T(0.5) * (T(0.5) + dy) * (T(0.5) + dy);
        s.wy[0] = T(1.0) - (s.wy[1] + s.wy[2]);
    } else {  // ScatterGatherImpl.hh:424-447
        T ax = lc == 0 ? T(0.5) : T(0), ay = lc == 1 ? T(0.5) : T(0);
        int lx = (int)floor_(ppos.x + ax), ly = (int)floor_(ppos.y + ay);
        if (lc == 0) lx -= 1; else ly -= 1;
        T gx = (T)lx, gy = (T)ly;
        if (lc == 0) gx += T(0.5); else gy += T(0.5);
        const T dx = ppos.x - gx, dy = ppos.y - gy;
        const T dx2 = dx * dx, dy2 = dy * dy;
        const T tt = T(2.0 / 3.0), os = T(1.0 / 6.0);
        s.I0 = lx - 1; s.J0 = ly - 1; s.n = 4;
        s.wx[1] = tt - dx2 * (1 - T(0.5) * dx);
        s.wx[2] = tt - T(0.5) * (1 - dx) * (1 - dx2);
        s.wx[3] = os * dx2 * dx;
        s.wx[0] = T(1.0) - (s.wx[1] + s.wx[2] + s.wx[3]);
        s.wy[1] = tt - dy2 * (1 - T(0.5) * dy);
        s.wy[2] = tt - T(0.5) * (1 - dy) * (1 - dy2);
        s.wy[3] = os * dy2 * dy;
        s.wy[0] = T(1.0) - (s.wy[1] + s.wy[2] + s.wy[3]);
    }
    return s;
}

// cell-centred (Hz at (i+1/2, j+1/2))
template <class T>
PM_HD Stencil<T> cell_stencil(int order, vec2<T> ppos)
{
    Stencil<T> s;
    if (order == 0) {  // ScatterGatherImpl.hh:118-132
        int lx = (int)floor_(ppos.x - T(0.5)), ly = (int)floor_(ppos.y - T(0.5));
        const T dx = ppos.x - ((T)lx + T(0.5)), dy = ppos.y - ((T)ly + T(0.5));
        s.I0 = lx; s.J0 = ly; s.n = 2;
        s.wx[0] = 1 - dx; s.wx[1] = dx;
        s.wy[0] = 1 - dy; s.wy[1] = dy;
    } else if (order == 1) {  // ScatterGatherImpl.hh:269-286
        int lx = (int)floor_(ppos.x), ly = (int)floor_(ppos.y);
        const T dx = ppos.x - ((T)lx + T(0.5)), dy = ppos.y - ((T)ly + T(0.5));
        s.I0 = lx - 1; s.J0 = ly - 1; s.n = 3;
        s.wx[1] = T(0.75) - dx * dx;
        s.wx[2] = T(0.5) * (T(0.5) + dx) * (T(0.5) + dx);
        s.wx[0] = T(1.0) - (s.wx[1] + s.wx[2]);
        s.wy[1] = T(0.75) - dy * dy;
        s.wy[2] = T(0.5) * (T(0.5) + dy) * (T(0.5) + dy);
        s.wy[0] = T(1.0) - (s.wy[1] + s.wy[2]);
    } else {  // ScatterGatherImpl.hh:486-505
        int lx = (int)floor_(ppos.x - T(0.5)), ly = (int)floor_(ppos.y - T(0.5));
        const T dx = ppos.x - ((T)lx + T(0.5)), dy = ppos.y - ((T)ly + T(0.5));
        const T dx2 = dx * dx, dy2 = dy * dy;
        const T tt = T(2.0 / 3.0), os = T(1.0 / 6.0);
        s.I0 = lx - 1; s.J0 = ly - 1; s.n = 4;
        s.wx[1] = tt - dx2 * (1 - T(0.5) * dx);
        s.wx[2] = tt - T(0.5) * (1 - dx) * (1 - dx2);
        s.wx[3] = os * dx2 * dx;
        s.wx[0] = T(1.0) - (s.wx[1] + s.wx[2] + s.wx[3]);
        s.wy[1] = tt - dy2 * (1 - T(0.5) * dy);
        s.wy[2] = tt - T(0.5) * (1 - dy) * (1 - dy2);
        s.wy[3] = os * dy2 * dy;
        s.wy[0] = T(1.0) - (s.wy[1] + s.wy[2] + s.wy[3]);
    }
    return s;
}

// Gather with the reference's nesting: sum over rows of wy[b] * (sum over
// columns of wx[a] * f), rows and columns accumulated left to right.
// s2bug reproduces ScatterGatherImpl.hh:467/525: the last term of the last row
// reads column I0+2 instead of I0+3.
template <class T, class F>
PM_HD T gather(const Stencil<T> &s, bool s2bug, F f)
{
    T acc = T(0);
    for (int b = 0; b < s.n; ++b) {
        const long long J = s.J0 + b;
        T row = s.wx[0] * f(s.I0, J);
        for (int a = 1; a < s.n; ++a) {
            long long I = s.I0 + a;
            if (s2bug && s.n == 4 && b == 3 && a == 3) I = s.I0 + 2;
            row = row + s.wx[a] * f(I, J);
        }
        const T t = s.wy[b] * row;
        acc = (b == 0) ? t : acc + t;
    }
    return acc;
}

// Same arithmetic with the stencil size fixed at compile time (N = order + 2),
// so the weights stay in registers.
template <int N, class T, class F>
PM_HD T gather_n(const Stencil<T> &s, bool s2bug, F f)
{
    T acc = T(0);
#pragma unroll
    for (int b = 0; b < N; ++b) {
        T row = s.wx[0] * f(0, b);
#pragma unroll
        for (int a = 1; a < N; ++a) {
            const int aa = (s2bug && N == 4 && b == 3 && a == 3) ? 2 : a;
            row = row + s.wx[a] * f(aa, b);
        }
        const T t = s.wy[b] * row;
        acc = (b == 0) ? t : acc + t;
    }
    return acc;
}

// Scatter weight of node (a, b) as the reference forms it: (wx*wy)*v for
// S1/S2, for CIC (wx*wy)*v with the same pairing.
template <class T>
PM_HD T scatter_w(const Stencil<T> &s, int a, int b, T v) { return s.wx[a] * s.wy[b] * v; }

// ---------------------------------------------------------------------------
// Current of one trajectory segment.
//
// S0 (ZerothOrderShapeFunction.cpp:239-251): 2 nodes.
// S1 (FirstOrderShapeFunction.cpp:540-683): 6 nodes, Jx 2x3, Jy 3x2.
// S2 (SecondOrderShapeFunction.cpp:583-826): 12 nodes, Jx 3x4, Jy 4x3.
// Node order: x fastest; the first node is (I + di0, J + dj0).
template <class T>
struct SegCurrent {
    int I0, J0;   // first node (absolute)
    int nx, ny;   // nodes per axis
    T w[12];
};

template <class T>
PM_HD void seg_s0(SegCurrent<T> &o, int comp, vec2<T> xi, vec2<T> xf, T q, vec2<T> odt)
{
    const vec2<T> D = mk(q * (xf.x - xi.x) * odt.x, q * (xf.y - xi.y) * odt.y);
    const vec2<T> c = mk((xi.x + xf.x) * T(0.5), (xi.y + xf.y) * T(0.5));
    const int I = (int)floor_(c.x), J = (int)floor_(c.y);
    o.I0 = I; o.J0 = J;
    if (comp == 0) {
        o.nx = 1; o.ny = 2;
        o.w[0] = D.x * ((T)(J + 1) - c.y);
        o.w[1] = D.x * (c.y - (T)J);
    } else {
        o.nx = 2; o.ny = 1;
        o.w[0] = D.y * ((T)(I + 1) - c.x);
        o.w[1] = D.y * (c.x - (T)I);
    }
}

template <class T>
PM_HD void seg_s1(SegCurrent<T> &o, int comp, T tau, vec2<T> a, vec2<T> b, vec2<T> qodt, vec2<T> d)
{
    const T h = T(0.5), o3 = T(1.0 / 3.0), o4 = h * h;
    const T tp2 = tau * tau * h;
    const T tp3 = tp2 * tau * 2 * o3;
    if (comp == 0) {
        const T cd = qodt.y * d.x;
        const T A[2] = {a.x, T(1.0) - a.x};
        const T B[2] = {-d.x * tau, d.x * tau};
        const T Ay[3] = {h * (h + b.y) * (h + b.y) * tau, (T(0.75) - b.y * b.y) * tau, h * (h - b.y) * (h - b.y) * tau};
        const T By[3] = {-(h + b.y) * d.y * tp2, 2 * b.y * d.y * tp2, (h - b.y) * d.y * tp2};
        const T Cy[3] = {h * d.y * d.y * tp3, -d.y * d.y * tp3, h * d.y * d.y * tp3};
        o.nx = 2; o.ny = 3;
        for (int jj = 0; jj < 3; ++jj)
            for (int ii = 0; ii < 2; ++ii)
                o.w[jj * 2 + ii] = cd * ((A[ii] + B[ii]) * (Ay[jj] + By[jj] + Cy[jj]) -
                                         B[ii] * (h * Ay[jj] + o3 * By[jj] + o4 * Cy[jj]));
    } else {
        const T cd = qodt.x * d.y;
        const T Ax[3] = {h * (h + b.x) * (h + b.x) * tau, (T(0.75) - b.x * b.x) * tau, h * (h - b.x) * (h - b.x) * tau};
        const T Bx[3] = {-(h + b.x) * d.x * tp2, 2 * b.x * d.x * tp2, (h - b.x) * d.x * tp2};
        const T Cx[3] = {h * d.x * d.x * tp3, -d.x * d.x * tp3, h * d.x * d.x * tp3};
        const T A[2] = {a.y, T(1.0) - a.y};
        const T B[2] = {-d.y * tau, d.y * tau};
        o.nx = 3; o.ny = 2;
        for (int jj = 0; jj < 2; ++jj)
            for (int ii = 0; ii < 3; ++ii)
                o.w[jj * 3 + ii] = cd * ((A[jj] + B[jj]) * (Ax[ii] + Bx[ii] + Cx[ii]) -
                                         B[jj] * (h * Ax[ii] + o3 * Bx[ii] + o4 * Cx[ii]));
    }
}

template <class T>
PM_HD void cubic_terms(T bb, T dl, T tau, T tp2, T tp3, T tp4, T A3[4], T B3[4], T C3[4], T D3[4])
{
    const T h = T(0.5), o6 = h * T(1.0 / 3.0);
    const T dp2 = dl * dl, dp3 = dp2 * dl;
    A3[0] = (1 + bb) * (1 + bb) * (1 + bb) * o6 * tau;
    A3[1] = (4 - 3 * bb * bb * (2 + bb)) * o6 * tau;
    A3[2] = (4 - 3 * (bb + 1) * (bb + 1) * (1 - bb)) * o6 * tau;
    A3[3] = -bb * bb * bb * o6 * tau;
    B3[0] = -(1 + bb) * (1 + bb) * dl * h * tp2;
    B3[1] = bb * dl * (3 * bb + 4) * h * tp2;
    B3[2] = (1 - 2 * bb - 3 * bb * bb) * dl * h * tp2;
    B3[3] = bb * bb * dl * h * tp2;
    C3[0] = (1 + bb) * dp2 * h * tp3;
    C3[1] = -dp2 * (2 + 3 * bb) * h * tp3;
    C3[2] = (1 + 3 * bb) * dp2 * h * tp3;
    C3[3] = -bb * dp2 * h * tp3;
    D3[0] = -dp3 * o6 * tp4;
    D3[1] = dp3 * h * tp4;
    D3[2] = -dp3 * h * tp4;
    D3[3] = dp3 * o6 * tp4;
}

template <class T>
PM_HD void seg_s2(SegCurrent<T> &o, int comp, T tau, vec2<T> a, vec2<T> b, vec2<T> qodt, vec2<T> d)
{
    const T h = T(0.5), o3 = T(1.0 / 3.0), o4 = T(0.25), o5 = T(0.2);
    const T tp2 = tau * tau / 2;
    const T tp3 = tp2 * tau * 2 / 3;
    const T tp4 = tp2 * tp2;
    T A3[4], B3[4], C3[4], D3[4];
    if (comp == 0) {
        const T cd = qodt.y * d.x;
        const T dp2 = d.x * d.x;
        const T A2[3] = {h * (h + a.x) * (h + a.x), T(0.75) - a.x * a.x, h * (T(0.5) - a.x) * (T(0.5) - a.x)};
        const T B2[3] = {-(h + a.x) * d.x * tau, 2 * a.x * d.x * tau, (h - a.x) * d.x * tau};
        const T C2[3] = {h * dp2 * tp2, -dp2 * tp2, h * dp2 * tp2};
        cubic_terms(b.y, d.y, tau, tp2, tp3, tp4, A3, B3, C3, D3);
        o.nx = 3; o.ny = 4;
        for (int jj = 0; jj < 4; ++jj)
            for (int ii = 0; ii < 3; ++ii)
                o.w[jj * 3 + ii] = cd * ((A2[ii] + B2[ii] + C2[ii]) * (D3[jj] + C3[jj] + B3[jj] + A3[jj]) +
                                         (C2[ii] * (D3[jj] - A3[jj]) - B2[ii] * B3[jj]) * o3 +
                                         (C2[ii] * C3[jj] - B2[ii] * D3[jj]) * o5 -
                                         B2[ii] * C3[jj] * o4 - B2[ii] * A3[jj] * h);
    } else {
        const T cd = qodt.x * d.y;
        const T dp2 = d.y * d.y;
        cubic_terms(b.x, d.x, tau, tp2, tp3, tp4, A3, B3, C3, D3);
        const T A2[3] = {(h + a.y) * (h + a.y) * h, T(0.75) - a.y * a.y, (h - a.y) * (h - a.y) * h};
        const T B2[3] = {-(h + a.y) * d.y * tau, 2 * a.y * d.y * tau, (h - a.y) * d.y * tau};
        const T C2[3] = {dp2 * h * tp2, -dp2 * tp2, dp2 * h * tp2};
        o.nx = 4; o.ny = 3;
        for (int jj = 0; jj < 3; ++jj)
            for (int ii = 0; ii < 4; ++ii)
                o.w[jj * 4 + ii] = cd * ((A2[jj] + B2[jj] + C2[jj]) * (D3[ii] + C3[ii] + B3[ii] + A3[ii]) +
                                         (C2[jj] * (D3[ii] - A3[ii]) - B2[jj] * B3[ii]) * o3 +
                                         (C2[jj] * C3[ii] - B2[jj] * D3[ii]) * o5 -
                                         B2[jj] * C3[ii] * o4 - B2[jj] * A3[ii] * h);
    }
}

// Segment k of a crossing list in mesh units (xi, xf), dispatch by order.
// For S1 the node anchor is I = floor(c + 1/2) and the Jx stencil spans
// I-1..I (x), J-1..J+1 (y); Jy spans I-1..I+1, J-1..J.  For S2 the anchor is
// I = floor(c); Jx spans I-1..I+1, J-1..J+2; Jy spans I-1..I+2, J-1..J+1.
template <class T>
PM_HD void segment_current(SegCurrent<T> &o, int order, int comp, vec2<T> xi, vec2<T> xf, vec2<T> seg,
                           T covered, T q, vec2<T> odt, vec2<T> delta)
{
    if (order == 0) {
        seg_s0(o, comp, xi, xf, q, odt);
        return;
    }
    const T tau = sqrt_(seg.x * seg.x + seg.y * seg.y) / covered;
    const vec2<T> c = order == 1 ? mk((xi.x + xf.x) * T(0.5), (xi.y + xf.y) * T(0.5))
                                 : mk(T(0.5) * (xi.x + xf.x), T(0.5) * (xi.y + xf.y));
    const int I = order == 1 ? (int)floor_(c.x + T(0.5)) : (int)floor_(c.x);
    const int J = order == 1 ? (int)floor_(c.y + T(0.5)) : (int)floor_(c.y);
    const vec2<T> a = mk((T)I + T(0.5) - xi.x, (T)J + T(0.5) - xi.y);
    const vec2<T> b = mk((T)I - xi.x, (T)J - xi.y);
    const vec2<T> qodt = mk(q * odt.x, q * odt.y);
    if (order == 1) {
        seg_s1(o, comp, tau, a, b, qodt, delta);
        o.I0 = I - 1; o.J0 = J - 1;
    } else {
        seg_s2(o, comp, tau, a, b, qodt, delta);
        o.I0 = I - 1; o.J0 = J - 1;
    }
}

// ---------------------------------------------------------------------------
// Vertex charge weights of getChargeDensityImpl: ORDER 0 CIC
// (ZerothOrderShapeFunction.cpp:344-357), 1 TSC about the nearest vertex
// (FirstOrderShapeFunction.cpp:389-413), 2 cubic B-spline
// (SecondOrderShapeFunction.cpp:393-427).  ne = (R - origin) / h.  The node
// (I0 + a, J0 + b) receives charge_w<ORDER>(q, wx[a], wy[b]); each order
// keeps its reference's association of the product.
template <int ORDER, class T>
PM_HD void charge_weights(vec2<T> ne, T q, T (&wx)[ORDER + 2], T (&wy)[ORDER + 2], int &I0, int &J0)
{
    if constexpr (ORDER == 0) {
        const int I = (int)floor_(ne.x), J = (int)floor_(ne.y);
        const T tx = ne.x - (T)I, ty = ne.y - (T)J;
        I0 = I; J0 = J;
        wx[0] = T(1.0) - tx; wx[1] = tx;
        wy[0] = T(1.0) - ty; wy[1] = ty;
    } else if constexpr (ORDER == 1) {
        const int I = (int)floor_(ne.x + T(0.5)), J = (int)floor_(ne.y + T(0.5));
        const T a = ne.x - (T)I, b = ne.y - (T)J;
        const T sxm = q * (T(0.5) - a) * (T(0.5) - a) * T(0.5);
        const T sxp = q * (T(0.5) + a) * (T(0.5) + a) * T(0.5);
        wx[0] = sxm; wx[1] = q - sxm - sxp; wx[2] = sxp;
        const T sym = (T(0.5) - b) * (T(0.5) - b) * T(0.5);
        const T syp = (T(0.5) + b) * (T(0.5) + b) * T(0.5);
        wy[0] = sym; wy[1] = T(1.0) - sym - syp; wy[2] = syp;
        I0 = I - 1; J0 = J - 1;
    } else {
        const T o6 = T(1.0 / 6.0);
        const int I = (int)floor_(ne.x), J = (int)floor_(ne.y);
        const T a = ne.x - (T)I, b = ne.y - (T)J;
        wx[0] = o6 * ((T(1.0) - a) * (T(1.0) - a) * (T(1.0) - a));
        wx[1] = o6 * (T(4.0) - T(3.0) * a * a * (T(2.0) - a));
        wx[2] = o6 * (T(4.0) - T(3.0) * (T(1.0) - a) * (T(1.0) - a) * (T(1.0) + a));
        wx[3] = o6 * (a * a * a);
        wy[0] = o6 * ((T(1.0) - b) * (T(1.0) - b) * (T(1.0) - b));
        wy[1] = o6 * (T(4.0) - T(3.0) * b * b * (T(2.0) - b));
        wy[2] = o6 * (T(4.0) - T(3.0) * (T(1.0) - b) * (T(1.0) - b) * (T(1.0) + b));
        wy[3] = o6 * (b * b * b);
        I0 = I - 1; J0 = J - 1;
    }
}
template <int ORDER, class T>
PM_HD T charge_w(T q, T wx, T wy)
{
    if constexpr (ORDER == 0) return wx * wy * q;
    else if constexpr (ORDER == 1) return wx * wy;
    else return q * wx * wy;
}

// ---------------------------------------------------------------------------
// Kick and rotation of PartBunch.cpp:366-383 (not Boris: the rotation matrix
// is ((1+a^2) s + a s_perp)/(1+a^2), and B = mu0 Hz + Bext/2).
template <class T>
PM_HD vec2<T> kick_rotate(vec2<T> P, T ex, T ey, T hf, T bext, T fact1, T fact2)
{
    const T B = T(kMu0) * hf + T(0.5) * bext;
    const vec2<T> kick = mk(fact1 * ex, fact1 * ey);
    vec2<T> um = mk(P.x + kick.x, P.y + kick.y);
    const T a = fact2 * B * (T)q_rsqrt((float)(T(1.0) + (um.x * um.x + um.y * um.y)));
    const vec2<T> s = mk(um.x + a * um.y, um.y + a * -um.x);
    const T tmp = T(1.0) + a * a;
    um.x = ((T(1.0) + a * a) * s.x + a * s.y) / tmp;
    um.y = ((T(1.0) + a * a) * s.y - a * s.x) / tmp;
    return mk(um.x + kick.x, um.y + kick.y);
}

// 1/gamma as the reference forms it: Q_rsqrt((float)(1 + P.P)).
template <class T>
PM_HD T inv_gamma(vec2<T> P) { return (T)q_rsqrt((float)(T(1.0) + (P.x * P.x + P.y * P.y))); }

}  // namespace pm
