/*
 * csrst_oracle.cpp -- CPU FP64 oracle for the CSRST particle-in-cell time step.
 *
 * TEST INFRASTRUCTURE ONLY: a plain C++ restatement of the reference algorithm
 * (kreischtauf/CSRST src/), used to check the HIP product path.  It
 * is never linked into, loaded by or called from the product library.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * written in the reference's evaluation order (IPPL expression templates
 * evaluate scalar sub-expressions first, then element-wise left to right) and
 * the file must be compiled with -ffp-contract=off.
 */
#include "csrst_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace {

/* Physics.cpp:41-51 */
const double C_LIGHT = 299792458.0;
const double MU_0 = 1.256637061e-06;
const double EPS_0 = 8.854187817e-12;
const double M_E_GEV = 0.51099906e-03;
const double PI = 3.14159265358979323846;
/* defs.hh:30-33 */
const double SPACE_EPS = 1e-10;
const double MINAREA = 1e-20;

struct V2 {
    double x, y;
};
inline V2 v2(double a, double b) { V2 r; r.x = a; r.y = b; return r; }
inline double dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }

/* utils.hh:491-504: float fast inverse square root with Kadlec constants. */
float q_rsqrt(float x)
{
    float xhalf = 0.5f * x;
    float xqhalf = 0.5010936f * x;
    uint32_t i;
    std::memcpy(&i, &x, sizeof(i));
    i = 0x5f3759df - (i >> 1);
    std::memcpy(&x, &i, sizeof(i));
    x = x * (1.5f - xhalf * x * x);
    x = x * (1.5010936f - xqhalf * x * x);
    return x;
}

/* ---------------------------------------------------------------------- */
/* Crossings, utils.hh:364-489.  The sort is libstdc++'s std::sort, which for
 * ranges of at most 16 elements is a stable insertion sort; std::unique keeps
 * the first element of each run, comparing against the last kept element,
 * with the predicate |fa - fb| < 1e-6 (utils.hh:80-83).                     */
struct Xing {
    double f;
    V2 s;
};

int finish_crossings(Xing *x, int next, V2 oldR, V2 newR)
{
    /* insertion sort of x[1 .. next-2] by fraction (stable) */
    for (int a = 2; a < next - 1; ++a) {
        Xing v = x[a];
        int b = a;
        while (b > 1 && v.f < x[b - 1].f) {
            x[b] = x[b - 1];
            --b;
        }
        x[b] = v;
    }
    /* unique */
    int dest = 0;
    for (int k = 1; k < next; ++k) {
        if (!(std::fabs(x[dest].f - x[k].f) < 1e-6)) {
            ++dest;
            x[dest] = x[k];
        }
    }
    int num = dest + 1;
    x[0].s = oldR;
    x[num - 1].s = newR;
    return num;
}

int cell_crossings(Xing *x, V2 oldR, V2 newR, V2 sp, V2 origin)
{
    oldR.x -= origin.x; oldR.y -= origin.y;
    newR.x -= origin.x; newR.y -= origin.y;
    const int Ii = (int)std::floor(oldR.x / sp.x);
    const int Ji = (int)std::floor(oldR.y / sp.y);
    const int If = (int)std::floor(newR.x / sp.x);
    const int Jf = (int)std::floor(newR.y / sp.y);
    int I = Ii, J = Ji;
    const V2 Dx = v2(newR.x - oldR.x, newR.y - oldR.y);
    int next = 0;
    x[next++].f = 0.0;
    while (I < If) {
        x[next].f = (sp.x * (I + 1) - oldR.x) / Dx.x;
        x[next].s = v2(sp.x * (I + 1), oldR.y + x[next].f * Dx.y);
        ++I; ++next;
    }
    while (I > If) {
        x[next].f = (sp.x * I - oldR.x) / Dx.x;
        x[next].s = v2(sp.x * I, oldR.y + x[next].f * Dx.y);
        --I; ++next;
    }
    while (J < Jf) {
        x[next].f = (sp.y * (J + 1) - oldR.y) / Dx.y;
        x[next].s = v2(oldR.x + x[next].f * Dx.x, sp.y * (J + 1));
        ++J; ++next;
    }
    while (J > Jf) {
        x[next].f = (sp.y * J - oldR.y) / Dx.y;
        x[next].s = v2(oldR.x + x[next].f * Dx.x, sp.y * J);
        --J; ++next;
    }
    x[next++].f = 1.0;
    return finish_crossings(x, next, oldR, newR);
}

int dual_crossings(Xing *x, V2 oldR, V2 newR, V2 sp, V2 origin)
{
    oldR.x -= origin.x; oldR.y -= origin.y;
    newR.x -= origin.x; newR.y -= origin.y;
    int Ii = (int)std::floor(oldR.x / sp.x - 0.5);
    int Ji = (int)std::floor(oldR.y / sp.y - 0.5);
    int If = (int)std::floor(newR.x / sp.x - 0.5);
    int Jf = (int)std::floor(newR.y / sp.y - 0.5);
    int I = Ii, J = Ji;
    V2 Dx = v2(newR.x - oldR.x, newR.y - oldR.y);
    int next = 0;
    x[next++].f = 0.0;
    while (I < If) {
        x[next].f = (sp.x * (I + 1.5) - oldR.x) / Dx.x;
        x[next].s = v2(sp.x * (I + 1.5), oldR.y + x[next].f * Dx.y);
        ++I; ++next;
    }
    while (I > If) {
        x[next].f = (sp.x * (I + 0.5) - oldR.x) / Dx.x;
        x[next].s = v2(sp.x * (I + 0.5), oldR.y + x[next].f * Dx.y);
        --I; ++next;
    }
    while (J < Jf) {
        x[next].f = (sp.y * (J + 1.5) - oldR.y) / Dx.y;
        x[next].s = v2(oldR.x + x[next].f * Dx.x, sp.y * (J + 1.5));
        ++J; ++next;
    }
    while (J > Jf) {
        x[next].f = (sp.y * (J + 0.5) - oldR.y) / Dx.y;
        x[next].s = v2(oldR.x + x[next].f * Dx.x, sp.y * (J + 0.5));
        --J; ++next;
    }
    x[next++].f = 1.0;
    return finish_crossings(x, next, oldR, newR);
}

/* ---------------------------------------------------------------------- */
/* ScatterGatherImpl.hh.  A patch is addressed with relative indices (I, J);
 * `Fetch` returns the value at relative index.                            */
template <class Fetch>
double gather_edge(int order, int comp, int s2bug, Fetch fp, V2 origin, V2 dx, V2 R)
{
    const V2 od = v2(1.0 / dx.x, 1.0 / dx.y);
    const V2 ppos = v2((R.x - origin.x) * od.x, (R.y - origin.y) * od.y);
    const int lc = comp % 2;
    if (order == 0) { /* ScatterGatherImpl.hh:70-104 */
        double add[2] = {0.0, 0.0};
        add[lc] = 0.5;
        long long lgp[2] = {(long long)std::floor(ppos.x + add[0]), (long long)std::floor(ppos.y + add[1])};
        lgp[lc] -= 1;
        double gpos[2] = {(double)lgp[0], (double)lgp[1]};
        gpos[lc] += 0.5;
        const V2 d = v2(ppos.x - gpos[0], ppos.y - gpos[1]);
        const long long I = lgp[0], J = lgp[1];
        return ((1 - d.y) * ((1 - d.x) * fp(I, J) + d.x * fp(I + 1, J)) +
                d.y * ((1 - d.x) * fp(I, J + 1) + d.x * fp(I + 1, J + 1)));
    } else if (order == 1) { /* ScatterGatherImpl.hh:202-253 */
        double add[2] = {0.5, 0.5};
        add[lc] = 0.0;
        long long lgp[2] = {(long long)std::floor(ppos.x + add[0]), (long long)std::floor(ppos.y + add[1])};
        double gpos[2] = {(double)lgp[0], (double)lgp[1]};
        gpos[lc] += 0.5;
        const V2 d = v2(ppos.x - gpos[0], ppos.y - gpos[1]);
        const double bx = 0.75 - d.x * d.x;
        const double cx = 0.5 * (0.5 + d.x) * (0.5 + d.x);
        const double ax = 1.0 - (bx + cx);
        const double by = 0.75 - d.y * d.y;
        const double cy = 0.5 * (0.5 + d.y) * (0.5 + d.y);
        const double ay = 1.0 - (by + cy);
        const long long I = lgp[0], J = lgp[1];
        return (ay * (ax * fp(I - 1, J - 1) + bx * fp(I, J - 1) + cx * fp(I + 1, J - 1)) +
                by * (ax * fp(I - 1, J) + bx * fp(I, J) + cx * fp(I + 1, J)) +
                cy * (ax * fp(I - 1, J + 1) + bx * fp(I, J + 1) + cx * fp(I + 1, J + 1)));
    } else { /* ScatterGatherImpl.hh:407-469 */
        const double twothirds = 2.0 / 3.0, onesixth = 1.0 / 6.0;
        double add[2] = {0.0, 0.0};
        add[lc] = 0.5;
        long long lgp[2] = {(long long)std::floor(ppos.x + add[0]), (long long)std::floor(ppos.y + add[1])};
        lgp[lc] -= 1;
        double gpos[2] = {(double)lgp[0], (double)lgp[1]};
        gpos[lc] += 0.5;
        const V2 d = v2(ppos.x - gpos[0], ppos.y - gpos[1]);
        const V2 d2 = v2(d.x * d.x, d.y * d.y);
        const double bx = twothirds - d2.x * (1 - 0.5 * d.x);
        const double cx = twothirds - 0.5 * (1 - d.x) * (1 - d2.x);
        const double ex = onesixth * d2.x * d.x;
        const double ax = 1.0 - (bx + cx + ex);
        const double by = twothirds - d2.y * (1 - 0.5 * d.y);
        const double cy = twothirds - 0.5 * (1 - d.y) * (1 - d2.y);
        const double ey = onesixth * d2.y * d.y;
        const double ay = 1.0 - (by + cy + ey);
        const long long I = lgp[0], J = lgp[1];
        const long long last = s2bug ? I + 1 : I + 2;
        return (ay * (ax * fp(I - 1, J - 1) + bx * fp(I, J - 1) + cx * fp(I + 1, J - 1) + ex * fp(I + 2, J - 1)) +
                by * (ax * fp(I - 1, J) + bx * fp(I, J) + cx * fp(I + 1, J) + ex * fp(I + 2, J)) +
                cy * (ax * fp(I - 1, J + 1) + bx * fp(I, J + 1) + cx * fp(I + 1, J + 1) + ex * fp(I + 2, J + 1)) +
                ey * (ax * fp(I - 1, J + 2) + bx * fp(I, J + 2) + cx * fp(I + 1, J + 2) + ex * fp(last, J + 2)));
    }
}

template <class Fetch>
double gather_cell(int order, int s2bug, Fetch fp, V2 origin, V2 dx, V2 R)
{
    const V2 od = v2(1.0 / dx.x, 1.0 / dx.y);
    const V2 ppos = v2((R.x - origin.x) * od.x, (R.y - origin.y) * od.y);
    if (order == 0) { /* ScatterGatherImpl.hh:106-134 */
        long long lgp[2] = {(long long)std::floor(ppos.x - 0.5), (long long)std::floor(ppos.y - 0.5)};
        const V2 d = v2(ppos.x - (lgp[0] + 0.5), ppos.y - (lgp[1] + 0.5));
        const long long I = lgp[0], J = lgp[1];
        return ((1 - d.y) * ((1 - d.x) * fp(I, J) + d.x * fp(I + 1, J)) +
                d.y * ((1 - d.x) * fp(I, J + 1) + d.x * fp(I + 1, J + 1)));
    } else if (order == 1) { /* ScatterGatherImpl.hh:255-301 */
        long long lgp[2] = {(long long)std::floor(ppos.x), (long long)std::floor(ppos.y)};
        const V2 d = v2(ppos.x - (lgp[0] + 0.5), ppos.y - (lgp[1] + 0.5));
        const double bx = 0.75 - d.x * d.x;
        const double cx = 0.5 * (0.5 + d.x) * (0.5 + d.x);
        const double ax = 1.0 - (bx + cx);
        const double by = 0.75 - d.y * d.y;
        const double cy = 0.5 * (0.5 + d.y) * (0.5 + d.y);
        const double ay = 1.0 - (by + cy);
        const long long I = lgp[0], J = lgp[1];
        return (ay * (ax * fp(I - 1, J - 1) + bx * fp(I, J - 1) + cx * fp(I + 1, J - 1)) +
                by * (ax * fp(I - 1, J) + bx * fp(I, J) + cx * fp(I + 1, J)) +
                cy * (ax * fp(I - 1, J + 1) + bx * fp(I, J + 1) + cx * fp(I + 1, J + 1)));
    } else { /* ScatterGatherImpl.hh:471-527 */
        const double twothirds = 2.0 / 3.0, onesixth = 1.0 / 6.0;
        long long lgp[2] = {(long long)std::floor(ppos.x - 0.5), (long long)std::floor(ppos.y - 0.5)};
        const V2 d = v2(ppos.x - (lgp[0] + 0.5), ppos.y - (lgp[1] + 0.5));
        const V2 d2 = v2(d.x * d.x, d.y * d.y);
        const double bx = twothirds - d2.x * (1 - 0.5 * d.x);
        const double cx = twothirds - 0.5 * (1 - d.x) * (1 - d2.x);
        const double ex = onesixth * d2.x * d.x;
        const double ax = 1.0 - (bx + cx + ex);
        const double by = twothirds - d2.y * (1 - 0.5 * d.y);
        const double cy = twothirds - 0.5 * (1 - d.y) * (1 - d2.y);
        const double ey = onesixth * d2.y * d.y;
        const double ay = 1.0 - (by + cy + ey);
        const long long I = lgp[0], J = lgp[1];
        const long long last = s2bug ? I + 1 : I + 2;
        return (ay * (ax * fp(I - 1, J - 1) + bx * fp(I, J - 1) + cx * fp(I + 1, J - 1) + ex * fp(I + 2, J - 1)) +
                by * (ax * fp(I - 1, J) + bx * fp(I, J) + cx * fp(I + 1, J) + ex * fp(I + 2, J)) +
                cy * (ax * fp(I - 1, J + 1) + bx * fp(I, J + 1) + cx * fp(I + 1, J + 1) + ex * fp(I + 2, J + 1)) +
                ey * (ax * fp(I - 1, J + 2) + bx * fp(I, J + 2) + cx * fp(I + 1, J + 2) + ex * fp(last, J + 2)));
    }
}

/* scatter: ScatterGatherImpl.hh:136-194, 303-392, 529-641 */
template <class Acc>
void scatter_edge(int order, int comp, Acc acc, V2 origin, V2 dx, V2 R, double v)
{
    const V2 od = v2(1.0 / dx.x, 1.0 / dx.y);
    const V2 ppos = v2((R.x - origin.x) * od.x, (R.y - origin.y) * od.y);
    const int lc = comp % 2;
    if (order == 0) {
        double add[2] = {0.0, 0.0};
        add[lc] = 0.5;
        long long lgp[2] = {(long long)std::floor(ppos.x + add[0]), (long long)std::floor(ppos.y + add[1])};
        lgp[lc] -= 1;
        double gpos[2] = {(double)lgp[0], (double)lgp[1]};
        gpos[lc] += 0.5;
        const V2 d = v2(ppos.x - gpos[0], ppos.y - gpos[1]);
        const long long I = lgp[0], J = lgp[1];
        acc(I, J, (1 - d.x) * (1 - d.y) * v);
        acc(I + 1, J, d.x * (1 - d.y) * v);
        acc(I, J + 1, (1 - d.x) * d.y * v);
        acc(I + 1, J + 1, d.x * d.y * v);
    } else if (order == 1) {
        double add[2] = {0.5, 0.5};
        add[lc] = 0.0;
        long long lgp[2] = {(long long)std::floor(ppos.x + add[0]), (long long)std::floor(ppos.y + add[1])};
        double gpos[2] = {(double)lgp[0], (double)lgp[1]};
        gpos[lc] += 0.5;
        const V2 d = v2(ppos.x - gpos[0], ppos.y - gpos[1]);
        const double bx = 0.75 - d.x * d.x;
        const double cx = 0.5 * (0.5 + d.x) * (0.5 + d.x);
        const double ax = 1.0 - (bx + cx);
        const double by = 0.75 - d.y * d.y;
        const double cy = 0.5 * (0.5 + d.y) * (0.5 + d.y);
        const double ay = 1.0 - (by + cy);
        const long long I = lgp[0], J = lgp[1];
        const double wx[3] = {ax, bx, cx}, wy[3] = {ay, by, cy};
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) acc(I - 1 + a, J - 1 + b, wx[a] * wy[b] * v);
    } else {
        const double twothirds = 2.0 / 3.0, onesixth = 1.0 / 6.0;
        double add[2] = {0.0, 0.0};
        add[lc] = 0.5;
        long long lgp[2] = {(long long)std::floor(ppos.x + add[0]), (long long)std::floor(ppos.y + add[1])};
        lgp[lc] -= 1;
        double gpos[2] = {(double)lgp[0], (double)lgp[1]};
        gpos[lc] += 0.5;
        const V2 d = v2(ppos.x - gpos[0], ppos.y - gpos[1]);
        const V2 d2 = v2(d.x * d.x, d.y * d.y);
        const double bx = twothirds - d2.x * (1 - 0.5 * d.x);
        const double cx = twothirds - 0.5 * (1 - d.x) * (1 - d2.x);
        const double ex = onesixth * d2.x * d.x;
        const double ax = 1.0 - (bx + cx + ex);
        const double by = twothirds - d2.y * (1 - 0.5 * d.y);
        const double cy = twothirds - 0.5 * (1 - d.y) * (1 - d2.y);
        const double ey = onesixth * d2.y * d.y;
        const double ay = 1.0 - (by + cy + ey);
        const long long I = lgp[0], J = lgp[1];
        const double wx[4] = {ax, bx, cx, ex}, wy[4] = {ay, by, cy, ey};
        for (int b = 0; b < 4; ++b)
            for (int a = 0; a < 4; ++a) acc(I - 1 + a, J - 1 + b, wx[a] * wy[b] * v);
    }
}

template <class Acc>
void scatter_cell(int order, Acc acc, V2 origin, V2 dx, V2 R, double v)
{
    const V2 od = v2(1.0 / dx.x, 1.0 / dx.y);
    const V2 ppos = v2((R.x - origin.x) * od.x, (R.y - origin.y) * od.y);
    if (order == 0) {
        long long lgp[2] = {(long long)std::floor(ppos.x - 0.5), (long long)std::floor(ppos.y - 0.5)};
        const V2 d = v2(ppos.x - (lgp[0] + 0.5), ppos.y - (lgp[1] + 0.5));
        const long long I = lgp[0], J = lgp[1];
        acc(I, J, (1 - d.x) * (1 - d.y) * v);
        acc(I + 1, J, d.x * (1 - d.y) * v);
        acc(I, J + 1, (1 - d.x) * d.y * v);
        acc(I + 1, J + 1, d.x * d.y * v);
    } else if (order == 1) { /* the cell scatter uses the explicit ax form, :372 */
        long long lgp[2] = {(long long)std::floor(ppos.x), (long long)std::floor(ppos.y)};
        const V2 d = v2(ppos.x - (lgp[0] + 0.5), ppos.y - (lgp[1] + 0.5));
        const double ax = 0.5 * (0.5 - d.x) * (0.5 - d.x);
        const double bx = 0.75 - d.x * d.x;
        const double cx = 0.5 * (0.5 + d.x) * (0.5 + d.x);
        const double ay = 0.5 * (0.5 - d.y) * (0.5 - d.y);
        const double by = 0.75 - d.y * d.y;
        const double cy = 0.5 * (0.5 + d.y) * (0.5 + d.y);
        const long long I = lgp[0], J = lgp[1];
        const double wx[3] = {ax, bx, cx}, wy[3] = {ay, by, cy};
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) acc(I - 1 + a, J - 1 + b, wx[a] * wy[b] * v);
    } else {
        const double twothirds = 2.0 / 3.0, onesixth = 1.0 / 6.0;
        long long lgp[2] = {(long long)std::floor(ppos.x - 0.5), (long long)std::floor(ppos.y - 0.5)};
        const V2 d = v2(ppos.x - (lgp[0] + 0.5), ppos.y - (lgp[1] + 0.5));
        const V2 d2 = v2(d.x * d.x, d.y * d.y);
        const double bx = twothirds - d2.x * (1 - 0.5 * d.x);
        const double cx = twothirds - 0.5 * (1 - d.x) * (1 - d2.x);
        const double ex = onesixth * d2.x * d.x;
        const double ax = 1.0 - (bx + cx + ex);
        const double by = twothirds - d2.y * (1 - 0.5 * d.y);
        const double cy = twothirds - 0.5 * (1 - d.y) * (1 - d2.y);
        const double ey = onesixth * d2.y * d.y;
        const double ay = 1.0 - (by + cy + ey);
        const long long I = lgp[0], J = lgp[1];
        const double wx[4] = {ax, bx, cx, ex}, wy[4] = {ay, by, cy, ey};
        for (int b = 0; b < 4; ++b)
            for (int a = 0; a < 4; ++a) acc(I - 1 + a, J - 1 + b, wx[a] * wy[b] * v);
    }
}

/* ---------------------------------------------------------------------- */
/* Current deposition.  `acc(I, J, v)` adds v to the absolute node (I, J).  */

/* ZerothOrderShapeFunction.cpp:202-294 */
template <class Acc>
void deposit_s0(int comp, Acc acc, V2 origin, V2 dx, double dt, V2 oR, V2 nR, double q)
{
    const V2 od = v2(1.0 / dx.x, 1.0 / dx.y);
    const V2 odt = v2(1.0 / (dx.y * dt), 1.0 / (dx.x * dt));
    Xing xs[10];
    const int num = cell_crossings(xs, oR, nR, dx, origin);
    for (int j = 0; j < num - 1; ++j) {
        const V2 xi = v2(xs[j].s.x * od.x, xs[j].s.y * od.y);
        const V2 xf = v2(xs[j + 1].s.x * od.x, xs[j + 1].s.y * od.y);
        const V2 D = v2(q * (xf.x - xi.x) * odt.x, q * (xf.y - xi.y) * odt.y);
        const V2 c = v2((xi.x + xf.x) * 0.5, (xi.y + xf.y) * 0.5);
        const int I = (int)std::floor(c.x);
        const int J = (int)std::floor(c.y);
        if (comp == 0) {
            acc(I, J, D.x * (J + 1 - c.y));
            acc(I, J + 1, D.x * (c.y - J));
        } else {
            acc(I, J, D.y * (I + 1 - c.x));
            acc(I + 1, J, D.y * (c.x - I));
        }
    }
}

/* FirstOrderShapeFunction.cpp:540-683 */
void s1_current(int comp, double *J, double tau, V2 a, V2 b, V2 qodt, V2 delta)
{
    const V2 cd = v2(qodt.y * delta.x, qodt.x * delta.y);
    const double half = 0.5, onethird = 1.0 / 3.0, onefourth = half * half;
    const double taupow2 = tau * tau * half;
    const double taupow3 = taupow2 * tau * 2 * onethird;
    if (comp == 0) {
        const double A[2] = {a.x, 1.0 - a.x};
        const double B[2] = {-delta.x * tau, delta.x * tau};
        const double Ay[3] = {half * (half + b.y) * (half + b.y) * tau, (0.75 - b.y * b.y) * tau,
                              half * (half - b.y) * (half - b.y) * tau};
        const double By[3] = {-(half + b.y) * delta.y * taupow2, 2 * b.y * delta.y * taupow2,
                              (half - b.y) * delta.y * taupow2};
        const double Cy[3] = {half * delta.y * delta.y * taupow3, -delta.y * delta.y * taupow3,
                              half * delta.y * delta.y * taupow3};
        int k = 0;
        for (int jj = 0; jj < 3; ++jj)
            for (int ii = 0; ii < 2; ++ii)
                J[k++] = cd.x * ((A[ii] + B[ii]) * (Ay[jj] + By[jj] + Cy[jj]) -
                                 B[ii] * (half * Ay[jj] + onethird * By[jj] + onefourth * Cy[jj]));
    } else {
        const double Ax[3] = {half * (half + b.x) * (half + b.x) * tau, (0.75 - b.x * b.x) * tau,
                              half * (half - b.x) * (half - b.x) * tau};
        const double Bx[3] = {-(half + b.x) * delta.x * taupow2, 2 * b.x * delta.x * taupow2,
                              (half - b.x) * delta.x * taupow2};
        const double Cx[3] = {half * delta.x * delta.x * taupow3, -delta.x * delta.x * taupow3,
                              half * delta.x * delta.x * taupow3};
        const double A[2] = {a.y, 1.0 - a.y};
        const double B[2] = {-delta.y * tau, delta.y * tau};
        int k = 0;
        for (int jj = 0; jj < 2; ++jj)
            for (int ii = 0; ii < 3; ++ii)
                J[k++] = cd.y * ((A[jj] + B[jj]) * (Ax[ii] + Bx[ii] + Cx[ii]) -
                                 B[jj] * (half * Ax[ii] + onethird * Bx[ii] + onefourth * Cx[ii]));
    }
}

/* FirstOrderShapeFunction.cpp:179-329 */
template <class Acc>
void deposit_s1(int comp, Acc acc, V2 origin, V2 dx, double dt, V2 oR, V2 nR, double q)
{
    const V2 od = v2(1.0 / dx.x, 1.0 / dx.y);
    const V2 odt = v2(od.x / dt, od.y / dt);
    const V2 Dx = v2(nR.x - oR.x, nR.y - oR.y);
    const V2 delta = v2(Dx.x * od.x, Dx.y * od.y);
    const double covered = std::sqrt(dot(Dx, Dx));
    Xing xs[10];
    const int num = dual_crossings(xs, oR, nR, dx, origin);
    double cur[6];
    for (int j = 0; j < num - 1; ++j) {
        const V2 xi = v2(xs[j].s.x * od.x, xs[j].s.y * od.y);
        const V2 xf = v2(xs[j + 1].s.x * od.x, xs[j + 1].s.y * od.y);
        const V2 seg = v2(xs[j + 1].s.x - xs[j].s.x, xs[j + 1].s.y - xs[j].s.y);
        const double tau = std::sqrt(dot(seg, seg)) / covered;
        const V2 c = v2((xi.x + xf.x) * 0.5, (xi.y + xf.y) * 0.5);
        const int I = (int)std::floor(c.x + 0.5);
        const int J = (int)std::floor(c.y + 0.5);
        const V2 a = v2(I + 0.5 - xi.x, J + 0.5 - xi.y);
        const V2 b = v2(I - xi.x, J - xi.y);
        s1_current(comp, cur, tau, a, b, v2(q * odt.x, q * odt.y), delta);
        if (comp == 0) {
            int k = 0;
            for (int jj = -1; jj <= 1; ++jj)
                for (int ii = -1; ii <= 0; ++ii) acc(I + ii, J + jj, cur[k++]);
        } else {
            int k = 0;
            for (int jj = -1; jj <= 0; ++jj)
                for (int ii = -1; ii <= 1; ++ii) acc(I + ii, J + jj, cur[k++]);
        }
    }
}

/* SecondOrderShapeFunction.cpp:583-826 */
void s2_current(int comp, double *J, double tau, V2 a, V2 b, V2 qodt, V2 delta)
{
    const V2 cd = v2(qodt.y * delta.x, qodt.x * delta.y);
    const V2 dp2 = v2(delta.x * delta.x, delta.y * delta.y);
    const V2 dp3 = v2(dp2.x * delta.x, dp2.y * delta.y);
    const double taupow2 = tau * tau / 2;
    const double taupow3 = taupow2 * tau * 2 / 3;
    const double taupow4 = taupow2 * taupow2;
    const double half = 0.5, onethird = 1.0 / 3.0, onefourth = 0.25, onefifth = 0.2;
    const double onesixth = half * onethird;
    if (comp == 0) {
        const double A2[3] = {half * (half + a.x) * (half + a.x), 0.75 - a.x * a.x, half * (0.5 - a.x) * (0.5 - a.x)};
        const double B2[3] = {-(half + a.x) * delta.x * tau, 2 * a.x * delta.x * tau, (half - a.x) * delta.x * tau};
        const double C2[3] = {half * dp2.x * taupow2, -dp2.x * taupow2, half * dp2.x * taupow2};
        const double bb = b.y;
        const double A3[4] = {(1 + bb) * (1 + bb) * (1 + bb) * onesixth * tau,
                              (4 - 3 * bb * bb * (2 + bb)) * onesixth * tau,
                              (4 - 3 * (bb + 1) * (bb + 1) * (1 - bb)) * onesixth * tau,
                              -bb * bb * bb * onesixth * tau};
        const double B3[4] = {-(1 + bb) * (1 + bb) * delta.y * half * taupow2,
                              bb * delta.y * (3 * bb + 4) * half * taupow2,
                              (1 - 2 * bb - 3 * bb * bb) * delta.y * half * taupow2,
                              bb * bb * delta.y * half * taupow2};
        const double C3[4] = {(1 + bb) * dp2.y * half * taupow3, -dp2.y * (2 + 3 * bb) * half * taupow3,
                              (1 + 3 * bb) * dp2.y * half * taupow3, -bb * dp2.y * half * taupow3};
        const double D3[4] = {-dp3.y * onesixth * taupow4, dp3.y * half * taupow4, -dp3.y * half * taupow4,
                              dp3.y * onesixth * taupow4};
        int k = 0;
        for (int jj = 0; jj < 4; ++jj)
            for (int ii = 0; ii < 3; ++ii)
                J[k++] = cd.x * ((A2[ii] + B2[ii] + C2[ii]) * (D3[jj] + C3[jj] + B3[jj] + A3[jj]) +
                                 (C2[ii] * (D3[jj] - A3[jj]) - B2[ii] * B3[jj]) * onethird +
                                 (C2[ii] * C3[jj] - B2[ii] * D3[jj]) * onefifth -
                                 B2[ii] * C3[jj] * onefourth - B2[ii] * A3[jj] * half);
    } else {
        const double bb = b.x;
        const double A3[4] = {(1 + bb) * (1 + bb) * (1 + bb) * onesixth * tau,
                              (4 - 3 * bb * bb * (2 + bb)) * onesixth * tau,
                              (4 - 3 * (bb + 1) * (bb + 1) * (1 - bb)) * onesixth * tau,
                              -bb * bb * bb * onesixth * tau};
        const double B3[4] = {-(1 + bb) * (1 + bb) * delta.x * half * taupow2,
                              bb * delta.x * (3 * bb + 4) * half * taupow2,
                              (1 - 2 * bb - 3 * bb * bb) * delta.x * half * taupow2,
                              bb * bb * delta.x * half * taupow2};
        const double C3[4] = {(1 + bb) * dp2.x * half * taupow3, -dp2.x * (2 + 3 * bb) * half * taupow3,
                              (1 + 3 * bb) * dp2.x * half * taupow3, -bb * dp2.x * half * taupow3};
        const double D3[4] = {-dp3.x * onesixth * taupow4, dp3.x * half * taupow4, -dp3.x * half * taupow4,
                              dp3.x * onesixth * taupow4};
        const double A2[3] = {(half + a.y) * (half + a.y) * half, 0.75 - a.y * a.y, (half - a.y) * (half - a.y) * half};
        const double B2[3] = {-(half + a.y) * delta.y * tau, 2 * a.y * delta.y * tau, (half - a.y) * delta.y * tau};
        const double C2[3] = {dp2.y * half * taupow2, -dp2.y * taupow2, dp2.y * half * taupow2};
        int k = 0;
        for (int jj = 0; jj < 3; ++jj)
            for (int ii = 0; ii < 4; ++ii)
                J[k++] = cd.y * ((A2[jj] + B2[jj] + C2[jj]) * (D3[ii] + C3[ii] + B3[ii] + A3[ii]) +
                                 (C2[jj] * (D3[ii] - A3[ii]) - B2[jj] * B3[ii]) * onethird +
                                 (C2[jj] * C3[ii] - B2[jj] * D3[ii]) * onefifth -
                                 B2[jj] * C3[ii] * onefourth - B2[jj] * A3[ii] * half);
    }
}

/* SecondOrderShapeFunction.cpp:178-345 */
template <class Acc>
void deposit_s2(int comp, Acc acc, V2 origin, V2 dx, double dt, V2 oR, V2 nR, double q)
{
    const V2 od = v2(1.0 / dx.x, 1.0 / dx.y);
    const V2 odt = v2(od.x / dt, od.y / dt);
    Xing xs[10];
    const int num = cell_crossings(xs, oR, nR, dx, origin);
    const V2 Dx = v2(nR.x - oR.x, nR.y - oR.y);
    const V2 rDx = v2(Dx.x * od.x, Dx.y * od.y);
    const double covered = std::sqrt(dot(Dx, Dx));
    double cur[12];
    for (int j = 0; j < num - 1; ++j) {
        const V2 xi = v2(xs[j].s.x * od.x, xs[j].s.y * od.y);
        const V2 xf = v2(xs[j + 1].s.x * od.x, xs[j + 1].s.y * od.y);
        const V2 c = v2(0.5 * (xi.x + xf.x), 0.5 * (xi.y + xf.y));
        const int I = (int)(long)std::floor(c.x);
        const int J = (int)(long)std::floor(c.y);
        const V2 seg = v2(xs[j + 1].s.x - xs[j].s.x, xs[j + 1].s.y - xs[j].s.y);
        const double tau = std::sqrt(dot(seg, seg)) / covered;
        const V2 a = v2(I + 0.5 - xi.x, J + 0.5 - xi.y);
        const V2 b = v2(I - xi.x, J - xi.y);
        s2_current(comp, cur, tau, a, b, v2(q * odt.x, q * odt.y), rDx);
        if (comp == 0) {
            int k = 0;
            for (int jj = -1; jj <= 2; ++jj)
                for (int ii = -1; ii <= 1; ++ii) acc(I + ii, J + jj, cur[k++]);
        } else {
            int k = 0;
            for (int jj = -1; jj <= 1; ++jj)
                for (int ii = -1; ii <= 2; ++ii) acc(I + ii, J + jj, cur[k++]);
        }
    }
}

template <class Acc>
void deposit_one(int order, int comp, Acc acc, V2 origin, V2 dx, double dt, V2 oR, V2 nR, double q)
{
    if (order == 0) deposit_s0(comp, acc, origin, dx, dt, oR, nR, q);
    else if (order == 1) deposit_s1(comp, acc, origin, dx, dt, oR, nR, q);
    else deposit_s2(comp, acc, origin, dx, dt, oR, nR, q);
}

/* Charge density, {Zeroth,First,Second}OrderShapeFunction.cpp getChargeDensityImpl.
 * The footprint of every order is the full stencil (bbox + vertex margin). */
template <class Acc>
void charge_one(int order, Acc acc, V2 origin, V2 dx, V2 R, double q)
{
    if (order == 0) { /* ZerothOrderShapeFunction.cpp:344-357 */
        const V2 xi = v2((R.x - origin.x) / dx.x, (R.y - origin.y) / dx.y);
        const int I = (int)std::floor(xi.x), J = (int)std::floor(xi.y);
        const V2 t = v2(xi.x - I, xi.y - J);
        acc(I, J, (1.0 - t.x) * (1.0 - t.y) * q);
        acc(I + 1, J, t.x * (1.0 - t.y) * q);
        acc(I, J + 1, (1.0 - t.x) * t.y * q);
        acc(I + 1, J + 1, t.x * t.y * q);
    } else if (order == 1) { /* FirstOrderShapeFunction.cpp:389-413 */
        const V2 ne = v2((R.x - origin.x) / dx.x, (R.y - origin.y) / dx.y);
        const int I = (int)std::floor(ne.x + 0.5), J = (int)std::floor(ne.y + 0.5);
        const double a = ne.x - I, b = ne.y - J;
        const double sxm = q * (0.5 - a) * (0.5 - a) * 0.5;
        const double sxp = q * (0.5 + a) * (0.5 + a) * 0.5;
        const double sx = q - sxm - sxp;
        const double sym = (0.5 - b) * (0.5 - b) * 0.5;
        const double syp = (0.5 + b) * (0.5 + b) * 0.5;
        const double sy = 1.0 - sym - syp;
        const double wx[3] = {sxm, sx, sxp}, wy[3] = {sym, sy, syp};
        for (int bj = 0; bj < 3; ++bj)
            for (int ai = 0; ai < 3; ++ai) acc(I - 1 + ai, J - 1 + bj, wx[ai] * wy[bj]);
    } else { /* SecondOrderShapeFunction.cpp:393-427 */
        const double onesixth = 1.0 / 6.0;
        const V2 ne = v2((R.x - origin.x) / dx.x, (R.y - origin.y) / dx.y);
        const int I = (int)std::floor(ne.x), J = (int)std::floor(ne.y);
        const double a = ne.x - I, b = ne.y - J;
        const double sxm1 = onesixth * ((1.0 - a) * (1.0 - a) * (1.0 - a));
        const double sx = onesixth * (4.0 - 3.0 * a * a * (2.0 - a));
        const double sxp1 = onesixth * (4.0 - 3.0 * (1.0 - a) * (1.0 - a) * (1.0 + a));
        const double sxp2 = onesixth * (a * a * a);
        const double sym1 = onesixth * ((1.0 - b) * (1.0 - b) * (1.0 - b));
        const double sy = onesixth * (4.0 - 3.0 * b * b * (2.0 - b));
        const double syp1 = onesixth * (4.0 - 3.0 * (1.0 - b) * (1.0 - b) * (1.0 + b));
        const double syp2 = onesixth * (b * b * b);
        const double wx[4] = {sxm1, sx, sxp1, sxp2}, wy[4] = {sym1, sy, syp1, syp2};
        for (int bj = 0; bj < 4; ++bj)
            for (int ai = 0; ai < 4; ++ai) acc(I - 1 + ai, J - 1 + bj, q * wx[ai] * wy[bj]);
    }
}

/* Shape-function margins, *OrderShapeFunction.hh */
struct Margins {
    int edge[2][2];  /* getExtraMarginEdge: [d][first,last] */
    int cell[2];     /* getExtraMarginCell */
    int cur[2][2];   /* getExtraMarginCurrent */
};
Margins margins(int order)
{
    Margins m;
    if (order == 0) { /* ZerothOrderShapeFunction.hh:99-127 */
        int e[2][2] = {{-1, 1}, {0, 1}}; std::memcpy(m.edge, e, sizeof(e));
        m.cell[0] = -1; m.cell[1] = 1;
        int c[2][2] = {{0, 0}, {0, 1}}; std::memcpy(m.cur, c, sizeof(c));
    } else if (order == 1) { /* FirstOrderShapeFunction.hh:123-144 */
        int e[2][2] = {{-1, 1}, {-1, 2}}; std::memcpy(m.edge, e, sizeof(e));
        m.cell[0] = -1; m.cell[1] = 1;
        int c[2][2] = {{-1, 1}, {-1, 2}}; std::memcpy(m.cur, c, sizeof(c));
    } else { /* SecondOrderShapeFunction.hh:128-149 */
        int e[2][2] = {{-2, 2}, {-1, 2}}; std::memcpy(m.edge, e, sizeof(e));
        m.cell[0] = -2; m.cell[1] = 2;
        int c[2][2] = {{-1, 1}, {-1, 2}}; std::memcpy(m.cur, c, sizeof(c));
    }
    return m;
}

/* ---------------------------------------------------------------------- */
/* A 2-component field over [-g, n-1+g] x [-g, m-1+g].                     */
struct VField {
    int n = 0, m = 0, g = 1;
    std::vector<double> c[2];
    void alloc(int n_, int m_, int g_ = 1)
    {
        n = n_; m = m_; g = g_;
        const size_t sz = (size_t)(n + 2 * g) * (m + 2 * g);
        c[0].assign(sz, 0.0);
        c[1].assign(sz, 0.0);
    }
    bool in(int i, int j) const { return i >= -g && i < n + g && j >= -g && j < m + g; }
    size_t idx(int i, int j) const
    {
        if (!in(i, j)) throw std::out_of_range("field index");
        return (size_t)(j + g) * (n + 2 * g) + (i + g);
    }
    double &at(int k, int i, int j) { return c[k][idx(i, j)]; }
    double at(int k, int i, int j) const { return c[k][idx(i, j)]; }
};

/* ---------------------------------------------------------------------- */
/* Bend geometry, Bend.cpp:60-124, 126-235, 237-353, 427-799, 865-907.     */
struct BCell {
    int i, j;
    double lam[2];
    double area;
    double lamN[2];
};

struct Bend {
    double mass = 0.511e6;
    double lenAfter, lenBefore, width, ekin, bz;
    double R, phi;
    double hr[2], origin[2], center[2];
    double dlen, dwid;
    int Nx, Ny;
    double cosphi_s, sinphi_s, rmin_sq, rmax_sq; /* isInside's statics */
    std::vector<int8_t> vert;    /* internal vertices over [-2, N+1] */
    std::vector<uint8_t> mask;   /* cells [0,N) */
    std::vector<BCell> bcells;   /* after deletion and limitArea */
    std::vector<int> bmap;       /* cell -> bcell index or -1 */

    int vidx(int i, int j) const { return (j + 2) * (Nx + 4) + (i + 2); }
    int vget(int i, int j) const
    {
        if (i < -2 || i > Nx + 1 || j < -2 || j > Ny + 1) return 0;
        return vert[vidx(i, j)];
    }

    Bend(double lb, double la, double w, double phiDeg, double ekinMeV, double bz_, int nx, int ny)
    {
        lenAfter = la; lenBefore = lb; width = w; ekin = ekinMeV * 1e6; bz = bz_;
        Nx = nx; Ny = ny;
        /* Bend.cpp:60-83 */
        if (std::fabs(bz) > 1e-8) {
            R = std::sqrt(ekin * ekin + 2. * ekin * mass) / (C_LIGHT * std::fabs(bz));
            phi = -bz / std::fabs(bz) * std::fabs(phiDeg) * PI / 180.;
        } else {
            R = 9999999.9;
            phi = 0.0;
        }
        init();
    }

    void init()
    {
        /* Bend.cpp:85-124; the geometry layout is the vertex mesh [0,Nx)x[0,Ny) */
        const double cphi = std::cos(phi);
        const double sphi = std::fabs(std::sin(phi));
        if (std::fabs(bz) > 1e-8) {
            dlen = (lenBefore + lenAfter * cphi + (R + 0.5 * width) * sphi);
            dwid = ((R - 0.5 * width) * (1 - cphi) + width + lenAfter * sphi);
        } else {
            dlen = lenAfter + lenBefore;
            dwid = width;
        }
        hr[0] = dlen / (Nx - 2);
        hr[1] = dwid / (Ny - 2);
        if (bz <= 0.0) {
            origin[0] = 0 - hr[0];
            origin[1] = (width - dwid) - hr[1];
            center[0] = lenBefore; center[1] = 0.5 * width - R;
        } else {
            origin[0] = -hr[0];
            origin[1] = -hr[1];
            center[0] = lenBefore; center[1] = 0.5 * width + R;
        }
        dlen += hr[0];
        dwid += hr[1];
        cosphi_s = std::cos(phi);
        sinphi_s = std::sin(phi);
        rmin_sq = (R - 0.5 * width) * (R - 0.5 * width);
        rmax_sq = (R + 0.5 * width) * (R + 0.5 * width);
        calc_vertices();
        fill_cells();
    }

    /* Bend.cpp:194-235 */
    bool isInside(double x0, double x1) const
    {
        const double cosphi = cosphi_s, sinphi = sinphi_s;
        const double D0 = x0 - center[0], D1 = x1 - center[1];
        if (D0 < 0.0) {
            if (x0 < SPACE_EPS || x1 < SPACE_EPS || x1 > width - SPACE_EPS) return false;
        } else if (D1 * sinphi > D0 * cosphi) {
            const double l2 = D0 * D0 + D1 * D1;
            if (l2 <= rmin_sq || l2 >= rmax_sq) return false;
        } else {
            double p0, p1;
            if (std::fabs(phi) > SPACE_EPS) {
                /* dot(rotation, Dx) + center with rotation = (cos, -sin; sin, cos) */
                p0 = (cosphi * D0 + -sinphi * D1) + center[0];
                p1 = (sinphi * D0 + cosphi * D1) + center[1];
            } else {
                p0 = x0; p1 = x1;
            }
            if (p1 < SPACE_EPS || p1 > width - SPACE_EPS || p0 > lenBefore + lenAfter - SPACE_EPS) return false;
        }
        return true;
    }

    /* Bend.cpp:126-192 */
    double crossBoundary(double b0, double b1, double d0, double d1) const
    {
        const double cosphi = std::cos(phi);
        const double sinphi = std::sin(phi);
        const double x0 = b0 + origin[0], x1 = b1 + origin[1];
        const double D0 = x0 - center[0], D1 = x1 - center[1];
        if (D0 < 0.0) {
            double t = 99.9;
            if (x1 < width + SPACE_EPS && x1 + d1 > width - SPACE_EPS) {
                t = std::max(1.0, std::min(0.0, (width - x1) / d1));
            }
            if (x1 > -SPACE_EPS && x1 + d1 < SPACE_EPS) {
                t = std::min(t, -x1 / d1);
            }
            if (x0 > -SPACE_EPS && x0 + d0 < SPACE_EPS) {
                t = std::min(t, -x0 / d0);
            }
            return t;
        } else if (D1 * sinphi > D0 * cosphi) {
            const double rmin2 = (R - 0.5 * width) * (R - 0.5 * width);
            const double rmax2 = (R + 0.5 * width) * (R + 0.5 * width);
            const double dist1 = D0 * D0 + D1 * D1;
            const double E0 = D0 + d0, E1 = D1 + d1;
            const double dist2 = E0 * E0 + E1 * E1;
            const double a = d0 * d0 + d1 * d1;
            const double b = D0 * d0 + D1 * d1;
            double c = 0.0;
            if (dist1 < rmax2 && dist2 > rmax2) {
                c = (D0 * D0 + D1 * D1) - rmax2;
            } else if (dist1 > rmin2 && dist2 < rmin2) {
                c = (D0 * D0 + D1 * D1) - rmin2;
            } else {
                return 99.9;
            }
            double t = 99.9;
            double tau = (-b + std::sqrt(b * b - a * c)) / a;
            if (std::fabs(tau - 0.5) < 0.5) t = tau;
            tau = -tau - 2 * b / a;
            if (std::fabs(tau - 0.5) < 0.5) t = std::min(t, tau);
            return t;
        } else {
            double xr0, xr1, dr0, dr1;
            if (std::fabs(phi) > SPACE_EPS) {
                xr0 = cosphi * D0 - sinphi * D1 + center[0];
                xr1 = sinphi * D0 + cosphi * D1 + center[1];
                dr0 = cosphi * d0 - sinphi * d1;
                dr1 = sinphi * d0 + cosphi * d1;
            } else {
                xr0 = x0; xr1 = x1; dr0 = d0; dr1 = d1;
            }
            double t0 = 99.9, t1 = 99.9, t2 = 99.9;
            if (std::fabs(dr1) > SPACE_EPS) {
                t0 = -xr1 / dr1;
                if (t0 < 0.0) t0 = 99.9;
                t1 = (width - xr1) / dr1;
                if (t1 < 0.0) t1 = 99.9;
            }
            if (std::fabs(dr0) > SPACE_EPS) {
                t2 = (lenBefore + lenAfter - xr0) / dr0;
                if (t2 < 0.0) t2 = 99.9;
            }
            double t = std::min(t0, std::min(t1, t2));
            if (t > 1 + SPACE_EPS && t < 1.001) t = 1.0;
            if (!(t < 1 + SPACE_EPS) || !(t >= 0)) throw std::runtime_error("crossBoundary: t out of range");
            return std::min(1.0, std::max(0.0, t));
        }
    }

    /* Bend.cpp:237-258 */
    void calc_vertices()
    {
        vert.assign((size_t)(Nx + 4) * (Ny + 4), 0);
        for (int cy = 0; cy < Ny; ++cy) {
            const double x1 = origin[1] + cy * hr[1];
            for (int cx = 0; cx < Nx; ++cx) {
                const double x0 = origin[0] + cx * hr[0];
                vert[vidx(cx, cy)] = isInside(x0, x1) ? 1 : 0;
            }
        }
    }

    /* Bend.cpp:664-729 */
    static double areaCases(int ccase, const double *lam, const double *lamN)
    {
        double area = 0.0;
        switch (ccase) {
        case 1: area = lam[0] * lam[1] / 2; break;
        case 2: area = lam[0] * lamN[1] / 2; break;
        case 3: area = (lam[1] + lamN[1]) / 2; break;
        case 4: area = lamN[0] * lam[1] / 2; break;
        case 5: area = (lam[0] + lamN[0]) / 2; break;
        case 7: area = 1 - (1 - lamN[0]) * (1 - lamN[1]) / 2; break;
        case 8: area = lamN[0] * lamN[1] / 2; break;
        case 10: area = (lam[0] + lamN[0]) / 2; break;
        case 11: area = 1 - (1 - lam[1]) * (1 - lamN[0]) / 2; break;
        case 12: area = (lam[1] + lamN[1]) / 2; break;
        case 13: area = 1 - (1 - lam[0]) * (1 - lamN[1]) / 2; break;
        case 14: area = 1 - (1 - lam[0]) * (1 - lam[1]) / 2; break;
        default: break;
        }
        return area;
    }

    /* Bend.cpp:282-353, 427-545, 619-662; SCBoundaryCondition.hh:96-137 */
    void fill_cells()
    {
        const int lw = Nx + 4;
        std::vector<double> lam((size_t)2 * lw * (Ny + 4), 0.0);
        auto L = [&](int k, int i, int j) -> double & { return lam[2 * ((size_t)(j + 2) * lw + (i + 2)) + k]; };
        mask.assign((size_t)Nx * Ny, 0);
        std::vector<std::pair<int, int>> cand;
        const double h0 = hr[0], h1 = hr[1];
        for (int cy = 0; cy < Ny; ++cy) {
            for (int cx = 0; cx < Nx; ++cx) {
                const int ccase = vget(cx, cy) + (vget(cx + 1, cy) << 1) + (vget(cx, cy + 1) << 2) +
                                  (vget(cx + 1, cy + 1) << 3);
                if (ccase == 15) {
                    L(0, cx, cy) = 1.0; L(1, cx, cy) = 1.0;
                    mask[(size_t)cy * Nx + cx] = 1;
                    continue;
                }
                if (ccase == 0) continue;
                cand.push_back(std::make_pair(cx, cy));
                double &l0 = L(0, cx, cy), &l1 = L(1, cx, cy);
                switch (ccase) {
                case 1:
                    l0 = crossBoundary(cx * h0, cy * h1, h0, 0.0);
                    l1 = crossBoundary(cx * h0, cy * h1, 0.0, h1);
                    break;
                case 2:
                    l0 = crossBoundary((cx + 1) * h0, cy * h1, -h0, 0.0);
                    l1 = 0.0;
                    break;
                case 3:
                    l0 = 1.0;
                    l1 = crossBoundary(cx * h0, cy * h1, 0.0, h1);
                    break;
                case 4:
                    l0 = 0.0;
                    l1 = crossBoundary(cx * h0, (cy + 1) * h1, 0.0, -h1);
                    break;
                case 5:
                    l0 = crossBoundary(cx * h0, cy * h1, h0, 0.0);
                    l1 = 1.0;
                    break;
                case 7: l0 = 1.0; l1 = 1.0; break;
                case 8: l0 = 0.0; l1 = 0.0; break;
                case 10:
                    l0 = crossBoundary((cx + 1) * h0, cy * h1, -h0, 0.0);
                    l1 = 0.0;
                    break;
                case 11:
                    l0 = 1.0;
                    l1 = crossBoundary(cx * h0, cy * h1, 0.0, h1);
                    break;
                case 12:
                    l0 = 0.0;
                    l1 = crossBoundary(cx * h0, (cy + 1) * h1, 0.0, -h1);
                    break;
                case 13:
                    l0 = crossBoundary(cx * h0, cy * h1, h0, 0.0);
                    l1 = 1.0;
                    break;
                case 14:
                    l0 = crossBoundary((cx + 1) * h0, cy * h1, -h0, 0.0);
                    l1 = crossBoundary(cx * h0, (cy + 1) * h1, 0.0, -h1);
                    break;
                default: break;
                }
            }
        }
        bmap.assign((size_t)Nx * Ny, -1);
        for (const auto &p : cand) {
            const int cx = p.first, cy = p.second;
            const int ccase = vget(cx, cy) + (vget(cx + 1, cy) << 1) + (vget(cx, cy + 1) << 2) +
                              (vget(cx + 1, cy + 1) << 3);
            const double l[2] = {L(0, cx, cy), L(1, cx, cy)};
            const double ln[2] = {L(0, cx, cy + 1), L(1, cx + 1, cy)};
            const double area = areaCases(ccase, l, ln);
            if (area < SPACE_EPS) continue; /* toBeDeleted */
            BCell b;
            b.i = cx; b.j = cy;
            b.lam[0] = l[0]; b.lam[1] = l[1];
            b.lamN[0] = ln[0]; b.lamN[1] = ln[1];
            b.area = area;
            /* SCBoundaryCondition.hh:131-137 limitArea */
            b.area = std::max(b.area, b.lam[0] / 2);
            b.area = std::max(b.area, b.lam[1] / 2);
            b.area = std::max(b.area, b.lamN[0] / 2);
            b.area = std::max(b.area, b.lamN[1] / 2);
            bmap[(size_t)cy * Nx + cx] = (int)bcells.size();
            bcells.push_back(b);
        }
    }

    /* Bend.cpp:865-907 */
    double external_field(double x0, double x1) const
    {
        const double cph = std::cos(0.5 * phi);
        const double sph = std::sin(0.5 * phi);
        const double hw = 0.5 * width;
        const double bw = 2 * R * std::fabs(sph);
        const double D0 = x0 - lenBefore, D1 = x1 - hw;
        const double dist = cph * D0 - sph * D1;
        return (dist > 0.0 && dist < bw) ? bz : 0.0;
    }
};

/* ---------------------------------------------------------------------- */
/* The two-grid simulation, main.cpp:168-532 with TETMSolver.hh:110-266.   */

struct Rect {
    int lo[2], hi[2];
    bool empty() const { return lo[0] > hi[0] || lo[1] > hi[1]; }
};
inline Rect rect(int a0, int b0, int a1, int b1) { Rect r; r.lo[0] = a0; r.hi[0] = b0; r.lo[1] = a1; r.hi[1] = b1; return r; }

enum OpKind { OP_H = 0, OP_EX = 1, OP_EY = 2 };
/* Communicator.cpp:640-642 */
const unsigned SIDES[3] = {(1u << 4) | (1u << 5) | (1u << 2) | (1u << 3), (1u << 7) | (1u << 1), (1u << 6) | 1u};

double op_execute(OpKind k, const VField &f, int i, int j)
{
    if (k == OP_H) return f.at(0, i, j) + f.at(1, i, j);
    if (k == OP_EX) return f.at(0, i, j);
    return f.at(1, i, j);
}
void op_invert(OpKind k, VField &f, int i, int j, double v)
{
    if (k == OP_H) { f.at(0, i, j) = 0.5 * v; f.at(1, i, j) = 0.5 * v; }
    else if (k == OP_EX) f.at(0, i, j) = v;
    else f.at(1, i, j) = v;
}

/* ---------------------------------------------------------------------- */
/* Poisson initialisation, PoissonSolver.cpp:85-201, 204-233, 441-507,      */
/* 678-766.  The reference solves A phi = rho/(eps0 gamma) with Belos CG +  */
/* ML AMG from a random start; the oracle solves the same linear system     */
/* directly (banded Cholesky), so it carries only rounding error.           */

/* Node classes of PoissonSolver::cutoffStencil (PoissonSolver.cpp:678-766):
 * flag = insideMask || (boundary cell with area > 0.5); a row is a stencil
 * row when idx < maxX and flag, otherwise an identity row (phi = rhs). */
void poisson_classes(const Bend &b, std::vector<uint8_t> &flag, std::vector<uint8_t> &active)
{
    const long maxX = (long)(size_t)(b.lenBefore / b.hr[0]);   /* Bend.hh:288-291 */
    flag.assign((size_t)b.Nx * b.Ny, 0);
    active.assign((size_t)b.Nx * b.Ny, 0);
    for (int j = 0; j < b.Ny; ++j)
        for (int i = 0; i < b.Nx; ++i) {
            const size_t id = (size_t)j * b.Nx + i;
            const int bc = b.bmap[id];
            const bool f = b.mask[id] || (bc > -1 && b.bcells[bc].area > 0.5);
            flag[id] = f ? 1 : 0;
            active[id] = (f && i < maxX) ? 1 : 0;
        }
}

/* Solve the cutoffStencil system on an nx x ny node grid (row-major, x
 * fastest).  Stencil rows: -2(EW+NS) on the diagonal, EW = -1/hx^2 to x
 * neighbours and NS = -1/hy^2 to y neighbours whose flag is set (neighbours
 * off the grid count as unflagged).  Identity rows: phi = b.  A flagged
 * neighbour that is an identity row is a known value and moves to the
 * right-hand side.  The stencil rows form an SPD system, factored by a
 * banded Cholesky with the nodes numbered along the shorter side of their
 * bounding box (bandwidth = that side). */
int poisson_direct(int nx, int ny, double hx, double hy, const uint8_t *flag, const uint8_t *active,
                   const double *b, double *phi)
{
    const double EW = -1 / (hx * hx);
    const double NS = -1 / (hy * hy);
    int i0 = nx, i1 = -1, j0 = ny, j1 = -1;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            if (active[(size_t)j * nx + i]) {
                i0 = std::min(i0, i); i1 = std::max(i1, i);
                j0 = std::min(j0, j); j1 = std::max(j1, j);
            }
    for (size_t k = 0; k < (size_t)nx * ny; ++k) phi[k] = b[k];
    if (i1 < 0) return 0;
    const int bw = j1 - j0 + 1, bx = i1 - i0 + 1;
    const long n = (long)bw * bx;
    const bool colOrder = bw <= bx;                     /* number along the shorter side */
    const int p = colOrder ? bw : bx;                   /* bandwidth */
    std::vector<double> L((size_t)n * (p + 1), 0.0);    /* L[k][d] = L(k, k-d) */
    std::vector<double> rhs(n, 0.0);
    auto kid = [&](int i, int j) {
        return colOrder ? (long)(i - i0) * bw + (j - j0) : (long)(j - j0) * bx + (i - i0);
    };
    auto isflag = [&](int i, int j) { return i >= 0 && j >= 0 && i < nx && j < ny && flag[(size_t)j * nx + i]; };
    auto isact = [&](int i, int j) { return i >= 0 && j >= 0 && i < nx && j < ny && active[(size_t)j * nx + i]; };
    for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j) {
            const long k = kid(i, j);
            double *Lk = &L[(size_t)k * (p + 1)];
            for (int d = 0; d <= p; ++d) Lk[d] = 0.0;
            if (!active[(size_t)j * nx + i]) { Lk[0] = 1.0; rhs[k] = 0.0; continue; }
            Lk[0] = -2 * (EW + NS);
            rhs[k] = b[(size_t)j * nx + i];
            const int ni[4] = {i, i - 1, i + 1, i}, nj[4] = {j - 1, j, j, j + 1};
            const double w[4] = {NS, EW, EW, NS};
            for (int q = 0; q < 4; ++q) {
                if (!isflag(ni[q], nj[q])) continue;
                if (isact(ni[q], nj[q])) {
                    const long m = kid(ni[q], nj[q]);
                    if (m < k) Lk[k - m] = w[q];               /* lower triangle only */
                } else {
                    rhs[k] -= w[q] * b[(size_t)nj[q] * nx + ni[q]];
                }
            }
        }
    /* in-place banded Cholesky */
    for (long k = 0; k < n; ++k) {
        double *Lk = &L[(size_t)k * (p + 1)];
        const long m0 = std::max(0L, k - p);
        for (long m = m0; m <= k; ++m) {
            const double *Lm = &L[(size_t)m * (p + 1)];
            double s = Lk[k - m];
            const long t0 = std::max(m0, std::max(0L, m - p));
            for (long t = t0; t < m; ++t) s -= Lk[k - t] * Lm[m - t];
            if (m == k) {
                if (!(s > 0)) return -1;
                Lk[0] = std::sqrt(s);
            } else {
                Lk[k - m] = s / Lm[0];
            }
        }
    }
    for (long k = 0; k < n; ++k) {                      /* L y = rhs */
        const double *Lk = &L[(size_t)k * (p + 1)];
        double s = rhs[k];
        for (long t = std::max(0L, k - p); t < k; ++t) s -= Lk[k - t] * rhs[t];
        rhs[k] = s / Lk[0];
    }
    for (long k = n - 1; k >= 0; --k) {                 /* L^T x = y */
        double s = rhs[k];
        for (long m = k + 1; m <= std::min(n - 1, k + p); ++m) s -= L[(size_t)m * (p + 1) + (m - k)] * rhs[m];
        rhs[k] = s / L[(size_t)k * (p + 1)];
    }
    for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j)
            if (active[(size_t)j * nx + i]) phi[(size_t)j * nx + i] = rhs[kid(i, j)];
    return 0;
}

/* The per-particle parts of PartBunch::push_particles (PartBunch.cpp:215-383):
 * the half drift R += dt/2 c P / gamma (:228-232) ... */
inline void half_drift(double *R, const double *P, double half)
{
    const V2 Pk = v2(P[0], P[1]);
    const double g = (double)q_rsqrt((float)(1.0 + dot(Pk, Pk)));
    R[0] += half * Pk.x * g;
    R[1] += half * Pk.y * g;
}

/* ... then the gather at R (:300-358), the kick with fact1 = dt/2 c q/m and
 * the rotation with B = mu0 Hf + Bext/2 (:360-378), and the second half
 * drift with the new momentum (:379-383). */
template <class FX, class FY, class FH>
void kick_drift(int order, int s2bug, const Bend &geo, double dt, double qm, FX fEX, FY fEY, FH fHZ, V2 oEx, V2 oEy,
                V2 oHz, V2 dx, double *R, double *P)
{
    const double half = 0.5 * dt * C_LIGHT;
    const double fact1 = half * qm;
    const double fact2 = fact1 * C_LIGHT;
    const V2 r = v2(R[0], R[1]);
    const double ex = gather_edge(order, 0, s2bug, fEX, oEx, dx, r);
    const double ey = gather_edge(order, 1, s2bug, fEY, oEy, dx, r);
    const double hf = gather_cell(order, s2bug, fHZ, oHz, dx, r);
    const double bext = geo.external_field(r.x, r.y);
    const double B = MU_0 * hf + 0.5 * bext;
    const V2 kick = v2(fact1 * ex, fact1 * ey);
    V2 um = v2(P[0] + kick.x, P[1] + kick.y);
    const double a = fact2 * B * (double)q_rsqrt((float)(1.0 + dot(um, um)));
    const V2 s = v2(um.x + a * um.y, um.y + a * -um.x);
    const double tmp = 1.0 + a * a;
    um.x = ((1.0 + a * a) * s.x + a * s.y) / tmp;
    um.y = ((1.0 + a * a) * s.y - a * s.x) / tmp;
    P[0] = um.x + kick.x;
    P[1] = um.y + kick.y;
    half_drift(R, P, half);
}

struct Sim {
    orc_sim_params p;
    Bend *geo = nullptr;
    Margins mg;
    double dt = 0.0;
    double hr[2];
    /* global fields: E (Ex,Ey), H (Hzx,Hzy), J (Jx,Jy) on [0,Nx)x[0,Ny) */
    VField E, H, J;
    /* window fields */
    VField mE, mH, mJ;
    int W = 0, Hh = 0;
    int ddim[2][2];          /* window position in global index space: [d][first,last] */
    double mmOrigin[2];
    /* particles */
    int N = 0;
    std::vector<double> R, oldR, P, Q;
    double qm;               /* charge-to-mass ratio, main.cpp:225 */
    int gb[2][2];            /* PartBunch::_globalBounds (window index space) */
    int steps = 0;
    /* IntegrateElongCommand */
    std::vector<long double> wsum, wcor;
    int nsamp = 0;

    V2 Rp(int k) const { return v2(R[2 * k], R[2 * k + 1]); }

    /* PartBunch::getBounds / getBoundsInclOld, PartBunch.hh:247-287 */
    void bounds(bool inclOld, double mn[2], double mx[2]) const
    {
        for (int d = 0; d < 2; ++d) { mn[d] = R[d]; mx[d] = R[d]; }
        for (int k = 0; k < N; ++k)
            for (int d = 0; d < 2; ++d) {
                const double r = R[2 * k + d];
                if (r < mn[d]) mn[d] = r;
                if (inclOld && oldR[2 * k + d] < mn[d]) mn[d] = oldR[2 * k + d];
                if (r > mx[d]) mx[d] = r;
                if (inclOld && oldR[2 * k + d] > mx[d]) mx[d] = oldR[2 * k + d];
            }
    }
    /* getLocalPDomain / InclOld, PartBunch.cpp:811-897 */
    void pdomain(bool inclOld, int out[2][2]) const
    {
        double mn[2], mx[2];
        bounds(inclOld, mn, mx);
        for (int d = 0; d < 2; ++d) {
            out[d][0] = (int)std::floor((mn[d] - mmOrigin[d]) / hr[d]);
            out[d][1] = (int)std::floor((mx[d] - mmOrigin[d]) / hr[d]);
        }
    }

    void setup(const orc_sim_params *pp)
    {
        p = *pp;
        geo = new Bend(p.lenBefore, p.lenAfter, p.width, p.phiDeg, p.ekinMeV, p.bz, p.Nx, p.Ny);
        hr[0] = geo->hr[0]; hr[1] = geo->hr[1];
        dt = hr[0] / C_LIGHT;                       /* main.cpp:187 */
        mg = margins(p.order);
        qm = -1e-9 / M_E_GEV;                       /* main.cpp:225 */
        E.alloc(p.Nx, p.Ny); H.alloc(p.Nx, p.Ny); J.alloc(p.Nx, p.Ny);
    }

    /* O0 null: the -I start (makeReady); else the restart of main.cpp:147-153,
       228 with LoadH5Part::restoreBunch's oldR and no makeReady. */
    int init(const double *R0, const double *P0, const double *Q0, int n, const double *O0 = nullptr, int step0 = 0)
    {
        N = n;
        R.assign(R0, R0 + 2 * n);
        P.assign(P0, P0 + 2 * n);
        Q.assign(Q0, Q0 + n);
        if (O0)
            oldR.assign(O0, O0 + 2 * n);
        else
            oldR.assign(2 * n, 0.0);
        /* Distribution::getDomain, Distribution.cpp:181-207, then +-5 cells */
        double mn[2] = {R[0], R[1]}, mx[2] = {R[0], R[1]};
        for (int k = 0; k < N; ++k)
            for (int d = 0; d < 2; ++d) {
                mn[d] = std::min(mn[d], R[2 * k + d]);
                mx[d] = std::max(mx[d], R[2 * k + d]);
            }
        for (int d = 0; d < 2; ++d) {
            ddim[d][0] = (int)std::floor((mn[d] - geo->origin[d]) / hr[d]) - 5;
            ddim[d][1] = (int)std::floor((mx[d] - geo->origin[d]) / hr[d]) + 5;
        }
        mmOrigin[0] = geo->origin[0] + ddim[0][0] * hr[0];
        mmOrigin[1] = geo->origin[1] + ddim[1][0] * hr[1];
        W = ddim[0][1] - ddim[0][0] + 1;
        Hh = ddim[1][1] - ddim[1][0] + 1;
        if (ddim[0][0] < 1 || ddim[1][0] < 1 || ddim[0][1] > p.Nx - 2 || ddim[1][1] > p.Ny - 2) return -2;
        mE.alloc(W, Hh); mH.alloc(W, Hh); mJ.alloc(W, Hh);
        /* PartBunch::sortParticles, PartBunch.cpp:1817-1844 (stable for equal x) */
        {
            int pd[2][2];
            pdomain(false, pd);
            std::vector<int> idx(N);
            std::vector<int> slice(N);
            for (int k = 0; k < N; ++k) {
                idx[k] = k;
                slice[k] = (int)std::floor((R[2 * k + 1] - mmOrigin[1]) / hr[1]) - pd[1][0];
            }
            std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
                if (slice[a] != slice[b]) return slice[a] < slice[b];
                return R[2 * a] < R[2 * b];
            });
            std::vector<double> R2(2 * N), P2(2 * N), O2(2 * N), Q2(N);
            for (int k = 0; k < N; ++k) {
                const int s = idx[k];
                R2[2 * k] = R[2 * s]; R2[2 * k + 1] = R[2 * s + 1];
                P2[2 * k] = P[2 * s]; P2[2 * k + 1] = P[2 * s + 1];
                O2[2 * k] = oldR[2 * s]; O2[2 * k + 1] = oldR[2 * s + 1];
                Q2[k] = Q[s];
            }
            R.swap(R2); P.swap(P2); oldR.swap(O2); Q.swap(Q2);
        }
        /* makeReady(dt/2), PartBunch.cpp:164-177 */
        for (int k = 0; k < N && !O0; ++k) {
            const V2 Pk = v2(P[2 * k], P[2 * k + 1]);
            const double g = (double)q_rsqrt((float)(1 + dot(Pk, Pk)));
            const double h = dt / 2 * C_LIGHT;
            oldR[2 * k] = R[2 * k] - h * Pk.x * g;
            oldR[2 * k + 1] = R[2 * k + 1] - h * Pk.y * g;
        }
        /* copyFields global -> window (main.cpp:335-337) */
        copy_to_window();
        /* exchangeBoundaries H, Ex, Ey (main.cpp:550-559) */
        int pos[2] = {ddim[0][0], ddim[1][0]};
        exchange(OP_H, pos);
        exchange(OP_EX, pos);
        exchange(OP_EY, pos);
        /* IntegrateElongCommand: numSamples = W - 4 (first+2 .. last-2) */
        nsamp = W - 4;
        wsum.assign(nsamp, 0.0L);
        wcor.assign(nsamp, 0.0L);
        J.c[0].assign(J.c[0].size(), 0.0); J.c[1].assign(J.c[1].size(), 0.0);
        steps = step0;
        return 0;
    }

    /* Communicator::copy, Communicator.hh:211-272 (global -> window, E and H) */
    void copy_to_window()
    {
        const int s[2] = {ddim[0][0], ddim[1][0]};
        for (int f = 0; f < 2; ++f) {
            const VField &from = f == 0 ? E : H;
            VField &to = f == 0 ? mE : mH;
            const int lo0 = std::max(0, 0 + s[0]), hi0 = std::min(p.Nx - 1, W - 1 + s[0]);
            const int lo1 = std::max(0, 0 + s[1]), hi1 = std::min(p.Ny - 1, Hh - 1 + s[1]);
            for (int j = lo1; j <= hi1; ++j)
                for (int i = lo0; i <= hi0; ++i)
                    for (int k = 0; k < 2; ++k) to.at(k, i - s[0], j - s[1]) = from.at(k, i, j);
        }
    }

    /* Communicator::exchangeBoundaries, single rank, Communicator.hh:575-876 */
    void exchange(OpKind op, const int pos[2])
    {
        VField &S = (op == OP_H) ? H : E;
        VField &Em = (op == OP_H) ? mH : mE;
        const Rect gEmb = rect(0, W - 1, 0, Hh - 1);
        const Rect incr = rect(-1 + pos[0], W + pos[0], -1 + pos[1], Hh + pos[1]);
        Rect g[8];
        for (int d = 0; d < 2; ++d) {
            g[d] = gEmb; g[d].lo[d] = g[d].hi[d] = gEmb.lo[d];
            g[d + 2] = gEmb; g[d + 2].lo[d] = g[d + 2].hi[d] = gEmb.hi[d];
            g[d + 4] = incr; g[d + 4].lo[d] = g[d + 4].hi[d] = incr.lo[d];
            g[d + 6] = incr; g[d + 6].lo[d] = g[d + 6].hi[d] = incr.hi[d];
        }
        const Rect surrLocal = rect(0, p.Nx - 1, 0, p.Ny - 1);
        const Rect embLocal = gEmb;
        struct Pack { Rect dst; bool toSurr; std::vector<double> v; };
        std::vector<Pack> packs;
        for (int side = 0; side < 8; ++side) {
            if (!(SIDES[op] & (1u << side))) continue;
            Rect dom;
            const Rect &src = side >= 4 ? surrLocal : embLocal;
            for (int d = 0; d < 2; ++d) {
                dom.lo[d] = std::max(src.lo[d], g[side].lo[d]);
                dom.hi[d] = std::min(src.hi[d], g[side].hi[d]);
            }
            if (dom.empty()) continue;
            const int idx = side < 4 ? 0 : 1;
            const Rect &fld = idx == 0 ? surrLocal : embLocal;
            const int diff[2] = {idx == 0 ? -pos[0] : pos[0], idx == 0 ? -pos[1] : pos[1]};
            Rect ov;
            bool inside = true;
            for (int d = 0; d < 2; ++d) {
                ov.lo[d] = std::max(dom.lo[d], fld.lo[d] + diff[d] - 1);
                ov.hi[d] = std::min(dom.hi[d], fld.hi[d] + diff[d] + 1);
                if (ov.lo[d] > ov.hi[d]) inside = false;
            }
            if (!inside) continue;
            Pack pk;
            pk.toSurr = (side < 4);
            const VField &sf = side < 4 ? Em : S;
            for (int j = ov.lo[1]; j <= ov.hi[1]; ++j)
                for (int i = ov.lo[0]; i <= ov.hi[0]; ++i) pk.v.push_back(op_execute(op, sf, i, j));
            for (int d = 0; d < 2; ++d) {
                pk.dst.lo[d] = side < 4 ? ov.lo[d] + pos[d] : ov.lo[d] - pos[d];
                pk.dst.hi[d] = side < 4 ? ov.hi[d] + pos[d] : ov.hi[d] - pos[d];
            }
            packs.push_back(pk);
        }
        for (auto &pk : packs) {
            VField &df = pk.toSurr ? S : Em;
            size_t ii = 0;
            for (int j = pk.dst.lo[1]; j <= pk.dst.hi[1]; ++j)
                for (int i = pk.dst.lo[0]; i <= pk.dst.hi[0]; ++i) op_invert(op, df, i, j, pk.v[ii++]);
        }
    }

    /* Communicator::exchangeBoundariesMovingMesh, Communicator.cpp:123-638 */
    void exchange_moving(const int pos[2], const int move[2])
    {
        const Rect gEH = rect(0, W - 1, 0, Hh - 1); /* embedded E and H global domains */
        std::vector<Rect> L[6];                     /* HzE2S ExE2S EyE2S HzS2E ExS2E EyS2E */
        const int f0 = 0, l0 = W - 1, f1 = 0, l1 = Hh - 1;
        if (move[0] == 1) {
            L[0].push_back(rect(f0 - 1, f0 - 1, f1 - move[1], l1 - move[1] - 1));
            L[1].push_back(rect(f0 - 1, f0 - 1, f1 - move[1] + 1, l1 - move[1]));
            L[2].push_back(rect(f0 + 1 - 1, f0 + 1 - 1, f1 - move[1], l1 - move[1]));
            L[3].push_back(rect(l0 + pos[0], l0 + pos[0], f1 + pos[1], l1 + pos[1]));
            L[4].push_back(rect(l0 + pos[0], l0 + pos[0], f1 + pos[1], l1 + pos[1] + 1));
            L[5].push_back(rect(l0 + pos[0] + 1, l0 + pos[0] + 1, f1 + pos[1], l1 + pos[1]));
        }
        if (move[1] == 1) {
            L[0].push_back(rect(f0 - move[0], l0 - 1 - move[0], f1 - 1, f1 - 1));
            L[1].push_back(rect(f0 - move[0], l0 - move[0], f1 + 1 - 1, f1 + 1 - 1));
            L[2].push_back(rect(f0 + 1 - move[0], l0 - move[0], f1 - 1, f1 - 1));
            L[3].push_back(rect(f0 + pos[0], l0 + pos[0], l1 + pos[1], l1 + pos[1]));
            L[4].push_back(rect(f0 + pos[0], l0 + pos[0], l1 + 1 + pos[1], l1 + 1 + pos[1]));
            L[5].push_back(rect(f0 + pos[0], l0 + pos[0] + 1, l1 + pos[1], l1 + pos[1]));
        } else if (move[1] == -1) {
            L[0].push_back(rect(f0 - move[0], l0 - move[0] - 1, l1 - 1 - move[1], l1 - move[1] - 1));
            L[1].push_back(rect(f0 - move[0], l0 - move[0], l1 - move[1], l1 - move[1]));
            L[2].push_back(rect(f0 - move[0] + 1, l0 - move[0], l1 - move[1], l1 - move[1]));
            L[3].push_back(rect(f0 + pos[0], l0 + pos[0], f1 - 1 + pos[1], f1 - 1 + pos[1]));
            L[4].push_back(rect(f0 + pos[0], l0 + pos[0], f1 + pos[1], f1 + pos[1]));
            L[5].push_back(rect(f0 + pos[0], l0 + pos[0] + 1, f1 + pos[1], f1 + pos[1]));
        }
        (void)gEH;
        /* local domains: embedded E/H extended by one layer on each side */
        const Rect embExt = rect(-1, W, -1, Hh);
        const Rect surr = rect(0, p.Nx - 1, 0, p.Ny - 1);
        struct Pack { int type; Rect dst; std::vector<double> v; };
        std::vector<Pack> packs;
        for (int t = 0; t < 6; ++t) {
            const Rect &domain = t < 3 ? embExt : surr;
            const Rect &fld = t < 3 ? surr : embExt;
            const int idx = t < 3 ? 0 : 1;
            const int diff[2] = {idx == 0 ? -pos[0] : pos[0], idx == 0 ? -pos[1] : pos[1]};
            for (const Rect &layer : L[t]) {
                Rect dom;
                bool has = true;
                for (int d = 0; d < 2; ++d) {
                    dom.lo[d] = std::max(domain.lo[d], layer.lo[d]);
                    dom.hi[d] = std::min(domain.hi[d], layer.hi[d]);
                    if (dom.lo[d] > dom.hi[d]) has = false;
                }
                if (!has) continue;
                Rect ov;
                bool inside = true;
                for (int d = 0; d < 2; ++d) {
                    ov.lo[d] = std::max(dom.lo[d], fld.lo[d] + diff[d]);
                    ov.hi[d] = std::min(dom.hi[d], fld.hi[d] + diff[d]);
                    if (ov.lo[d] > ov.hi[d]) inside = false;
                }
                if (!inside) continue;
                Pack pk;
                pk.type = t;
                const OpKind k = (OpKind)(t % 3);
                const VField &sf = t == 0 ? mH : t < 3 ? mE : t == 3 ? H : E;
                for (int j = ov.lo[1]; j <= ov.hi[1]; ++j)
                    for (int i = ov.lo[0]; i <= ov.hi[0]; ++i) pk.v.push_back(op_execute(k, sf, i, j));
                for (int d = 0; d < 2; ++d) {
                    pk.dst.lo[d] = ov.lo[d] - diff[d];
                    pk.dst.hi[d] = ov.hi[d] - diff[d];
                }
                packs.push_back(pk);
            }
        }
        for (auto &pk : packs) {
            const OpKind k = (OpKind)(pk.type % 3);
            VField &df = pk.type == 0 ? H : pk.type < 3 ? E : pk.type == 3 ? mH : mE;
            size_t ii = 0;
            for (int j = pk.dst.lo[1]; j <= pk.dst.hi[1]; ++j)
                for (int i = pk.dst.lo[0]; i <= pk.dst.hi[0]; ++i) op_invert(k, df, i, j, pk.v[ii++]);
        }
    }

    /* TETMSolver.hh:142-149 (global masked + window) */
    void h_half()
    {
        const double etax = C_LIGHT * dt / hr[0];
        const double etay = C_LIGHT * dt / hr[1];
        const double Z0 = std::sqrt(MU_0 / EPS_0);
        const double sx = 0.5 * etax / Z0, sy = 0.5 * etay / Z0;
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i) {
                if (!geo->mask[(size_t)j * p.Nx + i]) continue;
                H.at(0, i, j) = H.at(0, i, j) - sx * (E.at(1, i + 1, j) - E.at(1, i, j));
            }
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i) {
                if (!geo->mask[(size_t)j * p.Nx + i]) continue;
                H.at(1, i, j) = H.at(1, i, j) + sy * (E.at(0, i, j + 1) - E.at(0, i, j));
            }
        for (int j = 0; j < Hh; ++j)
            for (int i = 0; i < W; ++i) {
                mH.at(0, i, j) = mH.at(0, i, j) - sx * (mE.at(1, i + 1, j) - mE.at(1, i, j));
                mH.at(1, i, j) = mH.at(1, i, j) + sy * (mE.at(0, i, j + 1) - mE.at(0, i, j));
            }
    }
    /* TETMSolver.hh:165-172 */
    void e_x()
    {
        const double etay = C_LIGHT * dt / hr[1];
        const double Z0 = std::sqrt(MU_0 / EPS_0);
        const double dteps = C_LIGHT * dt * Z0;
        const double s = etay * Z0;
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i) {
                if (!geo->mask[(size_t)j * p.Nx + i]) continue;
                E.at(0, i, j) = E.at(0, i, j) +
                    (s * (((H.at(0, i, j) - H.at(0, i, j - 1)) + H.at(1, i, j)) - H.at(1, i, j - 1)) - dteps * J.at(0, i, j));
            }
        for (int j = 0; j < Hh; ++j)
            for (int i = 0; i < W; ++i)
                mE.at(0, i, j) = mE.at(0, i, j) +
                    (s * (((mH.at(0, i, j) - mH.at(0, i, j - 1)) + mH.at(1, i, j)) - mH.at(1, i, j - 1)) - dteps * mJ.at(0, i, j));
    }
    /* TETMSolver.hh:241-249 */
    void e_y()
    {
        const double etax = C_LIGHT * dt / hr[0];
        const double Z0 = std::sqrt(MU_0 / EPS_0);
        const double dteps = C_LIGHT * dt * Z0;
        const double s = etax * Z0;
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i) {
                if (!geo->mask[(size_t)j * p.Nx + i]) continue;
                E.at(1, i, j) = E.at(1, i, j) -
                    (s * (((H.at(0, i, j) - H.at(0, i - 1, j)) + H.at(1, i, j)) - H.at(1, i - 1, j)) + dteps * J.at(1, i, j));
            }
        for (int j = 0; j < Hh; ++j)
            for (int i = 0; i < W; ++i)
                mE.at(1, i, j) = mE.at(1, i, j) -
                    (s * (((mH.at(0, i, j) - mH.at(0, i - 1, j)) + mH.at(1, i, j)) - mH.at(1, i - 1, j)) + dteps * mJ.at(1, i, j));
    }
    /* SCBoundaryCondition.hh:392-400 + TETMSolver.cpp:87-103 */
    void sc_hz()
    {
        const double etax = dt / (hr[0] * MU_0);
        const double etay = dt / (hr[1] * MU_0);
        for (const BCell &b : geo->bcells) {
            const int i = b.i, j = b.j;
            const double u0 = -0.5 * etax * (b.lamN[1] * E.at(1, i + 1, j) - b.lam[1] * E.at(1, i, j)) / b.area;
            const double u1 = +0.5 * etay * (b.lamN[0] * E.at(0, i, j + 1) - b.lam[0] * E.at(0, i, j)) / b.area;
            H.at(0, i, j) += u0;
            H.at(1, i, j) += u1;
        }
    }
    /* SCBoundaryCondition.hh:402-415 + TETMSolver.cpp:105-115 */
    void sc_ex()
    {
        const double dteps = dt / EPS_0;
        for (const BCell &b : geo->bcells) {
            if (b.area < MINAREA) continue;
            const int i = b.i, j = b.j;
            if (b.lam[0] > SPACE_EPS) {
                const double u0 = (H.at(0, i, j) - H.at(0, i, j - 1)) / hr[1];
                const double u1 = (H.at(1, i, j) - H.at(1, i, j - 1)) / hr[1];
                E.at(0, i, j) += dteps * (u0 + u1 - J.at(0, i, j));
            }
        }
    }
    /* SCBoundaryCondition.hh:417-431 + TETMSolver.cpp:117-127 */
    void sc_ey()
    {
        const double dteps = dt / EPS_0;
        for (const BCell &b : geo->bcells) {
            if (b.area < MINAREA) continue;
            const int i = b.i, j = b.j;
            if (b.lam[1] > SPACE_EPS) {
                const double u0 = -((H.at(0, i, j) - H.at(0, i - 1, j))) / hr[0];
                const double u1 = -((H.at(1, i, j) - H.at(1, i - 1, j))) / hr[0];
                E.at(1, i, j) += dteps * (u0 + u1 - J.at(1, i, j));
            }
        }
    }

    /* PartBunch::drift_particles, PartBunch.cpp:179-187 */
    void drift(double ddt)
    {
        const double full = ddt * C_LIGHT;
        for (int k = 0; k < N; ++k) {
            const V2 Pk = v2(P[2 * k], P[2 * k + 1]);
            const double g = (double)q_rsqrt((float)(1 + dot(Pk, Pk)));
            R[2 * k] += full * Pk.x * g;
            R[2 * k + 1] += full * Pk.y * g;
        }
    }

    /* PartBunch::push_particles, PartBunch.cpp:215-519, single rank */
    void push(int jcomp)
    {
        const double half = 0.5 * dt * C_LIGHT;
        for (int k = 0; k < N; ++k) {
            oldR[2 * k] = R[2 * k];
            oldR[2 * k + 1] = R[2 * k + 1];
            half_drift(&R[2 * k], &P[2 * k], half);
        }
        /* getFields: patch origins from the bunch domain, PartBunch.cpp:957-1074 */
        int pd[2][2];
        pdomain(false, pd);
        const V2 po = v2(mmOrigin[0] + pd[0][0] * hr[0], mmOrigin[1] + pd[1][0] * hr[1]);
        const V2 dx = v2(hr[0], hr[1]);
        const V2 oHz = v2(po.x + mg.cell[0] * dx.x, po.y + mg.cell[0] * dx.y);
        const V2 oEx = v2(po.x + mg.edge[0][0] * dx.x, po.y + mg.edge[1][0] * dx.y);
        const V2 oEy = v2(po.x + mg.edge[1][0] * dx.x, po.y + mg.edge[0][0] * dx.y);
        const int fHz[2] = {pd[0][0] + mg.cell[0], pd[1][0] + mg.cell[0]};
        const int fEx[2] = {pd[0][0] + mg.edge[0][0], pd[1][0] + mg.edge[1][0]};
        const int fEy[2] = {pd[0][0] + mg.edge[1][0], pd[1][0] + mg.edge[0][0]};
        for (int d = 0; d < 2; ++d) {
            if (pd[d][0] + std::min(mg.cell[0], std::min(mg.edge[0][0], mg.edge[1][0])) < -1 ||
                pd[d][1] + std::max(mg.cell[1], std::max(mg.edge[0][1], mg.edge[1][1])) > (d == 0 ? W : Hh))
                throw std::runtime_error("gather patch leaves the window");
        }
        auto fEX = [&](long long I, long long Jj) { return mE.at(0, fEx[0] + (int)I, fEx[1] + (int)Jj); };
        auto fEY = [&](long long I, long long Jj) { return mE.at(1, fEy[0] + (int)I, fEy[1] + (int)Jj); };
        auto fHZ = [&](long long I, long long Jj) {
            return mH.at(0, fHz[0] + (int)I, fHz[1] + (int)Jj) + mH.at(1, fHz[0] + (int)I, fHz[1] + (int)Jj);
        };
        for (int k = 0; k < N; ++k)
            kick_drift(p.order, p.s2bug, *geo, dt, qm, fEX, fEY, fHZ, oEx, oEy, oHz, dx, &R[2 * k], &P[2 * k]);
        deposit(jcomp);
    }

    /* getCurrentDensityBaseImpl + getCurrentDensityImpl: zero the window
     * component, accumulate the patch over particles in index order, write the
     * patch into the window domain.  PartBunch.cpp:393-508 */
    void deposit(int comp)
    {
        int pd[2][2];
        pdomain(true, pd);
        gb[0][0] = pd[0][0]; gb[0][1] = pd[0][1];
        gb[1][0] = pd[1][0]; gb[1][1] = pd[1][1];
        const int a0 = comp == 0 ? 0 : 1, a1 = comp == 0 ? 1 : 0;
        const int lo[2] = {pd[0][0] + mg.cur[a0][0], pd[1][0] + mg.cur[a1][0]};
        const int hi[2] = {pd[0][1] + mg.cur[a0][1], pd[1][1] + mg.cur[a1][1]};
        const int pnx = hi[0] - lo[0] + 1, pny = hi[1] - lo[1] + 1;
        std::vector<double> patch((size_t)pnx * pny, 0.0);
        auto acc = [&](int I, int Jj, double v) {
            const int ri = I - lo[0], rj = Jj - lo[1];
            if (ri < 0 || rj < 0 || ri >= pnx || rj >= pny) throw std::runtime_error("deposit outside patch");
            patch[(size_t)rj * pnx + ri] += v;
        };
        const V2 org = v2(mmOrigin[0], mmOrigin[1]);
        const V2 dx = v2(hr[0], hr[1]);
        for (int k = 0; k < N; ++k)
            deposit_one(p.order, comp, acc, org, dx, dt, v2(oldR[2 * k], oldR[2 * k + 1]), Rp(k), Q[k]);
        /* JFD[lDom](c) -= JFD[lDom](c) */
        for (int j = 0; j < Hh; ++j)
            for (int i = 0; i < W; ++i) mJ.at(comp, i, j) = mJ.at(comp, i, j) - mJ.at(comp, i, j);
        for (int j = std::max(0, lo[1]); j <= std::min(Hh - 1, hi[1]); ++j)
            for (int i = std::max(0, lo[0]); i <= std::min(W - 1, hi[0]); ++i)
                mJ.at(comp, i, j) = patch[(size_t)(j - lo[1]) * pnx + (i - lo[0])];
    }

    /* main.cpp:562-652 */
    void move_mesh()
    {
        int ideal[2][2];
        const int len[2] = {W, Hh};
        for (int d = 0; d < 2; ++d) {
            const int glen = gb[d][1] - gb[d][0] + 1;
            if (glen >= len[d]) throw std::runtime_error("bunch no longer fits the window");
            const unsigned diff = (unsigned)(len[d] - glen);
            const unsigned upper = (unsigned)std::floor(diff / 2.0);
            const unsigned lower = diff - upper;
            ideal[d][0] = (int)lower - 1;
            ideal[d][1] = len[d] - (int)upper - 2;
        }
        const int moveI = gb[0][1] - ideal[0][1] - 1;
        const int maxJ = std::max(0, gb[1][1] - ideal[1][1] - 1);
        const int minJ = std::min(0, gb[1][1] - ideal[1][1] - 1);
        if (moveI < 0 || moveI > 1 || maxJ > 1 || minJ < -1)
            throw std::runtime_error("window move outside {0,1}x{-1,0,1}");
        int iI[2] = {0 - moveI, W - 1};
        int iJ[2] = {0 - maxJ, Hh - 1 - minJ};
        if (minJ + maxJ == 0) { iJ[0] -= 1; iJ[1] += 1; }
        const int sj = minJ + maxJ;
        const VField oE = mE, oH = mH, oJ = mJ;
        for (int j = iJ[0]; j <= iJ[1]; ++j)
            for (int i = iI[0]; i <= iI[1]; ++i)
                for (int k = 0; k < 2; ++k) {
                    mE.at(k, i, j) = oE.at(k, i + moveI, j + sj);
                    mH.at(k, i, j) = oH.at(k, i + moveI, j + sj);
                    mJ.at(k, i, j) = oJ.at(k, i + moveI, j + sj);
                }
        ddim[0][0] += moveI; ddim[0][1] += moveI;
        ddim[1][0] += sj; ddim[1][1] += sj;
        mmOrigin[0] += moveI * hr[0];
        mmOrigin[1] += sj * hr[1];
        const int pos[2] = {ddim[0][0], ddim[1][0]};
        const int mv[2] = {moveI, sj};
        exchange_moving(pos, mv);
    }

    /* PartBunch::calcMoments rmean/pmean (Kahan in long double), PartBunch.cpp:659-767 */
    void moments(double rmean[2], double pmean[2]) const
    {
        long double s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
        for (int k = 0; k < N; ++k) {
            const double part[4] = {R[2 * k], P[2 * k], R[2 * k + 1], P[2 * k + 1]};
            for (int m = 0; m < 4; ++m) {
                const long double y = part[m] - c[m];
                const long double t = s[m] + y;
                c[m] = (t - s[m]) - y;
                s[m] = t;
            }
        }
        rmean[0] = (double)(s[0]) / N; pmean[0] = (double)(s[1]) / N;
        rmean[1] = (double)(s[2]) / N; pmean[1] = (double)(s[3]) / N;
    }

    /* IntegrateElongCommand::execute, IntegrateElong.cpp:38-89.  The IPPL
     * IntCIC gather is taken as ScatterGatherImpl<IntCIC> on the window E. */
    void wake_sample()
    {
        double Rm[2], Pm[2];
        moments(Rm, Pm);
        const double lenP = std::sqrt(Pm[0] * Pm[0] + Pm[1] * Pm[1]);
        const double cosP = Pm[0] / lenP, sinP = Pm[1] / lenP;
        const V2 org = v2(mmOrigin[0], mmOrigin[1]);
        const V2 dx = v2(hr[0], hr[1]);
        auto fEX = [&](long long I, long long Jj) { return mE.at(0, (int)I, (int)Jj); };
        auto fEY = [&](long long I, long long Jj) { return mE.at(1, (int)I, (int)Jj); };
        for (int i = 2; i <= W - 1 - 2; ++i) {
            const double l0 = mmOrigin[0] + i * hr[0] - Rm[0];
            /* dot(M, l) + R with M = (cosP, sinP; -sinP, cosP), l = (l0, 0) */
            const double x = (cosP * l0 + sinP * 0.0) + Rm[0];
            const double y = (-sinP * l0 + cosP * 0.0) + Rm[1];
            const int ci = (int)std::floor((x - mmOrigin[0]) / hr[0]);
            const int cj = (int)std::floor((y - mmOrigin[1]) / hr[1]);
            if (ci < 0 || ci > W - 1 || cj < 0 || cj > Hh - 1) continue;
            const double ex = gather_edge(0, 0, 0, fEX, org, dx, v2(x, y));
            const double ey = gather_edge(0, 1, 0, fEY, org, dx, v2(x, y));
            const int ii = i - 2;
            const long double yv = (long double)(cosP * ex + sinP * ey) - wcor[ii];
            const long double t = wsum[ii] + yv;
            wcor[ii] = (t - wsum[ii]) - yv;
            wsum[ii] = t;
        }
    }

    /* the global-grid part of one step (TETMSolver.hh:142-249 without the
     * window): used to check field sweeps on their own */
    void field_step()
    {
        h_half_global();
        sc_hz();
        e_x_global();
        sc_ex();
        h_half_global();
        sc_hz();
        e_y_global();
        sc_ey();
    }
    void h_half_global()
    {
        const double etax = C_LIGHT * dt / hr[0];
        const double etay = C_LIGHT * dt / hr[1];
        const double Z0 = std::sqrt(MU_0 / EPS_0);
        const double sx = 0.5 * etax / Z0, sy = 0.5 * etay / Z0;
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i) {
                if (!geo->mask[(size_t)j * p.Nx + i]) continue;
                H.at(0, i, j) = H.at(0, i, j) - sx * (E.at(1, i + 1, j) - E.at(1, i, j));
                H.at(1, i, j) = H.at(1, i, j) + sy * (E.at(0, i, j + 1) - E.at(0, i, j));
            }
    }
    void e_x_global()
    {
        const double etay = C_LIGHT * dt / hr[1];
        const double Z0 = std::sqrt(MU_0 / EPS_0);
        const double dteps = C_LIGHT * dt * Z0;
        const double s = etay * Z0;
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i) {
                if (!geo->mask[(size_t)j * p.Nx + i]) continue;
                E.at(0, i, j) = E.at(0, i, j) +
                    (s * (((H.at(0, i, j) - H.at(0, i, j - 1)) + H.at(1, i, j)) - H.at(1, i, j - 1)) - dteps * J.at(0, i, j));
            }
    }
    void e_y_global()
    {
        const double etax = C_LIGHT * dt / hr[0];
        const double Z0 = std::sqrt(MU_0 / EPS_0);
        const double dteps = C_LIGHT * dt * Z0;
        const double s = etax * Z0;
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i) {
                if (!geo->mask[(size_t)j * p.Nx + i]) continue;
                E.at(1, i, j) = E.at(1, i, j) -
                    (s * (((H.at(0, i, j) - H.at(0, i - 1, j)) + H.at(1, i, j)) - H.at(1, i - 1, j)) + dteps * J.at(1, i, j));
            }
    }

    /* PartBunch::scatterQ onto the global vertex grid (ShapeFunction
     * getChargeDensity, charge_one above); out-of-grid deposits are dropped. */
    void scatter_q(std::vector<double> &rho) const
    {
        rho.assign((size_t)p.Nx * p.Ny, 0.0);
        auto acc = [&](int I, int Jj, double v) {
            if (I < 0 || Jj < 0 || I >= p.Nx || Jj >= p.Ny) return;
            rho[(size_t)Jj * p.Nx + I] += v;
        };
        for (int k = 0; k < N; ++k)
            charge_one(p.order, acc, v2(geo->origin[0], geo->origin[1]), v2(hr[0], hr[1]), Rp(k), Q[k]);
        for (double &v : rho) v = 0.0 + v / (hr[0] * hr[1]);
    }

    /* PoissonSolver::computeField (PoissonSolver.cpp:85-201) followed by
     * SCBoundaryCondition::resetOutsideDomain (SCBoundaryCondition.hh:164-196),
     * as main.cpp:229-247 does.  gamma from the mean momentum, a plain sum
     * over particles / N (Distribution.cpp:44-52, PoissonSolver.cpp:64-66);
     * hx scaled by gamma (PoissonSolver.cpp:71). */
    int poisson(double gamma_override)
    {
        double pm[2] = {0.0, 0.0};
        for (int k = 0; k < N; ++k) { pm[0] += P[2 * k]; pm[1] += P[2 * k + 1]; }
        pm[0] /= N; pm[1] /= N;
        const double gamma = gamma_override > 0 ? gamma_override : std::sqrt(pm[0] * pm[0] + pm[1] * pm[1] + 1);
        const double hr0 = hr[0] * gamma, hr1 = hr[1];
        std::vector<uint8_t> flag, active;
        poisson_classes(*geo, flag, active);
        const size_t nn = (size_t)p.Nx * p.Ny;
        std::vector<double> rho, rhs(nn), phi(nn);
        const double cc = 1 / (EPS_0 * gamma);               /* fillRHS, :451 */
        auto solve = [&]() {
            for (size_t k = 0; k < nn; ++k) rhs[k] = rho[k] * cc;
            return poisson_direct(p.Nx, p.Ny, hr0, hr1, flag.data(), active.data(), rhs.data(), phi.data());
        };
        auto ph = [&](int i, int j) { return (i < p.Nx && j < p.Ny) ? phi[(size_t)j * p.Nx + i] : 0.0; };
        const std::vector<double> saved = R;                 /* saveR */
        /* solve 1: Ey from rho at t + dt/2 (:154-166, fillFields case 1) */
        drift(dt / 2);
        scatter_q(rho);
        drift(-dt / 2);
        if (solve()) return -1;
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i)
                E.at(1, i, j) = gamma * ph(i, j) / hr1 - gamma * ph(i, j + 1) / hr1;
        /* solve 2: Hz, bunch moved back half a cell (:167-178, case 2) */
        for (int k = 0; k < N; ++k) R[2 * k] = R[2 * k] + (-0.5 * hr0 / gamma);
        scatter_q(rho);
        if (solve()) return -1;
        const double betaGamma = std::sqrt(gamma * gamma - 1);
        const double den = 2 * C_LIGHT * MU_0 * hr1;
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i) {
                const double h = betaGamma * ph(i, j) / den - betaGamma * ph(i, j + 1) / den;
                H.at(0, i, j) = h;
                H.at(1, i, j) = h;
            }
        /* solve 3: Ex (:179-195, case 0) */
        R = saved;                                           /* restoreR */
        scatter_q(rho);
        if (solve()) return -1;
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i) E.at(0, i, j) = ph(i, j) / hr0 - ph(i + 1, j) / hr0;
        /* resetOutsideDomain */
        for (int j = 0; j < p.Ny; ++j)
            for (int i = 0; i < p.Nx; ++i) {
                const size_t id = (size_t)j * p.Nx + i;
                if (geo->mask[id]) continue;
                const int bc = geo->bmap[id];
                if (bc == -1) { E.at(0, i, j) = 0.0; E.at(1, i, j) = 0.0; continue; }
                const BCell &c = geo->bcells[bc];
                if (c.lam[0] < SPACE_EPS) E.at(0, i, j) = 0.0;
                if (c.lam[1] < SPACE_EPS) E.at(1, i, j) = 0.0;
                if (c.area < MINAREA) { E.at(0, i, j) = 0.0; E.at(1, i, j) = 0.0; }
            }
        /* the window copy and exchanges of main.cpp:335-341 */
        copy_to_window();
        const int pos[2] = {ddim[0][0], ddim[1][0]};
        exchange(OP_H, pos);
        exchange(OP_EX, pos);
        exchange(OP_EY, pos);
        return 0;
    }

    /* One part of the step (the order of step() below), for bisection:
     * 0 push Jx, 1 push Jy, 2 H half step, 3 Ex, 4 Ey, 5 exchange H,
     * 6 exchange Ex, 7 exchange Ey, 8 SC Hz, 9 SC Ex, 10 SC Ey,
     * 11 drift -dt/2, 12 wake sample, 13 J = 0, 14 moveMesh (+ step count). */
    int substep(int op)
    {
        const int pos[2] = {ddim[0][0], ddim[1][0]};
        switch (op) {
        case 0: push(0); break;
        case 1: push(1); break;
        case 2: h_half(); break;
        case 3: e_x(); break;
        case 4: e_y(); break;
        case 5: exchange(OP_H, pos); break;
        case 6: exchange(OP_EX, pos); break;
        case 7: exchange(OP_EY, pos); break;
        case 8: sc_hz(); break;
        case 9: sc_ex(); break;
        case 10: sc_ey(); break;
        case 11: drift(-0.5 * dt); break;
        case 12: wake_sample(); break;
        case 13: J.c[0].assign(J.c[0].size(), 0.0); J.c[1].assign(J.c[1].size(), 0.0); break;
        case 14: move_mesh(); ++steps; break;
        default: return -1;
        }
        return 0;
    }

    /* main.cpp:390-506, one iteration */
    void step()
    {
        const int pos[2] = {ddim[0][0], ddim[1][0]};
        /* updateField_firstStep, TETMSolver.hh:110-211 */
        push(0);
        h_half();
        exchange(OP_H, pos);
        sc_hz();
        e_x();
        exchange(OP_EX, pos);
        sc_ex();
        h_half();
        exchange(OP_H, pos);
        sc_hz();
        drift(-0.5 * dt);
        /* updateField_secondStep, TETMSolver.hh:213-266 */
        push(1);
        e_y();
        exchange(OP_EY, pos);
        sc_ey();
        drift(-0.5 * dt);
        if (p.statModulus > 0 && steps % p.statModulus == p.statModulus - 1) wake_sample();
        J.c[0].assign(J.c[0].size(), 0.0); J.c[1].assign(J.c[1].size(), 0.0);
        move_mesh();
        ++steps;
    }
};

template <class T>
T *as(void *p) { return reinterpret_cast<T *>(p); }

} // namespace

/* ======================================================================= */
// ---------------------------------------------------------------------------
// Beam parameters: PartBunch::calcMoments (PartBunch.cpp:659-756) and
// PartBunch::calcBeamParameters (PartBunch.cpp:584-657) on one rank, with
// the Kahan sums in long double of utils.hh:512-521 and the reference's mix
// of double and long double arithmetic.  out: Ekin (MeV), rrms(2), prms(2),
// normalised emittance(2), rmean(2), pmean(2), Dy, DDy, dE, rmin(2), rmax(2),
// rprms(2).
namespace {
struct KahanLD {
    long double sum = 0, corr = 0;
    void add(double v)
    {
        const long double y = v - corr;
        const long double t = sum + y;
        corr = (t - sum) - y;
        sum = t;
    }
};
}  // namespace

void beam_parameters(const double *R, const double *P, int n, double out[20])
{
    for (int i = 0; i < 20; ++i) out[i] = 0;
    if (n <= 0) return;
    const double N = n;
    const double m0 = 0.51099906e-03 * 1000;   // Physics.cpp:51, PartBunch.cpp:588
    // calcMoments: centroid sums and limits
    KahanLD sums[4];
    double lim[4] = {R[0], R[1], R[0], R[1]};   // min x, min y, max x, max y
    for (int k = 0; k < n; ++k) {
        const double part[4] = {R[2 * k], P[2 * k], R[2 * k + 1], P[2 * k + 1]};
        for (int d = 0; d < 2; ++d) {
            lim[d] = std::min(lim[d], R[2 * k + d]);
            lim[2 + d] = std::max(lim[2 + d], R[2 * k + d]);
        }
        for (int i = 0; i < 4; ++i) sums[i].add(part[i]);
    }
    long double centroid[4];
    for (int i = 0; i < 4; ++i) centroid[i] = sums[i].sum;
    const double avgR[2] = {(double)(centroid[0] / n), (double)(centroid[2] / n)};
    const double avgP[2] = {(double)(centroid[1] / n), (double)(centroid[3] / n)};
    const double avgmom = std::sqrt(avgP[0] * avgP[0] + avgP[1] * avgP[1]);
    const double cosphi = avgP[0] / avgmom, sinphi = avgP[1] / avgmom;
    KahanLD sq[16];
    for (int k = 0; k < n; ++k) {
        long double part[4];
        part[1] = cosphi * P[2 * k] + sinphi * P[2 * k + 1];
        part[3] = -sinphi * P[2 * k] + cosphi * P[2 * k + 1];
        part[0] = R[2 * k] - avgR[0];
        part[2] = R[2 * k + 1] - avgR[1];
        const double tmp = (double)(cosphi * part[0] + sinphi * part[2]);
        part[2] = -sinphi * part[0] + cosphi * part[2];
        part[0] = tmp;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j <= i; ++j) sq[i * 4 + j].add((double)(part[i] * part[j]));
    }
    auto mom = [&](int i, int j) -> long double { return i >= j ? sq[i * 4 + j].sum : sq[j * 4 + i].sum; };
    // calcBeamParameters
    double rmean[2], pmean[2];
    for (int i = 0; i < 2; ++i) {
        rmean[i] = (double)(centroid[2 * i] / N);
        pmean[i] = (double)(centroid[2 * i + 1] / N);
    }
    centroid[1] = pmean[0] * pmean[0] + pmean[1] * pmean[1];
    centroid[3] = 0.0;
    double rsq[2], psq[2], rp[2];
    for (int i = 0; i < 2; ++i) {
        rsq[i] = (double)mom(2 * i, 2 * i);
        psq[i] = (double)(mom(2 * i + 1, 2 * i + 1) - N * centroid[2 * i + 1]);
        if (psq[i] < 0) psq[i] = 0;
        rp[i] = (double)mom(2 * i, 2 * i + 1);
    }
    for (int i = 0; i < 2; ++i) {
        const double eps2 = (rsq[i] * psq[i] - rp[i] * rp[i]) / (N * N);
        const double rpm = rp[i] / N;
        const double rrms = std::sqrt(rsq[i] / N), prms = std::sqrt(psq[i] / N);
        out[1 + i] = rrms;
        out[3 + i] = prms;
        out[5 + i] = std::sqrt(std::max(eps2, 0.0));
        const double t = rrms * prms;
        out[18 + i] = rpm * (t == 0 ? 0.0 : 1.0 / t);
        out[7 + i] = rmean[i];
        out[9 + i] = pmean[i];
    }
    out[11] = (double)(mom(2, 1) / N);
    out[12] = (double)(mom(3, 1) / N);
    double gamma = 0.0;
    for (int k = 0; k < n; ++k) gamma += std::sqrt(1.0 + (P[2 * k] * P[2 * k] + P[2 * k + 1] * P[2 * k + 1]));
    gamma /= N;
    out[0] = (gamma - 1.0) * m0;
    const double ek = out[0];
    out[13] = out[3] * std::sqrt(ek * (ek + 2. * m0) / (1. + ek * (ek + 2. * m0) / (m0 * m0)));
    out[14] = lim[0]; out[15] = lim[1]; out[16] = lim[2]; out[17] = lim[3];
}

extern "C" {

float orc_q_rsqrt(float x) { return q_rsqrt(x); }

int orc_crossings(int dual, const double oldR[2], const double newR[2], const double dx[2],
                  const double origin[2], double *frac, double *sub)
{
    Xing xs[10];
    const int n = dual ? dual_crossings(xs, v2(oldR[0], oldR[1]), v2(newR[0], newR[1]), v2(dx[0], dx[1]), v2(origin[0], origin[1]))
                       : cell_crossings(xs, v2(oldR[0], oldR[1]), v2(newR[0], newR[1]), v2(dx[0], dx[1]), v2(origin[0], origin[1]));
    for (int k = 0; k < n; ++k) {
        frac[k] = xs[k].f;
        sub[2 * k] = xs[k].s.x;
        sub[2 * k + 1] = xs[k].s.y;
    }
    return n;
}

void orc_gather_edge(int order, int comp, int s2bug, const double *patch, int pnx, int pny,
                     const double origin[2], const double dx[2], const double *R, int n, double *out)
{
    auto fp = [&](long long I, long long J) {
        if (I < 0 || J < 0 || I >= pnx || J >= pny) throw std::out_of_range("patch");
        return patch[(size_t)J * pnx + I];
    };
    for (int k = 0; k < n; ++k)
        out[k] = gather_edge(order, comp, s2bug, fp, v2(origin[0], origin[1]), v2(dx[0], dx[1]), v2(R[2 * k], R[2 * k + 1]));
}

void orc_gather_cell(int order, int s2bug, const double *patch, int pnx, int pny,
                     const double origin[2], const double dx[2], const double *R, int n, double *out)
{
    auto fp = [&](long long I, long long J) {
        if (I < 0 || J < 0 || I >= pnx || J >= pny) throw std::out_of_range("patch");
        return patch[(size_t)J * pnx + I];
    };
    for (int k = 0; k < n; ++k)
        out[k] = gather_cell(order, s2bug, fp, v2(origin[0], origin[1]), v2(dx[0], dx[1]), v2(R[2 * k], R[2 * k + 1]));
}

void orc_scatter_edge(int order, int comp, double *patch, int pnx, int pny, const double origin[2],
                      const double dx[2], const double *R, const double *val, int n)
{
    auto acc = [&](long long I, long long J, double v) {
        if (I < 0 || J < 0 || I >= pnx || J >= pny) throw std::out_of_range("patch");
        patch[(size_t)J * pnx + I] += v;
    };
    for (int k = 0; k < n; ++k)
        scatter_edge(order, comp, acc, v2(origin[0], origin[1]), v2(dx[0], dx[1]), v2(R[2 * k], R[2 * k + 1]), val[k]);
}

void orc_scatter_cell(int order, double *patch, int pnx, int pny, const double origin[2],
                      const double dx[2], const double *R, const double *val, int n)
{
    auto acc = [&](long long I, long long J, double v) {
        if (I < 0 || J < 0 || I >= pnx || J >= pny) throw std::out_of_range("patch");
        patch[(size_t)J * pnx + I] += v;
    };
    for (int k = 0; k < n; ++k)
        scatter_cell(order, acc, v2(origin[0], origin[1]), v2(dx[0], dx[1]), v2(R[2 * k], R[2 * k + 1]), val[k]);
}

int orc_deposit(int order, int comp, double *J, int i0, int j0, int pnx, int pny, const double origin[2],
                const double dx[2], double dt, const double *oldR, const double *R, const double *Q, int n)
{
    bool bad = false;
    auto acc = [&](int I, int Jj, double v) {
        const int ri = I - i0, rj = Jj - j0;
        if (ri < 0 || rj < 0 || ri >= pnx || rj >= pny) { bad = true; return; }
        J[(size_t)rj * pnx + ri] += v;
    };
    for (int k = 0; k < n; ++k)
        deposit_one(order, comp, acc, v2(origin[0], origin[1]), v2(dx[0], dx[1]), dt,
                    v2(oldR[2 * k], oldR[2 * k + 1]), v2(R[2 * k], R[2 * k + 1]), Q[k]);
    return bad ? -1 : 0;
}

int orc_beam_parameters(const double *R, const double *P, int n, double *out)
{
    if (!R || !P || !out || n < 0) return -1;
    beam_parameters(R, P, n, out);
    return 0;
}

int orc_charge_density(int order, double *rho, int i0, int j0, int pnx, int pny, const double origin[2],
                       const double dx[2], const double *R, const double *Q, int n)
{
    bool bad = false;
    std::vector<double> patch((size_t)pnx * pny, 0.0);
    auto acc = [&](int I, int Jj, double v) {
        const int ri = I - i0, rj = Jj - j0;
        if (ri < 0 || rj < 0 || ri >= pnx || rj >= pny) { bad = true; return; }
        patch[(size_t)rj * pnx + ri] += v;
    };
    for (int k = 0; k < n; ++k)
        charge_one(order, acc, v2(origin[0], origin[1]), v2(dx[0], dx[1]), v2(R[2 * k], R[2 * k + 1]), Q[k]);
    for (size_t m = 0; m < patch.size(); ++m) rho[m] = 0.0 + patch[m] / (dx[0] * dx[1]);
    return bad ? -1 : 0;
}

void *orc_bend_create(double lb, double la, double w, double phiDeg, double ekin, double bz, int Nx, int Ny)
{
    try {
        return new Bend(lb, la, w, phiDeg, ekin, bz, Nx, Ny);
    } catch (...) {
        return nullptr;
    }
}
void orc_bend_destroy(void *b) { delete as<Bend>(b); }
void orc_bend_info(void *bp, double *o)
{
    Bend *b = as<Bend>(bp);
    o[0] = b->hr[0]; o[1] = b->hr[1]; o[2] = b->origin[0]; o[3] = b->origin[1];
    o[4] = b->R; o[5] = b->phi; o[6] = b->center[0]; o[7] = b->center[1];
    o[8] = b->dlen; o[9] = b->dwid;
}
int orc_bend_is_inside(void *b, double x, double y) { return as<Bend>(b)->isInside(x, y) ? 1 : 0; }
void orc_bend_mask(void *bp, uint8_t *m)
{
    Bend *b = as<Bend>(bp);
    std::memcpy(m, b->mask.data(), b->mask.size());
}
int orc_bend_num_bcells(void *b) { return (int)as<Bend>(b)->bcells.size(); }
void orc_bend_bcells(void *bp, int *ij, double *data)
{
    Bend *b = as<Bend>(bp);
    for (size_t k = 0; k < b->bcells.size(); ++k) {
        const BCell &c = b->bcells[k];
        ij[2 * k] = c.i; ij[2 * k + 1] = c.j;
        data[5 * k] = c.lam[0]; data[5 * k + 1] = c.lam[1]; data[5 * k + 2] = c.area;
        data[5 * k + 3] = c.lamN[0]; data[5 * k + 4] = c.lamN[1];
    }
}
void orc_bend_external_field(void *bp, const double *R, int n, double *out)
{
    Bend *b = as<Bend>(bp);
    for (int k = 0; k < n; ++k) out[k] = b->external_field(R[2 * k], R[2 * k + 1]);
}

void *orc_sim_create(const orc_sim_params *p)
{
    try {
        Sim *s = new Sim();
        s->setup(p);
        return s;
    } catch (...) {
        return nullptr;
    }
}
void orc_sim_destroy(void *s)
{
    Sim *sm = as<Sim>(s);
    if (!sm) return;
    delete sm->geo;
    delete sm;
}
int orc_sim_init(void *s, const double *R, const double *P, const double *Q, int n)
{
    try {
        return as<Sim>(s)->init(R, P, Q, n);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "oracle init: %s\n", e.what());
        return -1;
    }
}
int orc_sim_restore(void *s, const double *R, const double *oldR, const double *P, const double *Q, int n, int step)
{
    try {
        return as<Sim>(s)->init(R, P, Q, n, oldR, step);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "oracle restore: %s\n", e.what());
        return -1;
    }
}
int orc_sim_step(void *s, int nsteps)
{
    try {
        for (int k = 0; k < nsteps; ++k) as<Sim>(s)->step();
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "oracle step: %s\n", e.what());
        return -1;
    }
}
void orc_sim_info(void *sp, double *info)
{
    Sim *s = as<Sim>(sp);
    info[0] = s->dt; info[1] = s->mmOrigin[0]; info[2] = s->mmOrigin[1];
    info[3] = s->ddim[0][0]; info[4] = s->ddim[1][0]; info[5] = s->W; info[6] = s->Hh;
    info[7] = s->steps; info[8] = s->nsamp;
}
static VField *pick(Sim *s, int which)
{
    switch (which) {
    case 0: return &s->E;
    case 1: return &s->H;
    case 2: return &s->mE;
    case 3: return &s->mH;
    case 4: return &s->mJ;
    case 5: return &s->J;
    default: return nullptr;
    }
}
int orc_sim_get_field(void *sp, int which, double *out)
{
    VField *f = pick(as<Sim>(sp), which);
    if (!f) return -1;
    const size_t sz = f->c[0].size();
    for (size_t m = 0; m < sz; ++m) {
        out[2 * m] = f->c[0][m];
        out[2 * m + 1] = f->c[1][m];
    }
    return (int)(2 * sz);
}
int orc_sim_set_field(void *sp, int which, const double *in)
{
    VField *f = pick(as<Sim>(sp), which);
    if (!f) return -1;
    const size_t sz = f->c[0].size();
    for (size_t m = 0; m < sz; ++m) {
        f->c[0][m] = in[2 * m];
        f->c[1][m] = in[2 * m + 1];
    }
    return (int)(2 * sz);
}
int orc_sim_field_steps(void *sp, int nsteps)
{
    try {
        for (int k = 0; k < nsteps; ++k) as<Sim>(sp)->field_step();
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "oracle field steps: %s\n", e.what());
        return -1;
    }
}

void orc_sim_get_particles(void *sp, double *R, double *P, double *oldR)
{
    Sim *s = as<Sim>(sp);
    std::memcpy(R, s->R.data(), sizeof(double) * 2 * s->N);
    std::memcpy(P, s->P.data(), sizeof(double) * 2 * s->N);
    if (oldR) std::memcpy(oldR, s->oldR.data(), sizeof(double) * 2 * s->N);
}
int orc_sim_get_wake(void *sp, double *out)
{
    Sim *s = as<Sim>(sp);
    for (int k = 0; k < s->nsamp; ++k) out[k] = (double)s->wsum[k];
    return s->nsamp;
}
void orc_sim_diag(void *sp, double *out)
{
    Sim *s = as<Sim>(sp);
    double Rm[2], Pm[2];
    s->moments(Rm, Pm);
    out[0] = Rm[0]; out[1] = Rm[1]; out[2] = Pm[0]; out[3] = Pm[1];
    long double qs = 0, qc = 0;
    for (int k = 0; k < s->N; ++k) {
        const long double y = s->Q[k] - qc;
        const long double t = qs + y;
        qc = (t - qs) - y;
        qs = t;
    }
    out[4] = (double)qs;
    /* CalcEnergyCommand: 1/2 eps0 E^2 + 1/2 mu0 H^2 over the global grid, per unit depth */
    long double en = 0;
    for (int j = 0; j < s->p.Ny; ++j)
        for (int i = 0; i < s->p.Nx; ++i) {
            const double ex = s->E.at(0, i, j), ey = s->E.at(1, i, j);
            const double hz = s->H.at(0, i, j) + s->H.at(1, i, j);
            en += 0.5 * EPS_0 * (ex * ex + ey * ey) + 0.5 * MU_0 * hz * hz;
        }
    out[5] = (double)(en * s->hr[0] * s->hr[1]);
}


/* CalcEnergyCommand::execute (CalcEnergy.cpp:60-124): sums of E.E and
 * (Hzx+Hzy)^2 over the global grid plus the window minus the global nodes
 * under the window, times the cell area */
void orc_sim_field_energy(void *sp, double *out)
{
    Sim *s = as<Sim>(sp);
    long double e2 = 0, h2 = 0;
    auto add = [&](const VField &E, const VField &H, int i0, int j0, int i1, int j1, int sign) {
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) {
                const double ex = E.at(0, i, j), ey = E.at(1, i, j);
                const double hz = H.at(0, i, j) + H.at(1, i, j);
                e2 += sign * (long double)(ex * ex + ey * ey);
                h2 += sign * (long double)(hz * hz);
            }
    };
    add(s->E, s->H, 0, 0, s->p.Nx - 1, s->p.Ny - 1, 1);
    add(s->mE, s->mH, 0, 0, s->W - 1, s->Hh - 1, 1);
    const int lo0 = std::max(0, s->ddim[0][0]), lo1 = std::max(0, s->ddim[1][0]);
    const int hi0 = std::min(s->p.Nx - 1, s->ddim[0][0] + s->W - 1), hi1 = std::min(s->p.Ny - 1, s->ddim[1][0] + s->Hh - 1);
    if (lo0 <= hi0 && lo1 <= hi1) add(s->E, s->H, lo0, lo1, hi0, hi1, -1);
    const long double area = (long double)s->hr[0] * s->hr[1];
    out[0] = (double)(e2 * area);
    out[1] = (double)(h2 * area);
}

int orc_poisson_classes(void *bp, uint8_t *flag, uint8_t *active)
{
    std::vector<uint8_t> f, a;
    Bend *b = as<Bend>(bp);
    poisson_classes(*b, f, a);
    std::memcpy(flag, f.data(), f.size());
    std::memcpy(active, a.data(), a.size());
    return (int)(size_t)(b->lenBefore / b->hr[0]);
}
int orc_poisson_solve(int nx, int ny, double hx, double hy, const uint8_t *flag, const uint8_t *active,
                      const double *b, double *phi)
{
    try {
        return poisson_direct(nx, ny, hx, hy, flag, active, b, phi);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "oracle poisson: %s\n", e.what());
        return -1;
    }
}
int orc_sim_poisson(void *s, double gamma)
{
    try {
        return as<Sim>(s)->poisson(gamma);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "oracle poisson: %s\n", e.what());
        return -1;
    }
}
/* SaveEMFieldsCommand::execute (SaveEMFields.cpp:34-160) on the global
 * fields, with getPartInside (:162-189) on a one-rank field domain.  The
 * state is PartBunch::calcBeamParameters (beam_parameters above): rmean,
 * pmean, rmin, rmax.  IPPL's ParticleAttrib::gather with IntOp_t is taken as
 * ScatterGatherImpl<IntOp_t> (edge semantics for E and J, cell semantics for
 * H, each split component gathered and the two H parts summed as :131).
 * Rows of 7: s, E along P, E across P, Hz, J along P, x, y (the file's
 * columns, :141-149).  Returns the number of rows (-1: cap too small). */
int orc_sim_line_fields(void *sp, double *out, int cap)
{
    Sim *s = as<Sim>(sp);
    double st[20];
    beam_parameters(s->R.data(), s->P.data(), s->N, st);
    const V2 R = v2(st[7], st[8]), P = v2(st[9], st[10]);
    const V2 h = v2(s->hr[0], s->hr[1]);
    const V2 origin = v2(s->geo->origin[0], s->geo->origin[1]);
    /* getCellContaining of rmin and rmax */
    const int gb0 = (int)std::floor((st[14] - origin.x) / h.x), gb1 = (int)std::floor((st[16] - origin.x) / h.x);
    V2 l, u;
    l.x = std::floor(1.5 * gb0 - 0.5 * gb1);
    u.x = 2 * gb1 - gb0;
    l.y = (R.y + (origin.x + l.x * h.x - R.x) / P.x * P.y - origin.y) / h.y;
    u.y = (R.y + (origin.x + u.x * h.x - R.x) / P.x * P.y - origin.y) / h.y;
    const long startIdx = (long)l.x;
    const unsigned numSamples = (unsigned)std::floor(u.x - l.x + 0.5) + 1;
    if ((int)numSamples > cap) return -1;
    std::vector<double> line(4 * (size_t)numSamples, 0.0);
    /* getPartInside: the first and last column whose line point lies in the field domain */
    V2 lc = l, uc = u;
    {
        const int lower = (int)std::floor(lc.x + 0.5), upper = (int)std::floor(uc.x + 0.5);
        int minL = upper, maxL = upper;
        const V2 proj = v2(uc.x - lc.x, uc.y - lc.y);
        for (int i = lower; i <= upper; ++i) {
            const double y = lc.y + (i - lower) / proj.x * proj.y;
            const int yi = (int)std::floor(y);
            if (i >= 0 && i < s->p.Nx && yi >= 0 && yi < s->p.Ny) {
                if (i < minL) minL = i;
                if (i > minL) maxL = i;        /* as written (:180) */
            }
        }
        const V2 tmp = v2(minL, lc.y + (minL - lower) / proj.x * proj.y);
        uc = v2(maxL, lc.y + (maxL - lower) / proj.x * proj.y);
        lc = tmp;
    }
    int dom[2][2] = {{(int)std::floor(lc.x + 0.5), (int)std::floor(uc.x + 0.5)},
                     {(int)std::floor(lc.y), (int)std::floor(uc.y)}};
    const int fdom[2][2] = {{0, s->p.Nx - 1}, {0, s->p.Ny - 1}};
    bool inside = true;
    for (int d = 0; d < 2; ++d) {
        const int lower = std::max(fdom[d][0], dom[d][0]), upper = std::min(fdom[d][1], dom[d][1]);
        if (lower > upper) { inside = false; break; }
        dom[d][0] = lower; dom[d][1] = upper;
    }
    if (inside) {
        auto fetch = [&](const VField &f, int k) {
            return [&f, k](long long I, long long J) {
                return f.in((int)I, (int)J) ? f.at(k, (int)I, (int)J) : 0.0;
            };
        };
        const double lenP = std::sqrt(dot(P, P));
        const double cosP = P.x / lenP, sinP = P.y / lenP;
        for (int i = dom[0][0]; i <= dom[0][1]; ++i) {
            const V2 pos = v2(origin.x + i * h.x, R.y + (origin.x + i * h.x - R.x) / P.x * P.y);
            const int o = s->p.order, bug = s->p.s2bug;
            const double ex = gather_edge(o, 0, bug, fetch(s->E, 0), origin, h, pos);
            const double ey = gather_edge(o, 1, bug, fetch(s->E, 1), origin, h, pos);
            const double h0 = gather_cell(o, bug, fetch(s->H, 0), origin, h, pos);
            const double h1 = gather_cell(o, bug, fetch(s->H, 1), origin, h, pos);
            const double jx = gather_edge(o, 0, bug, fetch(s->J, 0), origin, h, pos);
            const double jy = gather_edge(o, 1, bug, fetch(s->J, 1), origin, h, pos);
            const size_t ii = (size_t)(i - startIdx);
            line[4 * ii] = cosP * ex + sinP * ey;
            line[4 * ii + 1] = -sinP * ex + cosP * ey;
            line[4 * ii + 2] = h0 + h1;
            line[4 * ii + 3] = cosP * jx + sinP * jy;
        }
    }
    long ii = startIdx;
    for (unsigned i = 0; i < numSamples; ++i, ++ii) {
        const double c0 = origin.x + ii * h.x - R.x;
        const double c1 = c0 / P.x * P.y;
        double *row = out + 7 * (size_t)i;
        row[0] = std::sqrt(c0 * c0 + c1 * c1) * (c0 > 0 ? 1 : -1);
        row[1] = line[4 * i]; row[2] = line[4 * i + 1]; row[3] = line[4 * i + 2]; row[4] = line[4 * i + 3];
        row[5] = c0 + R.x;
        row[6] = c1 + R.y;
    }
    return (int)numSamples;
}

int orc_push_patch(int order, int s2bug, void *bend, double dt, const double *ex, const double *ey, const double *hz,
                   int pnx, int pny, const double origin[2], const double dx[2], double *R, double *P, double *oldR,
                   double *Rmid, int n, int driftBack)
{
    try {
        const Bend &geo = *as<Bend>(bend);
        const double qm = -1e-9 / M_E_GEV;           /* main.cpp:225 */
        const double half = 0.5 * dt * C_LIGHT;
        auto at = [&](const double *f, long long I, long long Jj) {
            if (I < 0 || Jj < 0 || I >= pnx || Jj >= pny) throw std::out_of_range("gather outside the patch");
            return f[(size_t)Jj * pnx + (size_t)I];
        };
        auto fEX = [&](long long I, long long Jj) { return at(ex, I, Jj); };
        auto fEY = [&](long long I, long long Jj) { return at(ey, I, Jj); };
        auto fHZ = [&](long long I, long long Jj) { return at(hz, I, Jj); };
        const V2 o = v2(origin[0], origin[1]), h = v2(dx[0], dx[1]);
        for (int k = 0; k < n; ++k) {
            if (oldR) { oldR[2 * k] = R[2 * k]; oldR[2 * k + 1] = R[2 * k + 1]; }
            half_drift(&R[2 * k], &P[2 * k], half);
            kick_drift(order, s2bug, geo, dt, qm, fEX, fEY, fHZ, o, o, o, h, &R[2 * k], &P[2 * k]);
            if (Rmid) { Rmid[2 * k] = R[2 * k]; Rmid[2 * k + 1] = R[2 * k + 1]; }
            if (driftBack) half_drift(&R[2 * k], &P[2 * k], -half);
        }
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "oracle push: %s\n", e.what());
        return -1;
    }
}

int orc_sim_substep(void *s, int op)
{
    try {
        return as<Sim>(s)->substep(op);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "oracle substep: %s\n", e.what());
        return -1;
    }
}
} /* extern "C" */
