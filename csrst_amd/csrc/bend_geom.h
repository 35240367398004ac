// bend_geom.h -- chamber geometry of the CSRST bend (straight section, arc,
// straight section) and the conformal cut-cell data derived from it.
//
// The functions are __host__ __device__: the two-grid step evaluates them on
// the host once at set-up; the moving-window step evaluates them on the device
// for every grid column that enters the window.  FP64 throughout (the wall
// positions need ~1e-9 relative resolution on a metre-scale bend radius).
#pragma once

#include "pic_math.h"

namespace geom {

// std::min / std::max semantics (first argument wins ties)
PM_HD double smin(double a, double b) { return (b < a) ? b : a; }
PM_HD double smax(double a, double b) { return (a < b) ? b : a; }

struct Bend {
    double width, lenBefore, lenAfter, bz;
    double R, phi;
    double cosphi, sinphi;      // cos(phi), sin(phi) (signed)
    double rmin_sq, rmax_sq;
    double center[2];
    double origin[2], hr[2];    // global mesh (two-grid step) -- Bend.cpp:85-124
    double dlen, dwid;
    // external field strip (Bend.cpp:865-907)
    double cph, sph, halfw, bendw;
};

// Bend.cpp:60-124.  Nx, Ny: global vertex mesh dimensions.
inline void bend_init(Bend &b, double lenBefore, double lenAfter, double width, double phiDeg, double ekinMeV,
                      double bz, int Nx, int Ny)
{
    const double mass = 0.511e6;
    const double ekin = ekinMeV * 1e6;
    b.width = width; b.lenBefore = lenBefore; b.lenAfter = lenAfter; b.bz = bz;
    if (fabs(bz) > 1e-8) {
        b.R = sqrt(ekin * ekin + 2. * ekin * mass) / (pm::kC * fabs(bz));
        b.phi = -bz / fabs(bz) * fabs(phiDeg) * pm::kPi / 180.;
    } else {
        b.R = 9999999.9;
        b.phi = 0.0;
    }
    const double cphi = cos(b.phi), sphi = fabs(sin(b.phi));
    if (fabs(bz) > 1e-8) {
        b.dlen = (lenBefore + lenAfter * cphi + (b.R + 0.5 * width) * sphi);
        b.dwid = ((b.R - 0.5 * width) * (1 - cphi) + width + lenAfter * sphi);
    } else {
        b.dlen = lenAfter + lenBefore;
        b.dwid = width;
    }
    b.hr[0] = b.dlen / (Nx - 2);
    b.hr[1] = b.dwid / (Ny - 2);
    if (bz <= 0.0) {
        b.origin[0] = 0 - b.hr[0];
        b.origin[1] = (width - b.dwid) - b.hr[1];
        b.center[0] = lenBefore; b.center[1] = 0.5 * width - b.R;
    } else {
        b.origin[0] = -b.hr[0];
        b.origin[1] = -b.hr[1];
        b.center[0] = lenBefore; b.center[1] = 0.5 * width + b.R;
    }
    b.dlen += b.hr[0];
    b.dwid += b.hr[1];
    b.cosphi = cos(b.phi);
    b.sinphi = sin(b.phi);
    b.rmin_sq = (b.R - 0.5 * width) * (b.R - 0.5 * width);
    b.rmax_sq = (b.R + 0.5 * width) * (b.R + 0.5 * width);
    b.cph = cos(0.5 * b.phi);
    b.sph = sin(0.5 * b.phi);
    b.halfw = 0.5 * width;
    b.bendw = 2 * b.R * fabs(b.sph);
}

// Bend.cpp:194-235
PM_HD bool is_inside(const Bend &b, double x0, double x1)
{
    const double D0 = x0 - b.center[0], D1 = x1 - b.center[1];
    if (D0 < 0.0) {
        if (x0 < pm::kSpaceEps || x1 < pm::kSpaceEps || x1 > b.width - pm::kSpaceEps) return false;
    } else if (D1 * b.sinphi > D0 * b.cosphi) {
        const double l2 = D0 * D0 + D1 * D1;
        if (l2 <= b.rmin_sq || l2 >= b.rmax_sq) return false;
    } else {
        double p0, p1;
        if (fabs(b.phi) > pm::kSpaceEps) {
            p0 = (b.cosphi * D0 + -b.sinphi * D1) + b.center[0];
            p1 = (b.sinphi * D0 + b.cosphi * D1) + b.center[1];
        } else {
            p0 = x0; p1 = x1;
        }
        if (p1 < pm::kSpaceEps || p1 > b.width - pm::kSpaceEps || p0 > b.lenBefore + b.lenAfter - pm::kSpaceEps)
            return false;
    }
    return true;
}

// Bend.cpp:126-192: fraction of the edge base -> base+dir inside the chamber.
// base is relative to the mesh origin `org`.  Returns -1 when the reference
// would abort (PAssert on t).
PM_HD double cross_boundary(const Bend &b, const double org[2], double b0, double b1, double d0, double d1)
{
    const double x0 = b0 + org[0], x1 = b1 + org[1];
    const double D0 = x0 - b.center[0], D1 = x1 - b.center[1];
    if (D0 < 0.0) {
        double t = 99.9;
        if (x1 < b.width + pm::kSpaceEps && x1 + d1 > b.width - pm::kSpaceEps) {
            const double v = (b.width - x1) / d1;
            t = smax(1.0, smin(0.0, v));   // reproduced as written: always 1.0
        }
        if (x1 > -pm::kSpaceEps && x1 + d1 < pm::kSpaceEps) t = smin(t, -x1 / d1);
        if (x0 > -pm::kSpaceEps && x0 + d0 < pm::kSpaceEps) t = smin(t, -x0 / d0);
        return t;
    } else if (D1 * b.sinphi > D0 * b.cosphi) {
        const double dist1 = D0 * D0 + D1 * D1;
        const double E0 = D0 + d0, E1 = D1 + d1;
        const double dist2 = E0 * E0 + E1 * E1;
        const double a = d0 * d0 + d1 * d1;
        const double bb = D0 * d0 + D1 * d1;
        double c;
        if (dist1 < b.rmax_sq && dist2 > b.rmax_sq) c = (D0 * D0 + D1 * D1) - b.rmax_sq;
        else if (dist1 > b.rmin_sq && dist2 < b.rmin_sq) c = (D0 * D0 + D1 * D1) - b.rmin_sq;
        else return 99.9;
        double t = 99.9;
        double tau = (-bb + sqrt(bb * bb - a * c)) / a;
        if (fabs(tau - 0.5) < 0.5) t = tau;
        tau = -tau - 2 * bb / a;
        if (fabs(tau - 0.5) < 0.5) t = smin(t, tau);
        return t;
    } else {
        double xr0, xr1, dr0, dr1;
        if (fabs(b.phi) > pm::kSpaceEps) {
            xr0 = b.cosphi * D0 - b.sinphi * D1 + b.center[0];
            xr1 = b.sinphi * D0 + b.cosphi * D1 + b.center[1];
            dr0 = b.cosphi * d0 - b.sinphi * d1;
            dr1 = b.sinphi * d0 + b.cosphi * d1;
        } else {
            xr0 = x0; xr1 = x1; dr0 = d0; dr1 = d1;
        }
        double t0 = 99.9, t1 = 99.9, t2 = 99.9;
        if (fabs(dr1) > pm::kSpaceEps) {
            t0 = -xr1 / dr1;
            if (t0 < 0.0) t0 = 99.9;
            t1 = (b.width - xr1) / dr1;
            if (t1 < 0.0) t1 = 99.9;
        }
        if (fabs(dr0) > pm::kSpaceEps) {
            t2 = (b.lenBefore + b.lenAfter - xr0) / dr0;
            if (t2 < 0.0) t2 = 99.9;
        }
        double t = smin(t0, smin(t1, t2));
        if (t > 1 + pm::kSpaceEps && t < 1.001) t = 1.0;
        if (!(t < 1 + pm::kSpaceEps) || !(t >= 0)) return -1.0;
        return smin(1.0, smax(0.0, t));
    }
}

// Bend.cpp:865-907 (applied at the particle, half of it enters the kick)
PM_HD double external_bz(const Bend &b, double x0, double x1)
{
    const double D0 = x0 - b.lenBefore, D1 = x1 - b.halfw;
    const double dist = b.cph * D0 - b.sph * D1;
    return (dist > 0.0 && dist < b.bendw) ? b.bz : 0.0;
}

// Vertex flag at global vertex (cx, cy) of a mesh with origin `org` and
// spacing `hr` (Bend.cpp:237-258).  `lim` bounds the vertex domain
// (vertices outside [0,lim0) x [0,lim1) are outside, as the guard cells of
// the reference are); pass INT_MAX for an unbounded mesh.
PM_HD int vertex_in(const Bend &b, const double org[2], const double hr[2], long long cx, long long cy, long long lim0,
                    long long lim1)
{
    if (cx < 0 || cy < 0 || cx >= lim0 || cy >= lim1) return 0;
    const double x0 = org[0] + (double)cx * hr[0];
    const double x1 = org[1] + (double)cy * hr[1];
    return is_inside(b, x0, x1) ? 1 : 0;
}

PM_HD int cell_case(const Bend &b, const double org[2], const double hr[2], long long cx, long long cy, long long l0,
                    long long l1)
{
    return vertex_in(b, org, hr, cx, cy, l0, l1) + (vertex_in(b, org, hr, cx + 1, cy, l0, l1) << 1) +
           (vertex_in(b, org, hr, cx, cy + 1, l0, l1) << 2) + (vertex_in(b, org, hr, cx + 1, cy + 1, l0, l1) << 3);
}

// Edge fractions of cell (cx, cy) as getLambdaField (Bend.cpp:427-545) fills
// them; cells whose vertex domain is outside give 0.  Returns false if
// crossBoundary would abort.
PM_HD bool cell_lambda(const Bend &b, const double org[2], const double hr[2], long long cx, long long cy, long long l0,
                       long long l1, double lam[2])
{
    lam[0] = 0.0; lam[1] = 0.0;
    if (cx < 0 || cy < 0 || cx >= l0 || cy >= l1) return true;
    const int cc = cell_case(b, org, hr, cx, cy, l0, l1);
    const double h0 = hr[0], h1 = hr[1];
    const double X = (double)cx * h0, Y = (double)cy * h1;
    const double X1 = (double)(cx + 1) * h0, Y1 = (double)(cy + 1) * h1;
    switch (cc) {
    case 15: lam[0] = 1.0; lam[1] = 1.0; break;
    case 0: break;
    case 1:
        lam[0] = cross_boundary(b, org, X, Y, h0, 0.0);
        lam[1] = cross_boundary(b, org, X, Y, 0.0, h1);
        break;
    case 2: lam[0] = cross_boundary(b, org, X1, Y, -h0, 0.0); lam[1] = 0.0; break;
    case 3: lam[0] = 1.0; lam[1] = cross_boundary(b, org, X, Y, 0.0, h1); break;
    case 4: lam[0] = 0.0; lam[1] = cross_boundary(b, org, X, Y1, 0.0, -h1); break;
    case 5: lam[0] = cross_boundary(b, org, X, Y, h0, 0.0); lam[1] = 1.0; break;
    case 7: lam[0] = 1.0; lam[1] = 1.0; break;
    case 8: lam[0] = 0.0; lam[1] = 0.0; break;
    case 10: lam[0] = cross_boundary(b, org, X1, Y, -h0, 0.0); lam[1] = 0.0; break;
    case 11: lam[0] = 1.0; lam[1] = cross_boundary(b, org, X, Y, 0.0, h1); break;
    case 12: lam[0] = 0.0; lam[1] = cross_boundary(b, org, X, Y1, 0.0, -h1); break;
    case 13: lam[0] = cross_boundary(b, org, X, Y, h0, 0.0); lam[1] = 1.0; break;
    case 14:
        lam[0] = cross_boundary(b, org, X1, Y, -h0, 0.0);
        lam[1] = cross_boundary(b, org, X, Y1, 0.0, -h1);
        break;
    default: break;
    }
    return !(lam[0] < 0.0 || lam[1] < 0.0);
}

// Bend.cpp:664-729
PM_HD double area_cases(int cc, const double lam[2], const double lamN[2])
{
    switch (cc) {
    case 1: return lam[0] * lam[1] / 2;
    case 2: return lam[0] * lamN[1] / 2;
    case 3: return (lam[1] + lamN[1]) / 2;
    case 4: return lamN[0] * lam[1] / 2;
    case 5: return (lam[0] + lamN[0]) / 2;
    case 7: return 1 - (1 - lamN[0]) * (1 - lamN[1]) / 2;
    case 8: return lamN[0] * lamN[1] / 2;
    case 10: return (lam[0] + lamN[0]) / 2;
    case 11: return 1 - (1 - lam[1]) * (1 - lamN[0]) / 2;
    case 12: return (lam[1] + lamN[1]) / 2;
    case 13: return 1 - (1 - lam[0]) * (1 - lamN[1]) / 2;
    case 14: return 1 - (1 - lam[0]) * (1 - lam[1]) / 2;
    default: return 0.0;
    }
}

// Classification of one cell: 0 outside, 1 inside (bulk update), 2 cut cell
// with the conformal data in `cut` (lambda, area after limitArea, lambdaN),
// 3 cut cell deleted (area < SPACE_EPS: neither bulk nor conformal update).
struct CutCell {
    double lam[2];
    double area;
    double lamN[2];
};

PM_HD int classify(const Bend &b, const double org[2], const double hr[2], long long cx, long long cy, long long l0,
                   long long l1, CutCell *cut, bool *ok)
{
    *ok = true;
    if (cx < 0 || cy < 0 || cx >= l0 || cy >= l1) return 0;
    const int cc = cell_case(b, org, hr, cx, cy, l0, l1);
    if (cc == 15) return 1;
    if (cc == 0) return 0;
    double lam[2], lx[2], ly[2];
    *ok &= cell_lambda(b, org, hr, cx, cy, l0, l1, lam);
    *ok &= cell_lambda(b, org, hr, cx, cy + 1, l0, l1, ly);   // lambdaPy
    *ok &= cell_lambda(b, org, hr, cx + 1, cy, l0, l1, lx);   // lambdaPx
    const double lamN[2] = {ly[0], lx[1]};
    const double area = area_cases(cc, lam, lamN);
    if (area < pm::kSpaceEps) return 3;
    cut->lam[0] = lam[0]; cut->lam[1] = lam[1];
    cut->lamN[0] = lamN[0]; cut->lamN[1] = lamN[1];
    double a = area;   // SCBoundaryCondition.hh:131-137 limitArea
    a = smax(a, lam[0] / 2);
    a = smax(a, lam[1] / 2);
    a = smax(a, lamN[0] / 2);
    a = smax(a, lamN[1] / 2);
    cut->area = a;
    return 2;
}

}  // namespace geom
