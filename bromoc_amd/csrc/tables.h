// tables.h -- EFPOT table construction (splinepot.f90:19-96 for sampled potentials,
// discretize.f90:19-120 for tables built from a pair's analytic parameters, both through
// squarespline, polynom.f90:20-169), shared by the host entry points (api.hip) and the
// device kernels (tables.hip): one code path, so a table built on the GPU is bit-identical
// to one built on the host, and both to the reference's.  Integer powers are expanded as
// the reference's compiler (amdflang) expands x**n -- exponents up to 8 multiplied out left
// to right, larger ones by binary powering; floating-point contraction is off so each
// expression rounds as written.
#pragma once

#include <hip/hip_runtime.h>

namespace bromoc {

#define TAB_HD __host__ __device__ inline

TAB_HD double tab_ipow(double x, unsigned n) {
#pragma clang fp contract(off)
    if (n == 0) return 1.0;
    if (n <= 8) {   // multiplied out left to right, as the compiler does small powers
        double r = x;
        for (unsigned k = 1; k < n; k++) r = r * x;
        return r;
    }
    double r = 0.0, sq = x;
    bool have = false;
    for (; n; n >>= 1) {
        if (n & 1u) { r = have ? r * sq : sq; have = true; }
        if (n > 1) sq = sq * sq;
    }
    return have ? r : 1.0;
}

// squarespline (polynom.f90:20-48): quadratic pieces q[3k..3k+2] = (a, b, c), U = a + b r +
// c r^2 on [x_k, x_{k+1}), from samples y at knots x, end derivatives d1, d2 and the second
// derivative dd2 past the last knot.  Scratch: yp, s0, s1 of n doubles each.
TAB_HD void square_spline(int n, const double *x, const double *y, double *q, double d1, double d2,
                          double dd2, double *yp, double *s0, double *s1) {
#pragma clang fp contract(off)
    for (int k = 0; k < 3 * n; k++) q[k] = 0.0;
    if (n == 1) { q[0] = y[0]; return; }
    if (n == 2) {
        const double b = (y[1] - y[0]) / (x[1] - x[0]);
        q[0] = q[3] = y[0] - b * x[0];
        q[1] = q[4] = b;
        return;
    }
    const double y0 = y[0], yn = y[n - 1];
#define QA(i) q[3 * (i)]
#define QB(i) q[3 * (i) + 1]
#define QC(i) q[3 * (i) + 2]
    // sqfit (polynom.f90:135-169): local quadratics through consecutive triples, averaged
    {
        const double tt = 1.0 / (x[1] - x[0]);
        QC(0) = ((y[1] - y[0]) * tt - d1) * tt;
        QB(0) = d1 - 2.0 * QC(0) * x[0];
        QA(0) = y[0] - x[0] * (QB(0) + QC(0) * x[0]);
    }
    for (int i = 1; i <= n - 2; i++) {
        const double tt = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
        QC(i) = (tt - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) / (x[i + 1] - x[i - 1]);
        QB(i) = tt - QC(i) * (x[i + 1] + x[i]);
        QA(i) = y[i] - x[i] * (QB(i) + QC(i) * x[i]);
        if (i >= 2) {
            QC(i - 1) = 0.5 * (QC(i - 1) + QC(i));
            QB(i - 1) = 0.5 * (QB(i - 1) + QB(i));
            QA(i - 1) = 0.5 * (QA(i - 1) + QA(i));
        }
    }
    {
        const int m = n - 1;
        const double tt = 1.0 / (x[m] - x[m - 1]);
        QC(m) = (d2 - (y[m] - y[m - 1]) * tt) * tt;
        QB(m) = d2 - 2.0 * QC(m) * x[m];
        QA(m) = y[m] - x[m] * (QB(m) + QC(m) * x[m]);
        QC(m - 1) = 0.5 * (QC(m - 1) + QC(m));
        QB(m - 1) = 0.5 * (QB(m - 1) + QB(m));
        QA(m - 1) = 0.5 * (QA(m - 1) + QA(m));
    }
    // fixder (polynom.f90:121-132): knot derivatives averaged across segments
    yp[0] = QB(0) + 2.0 * QC(0) * x[0];
    for (int i = 1; i <= n - 2; i++)
        yp[i] = 0.5 * (QB(i) + 2.0 * QC(i) * x[i] + QB(i - 1) + 2.0 * QC(i - 1) * x[i]);
    yp[n - 1] = QB(n - 1) + 2.0 * QC(n - 1) * x[n - 1];
    // linipol (polynom.f90:67-88): the derivative as a piecewise-linear function
    for (int i = 0; i <= n - 2; i++) {
        const double tt = 1.0 / (x[i + 1] - x[i]);
        s1[i] = (yp[i + 1] - yp[i]) * tt;
        s0[i] = (yp[i] * x[i + 1] - yp[i + 1] * x[i]) * tt;
    }
    s1[n - 1] = dd2;
    s0[n - 1] = yp[n - 1] - s1[n - 1] * x[n - 1];
    // intlin (polynom.f90:51-64): integrated, continuous at the knots
    QC(0) = s1[0] * 0.5;
    QB(0) = s0[0];
    QA(0) = y0 - x[0] * (QB(0) + QC(0) * x[0]);
    for (int i = 1; i < n; i++) {
        QC(i) = s1[i] * 0.5;
        QB(i) = s0[i];
        QA(i) = QA(i - 1) + x[i] * (QB(i - 1) - QB(i) + x[i] * (QC(i - 1) - QC(i)));
    }
    // fixdy (polynom.f90:91-118): restore U(x_n) with a correction that keeps the ends
    {
        const int m = n - 1;
        const double ynn = QA(m) + x[m] * (QB(m) + x[m] * QC(m));
        const double b = 2.0 * (yn - ynn) / (x[m] - x[0]);
        const double cc = -b * x[0] * 0.5;
        const double h = (x[1] - x[0]) / (double)(n - 1);
        for (int i = 0; i < m; i++) {
            QA(i) = QA(i) + cc;
            QB(i) = QB(i) + b;
            QC(i) = QC(i) + (-b * 0.5 / (x[i] + h * (double)i));
        }
        QA(m) = QA(m) + cc;
        QB(m) = QB(m) + b;
        QC(m) = QC(m) + (-b * 0.5 / x[m]);
    }
#undef QA
#undef QB
#undef QC
}

// The coefficients a table carries besides its knots.
struct TabEnds {
    double dmi, dm2lo, dm2hi, sc1, sc2, sc3;
};

// SPLINEPOT (splinepot.f90:19-96): n >= 2 samples u at increasing r; kf = fct of a charged
// pair, else 0.  Tail kf/r + sc3/r^6 matched at r_n, wall sc1/r^12 + sc2 through the first
// two samples.  Scratch: 3n doubles.
TAB_HD void table_splinepot(int n, const double *r, const double *u, double kf, double *ep,
                            TabEnds &o, double *scratch) {
#pragma clang fp contract(off)
    const double r1 = r[0], r2 = r[1], rn = r[n - 1];
    o.dmi = r1;
    o.dm2lo = r1 * r1;
    o.dm2hi = rn * rn;
    o.sc3 = (rn * u[n - 1] - kf) * tab_ipow(rn, 5);
    const double d2 = -kf / tab_ipow(rn, 2) - 6.0 * o.sc3 / tab_ipow(rn, 7);
    const double dd2 = 2.0 * kf / tab_ipow(rn, 3) + 42.0 * o.sc3 / tab_ipow(rn, 8);
    const double p1 = tab_ipow(r1, 12), p2 = tab_ipow(r2, 12);
    const double inv = 1.0 / (p1 - p2);
    o.sc1 = (u[1] - u[0]) * p2 * p1 * inv;
    o.sc2 = (u[0] * p1 - u[1] * p2) * inv;
    const double d1 = -12.0 * o.sc1 / tab_ipow(r1, 13);
    square_spline(n, r, u, ep, d1, d2, dd2, scratch, scratch + n, scratch + 2 * n);
}

// The transpose of table_splinepot's dependence on its samples: gu[k] = d/du_k of
// ep . gq + (sc1, sc2, sc3) . gsc for a table of n >= 2 samples at r.  The table is affine
// in u (the charge term kf only shifts it), so this is one reverse sweep through the
// stages of squarespline -- fixdy, intlin, linipol, fixder, the averaging of the local
// quadratics and the quadratics themselves -- and the end coefficients, O(n) in place of
// n table builds.  Scratch: 6n doubles.
TAB_HD void table_splinepot_vjp(int n, const double *x, const double *gq, const double *gsc,
                                double *gu, double *scratch) {
#pragma clang fp contract(off)
    for (int k = 0; k < n; k++) gu[k] = 0.0;
    double gd1 = 0.0, gd2 = 0.0, gdd2 = 0.0;
    if (n == 2) {
        // q[0] = q[3] = y0 - b x0, q[1] = q[4] = b, b = (y1 - y0) / (x1 - x0)
        const double gb = gq[1] + gq[4] - (gq[0] + gq[3]) * x[0];
        gu[0] += gq[0] + gq[3];
        gu[1] += gb / (x[1] - x[0]);
        gu[0] -= gb / (x[1] - x[0]);
    } else {
        const int m = n - 1;
        double *gA = scratch, *gB = scratch + n, *gC = scratch + 2 * n;
        double *gyp = scratch + 3 * n, *g0 = scratch + 4 * n, *g1 = scratch + 5 * n;
        for (int k = 0; k < n; k++) {
            gA[k] = gq[3 * k]; gB[k] = gq[3 * k + 1]; gC[k] = gq[3 * k + 2];
        }
        // fixdy: QA += cc, QB += b, QC -= b / 2 / (x_i + h i) (x_m at the last knot), with
        // b = 2 (y_m - ynn) / (x_m - x_0), cc = -b x_0 / 2, ynn = U(x_m) after intlin
        const double h = (x[1] - x[0]) / (double)(n - 1);
        double gcc = 0.0, gb = 0.0;
        for (int i = 0; i < m; i++) {
            gcc += gA[i];
            gb += gB[i];
            gb += gC[i] * (-0.5 / (x[i] + h * (double)i));
        }
        gcc += gA[m];
        gb += gB[m];
        gb += gC[m] * (-0.5 / x[m]);
        gb += gcc * (-x[0] * 0.5);
        const double gy = gb * 2.0 / (x[m] - x[0]);
        gu[m] += gy;
        gA[m] -= gy;
        gB[m] -= gy * x[m];
        gC[m] -= gy * x[m] * x[m];
        // intlin: QA(i) = QA(i-1) + x_i (QB(i-1) - QB(i) + x_i (QC(i-1) - QC(i))),
        // QA(0) = y_0 - x_0 (QB(0) + QC(0) x_0), QB = s0, QC = s1 / 2
        for (int i = m; i >= 1; i--) {
            const double g = gA[i];
            gA[i - 1] += g;
            gB[i - 1] += g * x[i];
            gB[i] -= g * x[i];
            gC[i - 1] += g * x[i] * x[i];
            gC[i] -= g * x[i] * x[i];
        }
        gu[0] += gA[0];
        gB[0] -= gA[0] * x[0];
        gC[0] -= gA[0] * x[0] * x[0];
        // linipol: s1_i = (yp_{i+1} - yp_i) / dx_i, s0_i = (yp_i x_{i+1} - yp_{i+1} x_i) / dx_i;
        // s1_m = dd2, s0_m = yp_m - s1_m x_m
        for (int k = 0; k < n; k++) {
            g0[k] = gB[k];
            g1[k] = 0.5 * gC[k];
            gyp[k] = 0.0;
        }
        gyp[m] += g0[m];
        g1[m] -= g0[m] * x[m];
        gdd2 += g1[m];
        for (int i = 0; i < m; i++) {
            const double tt = 1.0 / (x[i + 1] - x[i]);
            gyp[i + 1] += g1[i] * tt;
            gyp[i] -= g1[i] * tt;
            gyp[i] += g0[i] * x[i + 1] * tt;
            gyp[i + 1] -= g0[i] * x[i] * tt;
        }
        // fixder: yp from the averaged quadratics' QB + 2 QC x (only QB, QC reach here)
        for (int k = 0; k < n; k++) gB[k] = gC[k] = 0.0;
        gB[0] += gyp[0];
        gC[0] += 2.0 * x[0] * gyp[0];
        for (int i = 1; i < m; i++) {
            const double g = 0.5 * gyp[i];
            gB[i] += g;
            gC[i] += 2.0 * x[i] * g;
            gB[i - 1] += g;
            gC[i - 1] += 2.0 * x[i] * g;
        }
        gB[m] += gyp[m];
        gC[m] += 2.0 * x[m] * gyp[m];
        // sqfit's averaging: Q(i) = (L(i) + L(i+1)) / 2 for 0 < i < m, Q(0) = L(0), Q(m) = L(m)
        for (int k = 0; k < n; k++) g0[k] = g1[k] = 0.0;
        g0[0] += gB[0];
        g1[0] += gC[0];
        for (int i = 1; i < m; i++) {
            g0[i] += 0.5 * gB[i];
            g0[i + 1] += 0.5 * gB[i];
            g1[i] += 0.5 * gC[i];
            g1[i + 1] += 0.5 * gC[i];
        }
        g0[m] += gB[m];
        g1[m] += gC[m];
        // the local quadratics (QB in g0, QC in g1): the first, with the start slope d1
        {
            const double tt = 1.0 / (x[1] - x[0]);
            const double gc = g1[0] - 2.0 * x[0] * g0[0];
            gd1 += g0[0];
            gu[1] += gc * tt * tt;
            gu[0] -= gc * tt * tt;
            gd1 -= gc * tt;
        }
        for (int i = 1; i < m; i++) {
            const double gc = g1[i] - (x[i + 1] + x[i]) * g0[i];
            const double den = x[i + 1] - x[i - 1];
            const double gt = g0[i] + gc / den, gp = -gc / den;
            gu[i + 1] += gt / (x[i + 1] - x[i]);
            gu[i] -= gt / (x[i + 1] - x[i]);
            gu[i] += gp / (x[i] - x[i - 1]);
            gu[i - 1] -= gp / (x[i] - x[i - 1]);
        }
        {   // the last, with the end slope d2
            const double tt = 1.0 / (x[m] - x[m - 1]);
            const double gc = g1[m] - 2.0 * x[m] * g0[m];
            gd2 += g0[m] + gc * tt;
            gu[m] -= gc * tt * tt;
            gu[m - 1] += gc * tt * tt;
        }
    }
    // the ends: sc3 = (r_n u_n - kf) r_n^5, d2 = -kf / r_n^2 - 6 sc3 / r_n^7,
    // dd2 = 2 kf / r_n^3 + 42 sc3 / r_n^8, sc1 = (u_2 - u_1) p2 p1 / (p1 - p2),
    // sc2 = (u_1 p1 - u_2 p2) / (p1 - p2), d1 = -12 sc1 / r_1^13
    const double r1 = x[0], r2 = x[1], rn = x[n - 1];
    const double g3 = gsc[2] + gd2 * (-6.0 / tab_ipow(rn, 7)) + gdd2 * (42.0 / tab_ipow(rn, 8));
    gu[n - 1] += g3 * rn * tab_ipow(rn, 5);
    const double p1 = tab_ipow(r1, 12), p2 = tab_ipow(r2, 12);
    const double inv = 1.0 / (p1 - p2);
    const double gs1 = gsc[0] + gd1 * (-12.0 / tab_ipow(r1, 13));
    gu[1] += gs1 * p2 * p1 * inv;
    gu[0] -= gs1 * p2 * p1 * inv;
    gu[0] += gsc[1] * p1 * inv;
    gu[1] -= gsc[1] * p2 * inv;
}

// The analytic pair parameters DISCRETIZE samples (extramod/grandmod per pair type).
struct TabPair {
    int chr, lj, srpmf, pad;
    double fct, epp4, sgp2, c0, c1, c2, c3, c4;
};

// DISCRETIZE (discretize.f90:19-120): samples of Coulomb + LJ (+ SRPMF for r^2 <= rth) at
// r = k res, k = nnp..mnp, into x, y (n = mnp - nnp + 1 each), the force at the first knot
// as the start derivative, the tail matched at the last knot, then squarespline and a wall
// sc1/r^12 + sc2 with the table's own value and slope at the first knot.  dmi and dm2 are
// the EFPOT command's (shell_simul.f90:1850-1852), not set here.  Scratch: 5n doubles.
TAB_HD void table_discretize(const TabPair &p, int nnp, int mnp, double res, double rth,
                             double srpx, double srpk, double srpy, double *ep, TabEnds &o,
                             double *scratch) {
#pragma clang fp contract(off)
    const double pi = 3.14159265358979323846264338327950288419716939937510;
    const int n = mnp - nnp + 1;
    double *xx = scratch + 3 * n, *yy = scratch + 4 * n;
    const double kf = p.chr ? p.fct : 0.0;
    double d1 = 0.0;
    for (int k = nnp; k <= mnp; k++) {
        const double x1 = (double)k * res;
        const double x2 = 1.0 / x1;
        double scal = kf * x2;
        double scald = -(kf * tab_ipow(x2, 2));
        if (p.lj) {
            const double dist6 = tab_ipow(p.sgp2, 3) * tab_ipow(x2, 6);
            const double dist12 = dist6 * dist6;
            scal = scal + p.epp4 * (dist12 - dist6);
            scald = scald + p.epp4 * (dist6 - 2.0 * dist12) * 6.0 * x2;
        }
        if (p.srpmf && x1 * x1 <= rth) {
            const double cc1 = exp((p.c1 - x1) * p.c2);
            const double cc2 = cos(p.c3 * pi * (p.c1 - x1));
            const double cc3 = tab_ipow(p.c1 * x2, 6);
            double cc0 = p.c0 * cc1 * cc2 + p.c4 * cc3;
            double fdf = 0.0, fdv = 0.0;
            if (x1 >= srpx) {
                fdf = exp(-srpk * (x1 - srpx)) - srpy;
                fdv = cc0;
                cc0 = fdv * fdf;
            }
            scal = scal + cc0;
            double cc4 = (-p.c2 * cc2 + p.c3 * pi * sin(p.c3 * pi * (p.c1 - x1))) * p.c0 * cc1 -
                         6.0 * p.c4 * cc3 * x2;
            if (x1 >= srpx) cc4 = cc4 * fdf - fdv * srpk * (fdf + srpy);
            scald = scald + cc4;
        }
        xx[k - nnp] = x1;
        yy[k - nnp] = scal;
        if (k == nnp) d1 = scald;
    }
    const double xn = xx[n - 1];
    o.sc3 = (yy[n - 1] * xn - kf) * tab_ipow(xn, 5);
    const double d2 = -kf / tab_ipow(xn, 2) - 6.0 * o.sc3 / tab_ipow(xn, 7);
    const double dd2 = 2.0 * kf / tab_ipow(xn, 3) + 42.0 * o.sc3 / tab_ipow(xn, 8);
    square_spline(n, xx, yy, ep, d1, d2, dd2, scratch, scratch + n, scratch + 2 * n);
    o.sc1 = -(d1 * tab_ipow(xx[0], 13) / 12.0);
    const double y0 = ep[0] + xx[0] * (ep[1] + xx[0] * ep[2]);
    o.sc2 = y0 - o.sc1 / tab_ipow(xx[0], 12);
}

#undef TAB_HD

}  // namespace bromoc
