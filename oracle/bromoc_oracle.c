/*
 * bromoc_oracle.c -- CPU restatement of the BROMOC cg-gcmc-bd force / BD step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the product path
 * (bromoc_amd/), which fails loudly when its HIP library is missing.
 *
 * Every routine follows one reference routine loop-for-loop and in the same expression
 * order (file:line cited at each function).  Built with -O2 -ffp-contract=off so that each
 * Fortran expression rounds as written.  Indices are 0-based (Fortran index - 1).
 *
 * Parity pins (what checks each function; tests/test_oracle_vs_reference.py and
 * tests/test_oracle_math.py):
 *   getyd/gety, squarespline/splinepot, ENERGY, STATICF1, VDWGD1TRLN, VDWGD1SPLN,
 *   APFOR/CONTRANS, the PROXDIFF records, DYNAMICS0/0NUC/1/2/3, NOTRANS, FIXCOOR, COUNTIONS,
 *   SPLINE/SEVAL, RFPARION/RFPARIONJ, INTERACT, STATICF0, VDWGD0TRLN, VDWGD0SPLN, GRAND,
 *   METROPOLIS, MEMBRANE, RECT_RF1/RECT_RF0, the Qnmcden map choice: against the reference
 *   Fortran compiled by oracle/Makefile (oracle/_ref), plus closed forms and invariants (Newton's third law,
 *   F = -grad E by central differences, ENERGY = 1/2 sum INTERACT, exact trilinear
 *   interpolation of linear fields, the B-spline's reproduction of a linear map,
 *   closed-form BD step, acceptance rule, METROPOLIS energy bookkeeping, the reaction-field
 *   energy as a quadratic form of numpy's Legendre coefficients).
 *   philox4x64: bit-exact against numpy.random.Philox (a published library routine).
 *   The per-cycle statistics (rho, rDNA, g(r), prob2/prob3, virial pressure; inline in
 *   simul1.f90, no routine to compile): against hand-placed bins, numpy's truncated bins,
 *   the ideal-gas law and the virial recomputed from the forces.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/bromoc_gpu.h"

#define EXPORT __attribute__((visibility("default")))

/* constamod.f90:21-36 (binary64 under -fdefault-real-8) */
static const double PI = 3.14159265358979323846264338327950288419716939937510;
static double celec_value(void) {
    const double kcal = 4184.0, avogadro = 6.0221412927E23;
    const double kcalmol = kcal / avogadro;
    const double coulomb = 1.60217656535E-19, epermit = 8.854187817620E-12, angstrom = 1E-10;
    return coulomb * coulomb / (4.0 * PI * epermit * kcalmol * angstrom);
}
static const double RSMALL = 1E-3;

typedef struct orc_ctx {
    bromoc_system s;          /* scalars; pointer members re-pointed to owned copies */
    int cap;
    double *x, *y, *z, *fx, *fy, *fz;
    int32_t *typei;
    int32_t *typtyp, *namsite, *nwtype, *pairidx;
    int8_t *stfree;
    double *cg, *fact1a, *fact2a, *ep, *scal, *zcont;
    bromoc_pair *pairs;
    int32_t *lists[8];
    double *lvals[6];
    float *phix;
    int8_t *phiv;
    int32_t *warn, *nforward, *nbackward;
    double *noise;
    int noise_rows;
    double celec;
    /* GCMC */
    bromoc_gcmc g;
    int32_t *ibfftyp, *ntotat, *ibuffer;
    double *avnum, *mu, *kb, *lzmin, *lzmax, *rmin, *rmax;
    int8_t *bufferbias;
    /* DYNAMICS1/2 */
    double *diffusion, *ampl3, *p3, *xs, *ys, *bs, *cs, *ds;
    int seval_i, sevald_i;        /* the SAVEd intervals of seval and sevald (1-based) */
    uint64_t stream;              /* Philox stream of this system's Gaussians */
    /* RFPAR */
    float *gsrfen, *greff;
    double *srfe, *reff, *srfed, *reffd;   /* per ion slot; derivatives [3*cap] */
    /* PROXDIFF: nearest-site record dids(1:5,i) of ENERGY, [5*cap] */
    double *dids;
    /* APFOR / CONTRANS */
    int32_t *sn, *csn;
    double *af, *kx, *ky, *kz, *contrx, *contry, *contrz;
    /* MEMBRANE per atnam2 type */
    double *ampl1, *p1, *ampl2, *p2, *rcylinder;
    /* MMIJ reaction field (RECTBOX) */
    int32_t *lstpx, *lstpy, *lstpz;
    double *bnorm, *mmij, *coef, *mq, *lpx, *lpy, *lpz;
    /* per-cycle statistics of simul1.f90 (reals as the reference keeps them; prob integer) */
    bromoc_analysis an;
    double *rho, *rdna, *gr, pres;
    int32_t irmax;
    int64_t *prob2, *prob3, pres_steps;
} orc_ctx;

static void *dupmem(const void *p, size_t n) {
    void *q = calloc(1, n ? n : 1);
    if (p && n) memcpy(q, p, n);
    return q;
}

/* x**n for a constant integer n as the reference's compiler (amdflang, the one that builds
 * oracle/_ref) expands it: exponents up to 8 multiplied out left to right (x*x*...*x),
 * larger ones by binary powering (square-and-multiply from the low bit).  Measured on
 * 2000 random x for every n in 1..14. */
static double powi(double x, unsigned n) {
    if (n == 0) return 1.0;
    if (n <= 8) {
        double r = x;
        for (unsigned k = 1; k < n; k++) r = r * x;
        return r;
    }
    double r = 0.0, sq = x;
    int have = 0;
    while (n) {
        if (n & 1u) { r = have ? r * sq : sq; have = 1; }
        n >>= 1;
        if (n) sq = sq * sq;
    }
    return have ? r : 1.0;
}
static double pow3(double a) { return powi(a, 3); }
static double pow6(double a) { return powi(a, 6); }
static double pow5(double a) { return powi(a, 5); }
static double fsign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); } /* SIGN(a,b) */

/* nindex.f90:19-35 through the host-built pair table */
static int pair_of(const orc_ctx *c, int a2, int b2) { return c->pairidx[a2 * (c->s.nttyp + 1) + b2]; }

/* ---------------------------------------------------------------- table lookup */
/* gety, polynom.f90:172-194 */
static void gety(const orc_ctx *c, int is, double x2, double *y, double *x) {
    const bromoc_pair *p = &c->pairs[is];
    *x = 0.0;
    if (x2 < p->dm2lo) {
        double ix2 = 1.0 / x2;
        *y = p->sc1 * pow6(ix2) + p->sc2;
    } else if (x2 > p->dm2hi) {
        double ix2 = 1.0 / x2;
        if (p->chr) *y = p->sc3 * pow3(ix2) + p->fct * sqrt(ix2);
        else *y = p->sc3 * pow3(ix2);
    } else {
        *x = sqrt(x2);
        int i = (int)((*x - p->dmi) * c->s.ires) + p->nxi;
        const double *e = c->ep + 3 * (size_t)i;
        *y = e[0] + *x * (e[1] + *x * e[2]);
    }
}

/* getyd, polynom.f90:197-230 */
static void getyd(const orc_ctx *c, int is, double x2, double *y, double *yp, double *x) {
    const bromoc_pair *p = &c->pairs[is];
    *x = 0.0;
    if (x2 < p->dm2lo) {
        double ix2 = 1.0 / x2;
        double cc = p->sc1 * pow6(ix2);
        *y = cc + p->sc2;
        *yp = 12.0 * cc * ix2;
    } else if (x2 > p->dm2hi) {
        double ix2 = 1.0 / x2;
        double cc = p->sc3 * pow3(ix2);
        if (p->chr) {
            double f = p->fct * sqrt(ix2);
            *y = cc + f;
            *yp = (6.0 * cc + f) * ix2;
        } else {
            *y = cc;
            *yp = 6.0 * cc * ix2;
        }
    } else {
        *x = sqrt(x2);
        int i = (int)((*x - p->dmi) * c->s.ires) + p->nxi;
        const double *e = c->ep + 3 * (size_t)i;
        *y = e[0] + *x * (e[1] + *x * e[2]);
        *yp = -e[1] / *x - 2.0 * e[2];
    }
}

/* -------------------------------------------------------- table construction */
/* sqfit, polynom.f90:135-169 (p is ep(3,n), column-major: p[3*i+k]) */
static void sqfit(int n, const double *x, const double *y, double *p, double d1, double d2) {
#define P(k, i) p[3 * (i) + (k)]
    int i = 0;
    double tt = 1.0 / (x[i + 1] - x[i]);
    P(2, i) = ((y[i + 1] - y[i]) * tt - d1) * tt;
    P(1, i) = d1 - 2.0 * P(2, i) * x[i];
    P(0, i) = y[i] - x[i] * (P(1, i) + P(2, i) * x[i]);
    i = 1;
    tt = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
    P(2, i) = (tt - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) / (x[i + 1] - x[i - 1]);
    P(1, i) = tt - P(2, i) * (x[i + 1] + x[i]);
    P(0, i) = y[i] - x[i] * (P(1, i) + P(2, i) * x[i]);
    for (i = 2; i <= n - 2; i++) {
        tt = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
        P(2, i) = (tt - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) / (x[i + 1] - x[i - 1]);
        P(1, i) = tt - P(2, i) * (x[i + 1] + x[i]);
        P(0, i) = y[i] - x[i] * (P(1, i) + P(2, i) * x[i]);
        P(2, i - 1) = 0.5 * (P(2, i - 1) + P(2, i));
        P(1, i - 1) = 0.5 * (P(1, i - 1) + P(1, i));
        P(0, i - 1) = 0.5 * (P(0, i - 1) + P(0, i));
    }
    int m = n - 1;
    tt = 1.0 / (x[m] - x[m - 1]);
    P(2, m) = (d2 - (y[m] - y[m - 1]) * tt) * tt;
    P(1, m) = d2 - 2.0 * P(2, m) * x[m];
    P(0, m) = y[m] - x[m] * (P(1, m) + P(2, m) * x[m]);
    P(2, m - 1) = 0.5 * (P(2, m - 1) + P(2, m));
    P(1, m - 1) = 0.5 * (P(1, m - 1) + P(1, m));
    P(0, m - 1) = 0.5 * (P(0, m - 1) + P(0, m));
}

/* fixder, polynom.f90:121-132 */
static void fixder(int n, const double *x, const double *p, double *yp) {
    yp[0] = P(1, 0) + 2.0 * P(2, 0) * x[0];
    for (int i = 1; i <= n - 2; i++)
        yp[i] = 0.5 * (P(1, i) + 2.0 * P(2, i) * x[i] + P(1, i - 1) + 2.0 * P(2, i - 1) * x[i]);
    int i = n - 1;
    yp[i] = P(1, i) + 2.0 * P(2, i) * x[i];
}

/* linipol, polynom.f90:67-88 (pp is pp(2,n): pp[2*i+k]) */
static void linipol(int n, const double *x, const double *y, double *pp, double dd) {
    for (int i = 0; i <= n - 2; i++) {
        double tt = 1.0 / (x[i + 1] - x[i]);
        pp[2 * i + 1] = (y[i + 1] - y[i]) * tt;
        pp[2 * i + 0] = (y[i] * x[i + 1] - y[i + 1] * x[i]) * tt;
    }
    pp[2 * (n - 1) + 1] = dd;
    pp[2 * (n - 1) + 0] = y[n - 1] - pp[2 * (n - 1) + 1] * x[n - 1];
}

/* intlin, polynom.f90:51-64 */
static void intlin(int n, const double *x, const double *pp, double *p, double y0) {
    P(2, 0) = pp[1] * 0.5;
    P(1, 0) = pp[0];
    P(0, 0) = y0 - x[0] * (P(1, 0) + P(2, 0) * x[0]);
    for (int i = 1; i < n; i++) {
        P(2, i) = pp[2 * i + 1] * 0.5;
        P(1, i) = pp[2 * i + 0];
        P(0, i) = P(0, i - 1) + x[i] * (P(1, i - 1) - P(1, i) + x[i] * (P(2, i - 1) - P(2, i)));
    }
}

/* fixdy, polynom.f90:91-118 */
static void fixdy(int n, const double *x, double *p, double yn) {
    int i = n - 1;
    double ynn = P(0, i) + x[i] * (P(1, i) + x[i] * P(2, i));
    double b = 2.0 * (yn - ynn) / (x[n - 1] - x[0]);
    double c = -b * x[0] * 0.5;
    double h = (x[1] - x[0]) / (double)(n - 1);
    for (i = 0; i <= n - 2; i++) {
        P(0, i) = P(0, i) + c;
        P(1, i) = P(1, i) + b;
        double a = -b * 0.5 / (x[i] + h * (double)i);
        P(2, i) = P(2, i) + a;
    }
    i = n - 1;
    P(0, i) = P(0, i) + c;
    P(1, i) = P(1, i) + b;
    double a = -b * 0.5 / x[i];
    P(2, i) = P(2, i) + a;
}

/* squarespline, polynom.f90:20-48 */
static void squarespline(int n, const double *x, const double *y, double *p, double d1, double d2,
                         double dd2) {
    if (n == 1) {
        P(2, 0) = 0.0; P(1, 0) = 0.0; P(0, 0) = y[0];
    } else if (n == 2) {
        P(2, 0) = 0.0;
        P(1, 0) = (y[1] - y[0]) / (x[1] - x[0]);
        P(0, 0) = y[0] - P(1, 0) * x[0];
        P(2, 1) = P(2, 0); P(1, 1) = P(1, 0); P(0, 1) = P(0, 0);
    } else {
        double y0 = y[0], yn = y[n - 1];
        double *yp = (double *)malloc(sizeof(double) * n);
        double *pp = (double *)malloc(sizeof(double) * 2 * n);
        memcpy(yp, y, sizeof(double) * n);
        sqfit(n, x, yp, p, d1, d2);
        fixder(n, x, p, yp);
        linipol(n, x, yp, pp, dd2);
        intlin(n, x, pp, p, y0);
        fixdy(n, x, p, yn);
        free(yp);
        free(pp);
    }
}
#undef P

/* splinepot, splinepot.f90:19-96 */
EXPORT int orc_build_table(int32_t n, const double *xx, const double *yy, int32_t chr, double fct,
                           double *ep, bromoc_pair *pair) {
    if (n < 2 || !xx || !yy || !ep || !pair) return BROMOC_E_ARG;
    if (xx[n - 1] == 0.0 || xx[0] == 0.0) return BROMOC_E_ARG;
    pair->dmi = xx[0];
    pair->dm2lo = xx[0] * xx[0];
    pair->dm2hi = xx[n - 1] * xx[n - 1];
    double kf = chr ? fct : 0.0;
    double xn = xx[n - 1];
    pair->sc3 = (xn * yy[n - 1] - kf) * powi(xn, 5);
    double d2 = -kf / powi(xn, 2) - 6.0 * pair->sc3 / powi(xn, 7);
    double dd2 = 2.0 * kf / powi(xn, 3) + 42.0 * pair->sc3 / powi(xn, 8);
    double x1_12 = powi(xx[0], 12), x2_12 = powi(xx[1], 12);
    double dd1 = 1.0 / (x1_12 - x2_12);
    pair->sc1 = (yy[1] - yy[0]) * x2_12 * x1_12 * dd1;
    pair->sc2 = (yy[0] * x1_12 - yy[1] * x2_12) * dd1;
    double d1 = -12.0 * pair->sc1 / powi(xx[0], 13);
    squarespline(n, xx, yy, ep, d1, d2, dd2);
    pair->nknots = n;
    pair->efpot = 1;
    pair->chr = chr;
    pair->fct = fct;
    return BROMOC_OK;
}

/* DISCRETIZE, discretize.f90:19-120: the table of a pair from its analytic parameters
 * (fct/r when charged, LJ, SRPMF inside rth) sampled at x = k*res, k = nnp..mnp; sets ep
 * and sc(1:3) (dmi, dm2 are the EFPOT command's, shell_simul.f90:1850-1852) */
EXPORT int orc_discretize(const bromoc_pair *p, int32_t nnp, int32_t mnp, double res, double rth,
                          double srpx, double srpk, double srpy, double *ep, bromoc_pair *out) {
    if (!p || !ep || !out || mnp - nnp + 1 < 3 || nnp < 1 || !(res > 0.0)) return BROMOC_E_ARG;
    int nn = 1 - nnp + mnp;
    double *xx = (double *)malloc(sizeof(double) * nn), *yy = (double *)malloc(sizeof(double) * nn);
    double kf = p->chr ? p->fct : 0.0;
    double d1 = 0.0;
    for (int k = nnp; k <= mnp; k++) {
        double x1 = k * res;
        double x2 = 1.0 / x1;
        double scal = kf * x2;
        double scald = -kf * powi(x2, 2);
        if (p->lj) {
            double dist6 = powi(p->sgp2, 3) * powi(x2, 6);
            double dist12 = powi(dist6, 2);
            scal = scal + p->epp4 * (dist12 - dist6);
            scald = scald + p->epp4 * (dist6 - 2.0 * dist12) * 6.0 * x2;
        }
        if (p->srpmf) {
            if (powi(x1, 2) <= rth) {
                double cc1 = exp((p->c1 - x1) * p->c2);
                double cc2 = cos(p->c3 * PI * (p->c1 - x1));
                double cc3 = powi(p->c1 * x2, 6);
                double cc0 = p->c0 * cc1 * cc2 + p->c4 * cc3;
                double fdf = 0.0, fdv = 0.0;
                if (x1 >= srpx) {
                    fdf = exp(-srpk * (x1 - srpx)) - srpy;
                    fdv = cc0;
                    cc0 = fdv * fdf;
                }
                scal = scal + cc0;
                double cc4 = (-p->c2 * cc2 + p->c3 * PI * sin(p->c3 * PI * (p->c1 - x1))) * p->c0 * cc1 -
                             6.0 * p->c4 * cc3 * x2;
                if (x1 >= srpx) cc4 = cc4 * fdf - fdv * srpk * (fdf + srpy);
                scald = scald + cc4;
            }
        }
        xx[k - nnp] = x1;
        yy[k - nnp] = scal;
        if (k == nnp) d1 = scald;
    }
    out->sc3 = (yy[nn - 1] * xx[nn - 1] - kf) * powi(xx[nn - 1], 5);
    double d2 = -kf / powi(xx[nn - 1], 2) - 6.0 * out->sc3 / powi(xx[nn - 1], 7);
    double dd2 = 2.0 * kf / powi(xx[nn - 1], 3) + 42.0 * out->sc3 / powi(xx[nn - 1], 8);
    squarespline(nn, xx, yy, ep, d1, d2, dd2);
    out->sc1 = -d1 * powi(xx[0], 13) / 12.0;
    double y0 = ep[0] + xx[0] * (ep[1] + xx[0] * ep[2]);
    out->sc2 = y0 - out->sc1 / powi(xx[0], 12);
    free(xx);
    free(yy);
    return BROMOC_OK;
}

/* SPLINE, spline.f90:19-98 (Forsythe, Malcolm and Moler's cubic spline) */
EXPORT int orc_spline(int32_t n, const double *x, const double *y, double *b, double *c,
                      double *d) {
    if (n < 2) return BROMOC_E_ARG;
    const double *X = x - 1, *Y = y - 1;
    double *B = b - 1, *Cc = c - 1, *D = d - 1;
    int nm1 = n - 1;
    if (n > 2) {
        D[1] = X[2] - X[1];
        Cc[2] = (Y[2] - Y[1]) / D[1];
        for (int i = 2; i <= nm1; i++) {
            D[i] = X[i + 1] - X[i];
            B[i] = 2.0 * (D[i - 1] + D[i]);
            Cc[i + 1] = (Y[i + 1] - Y[i]) / D[i];
            Cc[i] = Cc[i + 1] - Cc[i];
        }
        B[1] = -D[1];
        B[n] = -D[n - 1];
        Cc[1] = 0.0;
        Cc[n] = 0.0;
        if (n > 3) {
            Cc[1] = Cc[3] / (X[4] - X[2]) - Cc[2] / (X[3] - X[1]);
            Cc[n] = Cc[n - 1] / (X[n] - X[n - 2]) - Cc[n - 2] / (X[n - 1] - X[n - 3]);
            Cc[1] = Cc[1] * powi(D[1], 2) / (X[4] - X[1]);
            Cc[n] = -Cc[n] * powi(D[n - 1], 2) / (X[n] - X[n - 3]);
        }
        for (int i = 2; i <= n; i++) {
            double t = D[i - 1] / B[i - 1];
            B[i] = B[i] - t * D[i - 1];
            Cc[i] = Cc[i] - t * Cc[i - 1];
        }
        Cc[n] = Cc[n] / B[n];
        for (int ib = 1; ib <= nm1; ib++) {
            int i = n - ib;
            Cc[i] = (Cc[i] - D[i] * Cc[i + 1]) / B[i];
        }
        B[n] = (Y[n] - Y[nm1]) / D[nm1] + D[nm1] * (Cc[nm1] + 2.0 * Cc[n]);
        for (int i = 1; i <= nm1; i++) {
            B[i] = (Y[i + 1] - Y[i]) / D[i] - D[i] * (Cc[i + 1] + 2.0 * Cc[i]);
            D[i] = (Cc[i + 1] - Cc[i]) / D[i];
            Cc[i] = 3.0 * Cc[i];
        }
        Cc[n] = 3.0 * Cc[n];
        D[n] = D[n - 1];
    } else {
        B[1] = (Y[2] - Y[1]) / (X[2] - X[1]);
        Cc[1] = 0.0;
        D[1] = 0.0;
        B[2] = B[1];
        Cc[2] = 0.0;
        D[2] = 0.0;
    }
    return BROMOC_OK;
}

/* ------------------------------------------------------------ context */
EXPORT void orc_destroy(orc_ctx *c) {
    if (!c) return;
    void *ps[] = {c->x, c->y, c->z, c->fx, c->fy, c->fz, c->typei, c->typtyp, c->namsite,
                  c->nwtype, c->pairidx, c->stfree, c->cg, c->fact1a, c->fact2a, c->ep,
                  c->scal, c->zcont, c->pairs, c->phix, c->phiv, c->warn, c->nforward,
                  c->nbackward, c->noise, c->ibfftyp, c->ntotat, c->ibuffer, c->avnum, c->mu,
                  c->kb, c->lzmin, c->lzmax, c->rmin, c->rmax, c->bufferbias, c->diffusion,
                  c->ampl3, c->p3, c->xs, c->ys, c->bs, c->cs, c->ds, c->gsrfen, c->greff, c->srfe, c->reff,
                  c->srfed, c->reffd, c->dids, c->sn, c->csn, c->af, c->kx, c->ky, c->kz,
                  c->contrx, c->contry, c->contrz, c->ampl1, c->p1, c->ampl2, c->p2,
                  c->rcylinder, c->rho, c->rdna, c->gr, c->prob2, c->prob3, c->lstpx,
                  c->lstpy, c->lstpz, c->bnorm, c->mmij, c->coef, c->mq, c->lpx, c->lpy,
                  c->lpz};
    for (size_t k = 0; k < sizeof(ps) / sizeof(ps[0]); k++) free(ps[k]);
    for (int k = 0; k < 8; k++) free(c->lists[k]);
    for (int k = 0; k < 6; k++) free(c->lvals[k]);
    free(c);
}

EXPORT orc_ctx *orc_create(const bromoc_system *s) {
    if (!s || s->ntot < 0 || s->capacity < s->ntot) return NULL;
    if ((s->flags & BROMOC_QMMIJ) &&
        (s->mmij_shape != 1 || s->xnpol < 3 || s->ynpol < 3 || s->znpol < 3 || s->xnpol > 64 ||
         s->ynpol > 64 || s->znpol > 64 || s->ntpol < 1))
        return NULL;
    if (((s->flags & BROMOC_QSPHERE) && (s->flags & BROMOC_QECYL)) ||
        ((s->flags & BROMOC_QSPHERE) && !(s->rsphe > 0.0)) ||
        ((s->flags & BROMOC_QECYL) && !(s->iecx > 0.0 && s->iecy > 0.0)))
        return NULL;
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->s = *s;
    int cap = s->capacity;
    c->cap = cap;
    size_t nd = sizeof(double) * (size_t)cap;
    c->x = (double *)calloc(cap + 1, sizeof(double));
    c->y = (double *)calloc(cap + 1, sizeof(double));
    c->z = (double *)calloc(cap + 1, sizeof(double));
    memcpy(c->x, s->x, sizeof(double) * s->ntot);
    memcpy(c->y, s->y, sizeof(double) * s->ntot);
    memcpy(c->z, s->z, sizeof(double) * s->ntot);
    c->fx = (double *)calloc(cap + 1, sizeof(double));
    c->fy = (double *)calloc(cap + 1, sizeof(double));
    c->fz = (double *)calloc(cap + 1, sizeof(double));
    (void)nd;
    c->typei = (int32_t *)calloc(cap + 1, sizeof(int32_t));
    if (s->typei) memcpy(c->typei, s->typei, sizeof(int32_t) * s->ntot);
    int ns = s->nsites;
    c->typtyp = (int32_t *)dupmem(s->typtyp, sizeof(int32_t) * ns);
    c->namsite = (int32_t *)dupmem(s->namsite, sizeof(int32_t) * ns);
    c->stfree = (int8_t *)dupmem(s->stfree, ns);
    c->nwtype = (int32_t *)dupmem(s->nwtype, sizeof(int32_t) * (s->ntype + 1));
    c->cg = (double *)dupmem(s->cg, sizeof(double) * (s->ntype + 1));
    c->fact1a = (double *)dupmem(s->fact1a, sizeof(double) * (s->ntype + 1));
    c->fact2a = (double *)dupmem(s->fact2a, sizeof(double) * (s->ntype + 1));
    c->pairidx = (int32_t *)dupmem(s->pairidx, sizeof(int32_t) * (s->nttyp + 1) * (s->nttyp + 1));
    c->pairs = (bromoc_pair *)dupmem(s->pairs, sizeof(bromoc_pair) * (s->npair + 1));
    c->ep = (double *)dupmem(s->ep, sizeof(double) * 3 * (size_t)s->nknots);
    c->scal = (double *)dupmem(s->scal, sizeof(double) * (s->nttyp + 1));
    c->zcont = (double *)dupmem(s->zcont, sizeof(double) * (s->cntpts > 0 ? s->cntpts : 0));
    int cnt[8] = {2 * s->nbond, 3 * s->nangle, 4 * s->ndihe, 2 * s->nstack,
                  2 * s->nbp, 2 * s->nex, 2 * s->nqq, 2 * s->nsolv};
    const int32_t *lp[8] = {s->sitebond, s->siteangle, s->sitedihe, s->sitestack,
                            s->sitebp, s->siteex, s->siteqq, s->siteslv};
    for (int k = 0; k < 8; k++) c->lists[k] = (int32_t *)dupmem(lp[k], sizeof(int32_t) * cnt[k]);
    int vcnt[6] = {s->nbond, s->nangle, s->ndihe, s->nstack, s->nbp, s->nex};
    const double *vp[6] = {s->distbond, s->valangle, s->valdihe, s->sgstack, s->sgbp, s->sgex};
    for (int k = 0; k < 6; k++) c->lvals[k] = (double *)dupmem(vp[k], sizeof(double) * vcnt[k]);
    if (s->flags & BROMOC_QPHIX) {
        size_t n = (size_t)s->phix.nclx * s->phix.ncly * s->phix.nclz;
        c->phix = (float *)dupmem(s->phix.data, sizeof(float) * n);
    }
    if (s->flags & BROMOC_QPHIV) {
        size_t n = (size_t)s->phiv.nclx * s->phiv.ncly * s->phiv.nclz;
        if (s->flags & BROMOC_QNMCDEN) {  /* readphi.f90:176-240 */
            const int nm = ((s->flags & BROMOC_QNUCL) ? 6 : 0) +
                           ((s->flags & BROMOC_QPAR) ? s->ntype - s->nold : 0);
            if (s->phiv.nmaps != nm) { orc_destroy(c); return NULL; }   /* caller's buffer */
            n *= (size_t)nm;
        }
        c->phiv = (int8_t *)dupmem(s->phiv.data, n);
    }
    c->warn = (int32_t *)calloc(s->nttyp + 2, sizeof(int32_t));
    int nt = s->ntype - s->nold;
    int np = s->cntpts > 0 ? s->cntpts : 1;
    c->nforward = (int32_t *)calloc((size_t)(nt > 0 ? nt : 1) * np, sizeof(int32_t));
    c->nbackward = (int32_t *)calloc((size_t)(nt > 0 ? nt : 1) * np, sizeof(int32_t));
    c->celec = celec_value();
    if (s->flags & (BROMOC_QDIFFUSE | BROMOC_QPROFILE))
        c->diffusion = (double *)dupmem(s->diffusion, sizeof(double) * (s->ntype + 1));
    if (s->flags & BROMOC_QDIFFUSE) {
        c->ampl3 = (double *)dupmem(s->ampl3, sizeof(double) * (s->ntype + 1));
        c->p3 = (double *)dupmem(s->p3, sizeof(double) * (s->ntype + 1));
    }
    if (s->flags & BROMOC_QPROFILE) {
        size_t n = sizeof(double) * (size_t)s->nspline;
        c->xs = (double *)dupmem(s->xs, n);
        c->ys = (double *)dupmem(s->ys, n);
        c->bs = (double *)dupmem(s->bs, n);
        c->cs = (double *)dupmem(s->cs, n);
        c->ds = (double *)dupmem(s->ds, n);
    }
    c->seval_i = c->sevald_i = 1;
    if (s->flags & BROMOC_QRFPAR) {
        size_t n = (size_t)s->rf.nclx * s->rf.ncly * s->rf.nclz;
        size_t ng = (s->flags & BROMOC_QRFPSIN) ? 1 : (size_t)(s->ntype - s->nold);
        c->gsrfen = (float *)dupmem(s->gsrfen, sizeof(float) * n * ng);
        c->greff = (float *)dupmem(s->greff, sizeof(float) * n * ng);
    }
    c->srfe = (double *)calloc(cap + 1, sizeof(double));
    c->reff = (double *)calloc(cap + 1, sizeof(double));
    c->srfed = (double *)calloc(3 * (size_t)cap + 3, sizeof(double));
    c->reffd = (double *)calloc(3 * (size_t)cap + 3, sizeof(double));
    c->dids = (double *)calloc(5 * (size_t)cap + 5, sizeof(double));
    if ((s->flags & BROMOC_QPROXDIFF) && !c->diffusion)
        c->diffusion = (double *)dupmem(s->diffusion, sizeof(double) * (s->ntype + 1));
    if (s->flags & BROMOC_QMEMB) {
        size_t n1 = (size_t)s->nttyp + 1;
        c->ampl1 = (double *)dupmem(s->ampl1, sizeof(double) * n1);
        c->ampl2 = (double *)dupmem(s->ampl2, sizeof(double) * n1);
        c->rcylinder = (double *)dupmem(s->rcylinder, sizeof(double) * n1);
        c->p1 = (double *)dupmem(s->p1, sizeof(double) * 2 * n1);
        c->p2 = (double *)dupmem(s->p2, sizeof(double) * 2 * n1);
    }
    if (s->flags & BROMOC_QMMIJ) {
        size_t n = (size_t)s->ntpol;
        c->lstpx = (int32_t *)dupmem(s->lstpx, sizeof(int32_t) * n);
        c->lstpy = (int32_t *)dupmem(s->lstpy, sizeof(int32_t) * n);
        c->lstpz = (int32_t *)dupmem(s->lstpz, sizeof(int32_t) * n);
        c->bnorm = (double *)dupmem(s->bnorm, sizeof(double) * n);
        c->mmij = (double *)dupmem(s->mmij, sizeof(double) * n * n);
        c->coef = (double *)calloc(n, sizeof(double));
        c->mq = (double *)calloc(n, sizeof(double));
        c->lpx = (double *)calloc(((size_t)cap + 1) * s->xnpol, sizeof(double));
        c->lpy = (double *)calloc(((size_t)cap + 1) * s->ynpol, sizeof(double));
        c->lpz = (double *)calloc(((size_t)cap + 1) * s->znpol, sizeof(double));
    }
    if (s->flags & BROMOC_QAPFOR) {
        c->sn = (int32_t *)dupmem(s->sn, sizeof(int32_t) * s->afn);
        c->af = (double *)dupmem(s->af, sizeof(double) * 3 * s->afn);
    }
    if (s->flags & BROMOC_QCONTRANS) {
        c->csn = (int32_t *)dupmem(s->csn, sizeof(int32_t) * s->ctn);
        c->kx = (double *)dupmem(s->kx, sizeof(double) * s->ctn);
        c->ky = (double *)dupmem(s->ky, sizeof(double) * s->ctn);
        c->kz = (double *)dupmem(s->kz, sizeof(double) * s->ctn);
        c->contrx = (double *)dupmem(s->contrx, sizeof(double) * (s->ctn + 1));
        c->contry = (double *)dupmem(s->contry, sizeof(double) * (s->ctn + 1));
        c->contrz = (double *)dupmem(s->contrz, sizeof(double) * (s->ctn + 1));
    }
    return c;
}

EXPORT int orc_set_positions(orc_ctx *c, int32_t ntot, const double *x, const double *y,
                             const double *z, const int32_t *typei) {
    if (ntot > c->cap || ntot < c->s.nsites + c->s.nfix) return BROMOC_E_ARG;
    c->s.ntot = ntot;
    memcpy(c->x, x, sizeof(double) * ntot);
    memcpy(c->y, y, sizeof(double) * ntot);
    memcpy(c->z, z, sizeof(double) * ntot);
    if (typei) memcpy(c->typei, typei, sizeof(int32_t) * ntot);
    return BROMOC_OK;
}

EXPORT int orc_get_positions(orc_ctx *c, double *x, double *y, double *z, int32_t *typei,
                             int32_t *ntot) {
    int n = c->s.ntot;
    if (x) memcpy(x, c->x, sizeof(double) * n);
    if (y) memcpy(y, c->y, sizeof(double) * n);
    if (z) memcpy(z, c->z, sizeof(double) * n);
    if (typei) memcpy(typei, c->typei, sizeof(int32_t) * n);
    if (ntot) *ntot = n;
    return BROMOC_OK;
}

EXPORT int orc_get_forces(orc_ctx *c, double *fx, double *fy, double *fz) {
    int n = c->s.ntot;
    memcpy(fx, c->fx, sizeof(double) * n);
    memcpy(fy, c->fy, sizeof(double) * n);
    memcpy(fz, c->fz, sizeof(double) * n);
    return BROMOC_OK;
}

EXPORT int orc_get_flux(orc_ctx *c, int32_t *nf, int32_t *nb) {
    int nt = c->s.ntype - c->s.nold, np = c->s.cntpts;
    if (nt <= 0 || np <= 0) return BROMOC_OK;
    memcpy(nf, c->nforward, sizeof(int32_t) * nt * np);
    memcpy(nb, c->nbackward, sizeof(int32_t) * nt * np);
    return BROMOC_OK;
}

EXPORT int orc_get_warn(orc_ctx *c, int32_t *warn) {
    memcpy(warn, c->warn, sizeof(int32_t) * (c->s.nttyp + 1));
    return BROMOC_OK;
}

static uint32_t supported_mask(void) {
    return BROMOC_QPAR | BROMOC_QNUCL | BROMOC_QNONBOND | BROMOC_QNOBOND | BROMOC_QPHIX |
           BROMOC_QPHIV | BROMOC_QSVDW | BROMOC_QSOLV | BROMOC_QDEBY | BROMOC_QCOUNTION |
           BROMOC_QDNAFREE | BROMOC_QDIFFUSE | BROMOC_QPROFILE | BROMOC_QRFPAR | BROMOC_QRFPSIN |
           BROMOC_QBSPLINE | BROMOC_QPROXDIFF | BROMOC_QNOTRANS | BROMOC_QAPFOR |
           BROMOC_QCONTRANS | BROMOC_QMEMB | BROMOC_QMMIJ | BROMOC_QDEBYHYB | BROMOC_QNMCDEN |
           BROMOC_QSPHERE | BROMOC_QECYL;
}

/* OUTBOX, grandmod.f90:74-97: outside the sphere, the elliptic cylinder or the box */
static int outbox(const bromoc_system *s, double xi, double yi, double zi) {
    if (s->flags & BROMOC_QSPHERE)
        return (xi - s->cx) * (xi - s->cx) + (yi - s->cy) * (yi - s->cy) +
               (zi - s->cz) * (zi - s->cz) > s->rsphe2;
    if (s->flags & BROMOC_QECYL) {
        const double ex = (xi - s->cx) * s->iecx, ey = (yi - s->cy) * s->iecy;
        return ex * ex + ey * ey > 1.0 || zi < s->lz2m || zi > s->lz2p;
    }
    return xi < s->lx2m || xi > s->lx2p || yi < s->ly2m || yi > s->ly2p || zi < s->lz2m ||
           zi > s->lz2p;
}

EXPORT int orc_check_flags(uint32_t flags) {
    return (flags & ~supported_mask()) ? BROMOC_E_UNSUPPORTED : BROMOC_OK;
}

/* ------------------------------------------------------------- grids */
/* STATICF1, staticf1.f90:19-118 */
static void staticf1(orc_ctx *c, int qforces, double *egsbpa) {
    const bromoc_system *s = &c->s;
    const bromoc_grid *g = &s->phix;
    const double one = 1.0, celec = c->celec;
    int ncyz = g->ncly * g->nclz;
    double dcel = g->dcel, idcel = 1.0 / dcel;
    double tranx = 0.5 * (g->nclx - 1) * dcel, trany = 0.5 * (g->ncly - 1) * dcel,
           tranz = 0.5 * (g->nclz - 1) * dcel;
    *egsbpa = 0.0;
    for (int i = 0; i < s->ntot; i++) {
        if (!(i < s->nsites || i >= s->nsites + s->nfix)) continue;
        int ok = 0;
        double chi = 0.0;
        if (i < s->nsites) {
            if (c->namsite[i] == BROMOC_SITE_P) { chi = s->cgnuc; ok = 1; }
        } else {
            chi = c->cg[abs(c->typei[i])];
            ok = 1;
        }
        if (ok)
            ok = c->x[i] <= g->xbcen + tranx && c->x[i] >= g->xbcen - tranx &&
                 c->y[i] <= g->ybcen + trany && c->y[i] >= g->ybcen - trany &&
                 c->z[i] <= g->zbcen + tranz && c->z[i] >= g->zbcen - tranz;
        if (!ok) continue;
        double rxnfx = 0.0, rxnfy = 0.0, rxnfz = 0.0;
        double xi = c->x[i] + tranx - g->xbcen;
        double yi = c->y[i] + trany - g->ybcen;
        double zi = c->z[i] + tranz - g->zbcen;
        int ix = (int)(xi * idcel), iy = (int)(yi * idcel), iz = (int)(zi * idcel);
        if (ix == g->nclx - 1) ix = g->nclx - 2;
        if (iy == g->ncly - 1) iy = g->ncly - 2;
        if (iz == g->nclz - 1) iz = g->nclz - 2;
        for (int n1 = ix; n1 <= ix + 1; n1++) {
            double ai = xi - n1 * dcel;
            double aisign = fsign(one, ai);
            ai = one - fabs(ai) * idcel;
            for (int n2 = iy; n2 <= iy + 1; n2++) {
                double bi = yi - n2 * dcel;
                double bisign = fsign(one, bi);
                bi = one - fabs(bi) * idcel;
                for (int n3 = iz; n3 <= iz + 1; n3++) {
                    double ci = zi - n3 * dcel;
                    double cisign = fsign(one, ci);
                    ci = one - fabs(ci) * idcel;
                    double fi = ai * bi * ci;
                    size_t in3 = (size_t)n1 * ncyz + (size_t)n2 * g->nclz + n3;
                    double ph = (double)c->phix[in3];
                    if (qforces) {
                        double prefac = ph * celec * chi * idcel;
                        if (ai < (1.0 - RSMALL) && ai > RSMALL) rxnfx = rxnfx + aisign * bi * ci * prefac;
                        if (bi < (1.0 - RSMALL) && bi > RSMALL) rxnfy = rxnfy + bisign * ai * ci * prefac;
                        if (ci < (1.0 - RSMALL) && ci > RSMALL) rxnfz = rxnfz + cisign * ai * bi * prefac;
                    }
                    *egsbpa = *egsbpa + (fi * chi * ph * celec);
                }
            }
        }
        if (qforces) {
            c->fx[i] = c->fx[i] + rxnfx;
            c->fy[i] = c->fy[i] + rxnfy;
            c->fz[i] = c->fz[i] + rxnfz;
        }
    }
}

/* Qnmcden map choice of particle i (vdwgd1trln.f90:61-80, vdwgd0trln.f90:56-75,
 * vdwgd1spln.f90:61-80, vdwgd0spln.f90:58-77): sets *ifir to the first node of the map and
 * returns 1, or returns 0 where none of the reference's branches applies (the reference
 * then keeps ifir as it was).  Maps: S, P, Ab, Tb, Cb, Gb with DNA, then ion types from
 * nold+1 (the reference's inuc+1 or 2*inuc+1 for one or two strands: nold = istrs*inuc). */
static int nmcden_map(const orc_ctx *c, int i, int itype, size_t ncel3, size_t *ifir) {
    const bromoc_system *s = &c->s;
    int qnucl = (s->flags & BROMOC_QNUCL) != 0, qpar = (s->flags & BROMOC_QPAR) != 0;
    if (qnucl && i < s->nsites) {
        *ifir = (size_t)c->namsite[i] * ncel3;   /* BROMOC_SITE_S..GB = 0..5 */
        return 1;
    }
    if (qnucl && qpar && i >= s->nsites + s->nfix) {
        int numb = itype - (s->nold + 1);
        *ifir = (size_t)(6 + numb) * ncel3;
        return 1;
    }
    if (!qnucl && qpar && i >= s->nfix) {
        *ifir = (size_t)(itype - 1) * ncel3;
        return 1;
    }
    return 0;
}

/* VDWGD1TRLN, vdwgd1trln.f90:19-161 */
static void vdwgd1trln(orc_ctx *c, int qforces, double *evdwgd) {
    const bromoc_system *s = &c->s;
    const bromoc_grid *g = &s->phiv;
    const double one = 1.0;
    int ncyz = g->ncly * g->nclz;
    double dcel = g->dcel, idcel = 1.0 / dcel;
    double tranx = 0.5 * (g->nclx - 1) * dcel, trany = 0.5 * (g->ncly - 1) * dcel,
           tranz = 0.5 * (g->nclz - 1) * dcel;
    double esvdw = s->svdw;
    int qsvdw = (s->flags & BROMOC_QSVDW) != 0;
    size_t ncel3 = (size_t)g->nclx * ncyz, ifir = 0;
    *evdwgd = 0.0;
    for (int i = 0; i < s->ntot; i++) {
        if (!(i < s->nsites || i >= s->nsites + s->nfix)) continue;
        int ok = c->x[i] <= g->xbcen + tranx && c->x[i] >= g->xbcen - tranx &&
                 c->y[i] <= g->ybcen + trany && c->y[i] >= g->ybcen - trany &&
                 c->z[i] <= s->vzmax && c->z[i] >= s->vzmin;
        if (!ok) continue;
        double xi = c->x[i] + tranx - g->xbcen;
        double yi = c->y[i] + trany - g->ybcen;
        double zi = c->z[i] + tranz - g->zbcen;
        int itype2 = (i < s->nsites) ? c->typtyp[i] : c->nwtype[abs(c->typei[i])];
        if (s->flags & BROMOC_QNMCDEN) nmcden_map(c, i, abs(c->typei[i]), ncel3, &ifir);
        else if (qsvdw) esvdw = s->svdw * c->scal[itype2];
        double vdwfx = 0.0, vdwfy = 0.0, vdwfz = 0.0;
        int ix = (int)(xi * idcel), iy = (int)(yi * idcel), iz = (int)(zi * idcel);
        if (ix == g->nclx - 1) ix = g->nclx - 2;
        if (iy == g->ncly - 1) iy = g->ncly - 2;
        if (iz == g->nclz - 1) iz = g->nclz - 2;
        double phisum = 0.0;
        for (int n1 = ix; n1 <= ix + 1; n1++) {
            double ai = xi - n1 * dcel;
            double aisign = fsign(one, ai);
            ai = 1.0 - fabs(ai) * idcel;
            for (int n2 = iy; n2 <= iy + 1; n2++) {
                double bi = yi - n2 * dcel;
                double bisign = fsign(one, bi);
                bi = 1.0 - fabs(bi) * idcel;
                for (int n3 = iz; n3 <= iz + 1; n3++) {
                    double ci = zi - n3 * dcel;
                    double cisign = fsign(one, ci);
                    ci = 1.0 - fabs(ci) * idcel;
                    double fi = ai * bi * ci;
                    size_t in3 = (size_t)n1 * ncyz + (size_t)n2 * g->nclz + n3;
                    double phis = (double)c->phiv[in3 + ifir];
                    phisum = phisum + phis;
                    if (qforces) {
                        double prefac = phis * esvdw * idcel;
                        if (ai < (1.0 - RSMALL) && ai > RSMALL) vdwfx = vdwfx + aisign * bi * ci * prefac;
                        if (bi < (1.0 - RSMALL) && bi > RSMALL) vdwfy = vdwfy + bisign * ai * ci * prefac;
                        if (ci < (1.0 - RSMALL) && ci > RSMALL) vdwfz = vdwfz + cisign * ai * bi * prefac;
                    }
                    *evdwgd = *evdwgd + fi * esvdw * phis;
                }
            }
        }
        if (phisum >= s->thold8) c->warn[itype2] = c->warn[itype2] + 1;
        if (qforces) {
            c->fx[i] = c->fx[i] + vdwfx;
            c->fy[i] = c->fy[i] + vdwfy;
            c->fz[i] = c->fz[i] + vdwfz;
        }
    }
}

/* M3 / DM3, m3.f90:19-42, dm3.f90:19-38 */
static double m3(double x) {
    double m2m = (x >= 0.0 && x <= 2.0) ? 1.0 - fabs(x - 1.0) : 0.0;
    double m2 = (x >= 1.0 && x <= 3.0) ? 1.0 - fabs(x - 2.0) : 0.0;
    return (x >= 0.0 && x <= 3.0) ? x * 0.5 * m2m + (3.0 - x) * 0.5 * m2 : 0.0;
}
static double dm3(double x) {
    double m2m = (x >= 0.0 && x <= 2.0) ? 1.0 - fabs(x - 1.0) : 0.0;
    double m2 = (x >= 1.0 && x <= 3.0) ? 1.0 - fabs(x - 2.0) : 0.0;
    return m2m - m2;
}

/* NINT: nearest integer, halves away from zero */
static int nint(double a) { return (int)round(a); }

/* VDWGD1SPLN, vdwgd1spln.f90:19-170 */
static void vdwgd1spln(orc_ctx *c, int qforces, double *evdwgd) {
    const bromoc_system *s = &c->s;
    const bromoc_grid *g = &s->phiv;
    int nclx2 = g->nclx, ncly2 = g->ncly, nclz2 = g->nclz;
    int ncyz = ncly2 * nclz2;
    double dcel2 = g->dcel, idcel2 = 1.0 / dcel2;
    double tranx2 = 0.5 * (nclx2 - 1) * dcel2, trany2 = 0.5 * (ncly2 - 1) * dcel2,
           tranz2 = 0.5 * (nclz2 - 1) * dcel2;
    double esvdw = s->svdw;
    size_t ncel3 = (size_t)nclx2 * ncyz, ifir = 0;
    *evdwgd = 0.0;
    for (int i = 0; i < s->ntot; i++) {
        if (!(i < s->nsites || i >= s->nsites + s->nfix)) continue;
        int ok = c->x[i] <= g->xbcen + tranx2 - dcel2 && c->x[i] >= g->xbcen - tranx2 + dcel2 &&
                 c->y[i] <= g->ybcen + trany2 - dcel2 && c->y[i] >= g->ybcen - trany2 + dcel2 &&
                 c->z[i] <= s->vzmax && c->z[i] >= s->vzmin;
        if (!ok) continue;
        double xi = c->x[i] + tranx2 - g->xbcen;
        double yi = c->y[i] + trany2 - g->ybcen;
        double zi = c->z[i] + tranz2 - g->zbcen;
        int itype2 = (i < s->nsites) ? c->typtyp[i] : c->nwtype[abs(c->typei[i])];
        if (s->flags & BROMOC_QNMCDEN) nmcden_map(c, i, abs(c->typei[i]), ncel3, &ifir);
        else if (s->flags & BROMOC_QSVDW) esvdw = s->svdw * c->scal[itype2];
        double vdwfx = 0.0, vdwfy = 0.0, vdwfz = 0.0;
        int ix = nint(xi * idcel2), iy = nint(yi * idcel2), iz = nint(zi * idcel2);
        if (ix == 0) ix = 1;
        if (iy == 0) iy = 1;
        if (iz == 0) iz = 1;
        if (ix == nclx2 - 1) ix = nclx2 - 2;
        if (iy == ncly2 - 1) iy = ncly2 - 2;
        if (iz == nclz2 - 1) iz = nclz2 - 2;
        double phisum = 0.0;
        for (int k = ix - 1; k <= ix + 1; k++) {
            size_t ipx = (size_t)k * ncyz;
            double xc = k * dcel2;
            double ai = 1.5 - (xi - xc) * idcel2;
            double dai = dm3(ai);
            ai = m3(ai);
            if (!(ai != 0.0 || dai != 0.0)) continue;
            for (int l = iy - 1; l <= iy + 1; l++) {
                size_t ipy = (size_t)l * nclz2 + ipx;
                double yc = l * dcel2;
                double bi = 1.5 - (yi - yc) * idcel2;
                double dbi = dm3(bi);
                bi = m3(bi);
                if (!(bi != 0.0 || dbi != 0.0)) continue;
                for (int m = iz - 1; m <= iz + 1; m++) {
                    size_t ipz = m + ipy;
                    double zc = m * dcel2;
                    double ci = 1.5 - (zi - zc) * idcel2;
                    double dci = dm3(ci);
                    ci = m3(ci);
                    if (!(ci != 0.0 || dci != 0.0)) continue;
                    double fi = ai * bi * ci;
                    double phivi = (double)c->phiv[ipz + ifir];
                    double phivsv = phivi * esvdw;
                    phisum = phisum + phivi;
                    if (qforces) {
                        double prefac = phivsv * idcel2;
                        vdwfx = vdwfx + dai * bi * ci * prefac;
                        vdwfy = vdwfy + ai * dbi * ci * prefac;
                        vdwfz = vdwfz + ai * bi * dci * prefac;
                    }
                    *evdwgd = *evdwgd + fi * phivsv;
                }
            }
        }
        if (phisum >= s->thold27) c->warn[itype2] = c->warn[itype2] + 1;
        if (qforces) {
            c->fx[i] = c->fx[i] + vdwfx;
            c->fy[i] = c->fy[i] + vdwfy;
            c->fz[i] = c->fz[i] + vdwfz;
        }
    }
}

/* SWITCH1, switch1.f90:19-47 */
static void switch1(double *sw, double *dsw, double z, double p1, double ip1, double zmemb1,
                    double zmemb2) {
    if (z >= (zmemb2 + p1) || z <= (zmemb1 - p1)) {
        *sw = 0.0; *dsw = 0.0;
    } else if (z > zmemb2) {
        double delz = z - zmemb2, delzp1 = delz * ip1;
        *sw = 1.0 + (2.0 * delzp1 - 3.0) * (delzp1 * delzp1);
        *dsw = -6.0 * (delzp1 - 1.0) * ip1 * delzp1;
    } else if (z < zmemb1) {
        double delz = zmemb1 - z, delzp1 = delz * ip1;
        *sw = 1.0 + (2.0 * delzp1 - 3.0) * (delzp1 * delzp1);
        *dsw = 6.0 * (delzp1 - 1.0) * ip1 * delzp1;
    } else {
        *sw = 1.0; *dsw = 0.0;
    }
}

/* SWITCH2, switch2.f90:19-38 (r is set in the wall branch only) */
static void switch2(double *sw, double *dsw, double r2, double radius, double p2, double ip2,
                    double *r) {
    if (r2 <= radius * radius) {
        *sw = 0.0; *dsw = 0.0;
    } else if (r2 < (radius + p2) * (radius + p2)) {
        *r = sqrt(r2);
        double delr = radius + p2 - *r, delrp2 = delr * ip2;
        *sw = 1.0 + (2.0 * delrp2 - 3.0) * (delrp2 * delrp2);
        *dsw = 6.0 * (delrp2 - 1.0) * delrp2 * ip2;
    } else {
        *sw = 1.0; *dsw = 0.0;
    }
}

/* MEMBRANE, membrane.f90:19-107 */
static void membrane(orc_ctx *c, int qforces, double *ememb) {
    const bromoc_system *s = &c->s;
    *ememb = 0.0;
    for (int iat = 0; iat < s->ntot; iat++) {
        int lgiat = (iat < s->nsites || iat >= s->nsites + s->nfix) && qforces;
        int itype;
        double charge;
        if (iat < s->nsites) {
            itype = c->typtyp[iat];
            charge = 0.0;
            if (c->namsite[iat] == BROMOC_SITE_P) charge = s->cgnuc;
        } else {
            int itype2 = abs(c->typei[iat]);
            charge = c->cg[itype2];
            itype = c->nwtype[itype2];
        }
        double x = c->x[iat], y = c->y[iat], z = c->z[iat];
        if (s->voltage != 0.0 && charge != 0.0) {
            if (z < s->zmemb1) {
                *ememb = *ememb + charge * s->afact * exp(s->ikappa * (z - s->zmemb1));
                if (lgiat) c->fz[iat] = c->fz[iat] - charge * s->afact * s->ikappa * exp(s->ikappa * (z - s->zmemb1));
            } else if (z >= s->zmemb1 && z <= s->zmemb2) {
                *ememb = *ememb + charge * s->afact * (s->ceps * s->ikappa * (z - s->zmemb1) + 1.0);
                if (lgiat) c->fz[iat] = c->fz[iat] - charge * s->afact * s->ceps * s->ikappa;
            } else if (z > s->zmemb2) {
                *ememb = *ememb + charge * (s->voltage - s->afact * exp(-s->ikappa * (z - s->zmemb2)));
                if (lgiat) c->fz[iat] = c->fz[iat] - charge * s->afact * s->ikappa * exp(-s->ikappa * (z - s->zmemb2));
            }
        }
        if (c->ampl1[itype] > 0.0) {
            double sw1, dsw1;
            switch1(&sw1, &dsw1, z, c->p1[2 * itype], c->p1[2 * itype + 1], s->zmemb1, s->zmemb2);
            if (sw1 != 0.0) {
                if (s->qpore) {
                    double r2 = x * x + y * y, sw2, dsw2, r = 0.0;
                    switch2(&sw2, &dsw2, r2, c->rcylinder[itype], c->p2[2 * itype], c->p2[2 * itype + 1], &r);
                    if (sw2 != 0.0) {
                        if (sw1 == 1.0) {
                            if (sw2 == 1.0) {
                                c->warn[itype] = c->warn[itype] + 1;
                                *ememb = *ememb + c->ampl2[itype];
                            } else {
                                if (lgiat) {
                                    double ir = 1.0 / r;
                                    c->fx[iat] = c->fx[iat] + c->ampl2[itype] * dsw2 * x * ir;
                                    c->fy[iat] = c->fy[iat] + c->ampl2[itype] * dsw2 * y * ir;
                                }
                                *ememb = *ememb + c->ampl2[itype] * sw2;
                            }
                        } else {
                            if (sw2 == 1.0) {
                                if (lgiat) c->fz[iat] = c->fz[iat] + c->ampl1[itype] * dsw1;
                                *ememb = *ememb + c->ampl1[itype] * sw1;
                            } else {
                                if (lgiat) {
                                    double ir = 1.0 / r;
                                    c->fx[iat] = c->fx[iat] + 0.5 * c->ampl2[itype] * dsw2 * x * ir;
                                    c->fy[iat] = c->fy[iat] + 0.5 * c->ampl2[itype] * dsw2 * y * ir;
                                    c->fz[iat] = c->fz[iat] + 0.5 * c->ampl1[itype] * dsw1;
                                }
                                *ememb = *ememb + 0.5 * (c->ampl2[itype] * sw2 + c->ampl1[itype] * sw1);
                            }
                        }
                    }
                } else {
                    if (sw1 == 1.0) c->warn[itype] = c->warn[itype] + 1;
                    *ememb = *ememb + c->ampl1[itype] * sw1;
                }
            }
        }
    }
}

/* the membrane part of INTERACT, interact.f90:58-119 */
static double membrane_interact(orc_ctx *c, double xj, double yj, double zj, int j, int jtype,
                                int jtype2, int beion, int alert) {
    const bromoc_system *s = &c->s;
    double ememb = 0.0, charge = 0.0;
    if (s->voltage != 0.0) {
        if (beion) charge = c->cg[jtype];
        else if (s->flags & BROMOC_QNUCL) charge = c->namsite[j] == BROMOC_SITE_P ? s->cgnuc : 0.0;
        if (charge != 0.0) {
            if (zj < s->zmemb1) ememb = ememb + charge * s->afact * exp(s->ikappa * (zj - s->zmemb1));
            else if (zj >= s->zmemb1 && zj <= s->zmemb2)
                ememb = ememb + charge * s->afact * (s->ceps * s->ikappa * (zj - s->zmemb1) + 1.0);
            else if (zj > s->zmemb2)
                ememb = ememb + charge * (s->voltage - s->afact * exp(-s->ikappa * (zj - s->zmemb2)));
        }
    }
    if (c->ampl1[jtype2] > 0.0) {
        double sw1, dsw1;
        switch1(&sw1, &dsw1, zj, c->p1[2 * jtype2], c->p1[2 * jtype2 + 1], s->zmemb1, s->zmemb2);
        if (sw1 != 0.0) {
            if (s->qpore) {
                double r2 = xj * xj + yj * yj, sw2, dsw2, r = 0.0;
                switch2(&sw2, &dsw2, r2, c->rcylinder[jtype2], c->p2[2 * jtype2], c->p2[2 * jtype2 + 1], &r);
                if (sw2 != 0.0) {
                    if (sw1 == 1.0) {
                        if (sw2 == 1.0) {
                            if (alert) c->warn[jtype2] = c->warn[jtype2] + 1;
                            ememb = ememb + 1.0e10;
                        } else {
                            ememb = ememb + c->ampl2[jtype2] * sw2;
                        }
                    } else if (sw2 == 1.0) {
                        ememb = ememb + c->ampl1[jtype2] * sw1;
                    } else {
                        ememb = ememb + 0.5 * (c->ampl2[jtype2] * sw2 + c->ampl1[jtype2] * sw1);
                    }
                }
            } else if (sw1 == 1.0) {
                if (alert) c->warn[jtype2] = c->warn[jtype2] + 1;
                ememb = ememb + 1.0e10;
            } else {
                ememb = ememb + c->ampl1[jtype2] * sw1;
            }
        }
    }
    return ememb;
}

/* ------------------------------------------------------------ ENERGY */
/* analytic pair branch, energy.mono.f90:135-178 (energies into e[0..2], returns de) */
static double pair_analytic(const orc_ctx *c, int is, double dist2, int qforces, double *eelec,
                            double *evdw, double *esrpmf) {
    const bromoc_pair *p = &c->pairs[is];
    const bromoc_system *s = &c->s;
    double idist2 = 1.0 / dist2, idist = 0.0, cofo = 0.0, dist6 = 0.0, dist12 = 0.0, dist = 0.0;
    double esrpmf1 = 0.0, esrpmf2 = 0.0, esrpmf3 = 0.0, fdf = 0.0, fdv = 0.0;
    if (p->chr || p->srpmf) idist = sqrt(idist2);
    if (p->chr) {
        cofo = p->fct * idist;
        *eelec = *eelec + cofo;
    }
    if (p->lj) {
        dist6 = pow3(p->sgp2 * idist2);
        dist12 = dist6 * dist6;
        *evdw = *evdw + p->epp4 * (dist12 - dist6);
    }
    if (p->srpmf && dist2 <= s->rth) {
        dist = 1.0 / idist;
        esrpmf1 = exp((p->c1 - dist) * p->c2);
        esrpmf2 = cos(p->c3 * PI * (p->c1 - dist));
        esrpmf3 = pow6(p->c1 * idist);
        double esrpmf0 = p->c0 * esrpmf1 * esrpmf2 + p->c4 * esrpmf3;
        if (dist >= s->srpx) {
            fdf = exp(-s->srpk * (dist - s->srpx)) - s->srpy;
            fdv = esrpmf0;
            esrpmf0 = fdv * fdf;
        }
        *esrpmf = *esrpmf + esrpmf0;
    }
    double de = 0.0;
    if (qforces) {
        if (p->chr) de = cofo * idist2;
        if (p->lj) de = de + p->epp4 * (2.0 * dist12 - dist6) * 6.0 * idist2;
        if (p->srpmf && dist2 <= s->rth) {
            double dc = (-p->c2 * esrpmf2 + p->c3 * PI * sin(p->c3 * PI * (p->c1 - dist))) * p->c0 * esrpmf1 -
                        6.0 * p->c4 * esrpmf3 * idist;
            if (dist >= s->srpx) dc = dc * fdf - fdv * s->srpk * (fdf + s->srpy);
            de = de - dc * idist;
        }
    }
    return de;
}

static void cross(const double *a, const double *b, double *r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* The `ok` guard of every DNA term (energy.mono.f90:216 etc.) parses as
 * (Qforces .and. stfree(a)) .or. stfree(b): with Qforces off the reference still adds
 * pair forces into fx, which it never zeroed in that call.  Forces are only defined after
 * a Qforces call (simul1.f90:318-322,363), so here a no-forces call leaves them alone. */

static void add_f(orc_ctx *c, int i, double sgn, const double *f) {
    c->fx[i] = c->fx[i] + sgn * f[0];
    c->fy[i] = c->fy[i] + sgn * f[1];
    c->fz[i] = c->fz[i] + sgn * f[2];
}

/* ------------------------------------------------------------ RFPAR */
static int rf_inbox(const bromoc_system *s, double x, double y, double z) {
    const bromoc_grid *g = &s->rf;
    double tx = 0.5 * (g->nclx - 1) * g->dcel, ty = 0.5 * (g->ncly - 1) * g->dcel,
           tz = 0.5 * (g->nclz - 1) * g->dcel;
    return x <= g->xbcen + tx && x >= g->xbcen - tx && y <= g->ybcen + ty &&
           y >= g->ybcen - ty && z <= g->zbcen + tz && z >= g->zbcen - tz;
}

/* srfe, reff (and gradients when qf) of one ion from the 8 corners of gsrfen and
 * sqrfac*greff, rfpar.f90:86-142 (the same loop as rfpar.f90:306-345); the innermost loop
 * is left at a zero gsrfen corner (`exit`, rfpar.f90:121). */
static void rf_point(const orc_ctx *c, double x, double y, double z, int itype, int qf,
                     double *srfe, double *reff, double *d1, double *d2) {
    const bromoc_system *s = &c->s;
    const bromoc_grid *g = &s->rf;
    double dcel = g->dcel, idcel = 1.0 / dcel;
    double tx = 0.5 * (g->nclx - 1) * dcel, ty = 0.5 * (g->ncly - 1) * dcel,
           tz = 0.5 * (g->nclz - 1) * dcel;
    long ncyz = (long)g->ncly * g->nclz, ncel3 = (long)g->nclx * ncyz;
    long ifir = (s->flags & BROMOC_QRFPSIN) ? 0 : (long)(itype - s->nold - 1) * ncel3;
    double one = 1.0, aux1 = 0.0, aux2 = 0.0;
    double a1x = 0.0, a1y = 0.0, a1z = 0.0, a2x = 0.0, a2y = 0.0, a2z = 0.0;
    double xi = x + tx - g->xbcen, yi = y + ty - g->ybcen, zi = z + tz - g->zbcen;
    int ix = (int)(xi * idcel), iy = (int)(yi * idcel), iz = (int)(zi * idcel);
    if (ix == g->nclx - 1) ix = g->nclx - 2;
    if (iy == g->ncly - 1) iy = g->ncly - 2;
    if (iz == g->nclz - 1) iz = g->nclz - 2;
    for (int n1 = ix; n1 <= ix + 1; n1++) {
        double ai = xi - n1 * dcel;
        double aisign = fsign(one, ai);
        ai = one - fabs(ai) * idcel;
        for (int n2 = iy; n2 <= iy + 1; n2++) {
            double bi = yi - n2 * dcel;
            double bisign = fsign(one, bi);
            bi = one - fabs(bi) * idcel;
            for (int n3 = iz; n3 <= iz + 1; n3++) {
                double ci = zi - n3 * dcel;
                double cisign = fsign(one, ci);
                ci = one - fabs(ci) * idcel;
                double fi = ai * bi * ci;
                long in3 = n1 * ncyz + (long)n2 * g->nclz + n3;
                double gaux1 = c->gsrfen[in3 + ifir];
                double gaux2 = s->sqrfac * c->greff[in3 + ifir];
                if (gaux1 == 0.0) break;
                double prefa1 = gaux1 * idcel, prefa2 = gaux2 * idcel;
                aux1 = aux1 + fi * gaux1;
                aux2 = aux2 + fi * gaux2;
                if (qf) {
                    if ((ai < (one - RSMALL)) && (ai > RSMALL)) {
                        double aux = aisign * bi * ci;
                        a1x = a1x - aux * prefa1;
                        a2x = a2x - aux * prefa2;
                    }
                    if ((bi < (one - RSMALL)) && (bi > RSMALL)) {
                        double aux = bisign * ai * ci;
                        a1y = a1y - aux * prefa1;
                        a2y = a2y - aux * prefa2;
                    }
                    if ((ci < (one - RSMALL)) && (ci > RSMALL)) {
                        double aux = cisign * ai * bi;
                        a1z = a1z - aux * prefa1;
                        a2z = a2z - aux * prefa2;
                    }
                }
            }
        }
    }
    *srfe = aux1;
    *reff = aux2;
    if (qf) {
        d1[0] = a1x; d1[1] = a1y; d1[2] = a1z;
        d2[0] = a2x; d2[1] = a2y; d2[2] = a2z;
    }
}

/* RFPARION, rfpar.f90:20-228 */
static void rfparion(orc_ctx *c, int qforces, double *erfpar) {
    const bromoc_system *s = &c->s;
    int ns = s->nsites, nf = s->nfix, nt = s->ntot;
    double E = 0.0;
    for (int i = ns; i < nt; i++) c->srfe[i] = c->reff[i] = 0.0;
    for (int i = ns; i < nt; i++) {
        if (!rf_inbox(s, c->x[i], c->y[i], c->z[i])) continue;
        int itype = abs(c->typei[i]);
        double *dsi = c->srfed + 3 * (size_t)i, *dri = c->reffd + 3 * (size_t)i;
        rf_point(c, c->x[i], c->y[i], c->z[i], itype, qforces, &c->srfe[i], &c->reff[i], dsi, dri);
        if (i < ns + nf) continue;
        double tau = c->celec * c->cg[itype];
        double aux = tau * c->cg[itype] * c->srfe[i];
        E = E + 0.5 * aux * c->srfe[i];
        if (qforces) {
            double de = -aux;
            c->fx[i] = c->fx[i] + de * dsi[0];
            c->fy[i] = c->fy[i] + de * dsi[1];
            c->fz[i] = c->fz[i] + de * dsi[2];
        }
        for (int j = ns; j < i; j++) {
            if (c->srfe[j] == 0.0) continue;
            int jtype = abs(c->typei[j]);
            double ddx = c->x[j] - c->x[i], ddy = c->y[j] - c->y[i], ddz = c->z[j] - c->z[i];
            double dist2 = ddx * ddx + ddy * ddy + ddz * ddz;
            double srfeij = c->srfe[j] * c->srfe[i];
            double reffij = c->reff[j] * c->reff[i];
            double rfdn = 1.0 / sqrt(reffij * reffij + dist2);
            double rfcf = reffij * rfdn;
            double aux0 = tau * c->cg[jtype];
            double aux1 = aux0 * rfcf;
            E = E + aux1 * srfeij;
            if (qforces) {
                double *dsj = c->srfed + 3 * (size_t)j, *drj = c->reffd + 3 * (size_t)j;
                double aux2 = aux0 * srfeij * (rfdn * (rfdn * rfdn));
                double de = -aux1 * c->srfe[i];
                c->fx[j] = c->fx[j] + de * dsj[0];
                c->fy[j] = c->fy[j] + de * dsj[1];
                c->fz[j] = c->fz[j] + de * dsj[2];
                de = -aux1 * c->srfe[j];
                c->fx[i] = c->fx[i] + de * dsi[0];
                c->fy[i] = c->fy[i] + de * dsi[1];
                c->fz[i] = c->fz[i] + de * dsi[2];
                double aux3 = -aux2 * (dist2 + c->reff[j] * c->reff[i] * 0.5);
                de = aux3 * c->reff[i];
                c->fx[j] = c->fx[j] + de * drj[0];
                c->fy[j] = c->fy[j] + de * drj[1];
                c->fz[j] = c->fz[j] + de * drj[2];
                de = aux3 * c->reff[j];
                c->fx[i] = c->fx[i] + de * dri[0];
                c->fy[i] = c->fy[i] + de * dri[1];
                c->fz[i] = c->fz[i] + de * dri[2];
                de = aux2 * reffij;
                c->fx[j] = c->fx[j] + de * (c->x[j] - c->x[i]);
                c->fy[j] = c->fy[j] + de * (c->y[j] - c->y[i]);
                c->fz[j] = c->fz[j] + de * (c->z[j] - c->z[i]);
                c->fx[i] = c->fx[i] - de * (c->x[j] - c->x[i]);
                c->fy[i] = c->fy[i] - de * (c->y[j] - c->y[i]);
                c->fz[i] = c->fz[i] - de * (c->z[j] - c->z[i]);
            }
        }
    }
    *erfpar = E;
}

/* RFPARIONJ, rfpar.f90:232-370, for an ion j (j == ntot: an insertion trial, for which the
 * reference raises ntot by one around the call, interact.f90:145-148) */
static double rfparionj(orc_ctx *c, double xj, double yj, double zj, int j, int jtype) {
    const bromoc_system *s = &c->s;
    int ns = s->nsites, n = (j == s->ntot) ? s->ntot + 1 : s->ntot;
    if (!rf_inbox(s, xj, yj, zj)) return 0.0;
    double *sr = (double *)calloc((size_t)n + 1, sizeof(double));
    double *rf = (double *)calloc((size_t)n + 1, sizeof(double));
    for (int i = ns; i < n; i++) {
        int itype, ok;
        double xi, yi, zi;
        if (i == j) {
            itype = jtype; xi = xj; yi = yj; zi = zj; ok = 1;
        } else {
            itype = abs(c->typei[i]); xi = c->x[i]; yi = c->y[i]; zi = c->z[i];
            ok = rf_inbox(s, xi, yi, zi);
        }
        if (ok) rf_point(c, xi, yi, zi, itype, 0, &sr[i], &rf[i], NULL, NULL);
    }
    double tau = c->celec * c->cg[jtype];
    double E = 0.0;
    E = E + 0.5 * tau * c->cg[jtype] * (sr[j] * sr[j]);
    for (int i = ns; i < n; i++) {
        if (i == j || sr[i] == 0.0) continue;
        int itype = abs(c->typei[i]);
        double ddx = c->x[i] - xj, ddy = c->y[i] - yj, ddz = c->z[i] - zj;
        double dist2 = ddx * ddx + ddy * ddy + ddz * ddz;
        double reffij = rf[j] * rf[i];
        E = E + tau * c->cg[itype] * reffij * sr[j] * sr[i] / sqrt(reffij * reffij + dist2);
    }
    free(sr);
    free(rf);
    return E;
}

/* ------------------------------------------------- MMIJ reaction field, RECTBOX */
/* Legendre polynomials P_0..P_{n-1} of the scaled coordinate xs into lp[0..n-1], the
 * recurrence of rect_rf1.f90:51-63 (n >= 3) */
static void legendre(double xs, int n, double *lp) {
    lp[0] = 1.0;
    lp[1] = xs;
    lp[2] = 0.5 * (3.0 * xs * xs - 1.0);
    for (int k = 3; k <= n - 1; k++)   /* Fortran n = k: lpol(n+1) from lpol(n), lpol(n-1) */
        lp[k] = ((2.0 * k - 1) * xs * lp[k - 1] - (k - 1.0) * lp[k - 2]) / k;
}

static int mmij_inside(const orc_ctx *c, double xg, double yg, double zg) {
    const bromoc_system *s = &c->s;
    double xmax = 1.0 / s->xscal, ymax = 1.0 / s->yscal, zmax = 1.0 / s->zscal;
    double xmin = -1.0 / s->xscal, ymin = -1.0 / s->yscal, zmin = -1.0 / s->zscal;
    return xg >= xmin && xg <= xmax && yg >= ymin && yg <= ymax && zg >= zmin && zg <= zmax;
}

/* RECT_RF1, rect_rf1.f90:19-163: the coefficients q_m over the moving ions inside the
 * region, MQ = M q rfscal, egsbpb = celec sum 0.5 q_m MQ_m and the forces */
static void rect_rf1(orc_ctx *c, int qforces, double *egsbpb) {
    const bromoc_system *s = &c->s;
    const int nt = s->ntpol, nx = s->xnpol, ny = s->ynpol, nz = s->znpol;
    const int ion0 = s->nsites + s->nfix;
    for (int ii = 0; ii < nt; ii++) c->coef[ii] = 0.0;
    for (int i = ion0; i < s->ntot; i++) {
        int itype = abs(c->typei[i]);
        double charge = c->cg[itype];
        double xg = c->x[i], yg = c->y[i], zg = c->z[i];
        if (!mmij_inside(c, xg, yg, zg)) continue;
        double *lx = c->lpx + (size_t)i * nx, *ly = c->lpy + (size_t)i * ny, *lz = c->lpz + (size_t)i * nz;
        legendre(s->xscal * xg, nx, lx);
        legendre(s->yscal * yg, ny, ly);
        legendre(s->zscal * zg, nz, lz);
        for (int ii = 0; ii < nt; ii++) {
            double norm = c->bnorm[ii];
            c->coef[ii] = c->coef[ii] + charge * norm * lx[c->lstpx[ii]] * ly[c->lstpy[ii]] * lz[c->lstpz[ii]];
        }
    }
    for (int ii = 0; ii < nt; ii++) {
        c->mq[ii] = 0.0;
        for (int jj = 0; jj < nt; jj++)
            c->mq[ii] = c->mq[ii] + c->mmij[(size_t)ii * nt + jj] * c->coef[jj] * s->rfscal;
    }
    double e = 0.0;
    for (int ii = 0; ii < nt; ii++) e = e + 0.5 * c->coef[ii] * c->mq[ii];
    *egsbpb = e * c->celec;
    if (!qforces) return;
    double dlx[64], dly[64], dlz[64];
    for (int mm = ion0; mm < s->ntot; mm++) {
        int itype = abs(c->typei[mm]);
        double charge = c->cg[itype];
        double xg = c->x[mm], yg = c->y[mm], zg = c->z[mm];
        if (!mmij_inside(c, xg, yg, zg)) continue;
        double rx = 0.0, ry = 0.0, rz = 0.0;
        double ccc = charge * c->celec;
        double xxs = s->xscal * xg, yys = s->yscal * yg, zs = s->zscal * zg;
        const double *lx = c->lpx + (size_t)mm * nx, *ly = c->lpy + (size_t)mm * ny, *lz = c->lpz + (size_t)mm * nz;
        dlx[0] = 0.0; dly[0] = 0.0; dlz[0] = 0.0;
        dlx[1] = 1.0; dly[1] = 1.0; dlz[1] = 1.0;
        dlx[2] = 3.0 * xxs; dly[2] = 3.0 * yys; dlz[2] = 3.0 * zs;
        for (int k = 3; k <= nx - 1; k++) dlx[k] = xxs * dlx[k - 1] + k * lx[k - 1];
        for (int k = 3; k <= ny - 1; k++) dly[k] = yys * dly[k - 1] + k * ly[k - 1];
        for (int k = 3; k <= nz - 1; k++) dlz[k] = zs * dlz[k - 1] + k * lz[k - 1];
        for (int k = 0; k < nx; k++) dlx[k] = dlx[k] * s->xscal;
        for (int k = 0; k < ny; k++) dly[k] = dly[k] * s->yscal;
        for (int k = 0; k < nz; k++) dlz[k] = dlz[k] * s->zscal;
        for (int ii = 0; ii < nt; ii++) {
            int xp = c->lstpx[ii], yp = c->lstpy[ii], zp = c->lstpz[ii];
            double norm = c->bnorm[ii];
            double xl = lx[xp], yl = ly[yp], zl = lz[zp];
            double dx = norm * dlx[xp] * yl * zl;
            double dy = norm * xl * dly[yp] * zl;
            double dz = norm * xl * yl * dlz[zp];
            rx = rx - dx * c->mq[ii];
            ry = ry - dy * c->mq[ii];
            rz = rz - dz * c->mq[ii];
        }
        c->fx[mm] = c->fx[mm] + rx * ccc;
        c->fy[mm] = c->fy[mm] + ry * ccc;
        c->fz[mm] = c->fz[mm] + rz * ccc;
    }
}

/* RECT_RF0, rect_rf0.f90:19-115: g(ntot) - g(ntot without j), particle j at (xj,yj,zj) of
 * type jtype (j = ntot: an insertion).  The reference reads j's polynomials whether or not
 * j is a moving ion inside the region (left unset otherwise); here such a j has no term. */
static double rect_rf0(orc_ctx *c, double xj, double yj, double zj, int j, int jtype) {
    const bromoc_system *s = &c->s;
    const int nt = s->ntpol, nx = s->xnpol, ny = s->ynpol, nz = s->znpol;
    const int ion0 = s->nsites + s->nfix;
    const int ntot = j == s->ntot ? s->ntot + 1 : s->ntot;
    int jin = 0;
    for (int ii = 0; ii < nt; ii++) c->coef[ii] = 0.0;
    for (int i = ion0; i < ntot; i++) {
        int itype;
        double xg, yg, zg;
        if (i == j) { itype = jtype; xg = xj; yg = yj; zg = zj; }
        else { itype = abs(c->typei[i]); xg = c->x[i]; yg = c->y[i]; zg = c->z[i]; }
        double charge = c->cg[itype];
        if (!mmij_inside(c, xg, yg, zg)) continue;
        if (i == j) jin = 1;
        double *lx = c->lpx + (size_t)i * nx, *ly = c->lpy + (size_t)i * ny, *lz = c->lpz + (size_t)i * nz;
        legendre(s->xscal * xg, nx, lx);
        legendre(s->yscal * yg, ny, ly);
        legendre(s->zscal * zg, nz, lz);
        for (int ii = 0; ii < nt; ii++) {
            double norm = c->bnorm[ii];
            c->coef[ii] = c->coef[ii] + charge * norm * lx[c->lstpx[ii]] * ly[c->lstpy[ii]] * lz[c->lstpz[ii]];
        }
    }
    if (!jin) return 0.0;
    double charge = c->cg[jtype];
    const double *lx = c->lpx + (size_t)j * nx, *ly = c->lpy + (size_t)j * ny, *lz = c->lpz + (size_t)j * nz;
    double *coef2 = c->mq;
    for (int ii = 0; ii < nt; ii++) {
        double norm = c->bnorm[ii];
        coef2[ii] = c->coef[ii] - charge * norm * lx[c->lstpx[ii]] * ly[c->lstpy[ii]] * lz[c->lstpz[ii]];
    }
    double e = 0.0;
    for (int ii = 0; ii < nt; ii++) {
        double ci = c->coef[ii], cj = coef2[ii];
        e = e + 0.5 * (ci * ci - cj * cj) * c->mmij[(size_t)ii * nt + ii];
        for (int jj = ii + 1; jj < nt; jj++)
            e = e + (ci * c->coef[jj] - cj * coef2[jj]) * c->mmij[(size_t)ii * nt + jj];
    }
    return e * c->celec * s->rfscal;
}

EXPORT int orc_energy(orc_ctx *c, int32_t qforces, bromoc_energies *out) {
    const bromoc_system *s = &c->s;
    if (orc_check_flags(s->flags)) return BROMOC_E_UNSUPPORTED;
    bromoc_energies E;
    memset(&E, 0, sizeof(E));
    int ns = s->nsites, nf = s->nfix, nt = s->ntot;
    uint32_t fl = s->flags;
    if (qforces) {
        for (int i = 0; i < nt; i++) c->fx[i] = c->fy[i] = c->fz[i] = 0.0;
    }
    if (fl & BROMOC_QMEMB) membrane(c, qforces, &E.ememb);   /* energy.mono.f90:81-83 */
    E.ener = E.ememb;
    if (fl & BROMOC_QMMIJ) {   /* energy.mono.f90:85-92 */
        rect_rf1(c, qforces, &E.egsbpb);
        E.ener = E.ener + E.egsbpb;
    }
    if (fl & BROMOC_QPHIX) {
        staticf1(c, qforces, &E.egsbpa);
        E.ener = E.ener + E.egsbpa;
    }
    if (fl & BROMOC_QPHIV) {
        if (fl & BROMOC_QBSPLINE) vdwgd1spln(c, qforces, &E.evdwgd);
        else vdwgd1trln(c, qforces, &E.evdwgd);
        E.ener = E.ener + E.evdwgd;
    }
    if (fl & BROMOC_QRFPAR) {   /* energy.mono.f90:111-115 */
        rfparion(c, qforces, &E.erfpar);
        E.ener = E.ener + E.erfpar;
        E.egsbpb = E.erfpar;
    }
    /* ion-ion, energy.mono.f90:117-194 */
    if ((fl & BROMOC_QPAR) && (fl & BROMOC_QNONBOND)) {
        for (int j = ns + nf; j < nt; j++) {
            int jtype2 = c->nwtype[abs(c->typei[j])];
            for (int i = ns; i < j; i++) {
                int itype2 = c->nwtype[abs(c->typei[i])];
                int is = pair_of(c, itype2, jtype2);
                double dx = c->x[i] - c->x[j], dy = c->y[i] - c->y[j], dz = c->z[i] - c->z[j];
                double dist2 = (dx * dx + dy * dy + dz * dz);
                double de = 0.0, eefp, dist;
                if (c->pairs[is].efpot) {
                    if (qforces) getyd(c, is, dist2, &eefp, &de, &dist);
                    else gety(c, is, dist2, &eefp, &dist);
                    E.eefpot = E.eefpot + eefp;
                } else {
                    de = pair_analytic(c, is, dist2, qforces, &E.eelec, &E.evdw, &E.esrpmf);
                }
                if (qforces && de != 0.0) {
                    if (i >= ns + nf) {
                        c->fx[i] = c->fx[i] + de * (c->x[i] - c->x[j]);
                        c->fy[i] = c->fy[i] + de * (c->y[i] - c->y[j]);
                        c->fz[i] = c->fz[i] + de * (c->z[i] - c->z[j]);
                    }
                    c->fx[j] = c->fx[j] - de * (c->x[i] - c->x[j]);
                    c->fy[j] = c->fy[j] - de * (c->y[i] - c->y[j]);
                    c->fz[j] = c->fz[j] - de * (c->z[i] - c->z[j]);
                }
            }
        }
    }
    /* CG-DNA, energy.mono.f90:196-508 */
    if (fl & BROMOC_QNUCL) {
        const int8_t *sf = c->stfree;
        double eps = s->epsnuc;
        if (fl & BROMOC_QNOBOND) {
            const int32_t *L = c->lists[0];
            for (int t = 0; t < s->nbond; t++) {
                int a = L[t], b = L[s->nbond + t];
                double v1[3] = {c->x[b] - c->x[a], c->y[b] - c->y[a], c->z[b] - c->z[a]};
                double dist = sqrt(dot3(v1, v1));
                double vard = dist - c->lvals[0][t];
                double vard2 = vard * vard;
                E.ebond = E.ebond + eps * vard2 * (1.0 + 100.0 * vard2);
                if (qforces && (sf[a] || sf[b])) {
                    double de = eps * vard * (2.0 + 400.0 * vard2) / dist;
                    double f1[3] = {de * v1[0], de * v1[1], de * v1[2]};
                    if (sf[a]) add_f(c, a, 1.0, f1);
                    if (sf[b]) add_f(c, b, -1.0, f1);
                }
            }
            L = c->lists[1];
            for (int t = 0; t < s->nangle; t++) {
                int a = L[t], b = L[s->nangle + t], d = L[2 * s->nangle + t];
                double v1[3] = {c->x[a] - c->x[b], c->y[a] - c->y[b], c->z[a] - c->z[b]};
                double v2[3] = {c->x[d] - c->x[b], c->y[d] - c->y[b], c->z[d] - c->z[b]};
                double m1 = dot3(v1, v1), m2 = dot3(v2, v2), m3 = dot3(v1, v2);
                double modval = 1.0 / sqrt(m1 * m2);
                double ang = acos(m3 * modval);
                double varang = ang - c->lvals[1][t];
                E.eang = E.eang + 700.0 * eps * (varang * varang);
                if (qforces) {
                    double de = 1400.0 * eps * varang * modval / sin(ang);
                    double f1[3] = {de * v1[0], de * v1[1], de * v1[2]};
                    double f2[3] = {de * v2[0], de * v2[1], de * v2[2]};
                    if (sf[a]) add_f(c, a, 1.0, f2);
                    if (sf[b]) {
                        double f12[3] = {f1[0] + f2[0], f1[1] + f2[1], f1[2] + f2[2]};
                        add_f(c, b, -1.0, f12);
                    }
                    if (sf[d]) add_f(c, d, 1.0, f1);
                }
            }
            L = c->lists[2];
            for (int t = 0; t < s->ndihe; t++) {
                int i1 = L[t], i2 = L[s->ndihe + t], i3 = L[2 * s->ndihe + t], i4 = L[3 * s->ndihe + t];
                double v1[3] = {c->x[i1] - c->x[i2], c->y[i1] - c->y[i2], c->z[i1] - c->z[i2]};
                double v2[3] = {c->x[i3] - c->x[i2], c->y[i3] - c->y[i2], c->z[i3] - c->z[i2]};
                double v3[3] = {c->x[i3] - c->x[i4], c->y[i3] - c->y[i4], c->z[i3] - c->z[i4]};
                double n1[3], n2[3];
                cross(v1, v2, n1);
                cross(v2, v3, n2);
                double v22 = dot3(v2, v2);
                double v2v = sqrt(v22);
                double av = dot3(v1, n2);
                double bv = dot3(n1, n2);
                double dihe = atan2(v2v * av, bv);
                double vardihe = dihe - c->lvals[2][t];
                E.edihe = E.edihe + 28.0 * eps * (1.0 - cos(vardihe));
                if (qforces) {
                    double n1v = dot3(n1, n1), n2v = dot3(n2, n2);
                    double v12 = dot3(v1, v2), v23 = dot3(v2, v3);
                    double de = 28.0 * eps * sin(vardihe) * v2v;
                    double f1[3], f4[3], f2[3], f3[3];
                    for (int k = 0; k < 3; k++) f1[k] = -de / n1v * n1[k];
                    for (int k = 0; k < 3; k++) f4[k] = de / n2v * n2[k];
                    double iv22 = 1.0 / v22;
                    v12 = v12 * iv22;
                    v23 = v23 * iv22;
                    for (int k = 0; k < 3; k++) f2[k] = -f1[k] + v12 * f1[k] - v23 * f4[k];
                    for (int k = 0; k < 3; k++) f3[k] = -f4[k] - v12 * f1[k] + v23 * f4[k];
                    if (sf[i1]) add_f(c, i1, 1.0, f1);
                    if (sf[i2]) add_f(c, i2, 1.0, f2);
                    if (sf[i3]) add_f(c, i3, 1.0, f3);
                    if (sf[i4]) add_f(c, i4, 1.0, f4);
                }
            }
        }
        if (fl & BROMOC_QNONBOND) {
            const int32_t *L = c->lists[3];
            for (int t = 0; t < s->nstack; t++) {
                int a = L[t], b = L[s->nstack + t];
                double v1[3] = {c->x[b] - c->x[a], c->y[b] - c->y[a], c->z[b] - c->z[a]};
                double dist2 = dot3(v1, v1);
                double idist2 = 1.0 / dist2;
                double sg = c->lvals[3][t];
                double cte1 = pow3(sg * sg * idist2);
                E.estack = E.estack + 4.0 * eps * cte1 * (cte1 - 1.0);
                if (qforces) {
                    double de = 24.0 * eps * cte1 * (2.0 * cte1 - 1.0) * idist2;
                    double f1[3] = {de * v1[0], de * v1[1], de * v1[2]};
                    if (sf[a]) add_f(c, a, -1.0, f1);
                    if (sf[b]) add_f(c, b, 1.0, f1);
                }
            }
            L = c->lists[4];
            for (int t = 0; t < s->nbp; t++) {
                int a = L[t], b = L[s->nbp + t];
                double epsln;
                if (c->namsite[a] == BROMOC_SITE_GB || c->namsite[a] == BROMOC_SITE_CB)
                    epsln = 2.532 * eps * s->scalepairing;
                else
                    epsln = 2.0 * eps * s->scalepairing;
                double v1[3] = {c->x[b] - c->x[a], c->y[b] - c->y[a], c->z[b] - c->z[a]};
                double dist2 = dot3(v1, v1);
                double idist2 = 1.0 / dist2;
                double sg = c->lvals[4][t];
                double cte1 = sg * sg * idist2;
                E.ebp = E.ebp + epsln * pow5(cte1) * (20.0 * cte1 - 24.0);
                if (qforces) {
                    double de = 240.0 * epsln * pow5(cte1) * (cte1 - 1.0) * idist2;
                    double f1[3] = {de * v1[0], de * v1[1], de * v1[2]};
                    if (sf[a]) add_f(c, a, -1.0, f1);
                    if (sf[b]) add_f(c, b, 1.0, f1);
                }
            }
            L = c->lists[5];
            for (int t = 0; t < s->nex; t++) {
                int a = L[t], b = L[s->nex + t];
                double v1[3] = {c->x[b] - c->x[a], c->y[b] - c->y[a], c->z[b] - c->z[a]};
                double dist2 = dot3(v1, v1);
                double idist2 = 1.0 / dist2;
                double sgex2 = c->lvals[5][t] * c->lvals[5][t];
                if (dist2 < sgex2) {
                    double cte1 = pow3(sgex2 * idist2);
                    E.eex = E.eex + 4.0 * eps * cte1 * (cte1 - 1.0) + eps;
                    if (qforces) {
                        double de = 24.0 * eps * cte1 * (2.0 * cte1 - 1.0) * idist2;
                        double f1[3] = {de * v1[0], de * v1[1], de * v1[2]};
                        if (sf[a]) add_f(c, a, -1.0, f1);
                        if (sf[b]) add_f(c, b, 1.0, f1);
                    }
                }
            }
            L = c->lists[6];
            for (int t = 0; t < s->nqq; t++) {
                int a = L[t], b = L[s->nqq + t];
                double v1[3] = {c->x[b] - c->x[a], c->y[b] - c->y[a], c->z[b] - c->z[a]};
                double dist = sqrt(dot3(v1, v1));
                double idist = 1.0 / dist;
                double de;
                /* energy.mono.f90:439-441: hybrid Debye screens pairs of two sites outside */
                int deby = (fl & BROMOC_QDEBY) != 0;
                if ((fl & BROMOC_QDEBYHYB) && outbox(s, c->x[a], c->y[a], c->z[a]) &&
                    outbox(s, c->x[b], c->y[b], c->z[b]))
                    deby = 1;
                if (deby) {
                    double idistkd = exp(-dist * s->ikappa);
                    E.eqq = E.eqq + s->fctn * idist * idistkd;
                    de = s->fctn * (dist + s->kappa) * pow3(idist) * s->ikappa * idistkd;
                } else {
                    E.eqq = E.eqq + s->fctn * idist;
                    de = s->fctn * pow3(idist);
                }
                if (qforces) {
                    double f1[3] = {de * v1[0], de * v1[1], de * v1[2]};
                    if (sf[a]) add_f(c, a, -1.0, f1);
                    if (sf[b]) add_f(c, b, 1.0, f1);
                }
            }
            if (fl & BROMOC_QSOLV) {
                L = c->lists[7];
                for (int t = 0; t < s->nsolv; t++) {
                    int a = L[t], b = L[s->nsolv + t];
                    double v1[3] = {c->x[b] - c->x[a], c->y[b] - c->y[a], c->z[b] - c->z[a]};
                    double dist = sqrt(dot3(v1, v1));
                    double cte1 = exp((13.38 - dist) * 0.1875);
                    E.esolv = E.esolv + s->epsolv * ((1.0 - cte1) * (1.0 - cte1)) - s->epsolv;
                    if (qforces) {
                        double de = 0.375 * s->epsolv * cte1 * (cte1 - 1.0) / dist;
                        double f1[3] = {de * v1[0], de * v1[1], de * v1[2]};
                        if (sf[a]) add_f(c, a, -1.0, f1);
                        if (sf[b]) add_f(c, b, 1.0, f1);
                    }
                }
            }
        }
    }
    if ((fl & BROMOC_QNUCL) && (fl & BROMOC_QAPFOR)) {   /* APFOR, energy.mono.f90:510-518 */
        for (int k = 0; k < s->afn; k++) {
            int j = c->sn[k];
            if (qforces && c->stfree[j]) {
                c->fx[j] = c->fx[j] + c->af[3 * k];
                c->fy[j] = c->fy[j] + c->af[3 * k + 1];
                c->fz[j] = c->fz[j] + c->af[3 * k + 2];
            }
        }
    }
    if ((fl & BROMOC_QNUCL) && (fl & BROMOC_QCONTRANS)) {   /* energy.mono.f90:520-580 */
        double insites = 1.0 / ns;
        int n1 = s->nsites1st;
        for (int k = 0; k < s->ctn; k++) {
            int j = c->csn[k];
            double *pos[3] = {c->x, c->y, c->z}, *frc[3] = {c->fx, c->fy, c->fz};
            const double *kk[3] = {c->kx, c->ky, c->kz};
            const double *cc[3] = {c->contrx, c->contry, c->contrz};
            for (int ax = 0; ax < 3; ax++) {
                double kv = kk[ax][k];
                if (kv == 0.0) continue;
                double *X = pos[ax], *F = frc[ax];
                if (j == 0) {
                    if (s->unsplit) {
                        double sum = 0.0;
                        for (int q = 0; q < n1; q++) sum += X[q];
                        double con = sum * insites * 2.0 - cc[ax][k];
                        for (int q = 0; q < n1; q++) F[q] = F[q] - kv * con * insites * 2.0;
                        E.econ = E.econ + 0.5 * kv * (con * con);
                        sum = 0.0;
                        for (int q = n1; q < ns; q++) sum += X[q];
                        con = sum * insites * 2.0 - cc[ax][s->ctn];
                        for (int q = n1; q < ns; q++) F[q] = F[q] - kv * con * insites * 2.0;
                        E.econ = E.econ + 0.5 * kv * (con * con);
                    } else {
                        double sum = 0.0;
                        for (int q = 0; q < ns; q++) sum += X[q];
                        double con = sum * insites - cc[ax][k];
                        for (int q = 0; q < ns; q++) F[q] = F[q] - kv * con * insites;
                        E.econ = E.econ + 0.5 * kv * (con * con);
                    }
                } else {
                    double d = X[j - 1] - cc[ax][k];
                    F[j - 1] = F[j - 1] - kv * d;
                    E.econ = E.econ + 0.5 * kv * (d * d);
                }
            }
        }
    }
    /* ion-DNA, energy.mono.f90:583-687; the nearest-site record dids (:583,603-620) */
    for (int i = ns; i < nt; i++)
        for (int q = 0; q < 5; q++) c->dids[5 * i + q] = 0.0;
    if ((fl & BROMOC_QPAR) && (fl & BROMOC_QNUCL) && (fl & BROMOC_QNONBOND)) {
        for (int j = 0; j < ns; j++) {
            int jtype2 = c->typtyp[j];
            for (int i = ns; i < nt; i++) {
                int itype2 = c->nwtype[abs(c->typei[i])];
                int is = pair_of(c, itype2, jtype2);
                double dx = c->x[i] - c->x[j], dy = c->y[i] - c->y[j], dz = c->z[i] - c->z[j];
                double dist2 = dx * dx + dy * dy + dz * dz;
                int ok2 = qforces && (i >= ns + nf || c->stfree[j]);
                double de = 0.0, eefp, dist;
                if (c->pairs[is].efpot) {
                    if (ok2) getyd(c, is, dist2, &eefp, &de, &dist);
                    else gety(c, is, dist2, &eefp, &dist);
                    E.eefpotmx = E.eefpotmx + eefp;
                    if ((fl & BROMOC_QPROXDIFF) && dist > 0.0) {
                        double *dd = c->dids + 5 * i;
                        double didstmp = dist - c->pairs[is].dmi;
                        if (j == 0 || didstmp < dd[0]) {
                            dd[0] = didstmp;
                            dd[1] = dist;
                            dd[2] = c->x[i] - c->x[j];
                            dd[3] = c->y[i] - c->y[j];
                            dd[4] = c->z[i] - c->z[j];
                        }
                    }
                } else {
                    de = pair_analytic(c, is, dist2, ok2, &E.eqqmx, &E.evdwmx, &E.esrpmfmx);
                }
                if (ok2 && de != 0.0) {
                    if (i >= ns + nf) {
                        c->fx[i] = c->fx[i] + de * (c->x[i] - c->x[j]);
                        c->fy[i] = c->fy[i] + de * (c->y[i] - c->y[j]);
                        c->fz[i] = c->fz[i] + de * (c->z[i] - c->z[j]);
                    }
                    if (c->stfree[j]) {
                        c->fx[j] = c->fx[j] - de * (c->x[i] - c->x[j]);
                        c->fy[j] = c->fy[j] - de * (c->y[i] - c->y[j]);
                        c->fz[j] = c->fz[j] - de * (c->z[i] - c->z[j]);
                    }
                }
            }
        }
    }
    /* energy.mono.f90:689 */
    E.ener = E.ener + E.eelec + E.evdw + E.esrpmf + E.esrpmfmx + E.ebond + E.eang + E.edihe +
             E.estack + E.ebp + E.eex + E.eqq + E.esolv + E.eqqmx + E.evdwmx + E.eefpot +
             E.eefpotmx + E.econ;
    if (out) *out = E;
    return BROMOC_OK;
}

/* ---------------------------------------------------------- dynamics */
/* FIXCOOR, fixcoor.f90:19-67: a particle outside the sphere is reflected through its
 * surface along the radius; outside the elliptic cylinder it is put back on the ellipse
 * (scale 0.999999999999999/sqrt(r2)) and its z clamped; in the box each coordinate is
 * clamped. */
static void fixcoor(const bromoc_system *s, double *xi, double *yi, double *zi) {
    if (s->flags & BROMOC_QSPHERE) {
        const double r2 = (*xi - s->cx) * (*xi - s->cx) + (*yi - s->cy) * (*yi - s->cy) +
                          (*zi - s->cz) * (*zi - s->cz);
        if (r2 > s->rsphe2) {
            const double lzrad = 2.0 * s->rsphe / sqrt(r2) - 1.0;
            *xi = lzrad * (*xi - s->cx) + s->cx;
            *yi = lzrad * (*yi - s->cy) + s->cy;
            *zi = lzrad * (*zi - s->cz) + s->cz;
        }
        return;
    }
    if (s->flags & BROMOC_QECYL) {
        const double ex = (*xi - s->cx) * s->iecx, ey = (*yi - s->cy) * s->iecy;
        const double r2 = ex * ex + ey * ey;
        if (r2 > 1.0) {
            const double r = 0.999999999999999 / sqrt(r2);
            *xi = r * (*xi - s->cx) + s->cx;
            *yi = r * (*yi - s->cy) + s->cy;
        }
        if (*zi < s->lz2m) *zi = s->lz2m;
        else if (*zi > s->lz2p) *zi = s->lz2p;
        return;
    }
    if (*xi < s->lx2m) *xi = s->lx2m;
    else if (*xi > s->lx2p) *xi = s->lx2p;
    if (*yi < s->ly2m) *yi = s->ly2m;
    else if (*yi > s->ly2p) *yi = s->ly2p;
    if (*zi < s->lz2m) *zi = s->lz2m;
    else if (*zi > s->lz2p) *zi = s->lz2p;
}

/* COUNTIONS, countions.f90:19-32 (it is 0-based ion-type row) */
static void countions(orc_ctx *c, double zold, double zi, int it) {
    int np = c->s.cntpts;
    for (int j = 0; j < np; j++) {
        if (zold < c->zcont[j] && zi >= c->zcont[j]) c->nforward[it * np + j] += 1;
        else if (zold >= c->zcont[j] && zi < c->zcont[j]) c->nbackward[it * np + j] += 1;
    }
}

/* SEVAL / SEVALD (sevalmod.f90:23-67) with their SAVEd intervals *ic (1-based) */
static double seval_orc(const orc_ctx *c, int *ic, double u, int deriv) {
    int n = c->s.nspline;
    const double *X = c->xs - 1;
    int i = *ic, j, k;
    if (i >= n) i = 1;
    if (!(u < X[i]) && u <= X[i + 1]) goto found;
    i = 1;
    j = n + 1;
    for (int it = 0; it < 64; it++) {
        k = (i + j) / 2;
        if (u < X[k]) j = k;
        if (u >= X[k]) i = k;
        if (!(j > i + 1)) break;
    }
found:
    *ic = i;
    double dx = u - X[i];
    if (deriv) return c->bs[i - 1] + dx * (c->cs[i - 1] + dx * c->ds[i - 1]);
    return c->ys[i - 1] + dx * (c->bs[i - 1] + dx * (c->cs[i - 1] + dx * c->ds[i - 1]));
}

/* drift and noise factors of an ion: DYNAMICS0 fact1a/fact2a; DYNAMICS1 pore switch
 * (dynamics1.mono.f90:36-70); DYNAMICS2 profile (dynamics2.mono.f90:38-74) */
static void ion_factors(orc_ctx *c, int itype, double zz, double *fact1, double *fact2,
                        double *delDz) {
    const bromoc_system *s = &c->s;
    double sw = 0.0, dsw = 0.0, idiffusion, didiffusion;
    if (s->flags & BROMOC_QDIFFUSE) {
        double pp3 = c->p3[itype];
        double pore1 = -s->plength2 + s->pcenter, pore2 = s->plength2 + s->pcenter;
        if (zz > (pore1 - pp3) && zz < (pore2 + pp3)) {
            if (zz >= pore1 && zz <= pore2) {
                sw = 0.0;
                dsw = 0.0;
            } else {
                double ipp3 = 1.0 / pp3, delz;
                if (zz > pore2) {
                    delz = (zz - pore2);
                    sw = 2.0 * powi(delz * ipp3, 3) - 3.0 * powi(delz * ipp3, 2);
                    dsw = 6.0 * (powi(delz * ipp3, 2) - delz * ipp3) * ipp3;
                } else if (zz < pore1) {
                    delz = (zz - pore1);
                    sw = -2.0 * powi(delz * ipp3, 3) - 3.0 * powi(delz * ipp3, 2);
                    dsw = -6.0 * (powi(delz * ipp3, 2) + delz * ipp3) * ipp3;
                }
                sw = -sw;
                dsw = -dsw;
            }
        } else {
            sw = 1.0;
            dsw = 0.0;
        }
        idiffusion = c->diffusion[itype] * (c->ampl3[itype] + (1.0 - c->ampl3[itype]) * sw);
        didiffusion = c->diffusion[itype] * (1.0 - c->ampl3[itype]) * dsw;
    } else if (s->flags & BROMOC_QPROFILE) {
        int n = s->nspline;
        const double *xs = c->xs, *ys = c->ys;
        if (xs[0] < xs[n - 1]) {
            if (zz < xs[0]) { sw = ys[0]; dsw = 0.0; }
            else if (zz > xs[n - 1]) { sw = ys[n - 1]; dsw = 0.0; }
            else { sw = seval_orc(c, &c->seval_i, zz, 0); dsw = seval_orc(c, &c->sevald_i, zz, 1); }
        } else if (xs[0] > xs[n - 1]) {
            if (zz > xs[0]) { sw = ys[0]; dsw = 0.0; }
            else if (zz < xs[n - 1]) { sw = ys[n - 1]; dsw = 0.0; }
            else { sw = seval_orc(c, &c->seval_i, zz, 0); dsw = seval_orc(c, &c->sevald_i, zz, 1); }
        }
        idiffusion = c->diffusion[itype] * sw;
        didiffusion = c->diffusion[itype] * dsw;
    } else {
        *fact1 = c->fact1a[itype];
        *fact2 = c->fact2a[itype];
        *delDz = 0.0;
        return;
    }
    *fact1 = idiffusion * s->kbtdt;
    *fact2 = sqrt(2.0 * s->dt * idiffusion);
    *delDz = didiffusion * s->dt;
}

/* DYNAMICS0/1/2 (dynamics0.mono.f90:31-46, dynamics1.mono.f90:71-86,
 * dynamics2.mono.f90:75-85) over moving ions, then DYNAMICS0NUC (dynamics0nuc.f90:31-44)
 * over sites; g = Gaussian for (particle, component) */
static void dynamics_pass(orc_ctx *c, const double *g) {
    const bromoc_system *s = &c->s;
    double bdmax = s->bdmax;
    int dyn12 = (s->flags & (BROMOC_QDIFFUSE | BROMOC_QPROFILE)) != 0;
    if ((s->flags & BROMOC_QPAR) && !dyn12 && (s->flags & BROMOC_QPROXDIFF)) {
        /* DYNAMICS3, dynamics3.mono.f90:37-72 */
        for (int i = s->nsites + s->nfix; i < s->ntot; i++) {
            int itype = abs(c->typei[i]);
            double zold = c->z[i];
            const double *dd = c->dids + 5 * i;
            double fact1, fact2, delDx, delDy, delDz;
            if (dd[0] > 0.0 && dd[0] <= s->diffcutoff) {
                double ebet = exp(-dd[0] * s->ibeta);
                double D = c->diffusion[itype];
                double idiffusion = D - (D - s->diff0) * ebet;
                double didiffusion = (D - s->diff0) * s->ibeta * ebet;
                double idist = 1.0 / dd[1];
                fact1 = idiffusion * s->kbtdt;
                fact2 = sqrt(2.0 * s->dt * idiffusion);
                delDx = didiffusion * s->dt * dd[2] * idist;
                delDy = didiffusion * s->dt * dd[3] * idist;
                delDz = didiffusion * s->dt * dd[4] * idist;
            } else {
                fact1 = c->fact1a[itype];
                fact2 = c->fact2a[itype];
                delDx = delDy = delDz = 0.0;
            }
            double delx = c->fx[i] * fact1;
            double dely = c->fy[i] * fact1;
            double delz = c->fz[i] * fact1;
            if (fabs(delx) > bdmax) delx = fsign(bdmax, delx);
            if (fabs(dely) > bdmax) dely = fsign(bdmax, dely);
            if (fabs(delz) > bdmax) delz = fsign(bdmax, delz);
            c->x[i] = c->x[i] + delx + fact2 * g[3 * i + 0] + delDx;
            c->y[i] = c->y[i] + dely + fact2 * g[3 * i + 1] + delDy;
            c->z[i] = c->z[i] + delz + fact2 * g[3 * i + 2] + delDz;
            if (s->flags & BROMOC_QCOUNTION) countions(c, zold, c->z[i], itype - s->nold - 1);
            fixcoor(s, &c->x[i], &c->y[i], &c->z[i]);
        }
    } else if (s->flags & BROMOC_QPAR) {
        for (int i = s->nsites + s->nfix; i < s->ntot; i++) {
            int itype = abs(c->typei[i]);
            double fact1, fact2, delDz;
            ion_factors(c, itype, c->z[i], &fact1, &fact2, &delDz);
            double delx = c->fx[i] * fact1;
            double dely = c->fy[i] * fact1;
            double delz = c->fz[i] * fact1;
            if (fabs(delx) > bdmax) delx = fsign(bdmax, delx);
            if (fabs(dely) > bdmax) dely = fsign(bdmax, dely);
            if (fabs(delz) > bdmax) delz = fsign(bdmax, delz);
            c->x[i] = c->x[i] + delx + fact2 * g[3 * i + 0];
            c->y[i] = c->y[i] + dely + fact2 * g[3 * i + 1];
            double zold = c->z[i];
            if (dyn12) c->z[i] = c->z[i] + delz + fact2 * g[3 * i + 2] + delDz;
            else c->z[i] = c->z[i] + delz + fact2 * g[3 * i + 2];
            if (s->flags & BROMOC_QCOUNTION) countions(c, zold, c->z[i], itype - s->nold - 1);
            fixcoor(s, &c->x[i], &c->y[i], &c->z[i]);
        }
    }
    if (s->flags & BROMOC_QNUCL) {
        for (int i = 0; i < s->nsites; i++) {
            if (!c->stfree[i]) continue;
            double delx = c->fx[i] * s->fact0n1;
            double dely = c->fy[i] * s->fact0n1;
            double delz = c->fz[i] * s->fact0n1;
            if (fabs(delx) > bdmax) delx = fsign(bdmax, delx);
            if (fabs(dely) > bdmax) dely = fsign(bdmax, dely);
            if (fabs(delz) > bdmax) delz = fsign(bdmax, delz);
            c->x[i] = c->x[i] + delx + s->fact0n2 * g[3 * i + 0];
            c->y[i] = c->y[i] + dely + s->fact0n2 * g[3 * i + 1];
            c->z[i] = c->z[i] + delz + s->fact0n2 * g[3 * i + 2];
            if (!(s->flags & BROMOC_QDNAFREE)) fixcoor(s, &c->x[i], &c->y[i], &c->z[i]);
        }
    }
}

/* ---------------------------------------------------------- Philox */
/* Philox4x64-10 (Salmon et al., SC'11), the generator of numpy.random.Philox. */
static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t *hi, uint64_t *lo) {
    __uint128_t p = (__uint128_t)a * b;
    *hi = (uint64_t)(p >> 64);
    *lo = (uint64_t)p;
}
EXPORT void orc_philox4x64(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4]) {
    uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint64_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo64(0xD2E7470EE14C6C93ull, c0, &hi0, &lo0);
        mulhilo64(0xCA5A826395121157ull, c2, &hi1, &lo1);
        uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B97F4A7C15ull;
        k1 += 0xBB67AE8584CAA73Bull;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Gaussians of particle i at step counter `step`: Box-Muller on two 53-bit uniforms
 * per pair of outputs, u1 in (0,1], u2 in [0,1).  Counter (step, i, stream, 0),
 * key (seed, 0). */
EXPORT void orc_philox_gauss3(uint64_t seed, uint64_t step, uint64_t stream, int64_t i, double g[3]) {
    uint64_t ctr[4] = {step, (uint64_t)i, stream, 0}, key[2] = {seed, 0}, r[4];
    orc_philox4x64(ctr, key, r);
    const double inv53 = 1.0 / 9007199254740992.0;
    double u1 = (double)((r[0] >> 11) + 1) * inv53;
    double u2 = (double)(r[1] >> 11) * inv53;
    double u3 = (double)((r[2] >> 11) + 1) * inv53;
    double u4 = (double)(r[3] >> 11) * inv53;
    double rad1 = sqrt(-2.0 * log(u1)), rad2 = sqrt(-2.0 * log(u3));
    const double twopi = 2.0 * PI;
    g[0] = rad1 * cos(twopi * u2);
    g[1] = rad1 * sin(twopi * u2);
    g[2] = rad2 * cos(twopi * u4);
}

EXPORT int orc_upload_noise(orc_ctx *c, int32_t nsteps, const double *noise) {
    free(c->noise);
    size_t n = (size_t)nsteps * 3 * (size_t)c->cap;
    c->noise = (double *)calloc(n ? n : 1, sizeof(double));
    /* stored with row stride 3*ntot at upload time, as given */
    memcpy(c->noise, noise, sizeof(double) * (size_t)nsteps * 3 * (size_t)c->s.ntot);
    c->noise_rows = nsteps;
    return BROMOC_OK;
}

static int fill_noise(orc_ctx *c, int mode, int row, uint64_t seed, uint64_t counter, double *g) {
    int nt = c->s.ntot;
    if (mode == BROMOC_NOISE_INJECTED) {
        if (!c->noise || row < 0 || row >= c->noise_rows) return BROMOC_E_STATE;
        memcpy(g, c->noise + (size_t)row * 3 * nt, sizeof(double) * 3 * nt);
    } else {
        for (int i = 0; i < nt; i++) orc_philox_gauss3(seed, counter, c->stream, i, g + 3 * i);
    }
    return BROMOC_OK;
}

/* Philox stream of the Gaussians (replica r of a batch draws from stream r) */
EXPORT void orc_set_stream(orc_ctx *c, uint64_t stream) { c->stream = stream; }

EXPORT int orc_dynamics(orc_ctx *c, int32_t mode, int32_t row, uint64_t seed, uint64_t counter) {
    double *g = (double *)malloc(sizeof(double) * 3 * (c->s.ntot + 1));
    int st = fill_noise(c, mode, row, seed, counter, g);
    if (st == BROMOC_OK) dynamics_pass(c, g);
    free(g);
    return st;
}

/* NOTRANS after DYNAMICS0NUC, simul1.f90:382-386 */
static void notrans(orc_ctx *c) {
    const bromoc_system *s = &c->s;
    int ns = s->nsites;
    if (!(s->flags & BROMOC_QNOTRANS) || !(s->flags & BROMOC_QNUCL) || ns < 1) return;
    double insites = 1.0 / ns;
    double *pos[3] = {c->x, c->y, c->z};
    const double notr[3] = {s->notrx, s->notry, s->notrz};
    for (int ax = 0; ax < 3; ax++) {
        if (!(s->notrans_axes & (1 << ax))) continue;
        double *X = pos[ax], sum = 0.0;
        for (int q = 0; q < ns; q++) sum += X[q];
        for (int q = 0; q < ns; q++) X[q] = X[q] - sum * insites + notr[ax];
    }
}

/* simul1.f90:363-390 with Qforces on: ENERGY then DYNAMICS0/1/2/3 then DYNAMICS0NUC and
 * NOTRANS */
/* the virial pressure of one step, simul1.f90:365-368 (i3 = 1/3, itvol = 1/tvol,
 * kba3bar = kboltz*1e25, constamod.f90:24,36) */
static void pressure(orc_ctx *c) {
    const int n = c->s.ntot;
    const double i3 = 1.0 / 3.0, kba3bar = 1.380648813E-23 * 1e25;
    const double itvol = 1.0 / c->an.tvol;
    double dx = 0.0, dy = 0.0, dz = 0.0;   /* dot_product(fx(1:ntot), x(1:ntot)) etc. */
    for (int i = 0; i < n; i++) dx += c->fx[i] * c->x[i];
    for (int i = 0; i < n; i++) dy += c->fy[i] * c->y[i];
    for (int i = 0; i < n; i++) dz += c->fz[i] * c->z[i];
    const double vir = dx + dy + dz;
    c->pres = c->pres + (n - vir * i3 * c->an.ikbt) * itvol * kba3bar * c->an.temp;
    c->pres_steps++;
}

EXPORT int orc_step(orc_ctx *c, int32_t nsteps, int32_t mode, int32_t first_row, uint64_t seed,
                    uint64_t counter0, bromoc_energies *energies) {
    double *g = (double *)malloc(sizeof(double) * 3 * (c->s.ntot + 1));
    int st = BROMOC_OK;
    for (int s = 0; s < nsteps && st == BROMOC_OK; s++) {
        st = orc_energy(c, 1, energies);
        if (st) break;
        if (c->an.what & BROMOC_AN_PRES) pressure(c);
        st = fill_noise(c, mode, first_row + s, seed, counter0 + (uint64_t)s, g);
        if (st) break;
        dynamics_pass(c, g);
        notrans(c);
    }
    free(g);
    return st;
}

/* -------------------------------------------------------- INTERACT */
/* STATICF0, staticf0.f90:19-82 */
static double staticf0(const orc_ctx *c, double xj, double yj, double zj, int j, int jtype) {
    const bromoc_system *s = &c->s;
    const bromoc_grid *g = &s->phix;
    double egsbpa = 0.0;
    int beion = (s->flags & BROMOC_QPAR) && j >= s->nsites;
    int besite = 0;
    if ((s->flags & BROMOC_QNUCL) && j < s->nsites) besite = c->namsite[j] == BROMOC_SITE_P;
    int ok = beion || besite;
    double chi = 0.0;
    if (beion) chi = c->cg[jtype];
    if (besite) chi = s->cgnuc;
    double dcel = g->dcel, idcel = 1.0 / dcel;
    double tranx = 0.5 * (g->nclx - 1) * dcel, trany = 0.5 * (g->ncly - 1) * dcel,
           tranz = 0.5 * (g->nclz - 1) * dcel;
    if (ok)
        ok = xj <= g->xbcen + tranx && xj >= g->xbcen - tranx && yj <= g->ybcen + trany &&
             yj >= g->ybcen - trany && zj <= g->zbcen + tranz && zj >= g->zbcen - tranz;
    if (!ok) return egsbpa;
    int ncyz = g->ncly * g->nclz;
    double xi = xj + tranx - g->xbcen, yi = yj + trany - g->ybcen, zi = zj + tranz - g->zbcen;
    int ix = (int)(xi * idcel), iy = (int)(yi * idcel), iz = (int)(zi * idcel);
    if (ix == g->nclx - 1) ix = g->nclx - 2;
    if (iy == g->ncly - 1) iy = g->ncly - 2;
    if (iz == g->nclz - 1) iz = g->nclz - 2;
    for (int n1 = ix; n1 <= ix + 1; n1++) {
        double ai = xi - n1 * dcel;
        ai = 1.0 - fabs(ai) * idcel;
        for (int n2 = iy; n2 <= iy + 1; n2++) {
            double bi = yi - n2 * dcel;
            bi = 1.0 - fabs(bi) * idcel;
            for (int n3 = iz; n3 <= iz + 1; n3++) {
                double ci = zi - n3 * dcel;
                ci = 1.0 - fabs(ci) * idcel;
                double fi = ai * bi * ci;
                size_t in3 = (size_t)n1 * ncyz + (size_t)n2 * g->nclz + n3;
                egsbpa = egsbpa + (fi * chi * (double)c->phix[in3] * c->celec);
            }
        }
    }
    return egsbpa;
}

/* VDWGD0TRLN, vdwgd0trln.f90:19-117 */
static double vdwgd0trln(orc_ctx *c, double xj, double yj, double zj, int j, int jtype, int jtype2,
                         int alert) {
    const bromoc_system *s = &c->s;
    const bromoc_grid *g = &s->phiv;
    double evdwgd = 0.0;
    double dcel = g->dcel, idcel = 1.0 / dcel;
    double tranx = 0.5 * (g->nclx - 1) * dcel, trany = 0.5 * (g->ncly - 1) * dcel,
           tranz = 0.5 * (g->nclz - 1) * dcel;
    int ok = xj <= g->xbcen + tranx && xj >= g->xbcen - tranx && yj <= g->ybcen + trany &&
             yj >= g->ybcen - trany && zj <= s->vzmax && zj >= s->vzmin;
    if (!ok) return evdwgd;
    int ncyz = g->ncly * g->nclz;
    double xi = xj + tranx - g->xbcen, yi = yj + trany - g->ybcen, zi = zj + tranz - g->zbcen;
    double esvdw = s->svdw;
    size_t ifir = 0;
    if (s->flags & BROMOC_QNMCDEN) nmcden_map(c, j, jtype, (size_t)g->nclx * ncyz, &ifir);
    else if (s->flags & BROMOC_QSVDW) esvdw = esvdw * c->scal[jtype2];
    int ix = (int)(xi * idcel), iy = (int)(yi * idcel), iz = (int)(zi * idcel);
    if (ix == g->nclx - 1) ix = g->nclx - 2;
    if (iy == g->ncly - 1) iy = g->ncly - 2;
    if (iz == g->nclz - 1) iz = g->nclz - 2;
    double phisum = 0.0;
    for (int n1 = ix; n1 <= ix + 1; n1++) {
        double ai = xi - n1 * dcel;
        ai = 1.0 - fabs(ai) * idcel;
        for (int n2 = iy; n2 <= iy + 1; n2++) {
            double bi = yi - n2 * dcel;
            bi = 1.0 - fabs(bi) * idcel;
            for (int n3 = iz; n3 <= iz + 1; n3++) {
                double ci = zi - n3 * dcel;
                ci = 1.0 - fabs(ci) * idcel;
                double fi = ai * bi * ci;
                size_t in3 = (size_t)n1 * ncyz + (size_t)n2 * g->nclz + n3;
                double phis = (double)c->phiv[in3 + ifir];
                phisum = phisum + phis;
                evdwgd = evdwgd + fi * esvdw * phis;
            }
        }
    }
    if (phisum >= s->thold8) {
        evdwgd = 1.0e10;
        if (alert) c->warn[jtype2] = c->warn[jtype2] + 1;
    }
    return evdwgd;
}

/* VDWGD0SPLN, vdwgd0spln.f90:19-130 */
static double vdwgd0spln(orc_ctx *c, double xj, double yj, double zj, int j, int jtype, int jtype2,
                         int alert) {
    const bromoc_system *s = &c->s;
    const bromoc_grid *g = &s->phiv;
    int nclx2 = g->nclx, ncly2 = g->ncly, nclz2 = g->nclz;
    int ncyz = ncly2 * nclz2;
    double dcel2 = g->dcel, idcel2 = 1.0 / dcel2;
    double tranx2 = 0.5 * (nclx2 - 1) * dcel2, trany2 = 0.5 * (ncly2 - 1) * dcel2,
           tranz2 = 0.5 * (nclz2 - 1) * dcel2;
    double evdwgd = 0.0;
    int ok = xj <= g->xbcen + tranx2 && xj >= g->xbcen - tranx2 && yj <= g->ybcen + trany2 &&
             yj >= g->ybcen - trany2 && zj <= s->vzmax && zj >= s->vzmin;
    if (!ok) return evdwgd;
    double xi = xj + tranx2 - g->xbcen, yi = yj + trany2 - g->ybcen, zi = zj + tranz2 - g->zbcen;
    double esvdw = s->svdw;
    size_t ifir = 0;
    if (s->flags & BROMOC_QNMCDEN) nmcden_map(c, j, jtype, (size_t)nclx2 * ncyz, &ifir);
    else if (s->flags & BROMOC_QSVDW) esvdw = esvdw * c->scal[jtype2];
    int ix = nint(xi * idcel2), iy = nint(yi * idcel2), iz = nint(zi * idcel2);
    if (ix == 0) ix = 1;
    if (iy == 0) iy = 1;
    if (iz == 0) iz = 1;
    if (ix == nclx2 - 1) ix = nclx2 - 2;
    if (iy == ncly2 - 1) iy = ncly2 - 2;
    if (iz == nclz2 - 1) iz = nclz2 - 2;
    double phisum = 0.0;
    for (int k = ix - 1; k <= ix + 1; k++) {
        size_t ipx = (size_t)k * ncyz;
        double xc = k * dcel2;
        double ai = m3(1.5 - (xi - xc) * idcel2);
        if (ai == 0.0) continue;
        for (int l = iy - 1; l <= iy + 1; l++) {
            size_t ipy = (size_t)l * nclz2 + ipx;
            double yc = l * dcel2;
            double bi = m3(1.5 - (yi - yc) * idcel2);
            if (bi == 0.0) continue;
            for (int m = iz - 1; m <= iz + 1; m++) {
                size_t ipz = m + ipy;
                double zc = m * dcel2;
                double ci = 1.5 - (zi - zc) * idcel2;
                double fi = ai * bi * m3(ci);
                if (fi != 0.0) {
                    double phis = (double)c->phiv[ipz + ifir];
                    phisum = phisum + phis;
                    evdwgd = evdwgd + fi * esvdw * phis;
                }
            }
        }
    }
    if (phisum >= s->thold27) {
        evdwgd = 1.0e10;
        if (alert) c->warn[jtype2] = c->warn[jtype2] + 1;
    }
    return evdwgd;
}

/* energy of one analytic pair as INTERACT writes it (interact.f90:184-204, 365-383):
 * SRPMF smoothing is e*exp(...)-srpy there, unlike ENERGY's e*(exp(...)-srpy). */
static double interact_analytic(const orc_ctx *c, int is, double dist2) {
    const bromoc_pair *p = &c->pairs[is];
    const bromoc_system *s = &c->s;
    double e = 0.0, idist = 0.0;
    double idist2 = 1.0 / dist2;
    if (p->chr || p->srpmf) idist = sqrt(idist2);
    if (p->chr) e = e + p->fct * idist;
    if (p->lj) {
        double dist6 = pow3(p->sgp2 * idist2);
        e = e + p->epp4 * dist6 * (dist6 - 1.0);
    }
    if (p->srpmf && dist2 <= s->rth) {
        double dist = 1.0 / idist;
        double es = p->c0 * exp((p->c1 - dist) * p->c2) * cos(p->c3 * PI * (p->c1 - dist)) +
                    p->c4 * pow6(p->c1 * idist);
        if (dist >= s->srpx) es = es * exp(-s->srpk * (dist - s->srpx)) - s->srpy;
        e = e + es;
    }
    return e;
}

EXPORT int orc_interact(orc_ctx *c, double xj, double yj, double zj, int32_t jtype, int32_t j,
                        int32_t alert, double *dener_out) {
    const bromoc_system *s = &c->s;
    if (orc_check_flags(s->flags)) return BROMOC_E_UNSUPPORTED;
    uint32_t fl = s->flags;
    double dener = 0.0;
    int ns = s->nsites, nt = s->ntot;
    int beion = (fl & BROMOC_QPAR) && j >= ns;
    int jtype2 = beion ? c->nwtype[jtype] : c->typtyp[j];
    if (fl & BROMOC_QMEMB) dener = membrane_interact(c, xj, yj, zj, j, jtype, jtype2, beion, alert);
    if (fl & BROMOC_QMMIJ) dener = dener + rect_rf0(c, xj, yj, zj, j, jtype);   /* interact.f90:121-141 */
    /* interact.f90:144-154; the reference also calls it for a site j, reading srfe
     * outside its array (rfpar.f90:356); sites get no reaction-field term here */
    if ((fl & BROMOC_QRFPAR) && beion) dener = dener + rfparionj(c, xj, yj, zj, j, jtype);
    if (fl & BROMOC_QPHIX) dener = dener + staticf0(c, xj, yj, zj, j, jtype);
    if (fl & BROMOC_QPHIV) {
        if (fl & BROMOC_QBSPLINE) dener = dener + vdwgd0spln(c, xj, yj, zj, j, jtype, jtype2, alert);
        else dener = dener + vdwgd0trln(c, xj, yj, zj, j, jtype, jtype2, alert);
    }
    if (beion && (fl & BROMOC_QNONBOND)) {
        for (int i = ns; i < nt; i++) {
            if (i == j) continue;
            int itype2 = c->nwtype[abs(c->typei[i])];
            int is = pair_of(c, itype2, jtype2);
            double dx = c->x[i] - xj, dy = c->y[i] - yj, dz = c->z[i] - zj;
            double dist2 = dx * dx + dy * dy + dz * dz;
            if (c->pairs[is].efpot) {
                double eefp, dist;
                gety(c, is, dist2, &eefp, &dist);
                dener = dener + eefp;
            } else {
                dener = dener + interact_analytic(c, is, dist2);
            }
        }
    }
    if ((fl & BROMOC_QNUCL) && !beion) {
        double eps = s->epsnuc;
        if (fl & BROMOC_QNOBOND) {
            const int32_t *L = c->lists[0];
            for (int t = 0; t < s->nbond; t++) {
                int a = L[t], b = L[s->nbond + t];
                if (a != j && b != j) continue;
                double dx = c->x[a] - c->x[b], dy = c->y[a] - c->y[b], dz = c->z[a] - c->z[b];
                double dist = sqrt(dx * dx + dy * dy + dz * dz);
                double vard = dist - c->lvals[0][t];
                dener = dener + eps * (vard * vard) * (1.0 + 100.0 * (vard * vard));
            }
            L = c->lists[1];
            for (int t = 0; t < s->nangle; t++) {
                int a = L[t], b = L[s->nangle + t], d = L[2 * s->nangle + t];
                if (a != j && b != j && d != j) continue;
                /* angbond.f90:19-39 */
                double v1[3] = {c->x[a] - c->x[b], c->y[a] - c->y[b], c->z[a] - c->z[b]};
                double v2[3] = {c->x[d] - c->x[b], c->y[d] - c->y[b], c->z[d] - c->z[b]};
                double r1 = sqrt(dot3(v1, v1)), r2 = sqrt(dot3(v2, v2));
                double ang = acos(dot3(v1, v2) / (r1 * r2));
                double varang = ang - c->lvals[1][t];
                dener = dener + 700.0 * eps * (varang * varang);
            }
            L = c->lists[2];
            for (int t = 0; t < s->ndihe; t++) {
                int i1 = L[t], i2 = L[s->ndihe + t], i3 = L[2 * s->ndihe + t], i4 = L[3 * s->ndihe + t];
                if (i1 != j && i2 != j && i3 != j && i4 != j) continue;
                /* angdih.f90:19-48 on bond vectors b1=r2-r1, b2=r3-r2, b3=r4-r3 */
                double b1[3] = {c->x[i2] - c->x[i1], c->y[i2] - c->y[i1], c->z[i2] - c->z[i1]};
                double b2[3] = {c->x[i3] - c->x[i2], c->y[i3] - c->y[i2], c->z[i3] - c->z[i2]};
                double b3[3] = {c->x[i4] - c->x[i3], c->y[i4] - c->y[i3], c->z[i4] - c->z[i3]};
                double n1[3], n2[3];
                cross(b1, b2, n1);
                cross(b2, b3, n2);
                double r2 = sqrt(dot3(b2, b2));
                double sinphi = r2 * dot3(b1, n2);
                double dihe = atan2(sinphi, dot3(n1, n2));
                double vardihe = dihe - c->lvals[2][t];
                dener = dener + 28.0 * eps * (1.0 - cos(vardihe));
            }
        }
        if (fl & BROMOC_QNONBOND) {
            const int32_t *L = c->lists[3];
            for (int t = 0; t < s->nstack; t++) {
                int a = L[t], b = L[s->nstack + t];
                if (a != j && b != j) continue;
                double dx = c->x[b] - c->x[a], dy = c->y[b] - c->y[a], dz = c->z[b] - c->z[a];
                double dist2 = dx * dx + dy * dy + dz * dz;
                double sg = c->lvals[3][t];
                double cte1 = pow3(sg * sg / dist2);
                dener = dener + 4.0 * eps * cte1 * (cte1 - 1.0);
            }
            L = c->lists[4];
            for (int t = 0; t < s->nbp; t++) {
                int a = L[t], b = L[s->nbp + t];
                if (a != j && b != j) continue;
                double epsln = (c->namsite[a] == BROMOC_SITE_GB || c->namsite[a] == BROMOC_SITE_CB)
                                   ? 2.532 * eps : 2.0 * eps;
                double dx = c->x[b] - c->x[a], dy = c->y[b] - c->y[a], dz = c->z[b] - c->z[a];
                double dist2 = dx * dx + dy * dy + dz * dz;
                double sg = c->lvals[4][t];
                double cte1 = sg * sg / dist2;
                dener = dener + epsln * pow5(cte1) * (20.0 * cte1 - 24.0);
            }
            L = c->lists[5];
            for (int t = 0; t < s->nex; t++) {
                int a = L[t], b = L[s->nex + t];
                if (a != j && b != j) continue;
                double dx = c->x[b] - c->x[a], dy = c->y[b] - c->y[a], dz = c->z[b] - c->z[a];
                double dist2 = dx * dx + dy * dy + dz * dz;
                double sgex2 = c->lvals[5][t] * c->lvals[5][t];
                if (dist2 < sgex2) {
                    double cte1 = pow3(sgex2 / dist2);
                    dener = dener + 4.0 * eps * cte1 * (cte1 - 1.0) + eps;
                }
            }
            L = c->lists[6];
            for (int t = 0; t < s->nqq; t++) {
                int a = L[t], b = L[s->nqq + t];
                if (a != j && b != j) continue;
                double dx = c->x[b] - c->x[a], dy = c->y[b] - c->y[a], dz = c->z[b] - c->z[a];
                double dist = sqrt(dx * dx + dy * dy + dz * dz);
                double idist = 1.0 / dist;
                int deby = (fl & BROMOC_QDEBY) != 0;   /* interact.f90:323-331 */
                if ((fl & BROMOC_QDEBYHYB) && outbox(s, c->x[a], c->y[a], c->z[a]) &&
                    outbox(s, c->x[b], c->y[b], c->z[b]))
                    deby = 1;
                if (!deby) dener = dener + s->fctn * idist;
                else dener = dener + s->fctn * idist * exp(-dist * s->ikappa);
            }
            if (fl & BROMOC_QSOLV) {
                L = c->lists[7];
                for (int t = 0; t < s->nsolv; t++) {
                    int a = L[t], b = L[s->nsolv + t];
                    if (a != j && b != j) continue;
                    double dx = c->x[b] - c->x[a], dy = c->y[b] - c->y[a], dz = c->z[b] - c->z[a];
                    double dist = sqrt(dx * dx + dy * dy + dz * dz);
                    double cte1 = exp((13.38 - dist) * 0.1875);
                    dener = dener + s->epsolv * ((1.0 - cte1) * (1.0 - cte1)) - s->epsolv;
                }
            }
        }
    }
    if ((fl & BROMOC_QPAR) && (fl & BROMOC_QNUCL) && (fl & BROMOC_QNONBOND)) {
        if (beion) {
            for (int i = 0; i < ns; i++) {
                int is = pair_of(c, c->typtyp[i], jtype2);
                double dx = c->x[i] - xj, dy = c->y[i] - yj, dz = c->z[i] - zj;
                double dist2 = dx * dx + dy * dy + dz * dz;
                if (c->pairs[is].efpot) {
                    double eefp, dist;
                    gety(c, is, dist2, &eefp, &dist);
                    dener = dener + eefp;
                } else {
                    dener = dener + interact_analytic(c, is, dist2);
                }
            }
        } else {
            for (int i = ns; i < nt; i++) {
                int is = pair_of(c, c->nwtype[abs(c->typei[i])], jtype2);
                double dx = c->x[i] - xj, dy = c->y[i] - yj, dz = c->z[i] - zj;
                double dist2 = dx * dx + dy * dy + dz * dz;
                if (c->pairs[is].efpot) {
                    double eefp, dist;
                    gety(c, is, dist2, &eefp, &dist);
                    dener = dener + eefp;
                } else {
                    dener = dener + interact_analytic(c, is, dist2);
                }
            }
        }
    }
    *dener_out = dener;
    return BROMOC_OK;
}

/* ----------------------------------------------------------- GRAND */
EXPORT int orc_set_gcmc(orc_ctx *c, const bromoc_gcmc *g, const int32_t *ibuffer) {
    int nb = g->nbuffer;
    if ((c->s.flags & BROMOC_QSPHERE) && (!g->rmin || !g->rmax)) return BROMOC_E_ARG;
    c->g = *g;
    free(c->ibfftyp); free(c->ntotat); free(c->avnum); free(c->mu); free(c->kb);
    free(c->lzmin); free(c->lzmax); free(c->bufferbias); free(c->ibuffer);
    free(c->rmin); free(c->rmax);
    c->rmin = g->rmin ? (double *)dupmem(g->rmin, sizeof(double) * (nb + 1)) : NULL;
    c->rmax = g->rmax ? (double *)dupmem(g->rmax, sizeof(double) * (nb + 1)) : NULL;
    c->ibfftyp = (int32_t *)dupmem(g->ibfftyp, sizeof(int32_t) * (nb + 1));
    c->ntotat = (int32_t *)dupmem(g->ntotat, sizeof(int32_t) * (nb + 1));
    c->avnum = (double *)dupmem(g->avnum, sizeof(double) * (nb + 1));
    c->mu = (double *)dupmem(g->mu, sizeof(double) * (nb + 1));
    c->kb = (double *)dupmem(g->kb, sizeof(double) * (nb + 1));
    c->lzmin = (double *)dupmem(g->lzmin, sizeof(double) * (nb + 1));
    c->lzmax = (double *)dupmem(g->lzmax, sizeof(double) * (nb + 1));
    c->bufferbias = (int8_t *)dupmem(g->bufferbias, nb + 1);
    c->ibuffer = (int32_t *)calloc(c->cap + 1, sizeof(int32_t));
    if (ibuffer) memcpy(c->ibuffer, ibuffer, sizeof(int32_t) * c->s.ntot);
    return BROMOC_OK;
}

EXPORT int orc_set_ibuffer(orc_ctx *c, const int32_t *ib) {
    if (!c->ibuffer) return BROMOC_E_STATE;
    memcpy(c->ibuffer, ib, sizeof(int32_t) * c->s.ntot);
    return BROMOC_OK;
}

EXPORT int orc_get_ibuffer(orc_ctx *c, int32_t *ib) {
    memcpy(ib, c->ibuffer, sizeof(int32_t) * c->s.ntot);
    return BROMOC_OK;
}

typedef struct { const double *u; int64_t n, k; int err; } ustream;
static double rndm(ustream *st) {
    if (st->k >= st->n) { st->err = 1; return 0.5; }
    return st->u[st->k++];
}

/* INSERT's acceptance, insert.f90:33-40: inside the spherical shell rmin..rmax of buffer
 * ib (strictly), inside the ellipse, or anywhere in the box slab. */
static int insert_ok(const orc_ctx *c, int ib, double x, double y, double z) {
    const bromoc_system *s = &c->s;
    const bromoc_gcmc *g = &c->g;
    if (s->flags & BROMOC_QSPHERE) {
        const double r2 = (x - g->cx) * (x - g->cx) + (y - g->cy) * (y - g->cy) +
                          (z - g->cz) * (z - g->cz);
        return r2 > c->rmin[ib] * c->rmin[ib] && r2 < c->rmax[ib] * c->rmax[ib];
    }
    if (s->flags & BROMOC_QECYL) {
        const double ex = (x - g->cx) * s->iecx, ey = (y - g->cy) * s->iecy;
        return ex * ex + ey * ey < 1.0;
    }
    return 1;
}

/* GRAND, grand.f90:40-118 with INSERT (insert.f90:29-41) and FIND
 * (find.f90:26-32); every rndm() is taken from `uniforms` in call order. */
EXPORT int orc_grand(orc_ctx *c, int32_t ngcmc, const double *uniforms, int64_t nuniforms,
                     int64_t *used, int32_t *nat, int32_t *prob, int64_t icycle, double *ener) {
    bromoc_system *s = &c->s;
    const bromoc_gcmc *g = &c->g;
    ustream st = {uniforms, nuniforms, 0, 0};
    int nbuffer = g->nbuffer;
    double nbufferr = (double)nbuffer;
    int pstride = c->cap + 1;
    for (int igrand = 0; igrand < ngcmc; igrand++) {
        if (rndm(&st) < 0.5) {
            int ib = (int)(nbufferr * rndm(&st)) + 1;
            if (ib <= nbuffer && c->avnum[ib] > 0.0) {
                int itype = c->ibfftyp[ib];
                double xnew, ynew, znew;
                for (;;) {   /* INSERT: draw until the point is in the buffer's region */
                    xnew = g->lx * (rndm(&st) - 0.5) + g->cx;
                    ynew = g->ly * (rndm(&st) - 0.5) + g->cy;
                    znew = c->lzmin[ib] + (c->lzmax[ib] - c->lzmin[ib]) * rndm(&st);
                    if (st.err || insert_ok(c, ib, xnew, ynew, znew)) break;
                }
                double dener;
                int rc = orc_interact(c, xnew, ynew, znew, itype, s->ntot, 0, &dener);
                if (rc) return rc;
                double rate;
                if (c->bufferbias[ib]) {
                    rate = (c->avnum[ib] + c->kb[ib] * (c->avnum[ib] - (double)(c->ntotat[ib] / icycle)));
                    rate = rate * exp(-(dener - c->mu[ib]) * g->ikbt) / (double)(nat[ib] + 1);
                } else {
                    rate = (c->avnum[ib] / (nat[ib] + 1)) * exp(-(dener - c->mu[ib]) * g->ikbt);
                }
                rate = rate / (1.0 + rate);
                if (rndm(&st) <= rate) {
                    if (s->ntot + 1 > c->cap) return BROMOC_E_CAPACITY;
                    *ener = *ener + dener;
                    nat[ib] = nat[ib] + 1;
                    int n = s->ntot;
                    s->ntot = s->ntot + 1;
                    c->x[n] = xnew; c->y[n] = ynew; c->z[n] = znew;
                    c->typei[n] = itype;
                    if (znew < g->cz) c->typei[n] = -itype;
                    c->ibuffer[n] = ib;
                }
            }
        } else {
            int ib = (int)(nbufferr * rndm(&st)) + 1;
            double natib = ib <= nbuffer ? (double)nat[ib] : 0.0; /* nat(nbuffer+1) is unused */
            int ip = (int)(natib * rndm(&st)) + 1;
            if (ib <= nbuffer && ip <= nat[ib]) {
                int itype = c->ibfftyp[ib];
                int iat = -1, ipx = 0;
                for (int a = s->nfix + s->nsites; a < s->ntot; a++) {
                    if (c->ibuffer[a] == ib) {
                        ipx++;
                        if (ip == ipx) { iat = a; break; }
                    }
                }
                if (iat < 0) return BROMOC_E_STATE;
                double dener;
                int rc = orc_interact(c, c->x[iat], c->y[iat], c->z[iat], itype, iat, 1, &dener);
                if (rc) return rc;
                double rate;
                if (c->bufferbias[ib]) {
                    rate = (c->avnum[ib] + c->kb[ib] * (c->avnum[ib] - (double)(c->ntotat[ib] / icycle)));
                    rate = rate * exp(-(dener - c->mu[ib]) * g->ikbt) / (double)nat[ib];
                } else {
                    rate = (c->avnum[ib] / nat[ib]) * exp(-(dener - c->mu[ib]) * g->ikbt);
                }
                rate = 1.0 / (1.0 + rate);
                if (rndm(&st) <= rate) {
                    *ener = *ener - dener;
                    int n = s->ntot - 1;
                    c->x[iat] = c->x[n]; c->y[iat] = c->y[n]; c->z[iat] = c->z[n];
                    c->typei[iat] = c->typei[n];
                    c->ibuffer[iat] = c->ibuffer[n];
                    nat[ib] = nat[ib] - 1;
                    s->ntot = s->ntot - 1;
                }
            }
        }
        for (int ib = 1; ib <= nbuffer; ib++) prob[(ib - 1) * pstride + nat[ib]] += 1;
        if (st.err) return BROMOC_E_ARG;
    }
    if (used) *used = st.k;
    return BROMOC_OK;
}

/* METROPOLIS, metropolis.f90:31-55, with MOVE (move.f90:32-47, rectangular box); every
 * rndm() from `uniforms` in call order.  A trial runs only when its MOVE draws and one more
 * uniform are available, else the call stops with *done < nmcm. */
EXPORT int orc_metropolis(orc_ctx *c, int32_t nmcm, double mcmax, double ikbt,
                          const double *u, int64_t nu, int64_t *used, int32_t *done,
                          double *ener) {
    bromoc_system *s = &c->s;
    int first = s->nsites + s->nfix;
    double natom2 = (double)(s->ntot - first);
    int64_t k = 0;
    int32_t nd = 0;
    if (natom2 <= 0.0 && nmcm > 0) return BROMOC_E_STATE;
    for (int im = 0; im < nmcm; im++) {
        int64_t kk = k;
        int iat = 0, out = 1;
        double xnew = 0.0, ynew = 0.0, znew = 0.0;
        while (out) {
            if (kk + 4 > nu) break;
            iat = first + (int)(natom2 * u[kk]);
            xnew = c->x[iat] + mcmax * (u[kk + 1] - 0.5);
            ynew = c->y[iat] + mcmax * (u[kk + 2] - 0.5);
            znew = c->z[iat] + mcmax * (u[kk + 3] - 0.5);
            kk += 4;
            /* MOVE's test (move.f90:39-46) is OUTBOX's */
            out = outbox(s, xnew, ynew, znew);
        }
        if (out || kk + 1 > nu) break;
        double eold, enew;
        int rc = orc_interact(c, c->x[iat], c->y[iat], c->z[iat], abs(c->typei[iat]), iat, 1, &eold);
        if (rc) return rc;
        rc = orc_interact(c, xnew, ynew, znew, abs(c->typei[iat]), iat, 0, &enew);
        if (rc) return rc;
        int acc = enew <= eold;
        if (!acc) {
            double bltz = exp(-(enew - eold) * ikbt);
            acc = u[kk++] < bltz;
        }
        if (acc) {
            c->x[iat] = xnew;
            c->y[iat] = ynew;
            c->z[iat] = znew;
            *ener = *ener + (enew - eold);
        }
        k = kk;
        nd = im + 1;
    }
    if (used) *used = k;
    if (done) *done = nd;
    return BROMOC_OK;
}

/* the nearest-site records dids(1:5, i) left by the last ENERGY, [5*ntot] */
EXPORT int orc_get_dids(orc_ctx *c, double *out) {
    memcpy(out, c->dids, sizeof(double) * 5 * (size_t)c->s.ntot);
    return BROMOC_OK;
}

/* REPWALLS, repwalls.f90:19-67 (one map, no Qnmcden): successive layers, in place */
EXPORT int orc_repwalls(int32_t nclx2, int32_t ncly2, int32_t nclz2, int8_t *phiv, int32_t walls) {
    size_t ncel3 = (size_t)nclx2 * ncly2 * nclz2;
    int8_t *phivt = (int8_t *)malloc(ncel3 ? ncel3 : 1);
    for (int n = 1; n <= walls; n++) {
        memcpy(phivt, phiv, ncel3);
        int ck = 0;
        for (int a = 1 + n; a <= nclx2 - n; a++)
            for (int b = 1 + n; b <= ncly2 - n; b++)
                for (int c = 1 + n; c <= nclz2 - n; c++) {
                    size_t pos = (size_t)(a - 1) * ncly2 * nclz2 + (size_t)(b - 1) * nclz2 + (c - 1);
                    if (phiv[pos] != n) continue;
                    ck = ck + 1;
                    int phisum = 0;
                    for (int i = a - 1; i <= a + 1; i++)
                        for (int j = b - 1; j <= b + 1; j++)
                            for (int k = c - 1; k <= c + 1; k++)
                                phisum += phiv[(size_t)(i - 1) * ncly2 * nclz2 + (size_t)(j - 1) * nclz2 + (k - 1)];
                    if (phisum == 27 * n) phivt[pos] = (int8_t)(n + 1);
                }
        memcpy(phiv, phivt, ncel3);
        if (ck == 0) break;
    }
    free(phivt);
    return BROMOC_OK;
}

/* ------------------------------------------------------ per-cycle statistics */
EXPORT int orc_set_analysis(orc_ctx *c, const bromoc_analysis *an) {
    free(c->rho); free(c->rdna); free(c->gr); free(c->prob2); free(c->prob3);
    c->an = *an;
    const size_t nt = (size_t)c->s.ntype, n2 = (size_t)c->s.nttyp;
    c->rho = (double *)calloc(nt * (an->nzmax > 0 ? an->nzmax : 1), sizeof(double));
    c->rdna = (double *)calloc(n2 * (an->nzmax > 0 ? an->nzmax : 1), sizeof(double));
    c->gr = (double *)calloc(nt * (an->npoints > 0 ? an->npoints : 1), sizeof(double));
    c->prob2 = (int64_t *)calloc(nt * ((size_t)(an->datom > 0 ? an->datom : 0) + 1), sizeof(int64_t));
    c->prob3 = (int64_t *)calloc(nt * ((size_t)(an->datom > 0 ? an->datom : 0) + 1), sizeof(int64_t));
    c->irmax = 0;
    c->pres = 0.0;
    c->pres_steps = 0;
    return BROMOC_OK;
}

/* int() of a real, with the samples whose int() overflows (the reference's undefined case)
 * reported as out of range */
static int int_of(double v, int *out) {
    if (!(v > -2147483648.0 && v < 2147483647.0)) return 0;
    *out = (int)v;
    return 1;
}

/* one cycle of simul1.f90:477-538 (Qrho, Qrdna, Qgr, Qprob), in the reference's loop order.
 * Samples with iz < 1 would be written below the reference's arrays: left out. */
EXPORT int orc_analyze(orc_ctx *c) {
    const bromoc_system *s = &c->s;
    const bromoc_analysis *a = &c->an;
    const int ion0 = s->nsites + s->nfix, nz = a->nzmax;
    int iz, ir;
    if (a->what & BROMOC_AN_RHO) {
        for (int iat = ion0; iat < s->ntot; iat++) {
            int itype = abs(c->typei[iat]);
            if (!int_of((c->z[iat] - a->zmini) * a->idelz, &iz)) continue;
            iz = iz + 1;
            if (iz <= nz && iz >= 1) c->rho[(size_t)(itype - 1) * nz + iz - 1] += 1.0;
        }
    }
    if (a->what & BROMOC_AN_RDNA) {
        for (int iat = 0; iat < s->nsites; iat++) {
            int itype = c->typtyp[iat];
            if (!int_of((c->z[iat] - a->zmini) * a->idelz, &iz)) continue;
            iz = iz + 1;
            if (iz <= nz && iz >= 1) c->rdna[(size_t)(itype - 1) * nz + iz - 1] += 1.0;
        }
    }
    if (a->what & BROMOC_AN_GR) {
        const int igr = a->igr;
        for (int iat = ion0; iat < s->ntot; iat++) {
            int itype = abs(c->typei[iat]);
            double dist = (c->x[iat] - c->x[igr]) * (c->x[iat] - c->x[igr]) +
                          (c->y[iat] - c->y[igr]) * (c->y[iat] - c->y[igr]) +
                          (c->z[iat] - c->z[igr]) * (c->z[iat] - c->z[igr]);
            dist = sqrt(dist);
            if (!int_of(dist / a->dr, &ir)) ir = 2147483646;
            ir = ir + 1;
            if (ir > c->irmax) c->irmax = ir;
            if (c->irmax > a->npoints) c->irmax = a->npoints;
            if (ir <= a->npoints) c->gr[(size_t)(itype - 1) * a->npoints + ir - 1] += 1.0;
        }
    }
    if (a->what & BROMOC_AN_PROB) {
        const int nt = s->ntype - s->nold;
        int *ncount = (int *)calloc((size_t)nt + 1, sizeof(int));
        int qmemb = (s->flags & BROMOC_QMEMB) != 0;
        int ok = !qmemb && a->czmin == 0.0 && a->czmax == 0.0;
        if (!ok) {
            for (int iat = ion0; iat < s->ntot; iat++) {
                int itype = abs(c->typei[iat]) - s->nold;
                if (qmemb) {
                    double r = sqrt(c->x[iat] * c->x[iat] + c->y[iat] * c->y[iat]);
                    ok = c->z[iat] >= s->zmemb1 && c->z[iat] <= s->zmemb2;
                    if (ok) {
                        double rc = c->rcylinder[c->nwtype[itype + s->nold]];
                        ok = rc > 0.0 && r <= rc;
                    }
                } else {
                    ok = c->z[iat] >= a->czmin && c->z[iat] <= a->czmax;
                }
                if (ok) ncount[itype - 1]++;
            }
            for (int itype = s->nold + 1; itype <= s->ntype; itype++) {
                int n = ncount[itype - s->nold - 1];
                if (n <= a->datom) c->prob2[(size_t)(itype - 1) * (a->datom + 1) + n] += 1;
            }
        }
        memset(ncount, 0, sizeof(int) * ((size_t)nt + 1));
        for (int iat = ion0; iat < s->ntot; iat++) {
            int itype = abs(c->typei[iat]) - s->nold;
            ncount[itype - 1]++;
        }
        for (int itype = s->nold + 1; itype <= s->ntype; itype++) {
            int n = ncount[itype - s->nold - 1];
            if (n <= a->datom) c->prob3[(size_t)(itype - 1) * (a->datom + 1) + n] += 1;
        }
        free(ncount);
    }
    return BROMOC_OK;
}

/* the statistics as accumulated (rho, rdna, gr as the reference's reals) */
EXPORT int orc_get_analysis(orc_ctx *c, double *rho, double *rdna, double *gr, int32_t *irmax,
                            int64_t *prob2, int64_t *prob3, double *pres, int64_t *pres_steps) {
    const size_t nt = (size_t)c->s.ntype, n2 = (size_t)c->s.nttyp;
    const bromoc_analysis *a = &c->an;
    if (rho) memcpy(rho, c->rho, sizeof(double) * nt * a->nzmax);
    if (rdna) memcpy(rdna, c->rdna, sizeof(double) * n2 * a->nzmax);
    if (gr) memcpy(gr, c->gr, sizeof(double) * nt * a->npoints);
    if (prob2) memcpy(prob2, c->prob2, sizeof(int64_t) * nt * ((size_t)a->datom + 1));
    if (prob3) memcpy(prob3, c->prob3, sizeof(int64_t) * nt * ((size_t)a->datom + 1));
    if (irmax) *irmax = c->irmax;
    if (pres) *pres = c->pres;
    if (pres_steps) *pres_steps = c->pres_steps;
    return BROMOC_OK;
}
