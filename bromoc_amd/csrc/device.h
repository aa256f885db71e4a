// device.h -- device-side helpers shared by the single-system kernels (kernels.hip) and
// the replica-batch kernels (batch.hip): pair potentials, grids, BD update, Philox.
#pragma once

#include <math.h>

#include "engine.h"

namespace bromoc {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;
constexpr int WPB = BLOCK / WAVE;

// feature bits of the k_step specialisations
constexpr int F_DNA = 1;     // CG-DNA sites and lists present
constexpr int F_PHIX = 2;    // static field grid
constexpr int F_PHIV = 4;    // repulsion grid
constexpr int F_UPDATE = 8;  // apply the BD update (else ENERGY only)
constexpr int F_TAB = 16;    // every pair type tabulated: the analytic pair path is compiled
                             // out (it would set the register budget of the whole kernel)
constexpr int F_PROX = 32;   // track each ion's nearest DNA site (PROXDIFF, DYNAMICS3)

// host-side flag values (include/bromoc_gpu.h)
constexpr uint32_t QPAR = 1u << 0, QNUCL = 1u << 1, QNONBOND = 1u << 2, QNOBOND = 1u << 3;
constexpr uint32_t QPHIX = 1u << 4, QPHIV = 1u << 5, QSVDW = 1u << 6, QSOLV = 1u << 7;
constexpr uint32_t QDEBY = 1u << 8, QCOUNTION = 1u << 9, QDNAFREE = 1u << 10;
constexpr uint32_t QBSPLINE = 1u << 18, QPROXDIFF = 1u << 21, QNOTRANS = 1u << 22;
constexpr uint32_t QAPFOR = 1u << 23, QCONTRANS = 1u << 24, QMEMB = 1u << 16;
constexpr uint32_t QDEBYHYB = 1u << 20, QNMCDEN = 1u << 19;

constexpr double PI = 3.14159265358979323846264338327950288419716939937510;
constexpr double RSMALL = 1E-3;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// Order-independent sums of doubles: a term v is cut at the 2^-32 grid into an integer part
// floor(v 2^32) and W - 1 further 32-bit digits below it (units 2^-64, 2^-96), each added
// with an integer atomic; integer addition is associative, so the total is the same
// whatever the order of the terms -- bit-reproducible -- and exact in units of 2^-32W for
// |total| < 2^31 and fewer than 2^32 terms (fix_value, engine.h, reads a sum back).
template <int W>
__device__ __forceinline__ void fix_add(unsigned long long *acc, double v) {
    double a = v * 4294967296.0;
    const double f = floor(a);
    atomicAdd(acc, (unsigned long long)(long long)f);
    a = a - f;
#pragma unroll
    for (int w = 1; w < W; w++) {
        a = a * 4294967296.0;
        const double d = floor(a);
        atomicAdd(acc + w, (unsigned long long)d);
        a = a - d;
    }
}
// A block's partial sum before it goes to a global one: the carries of the lower words
// moved up, so each lower word adds less than 2^32
template <int W>
__device__ __forceinline__ void fix_carry(unsigned long long *acc) {
#pragma unroll
    for (int w = W - 1; w > 0; w--) {
        acc[w - 1] += acc[w] >> 32;
        acc[w] &= 0xFFFFFFFFull;
    }
}

// small integer powers by repeated squaring (the step kernels agree with the reference to
// rounding, not bitwise; the bit-exact table construction has its own, tables.h)
__device__ __forceinline__ double pow3(double a) { double a2 = a * a; return a * a2; }
__device__ __forceinline__ double pow5(double a) { double a2 = a * a; double a4 = a2 * a2; return a * a4; }
__device__ __forceinline__ double pow6(double a) { double a2 = a * a; double a4 = a2 * a2; return a2 * a4; }
__device__ __forceinline__ double fsign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

// getyd, polynom.f90:197-230: energy y and force factor yp (= -dU/dr / r)
__device__ __forceinline__ void getyd(const DevPair &p, const double4 *__restrict__ ep,
                                      double ires, double x2, double &y, double &yp) {
    if (x2 < p.dm2lo) {
        double ix2 = 1.0 / x2;
        double c = p.sc1 * pow6(ix2);
        y = c + p.sc2;
        yp = 12.0 * c * ix2;
    } else if (x2 > p.dm2hi) {
        double ix2 = 1.0 / x2;
        double c = p.sc3 * pow3(ix2);
        if (p.flags & PF_CHR) {
            double f = p.fct * sqrt(ix2);
            y = c + f;
            yp = (6.0 * c + f) * ix2;
        } else {
            y = c;
            yp = 6.0 * c * ix2;
        }
    } else {
        double x = sqrt(x2);
        int i = (int)((x - p.dmi) * ires) + p.nxi;
        double4 e = ep[i];
        y = e.x + x * (e.y + x * e.z);
        yp = -e.y / x - 2.0 * e.z;
    }
}

// analytic pair branch, energy.mono.f90:135-178
__device__ __forceinline__ double pair_analytic(const DevPair &p, const StepParams &P, double dist2,
                                                double &eel, double &evd, double &esr) {
    double idist2 = 1.0 / dist2, idist = 0.0, cofo = 0.0, dist6 = 0.0, dist12 = 0.0, dist = 0.0;
    double e1 = 0.0, e2 = 0.0, e3 = 0.0, fdf = 0.0, fdv = 0.0;
    const bool srp = (p.flags & PF_SRPMF) && dist2 <= P.rth;
    if (p.flags & (PF_CHR | PF_SRPMF)) idist = sqrt(idist2);
    if (p.flags & PF_CHR) {
        cofo = p.fct * idist;
        eel += cofo;
    }
    if (p.flags & PF_LJ) {
        dist6 = pow3(p.sgp2 * idist2);
        dist12 = dist6 * dist6;
        evd += p.epp4 * (dist12 - dist6);
    }
    if (srp) {
        dist = 1.0 / idist;
        e1 = exp((p.c1 - dist) * p.c2);
        e2 = cos(p.c3 * PI * (p.c1 - dist));
        e3 = pow6(p.c1 * idist);
        double e0 = p.c0 * e1 * e2 + p.c4 * e3;
        if (dist >= P.srpx) {
            fdf = exp(-P.srpk * (dist - P.srpx)) - P.srpy;
            fdv = e0;
            e0 = fdv * fdf;
        }
        esr += e0;
    }
    double de = 0.0;
    if (p.flags & PF_CHR) de = cofo * idist2;
    if (p.flags & PF_LJ) de = de + p.epp4 * (2.0 * dist12 - dist6) * 6.0 * idist2;
    if (srp) {
        double dc = (-p.c2 * e2 + p.c3 * PI * sin(p.c3 * PI * (p.c1 - dist))) * p.c0 * e1 -
                    6.0 * p.c4 * e3 * idist;
        if (dist >= P.srpx) dc = dc * fdf - fdv * P.srpk * (fdf + P.srpy);
        de = de - dc * idist;
    }
    return de;
}

// Philox4x64-10 (Salmon et al. 2011) and Box-Muller on 53-bit uniforms
__device__ __forceinline__ void philox4x64(uint64_t c[4], uint64_t k0, uint64_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
        uint64_t hi0 = __umul64hi(M0, c[0]), lo0 = M0 * c[0];
        uint64_t hi1 = __umul64hi(M1, c[2]), lo1 = M1 * c[2];
        uint64_t n0 = hi1 ^ c[1] ^ k0, n1 = lo1, n2 = hi0 ^ c[3] ^ k1, n3 = lo0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += 0x9E3779B97F4A7C15ull;
        k1 += 0xBB67AE8584CAA73Bull;
    }
}

__device__ __forceinline__ void philox_gauss3(uint64_t seed, uint64_t step, uint64_t stream,
                                              int64_t i, double g[3]) {
    uint64_t c[4] = {step, (uint64_t)i, stream, 0};
    philox4x64(c, seed, 0);
    const double inv53 = 1.0 / 9007199254740992.0;
    double u1 = (double)((c[0] >> 11) + 1) * inv53;
    double u2 = (double)(c[1] >> 11) * inv53;
    double u3 = (double)((c[2] >> 11) + 1) * inv53;
    double u4 = (double)(c[3] >> 11) * inv53;
    double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    const double twopi = 2.0 * PI;
    g[0] = r1 * cos(twopi * u2);
    g[1] = r1 * sin(twopi * u2);
    g[2] = r2 * cos(twopi * u4);
}

__device__ __forceinline__ void fixcoor(const StepParams &P, double &x, double &y, double &z) {
    if (x < P.lx2m) x = P.lx2m;
    else if (x > P.lx2p) x = P.lx2p;
    if (y < P.ly2m) y = P.ly2m;
    else if (y > P.ly2p) y = P.ly2p;
    if (z < P.lz2m) z = P.lz2m;
    else if (z > P.lz2p) z = P.lz2p;
}

// One trilinear corner (lane c = 0..7) of staticf1/vdwgd1trln.  Adds the corner's force
// and energy to the lane's accumulators; returns the corner's grid value.
template <typename T>
__device__ __forceinline__ double tri_corner(const GridGeom &g, const T *__restrict__ data,
                                             double xi, double yi, double zi, int ix, int iy,
                                             int iz, int c, double scale, double &fx, double &fy,
                                             double &fz, double &e) {
    const double one = 1.0;
    int n1 = ix + (c >> 2), n2 = iy + ((c >> 1) & 1), n3 = iz + (c & 1);
    double ai = xi - n1 * g.dcel;
    double aisign = fsign(one, ai);
    ai = one - fabs(ai) * g.idcel;
    double bi = yi - n2 * g.dcel;
    double bisign = fsign(one, bi);
    bi = one - fabs(bi) * g.idcel;
    double ci = zi - n3 * g.dcel;
    double cisign = fsign(one, ci);
    ci = one - fabs(ci) * g.idcel;
    double fi = ai * bi * ci;
    size_t in3 = ((size_t)n1 * g.ny + n2) * g.nz + n3;
    double ph = (double)data[in3];
    double prefac = ph * scale * g.idcel;
    if (ai < (1.0 - RSMALL) && ai > RSMALL) fx += aisign * bi * ci * prefac;
    if (bi < (1.0 - RSMALL) && bi > RSMALL) fy += bisign * ai * ci * prefac;
    if (ci < (1.0 - RSMALL) && ci > RSMALL) fz += cisign * ai * bi * prefac;
    e += fi * scale * ph;
    return ph;
}

__device__ __forceinline__ void grid_cell(const GridGeom &g, double x, double y, double z,
                                          double &xi, double &yi, double &zi, int &ix, int &iy,
                                          int &iz) {
    xi = x + g.tranx - g.xbcen;
    yi = y + g.trany - g.ybcen;
    zi = z + g.tranz - g.zbcen;
    ix = (int)(xi * g.idcel);
    iy = (int)(yi * g.idcel);
    iz = (int)(zi * g.idcel);
    if (ix == g.nx - 1) ix = g.nx - 2;
    if (iy == g.ny - 1) iy = g.ny - 2;
    if (iz == g.nz - 1) iz = g.nz - 2;
}

// The trilinear grid force of one ion and its Jacobian J = dF/dx from the 8 corners in
// order (the step kernels' per-corner terms, tri_corner), added to F and J.  Within a cell
// the interpolant is linear in each coordinate, so J's diagonal is zero; an off-diagonal
// entry is the mixed second derivative of the force component it belongs to, under that
// component's near-node guard, so J is not symmetric where the guards differ.
// J[0..5] = dFx/dy, dFx/dz, dFy/dx, dFy/dz, dFz/dx, dFz/dy.
template <typename T>
__device__ inline void tri_force_jac(const GridGeom &g, const T *__restrict__ data, double scale,
                                     double x, double y, double z, double F[3], double J[6]) {
    double gx, gy, gz;
    int ix, iy, iz;
    grid_cell(g, x, y, z, gx, gy, gz, ix, iy, iz);
    const double one = 1.0;
    for (int c = 0; c < 8; c++) {
        const int n1 = ix + (c >> 2), n2 = iy + ((c >> 1) & 1), n3 = iz + (c & 1);
        double ai = gx - n1 * g.dcel;
        const double as = fsign(one, ai);
        ai = one - fabs(ai) * g.idcel;
        double bi = gy - n2 * g.dcel;
        const double bs = fsign(one, bi);
        bi = one - fabs(bi) * g.idcel;
        double ci = gz - n3 * g.dcel;
        const double cs = fsign(one, ci);
        ci = one - fabs(ci) * g.idcel;
        const double ph = (double)data[((size_t)n1 * g.ny + n2) * g.nz + n3];
        const double prefac = ph * scale * g.idcel;
        const double q = prefac * g.idcel;   // d(1 - |u|/dcel)/du = -sign(u)/dcel
        if (ai < (1.0 - RSMALL) && ai > RSMALL) {
            F[0] += as * bi * ci * prefac;
            J[0] -= as * bs * ci * q;
            J[1] -= as * bi * cs * q;
        }
        if (bi < (1.0 - RSMALL) && bi > RSMALL) {
            F[1] += bs * ai * ci * prefac;
            J[2] -= bs * as * ci * q;
            J[3] -= bs * ai * cs * q;
        }
        if (ci < (1.0 - RSMALL) && ci > RSMALL) {
            F[2] += cs * ai * bi * prefac;
            J[4] -= cs * as * bi * q;
            J[5] -= cs * ai * bs * q;
        }
    }
}

// Quadratic B-spline M3 and its derivative DM3 (m3.f90:19-42, dm3.f90:19-38).
__device__ __forceinline__ void bspline_m3(double x, double &m, double &dm) {
    const double m2m = (x >= 0.0 && x <= 2.0) ? 1.0 - fabs(x - 1.0) : 0.0;
    const double m2 = (x >= 1.0 && x <= 3.0) ? 1.0 - fabs(x - 2.0) : 0.0;
    m = (x >= 0.0 && x <= 3.0) ? x * 0.5 * m2m + (3.0 - x) * 0.5 * m2 : 0.0;
    dm = m2m - m2;
}

// Centre node of the 27-point stencil (vdwgd1spln.f90:84-96): nearest node, kept one node
// inside every face.
__device__ __forceinline__ void spline_cell(const GridGeom &g, double x, double y, double z,
                                            double &xi, double &yi, double &zi, int &ix,
                                            int &iy, int &iz) {
    xi = x + g.tranx - g.xbcen;
    yi = y + g.trany - g.ybcen;
    zi = z + g.tranz - g.zbcen;
    ix = (int)round(xi * g.idcel);
    iy = (int)round(yi * g.idcel);
    iz = (int)round(zi * g.idcel);
    if (ix == 0) ix = 1;
    if (iy == 0) iy = 1;
    if (iz == 0) iz = 1;
    if (ix == g.nx - 1) ix = g.nx - 2;
    if (iy == g.ny - 1) iy = g.ny - 2;
    if (iz == g.nz - 1) iz = g.nz - 2;
}

// First node of particle i's repulsion map (vdwgd1trln.f90:61-80, vdwgd0trln.f90:56-75):
// one map for all without NMCDEN.  With it, a DNA site reads its site name's map (S, P, Ab,
// Tb, Cb, Gb: maps 0..5), a moving ion of atnam type itype the map of its type, numbered
// from the first ion type (nold + 1) after the six site maps when there is DNA; a fixed
// ion, or an ion without PARTICLE, reads map 0.
__device__ __forceinline__ size_t vdw_map(const StepParams &P, int i, int itype) {
    if (!(P.flags & QNMCDEN)) return 0;
    const size_t ncel3 = (size_t)P.gv.nx * P.gv.ny * P.gv.nz;
    if ((P.flags & QNUCL) && i < P.nsites) return (size_t)P.namsite[i] * ncel3;
    if (!(P.flags & QPAR) || i < P.nsites + P.nfix) return 0;
    return (size_t)(((P.flags & QNUCL) ? 6 : 0) + itype - P.nold - 1) * ncel3;
}

// The same for a stored particle i (its type from typei).
__device__ __forceinline__ size_t vdw_map_of(const StepParams &P, int i) {
    if (!(P.flags & QNMCDEN)) return 0;
    return vdw_map(P, i, i < P.nsites ? 0 : abs(P.typei[i]));
}

// The grid forces of moving ion i at (x, y, z), atnam2 type t2, and their Jacobian:
// STATICF1 (staticf1.f90, charge cg) and the trilinear VDWGD1TRLN (vdwgd1trln.f90, svdw,
// scal with QSVDW, the ion's own map with NMCDEN), inside the same bounds as the step.
__device__ inline void grid_force_jac(const StepParams &P, int i, double x, double y, double z,
                                      int t2, double F[3], double J[6]) {
    for (int k = 0; k < 3; k++) F[k] = 0.0;
    for (int k = 0; k < 6; k++) J[k] = 0.0;
    if (P.flags & QPHIX) {
        const GridGeom &g = P.gx;
        if (x <= g.xbcen + g.tranx && x >= g.xbcen - g.tranx && y <= g.ybcen + g.trany &&
            y >= g.ybcen - g.trany && z <= g.zbcen + g.tranz && z >= g.zbcen - g.tranz)
            tri_force_jac(g, P.phix, P.celec * P.cg[abs(P.typei[i])], x, y, z, F, J);
    }
    if (P.flags & QPHIV) {
        const GridGeom &g = P.gv;
        const double esvdw = (P.flags & QSVDW) ? P.svdw * P.scal[t2] : P.svdw;
        if (x <= g.xbcen + g.tranx && x >= g.xbcen - g.tranx && y <= g.ybcen + g.trany &&
            y >= g.ybcen - g.trany && z <= P.vzmax && z >= P.vzmin)
            tri_force_jac(g, P.phiv + vdw_map_of(P, i), esvdw, x, y, z, F, J);
    }
}

// Repulsion map of one ENERGY target at (x, y, z), atnam2 type t2, by a whole wavefront:
// the trilinear stencil of vdwgd1trln.f90:50-161 on lanes 0..7, or with QBSPLINE the
// 27-point B-spline of vdwgd1spln.f90:50-170 on lanes 0..26 (a node counts only where
// none of its three factors vanishes with its derivative, vdwgd1spln.f90:106-124).  Adds
// the lane's force and energy shares; returns, on every lane, whether the map values
// summed over the stencil reach the wall threshold (warn).  map: the first node of the
// target's map (vdw_map).
__device__ inline bool vdw_grid(const StepParams &P, double x, double y, double z, int t2,
                                double &fx, double &fy, double &fz, double &e,
                                size_t map = 0) {
    const int lane = threadIdx.x & (WAVE - 1);
    const GridGeom &g = P.gv;
    const int8_t *phiv = P.phiv + map;
    double esvdw = P.svdw;
    if (P.flags & QSVDW) esvdw = P.svdw * P.scal[t2];
    double ph = 0.0;
    bool ok;
    if (P.flags & QBSPLINE) {
        ok = x <= g.xbcen + g.tranx - g.dcel && x >= g.xbcen - g.tranx + g.dcel &&
             y <= g.ybcen + g.trany - g.dcel && y >= g.ybcen - g.trany + g.dcel &&
             z <= P.vzmax && z >= P.vzmin;
        if (ok && lane < 27) {
            double xi, yi, zi;
            int ix, iy, iz;
            spline_cell(g, x, y, z, xi, yi, zi, ix, iy, iz);
            const int k = ix - 1 + lane / 9, l = iy - 1 + (lane / 3) % 3, m = iz - 1 + lane % 3;
            double ai, dai, bi, dbi, ci, dci;
            bspline_m3(1.5 - (xi - k * g.dcel) * g.idcel, ai, dai);
            bspline_m3(1.5 - (yi - l * g.dcel) * g.idcel, bi, dbi);
            bspline_m3(1.5 - (zi - m * g.dcel) * g.idcel, ci, dci);
            if ((ai != 0.0 || dai != 0.0) && (bi != 0.0 || dbi != 0.0) && (ci != 0.0 || dci != 0.0)) {
                ph = (double)phiv[((size_t)k * g.ny + l) * g.nz + m];
                const double phivsv = ph * esvdw;
                const double prefac = phivsv * g.idcel;
                fx += dai * bi * ci * prefac;
                fy += ai * dbi * ci * prefac;
                fz += ai * bi * dci * prefac;
                e += ai * bi * ci * phivsv;
            }
        }
    } else {
        ok = x <= g.xbcen + g.tranx && x >= g.xbcen - g.tranx && y <= g.ybcen + g.trany &&
             y >= g.ybcen - g.trany && z <= P.vzmax && z >= P.vzmin;
        if (ok && lane < 8) {
            double xi, yi, zi;
            int ix, iy, iz;
            grid_cell(g, x, y, z, xi, yi, zi, ix, iy, iz);
            ph = tri_corner(g, phiv, xi, yi, zi, ix, iy, iz, lane, esvdw, fx, fy, fz, e);
        }
    }
    // map values are integers: the sum is exact in any order
    const double sum = wave_sum(ph);
    return ok && sum >= ((P.flags & QBSPLINE) ? P.thold27 : P.thold8);
}

// SWITCH1 (switch1.f90:19-47): the membrane walls' switch along z and its derivative.
__device__ __forceinline__ void mswitch1(double z, double p1, double ip1, double zm1, double zm2,
                                         double &sw, double &dsw) {
    if (z >= zm2 + p1 || z <= zm1 - p1) {
        sw = 0.0; dsw = 0.0;
    } else if (z > zm2) {
        const double d = (z - zm2) * ip1;
        sw = 1.0 + (2.0 * d - 3.0) * (d * d);
        dsw = -6.0 * (d - 1.0) * ip1 * d;
    } else if (z < zm1) {
        const double d = (zm1 - z) * ip1;
        sw = 1.0 + (2.0 * d - 3.0) * (d * d);
        dsw = 6.0 * (d - 1.0) * ip1 * d;
    } else {
        sw = 1.0; dsw = 0.0;
    }
}

// SWITCH2 (switch2.f90:19-38): the pore wall's switch in r^2; r is set in the wall only.
__device__ __forceinline__ void mswitch2(double r2, double radius, double p2, double ip2,
                                         double &sw, double &dsw, double &r) {
    if (r2 <= radius * radius) {
        sw = 0.0; dsw = 0.0;
    } else if (r2 < (radius + p2) * (radius + p2)) {
        r = sqrt(r2);
        const double d = (radius + p2 - r) * ip2;
        sw = 1.0 + (2.0 * d - 3.0) * (d * d);
        dsw = 6.0 * (d - 1.0) * d * ip2;
    } else {
        sw = 1.0; dsw = 0.0;
    }
}

// MEMBRANE's term of one particle (membrane.f90:36-104): the Nernst potential of a charge
// and the membrane / pore walls of atnam2 type t2.  Returns the energy, adds the force to
// F, sets warn when the particle sits inside the membrane (counted by the caller).
__device__ inline double membrane_energy(const StepParams &P, double x, double y, double z,
                                         double charge, int t2, double F[3], bool &warn) {
    double em = 0.0;
    warn = false;
    if (P.voltage != 0.0 && charge != 0.0) {
        if (z < P.zmemb1) {
            const double ex = exp(P.ikappa * (z - P.zmemb1));
            em = em + charge * P.afact * ex;
            F[2] = F[2] - charge * P.afact * P.ikappa * ex;
        } else if (z >= P.zmemb1 && z <= P.zmemb2) {
            em = em + charge * P.afact * (P.ceps * P.ikappa * (z - P.zmemb1) + 1.0);
            F[2] = F[2] - charge * P.afact * P.ceps * P.ikappa;
        } else if (z > P.zmemb2) {
            const double ex = exp(-P.ikappa * (z - P.zmemb2));
            em = em + charge * (P.voltage - P.afact * ex);
            F[2] = F[2] - charge * P.afact * P.ikappa * ex;
        }
    }
    const double a1 = P.ampl1[t2];
    if (a1 > 0.0) {
        double sw1, dsw1;
        mswitch1(z, P.p1[2 * t2], P.p1[2 * t2 + 1], P.zmemb1, P.zmemb2, sw1, dsw1);
        if (sw1 != 0.0) {
            if (P.qpore) {
                double sw2, dsw2, r = 0.0;
                const double a2 = P.ampl2[t2];
                mswitch2(x * x + y * y, P.rcylinder[t2], P.p2[2 * t2], P.p2[2 * t2 + 1], sw2, dsw2, r);
                if (sw2 != 0.0) {
                    if (sw1 == 1.0) {
                        if (sw2 == 1.0) {
                            warn = true;
                            em = em + a2;
                        } else {
                            const double ir = 1.0 / r;
                            F[0] = F[0] + a2 * dsw2 * x * ir;
                            F[1] = F[1] + a2 * dsw2 * y * ir;
                            em = em + a2 * sw2;
                        }
                    } else if (sw2 == 1.0) {
                        F[2] = F[2] + a1 * dsw1;
                        em = em + a1 * sw1;
                    } else {
                        const double ir = 1.0 / r;
                        F[0] = F[0] + 0.5 * a2 * dsw2 * x * ir;
                        F[1] = F[1] + 0.5 * a2 * dsw2 * y * ir;
                        F[2] = F[2] + 0.5 * a1 * dsw1;
                        em = em + 0.5 * (a2 * sw2 + a1 * sw1);
                    }
                }
            } else {
                if (sw1 == 1.0) warn = true;
                em = em + a1 * sw1;
            }
        }
    }
    return em;
}

// INTERACT's membrane term (interact.f90:58-119): as ENERGY's, but a particle inside the
// membrane costs 1e10 (and a warning when alert is set).
__device__ inline double membrane_interact(const StepParams &P, double x, double y, double z,
                                           double charge, int t2, bool &warn) {
    double em = 0.0;
    warn = false;
    if (P.voltage != 0.0 && charge != 0.0) {
        if (z < P.zmemb1) em = em + charge * P.afact * exp(P.ikappa * (z - P.zmemb1));
        else if (z >= P.zmemb1 && z <= P.zmemb2)
            em = em + charge * P.afact * (P.ceps * P.ikappa * (z - P.zmemb1) + 1.0);
        else if (z > P.zmemb2)
            em = em + charge * (P.voltage - P.afact * exp(-P.ikappa * (z - P.zmemb2)));
    }
    const double a1 = P.ampl1[t2];
    if (a1 > 0.0) {
        double sw1, dsw1;
        mswitch1(z, P.p1[2 * t2], P.p1[2 * t2 + 1], P.zmemb1, P.zmemb2, sw1, dsw1);
        if (sw1 != 0.0) {
            if (P.qpore) {
                double sw2, dsw2, r = 0.0;
                const double a2 = P.ampl2[t2];
                mswitch2(x * x + y * y, P.rcylinder[t2], P.p2[2 * t2], P.p2[2 * t2 + 1], sw2, dsw2, r);
                if (sw2 != 0.0) {
                    if (sw1 == 1.0) {
                        if (sw2 == 1.0) { warn = true; em = em + 1.0e10; }
                        else em = em + a2 * sw2;
                    } else if (sw2 == 1.0) {
                        em = em + a1 * sw1;
                    } else {
                        em = em + 0.5 * (a2 * sw2 + a1 * sw1);
                    }
                }
            } else if (sw1 == 1.0) {
                warn = true;
                em = em + 1.0e10;
            } else {
                em = em + a1 * sw1;
            }
        }
    }
    return em;
}

// Diffusion profile of DYNAMICS2 (dynamics2.mono.f90:52-74): the end values outside the
// knots, else the cubic spline and its derivative (seval/sevald, sevalmod.f90:23-67).  The
// interval is found by the same bisection; seval's SAVEd interval is not kept (it only
// changes the chosen interval when z equals a knot exactly).
__device__ __forceinline__ void diff_profile(const StepParams &P, double zz, double &sw,
                                             double &dsw) {
    const int n = P.nspline;
    const double *xs = P.xs;
    const bool up = xs[0] < xs[n - 1];
    if (up ? zz < xs[0] : zz > xs[0]) { sw = P.ys[0]; dsw = 0.0; return; }
    if (up ? zz > xs[n - 1] : zz < xs[n - 1]) { sw = P.ys[n - 1]; dsw = 0.0; return; }
    int i = 1, j = n + 1;   // 1-based, as the reference's bisection
    for (int it = 0; it < 64 && j > i + 1; it++) {
        const int k = (i + j) / 2;
        if (zz < xs[k - 1]) j = k;
        if (zz >= xs[k - 1]) i = k;
    }
    const double dx = zz - xs[i - 1];
    sw = P.ys[i - 1] + dx * (P.bs[i - 1] + dx * (P.cs[i - 1] + dx * P.ds[i - 1]));
    dsw = P.bs[i - 1] + dx * (P.cs[i - 1] + dx * P.ds[i - 1]);
}

// Drift and noise factors of an ion of type itype at height z: fact1a/fact2a for DYNAMICS0;
// for DYNAMICS1 the pore switch (dynamics1.mono.f90:38-70) and for DYNAMICS2 the profile
// scale D, with the drift dD/dz * dt on z.
__device__ __forceinline__ void ion_factors(const StepParams &P, int itype, double z,
                                            double &fac1, double &fac2, double &ddz) {
    if (P.dmode == 0) {
        fac1 = P.fact1a[itype];
        fac2 = P.fact2a[itype];
        ddz = 0.0;
        return;
    }
    const double D = P.diffusion[itype];
    double idf, didf;
    if (P.dmode == 1) {
        const double pp3 = P.p3[itype];
        const double pore1 = -P.plength2 + P.pcenter, pore2 = P.plength2 + P.pcenter;
        double sw = 1.0, dsw = 0.0;
        if (z > pore1 - pp3 && z < pore2 + pp3) {
            sw = 0.0;
            if (!(z >= pore1 && z <= pore2)) {
                const double ipp3 = 1.0 / pp3;
                if (z > pore2) {
                    const double t = (z - pore2) * ipp3;
                    sw = 2.0 * (t * t * t) - 3.0 * (t * t);
                    dsw = 6.0 * (t * t - t) * ipp3;
                } else {
                    const double t = (z - pore1) * ipp3;
                    sw = -2.0 * (t * t * t) - 3.0 * (t * t);
                    dsw = -6.0 * (t * t + t) * ipp3;
                }
                sw = -sw;
                dsw = -dsw;
            }
        }
        const double a = P.ampl3[itype];
        idf = D * (a + (1.0 - a) * sw);
        didf = D * (1.0 - a) * dsw;
    } else {
        double sw, dsw;
        diff_profile(P, z, sw, dsw);
        idf = D * sw;
        didf = D * dsw;
    }
    fac1 = idf * P.kbtdt;
    fac2 = sqrt(2.0 * P.dt * idf);
    ddz = didf * P.dt;
}

// Ermak-McCammon update of moving particle i (position pa, force F, Gaussians g):
// DYNAMICS0/1/2/3 for ions (dynamics0.mono.f90:31-46, dynamics1.mono.f90:71-86,
// dynamics2.mono.f90:75-85, dynamics3.mono.f90:37-72) with countions, DYNAMICS0NUC
// (dynamics0nuc.f90:31-44) for sites; fixcoor unless the site is free of the box
// (dnafree).  dids (DYNAMICS3 only): the ion's nearest-site record of ENERGY
// (energy.mono.f90:603-620): dist - dmi, dist, and the ion-minus-site vector.
__device__ __forceinline__ double4 bd_move(const StepParams &P, bool is_site, int itype,
                                           const double4 &pa, const double F[3],
                                           const double *g, int32_t *nfwd, int32_t *nbwd,
                                           const double *dids = nullptr) {
    double fac1, fac2, ddz = 0.0;
    double ddx3 = 0.0, ddy3 = 0.0, ddz3 = 0.0;
    const bool prox = !is_site && P.dmode == 3;
    if (is_site) {
        fac1 = P.fact0n1;
        fac2 = P.fact0n2;
    } else if (prox) {
        if (dids[0] > 0.0 && dids[0] <= P.diffcutoff) {
            const double D = P.diffusion[itype];
            const double ebet = exp(-dids[0] * P.ibeta);
            const double idf = D - (D - P.diff0) * ebet;
            const double didf = (D - P.diff0) * P.ibeta * ebet;
            const double idist = 1.0 / dids[1];
            fac1 = idf * P.kbtdt;
            fac2 = sqrt(2.0 * P.dt * idf);
            ddx3 = didf * P.dt * dids[2] * idist;
            ddy3 = didf * P.dt * dids[3] * idist;
            ddz3 = didf * P.dt * dids[4] * idist;
        } else {
            fac1 = P.fact1a[itype];
            fac2 = P.fact2a[itype];
        }
    } else {
        ion_factors(P, itype, pa.z, fac1, fac2, ddz);
    }
    const double bdmax = P.bdmax;
    double delx = F[0] * fac1, dely = F[1] * fac1, delz = F[2] * fac1;
    if (fabs(delx) > bdmax) delx = fsign(bdmax, delx);
    if (fabs(dely) > bdmax) dely = fsign(bdmax, dely);
    if (fabs(delz) > bdmax) delz = fsign(bdmax, delz);
    double xn = pa.x + delx + fac2 * g[0];
    double yn = pa.y + dely + fac2 * g[1];
    double zn = pa.z + delz + fac2 * g[2];
    if (prox) {
        xn = xn + ddx3;
        yn = yn + ddy3;
        zn = zn + ddz3;
    } else if (!is_site && P.dmode != 0) {
        zn = zn + ddz;
    }
    if (!is_site && (P.flags & QCOUNTION)) {
        const int row = itype - P.nold - 1;
        for (int c = 0; c < P.cntpts; c++) {
            const double zc = P.zcont[c];
            if (pa.z < zc && zn >= zc) atomicAdd(&nfwd[row * P.cntpts + c], 1);
            else if (pa.z >= zc && zn < zc) atomicAdd(&nbwd[row * P.cntpts + c], 1);
        }
    }
    if (!is_site || !(P.flags & QDNAFREE)) fixcoor(P, xn, yn, zn);
    return make_double4(xn, yn, zn, pa.w);
}

__device__ __forceinline__ double4 bd_update(const StepParams &P, int i, bool is_site,
                                             const double4 &pa, const double F[3],
                                             const double *g, const double *dids = nullptr) {
    return bd_move(P, is_site, is_site ? 0 : abs(P.typei[i]), pa, F, g, P.nfwd, P.nbwd, dids);
}

// Reaction-field parameters of an ion at (x, y, z) of atnam type itype: srfe and reff
// from the 8 corners of gsrfen / sqrfac*greff (rfpar.f90:75-142), with their gradients
// when grad is set.  The corner loop leaves the innermost (z) loop at a zero gsrfen corner,
// as the reference's `exit` does (rfpar.f90:121).  Returns false outside the grid box.
__device__ inline bool rf_params(const StepParams &P, double x, double y, double z, int itype,
                          bool grad, double &srfe, double &reff, double ds[3], double dr[3]) {
    const GridGeom &g = P.gr;
    srfe = 0.0; reff = 0.0;
    ds[0] = ds[1] = ds[2] = dr[0] = dr[1] = dr[2] = 0.0;
    if (!(x <= g.xbcen + g.tranx && x >= g.xbcen - g.tranx && y <= g.ybcen + g.trany &&
          y >= g.ybcen - g.trany && z <= g.zbcen + g.tranz && z >= g.zbcen - g.tranz))
        return false;
    const size_t ncyz = (size_t)g.ny * g.nz, ncel3 = ncyz * g.nx;
    const size_t ifir = P.rfsingle ? 0 : (size_t)(itype - P.nold - 1) * ncel3;
    double xi, yi, zi;
    int ix, iy, iz;
    grid_cell(g, x, y, z, xi, yi, zi, ix, iy, iz);
    ix = min(max(ix, 0), g.nx - 2);
    iy = min(max(iy, 0), g.ny - 2);
    iz = min(max(iz, 0), g.nz - 2);
    const double one = 1.0;
    double aux1 = 0.0, aux2 = 0.0;
    for (int n1 = ix; n1 <= ix + 1; n1++) {
        double ai = xi - n1 * g.dcel;
        const double aisign = fsign(one, ai);
        ai = one - fabs(ai) * g.idcel;
        for (int n2 = iy; n2 <= iy + 1; n2++) {
            double bi = yi - n2 * g.dcel;
            const double bisign = fsign(one, bi);
            bi = one - fabs(bi) * g.idcel;
            for (int n3 = iz; n3 <= iz + 1; n3++) {
                double ci = zi - n3 * g.dcel;
                const double cisign = fsign(one, ci);
                ci = one - fabs(ci) * g.idcel;
                const double fi = ai * bi * ci;
                const size_t in3 = n1 * ncyz + (size_t)n2 * g.nz + n3 + ifir;
                const double gaux1 = (double)P.gsrfen[in3];
                const double gaux2 = P.sqrfac * (double)P.greff[in3];
                if (gaux1 == 0.0) break;
                const double prefa1 = gaux1 * g.idcel, prefa2 = gaux2 * g.idcel;
                aux1 = aux1 + fi * gaux1;
                aux2 = aux2 + fi * gaux2;
                if (grad) {
                    if (ai < (one - RSMALL) && ai > RSMALL) {
                        const double aux = aisign * bi * ci;
                        ds[0] = ds[0] - aux * prefa1;
                        dr[0] = dr[0] - aux * prefa2;
                    }
                    if (bi < (one - RSMALL) && bi > RSMALL) {
                        const double aux = bisign * ai * ci;
                        ds[1] = ds[1] - aux * prefa1;
                        dr[1] = dr[1] - aux * prefa2;
                    }
                    if (ci < (one - RSMALL) && ci > RSMALL) {
                        const double aux = cisign * ai * bi;
                        ds[2] = ds[2] - aux * prefa1;
                        dr[2] = dr[2] - aux * prefa2;
                    }
                }
            }
        }
    }
    srfe = aux1;
    reff = aux2;
    return true;
}

struct V3 { double x, y, z; };
__device__ __forceinline__ V3 ld3(const StepParams &P, int i) {
    const double4 q = P.pos[i];
    return {q.x, q.y, q.z};
}
__device__ __forceinline__ V3 sub3(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross3(V3 u, V3 v) {
    return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}
__device__ __forceinline__ V3 scl3(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }

// slots of the DNA/grid energy accumulator and the energy categories they feed
constexpr int ND = 10;
__device__ constexpr int ND_CAT[ND] = {E_BOND, E_ANG, E_DIHE, E_STACK, E_BP, E_EX, E_QQ, E_SOLV,
                                       E_GSBPA, E_VDWGD};
enum { D_BOND = 0, D_ANG, D_DIHE, D_STACK, D_BP, D_EX, D_QQ, D_SOLV, D_GSBPA, D_VDWGD };

// OUTBOX (grandmod.f90:74-97): outside the box, the sphere or the elliptic cylinder
// One form for the three shapes, without branches: q = ((x-cx) sx)^2 + ((y-cy) sy)^2 +
// ((z-cz) sz)^2 against thr -- the sphere's r2 > rsphe2 (sx = sy = sz = 1), the ellipse's
// > 1 (sx = iecx, sy = iecy, sz = 0), never in a box (thr = +inf) -- or outside the clamp
// box, which is open along the curved axes (all three for the sphere, x and y for the
// cylinder).  The products by 1 and 0 are exact, so each shape's arithmetic is the
// reference's.
__device__ __forceinline__ bool outbox(const StepParams &P, const V3 &r) {
#pragma clang fp contract(off)
    const double a = (r.x - P.cv.cx) * P.cv.sx, b = (r.y - P.cv.cy) * P.cv.sy;
    const double c = (r.z - P.cv.cz) * P.cv.sz;
    return a * a + b * b + c * c > P.cv.thr || r.x < P.lx2m || r.x > P.lx2p ||
           r.y < P.ly2m || r.y > P.ly2p || r.z < P.lz2m || r.z > P.lz2p;
}

// Debye screening of a P-P pair (energy.mono.f90:439-441): always with QDEBY; with the
// hybrid option only when both sites are outside the box
__device__ __forceinline__ bool qq_debye(const StepParams &P, const V3 &ra, const V3 &rb) {
    return (P.flags & QDEBY) || ((P.flags & QDEBYHYB) && outbox(P, ra) && outbox(P, rb));
}

// A two-site DNA term (bond, stacking, base pair, excluded volume, P-P, solvent) incident
// on site `me` in role `role` (0: site a, 1: site b), given v1 = r_b - r_a, the term's
// list value (equilibrium distance or sigma) and, for a base pair, whether site a is a C
// or G base: force on `me` (added to f) and the term's energy (added to e when role == 0).
// energy.mono.f90:205-232, 335-507.
__device__ inline void dna_pair(const StepParams &P, int kind, int role, V3 v1, double par,
                                bool wide_bp, bool deby, bool free_me, V3 &f, double *e) {
    const double eps = P.epsnuc;
    if (kind == T_BOND) {
        double dist = sqrt(dot3(v1, v1));
        double vard = dist - par;
        double vard2 = vard * vard;
        if (role == 0) e[D_BOND] += eps * vard2 * (1.0 + 100.0 * vard2);
        if (free_me) {
            double de = eps * vard * (2.0 + 400.0 * vard2) / dist;
            V3 f1 = scl3(de, v1);
            double s = role == 0 ? 1.0 : -1.0;
            f.x += s * f1.x; f.y += s * f1.y; f.z += s * f1.z;
        }
        return;
    }
    double de = 0.0;
    bool force = free_me;
    if (kind == T_STACK) {
        double dist2 = dot3(v1, v1);
        double idist2 = 1.0 / dist2;
        double sg = par;
        double cte1 = pow3(sg * sg * idist2);
        if (role == 0) e[D_STACK] += 4.0 * eps * cte1 * (cte1 - 1.0);
        de = 24.0 * eps * cte1 * (2.0 * cte1 - 1.0) * idist2;
    } else if (kind == T_BP) {
        double epsln = wide_bp ? 2.532 * eps * P.scalepairing : 2.0 * eps * P.scalepairing;
        double dist2 = dot3(v1, v1);
        double idist2 = 1.0 / dist2;
        double sg = par;
        double cte1 = sg * sg * idist2;
        if (role == 0) e[D_BP] += epsln * pow5(cte1) * (20.0 * cte1 - 24.0);
        de = 240.0 * epsln * pow5(cte1) * (cte1 - 1.0) * idist2;
    } else if (kind == T_EX) {
        double dist2 = dot3(v1, v1);
        double idist2 = 1.0 / dist2;
        double sgex2 = par * par;
        if (dist2 < sgex2) {
            double cte1 = pow3(sgex2 * idist2);
            if (role == 0) e[D_EX] += 4.0 * eps * cte1 * (cte1 - 1.0) + eps;
            de = 24.0 * eps * cte1 * (2.0 * cte1 - 1.0) * idist2;
        } else {
            force = false;
        }
    } else if (kind == T_QQ) {
        double dist = sqrt(dot3(v1, v1));
        double idist = 1.0 / dist;
        if (deby) {
            double idistkd = exp(-dist * P.ikappa);
            if (role == 0) e[D_QQ] += P.fctn * idist * idistkd;
            de = P.fctn * (dist + P.kappa) * pow3(idist) * P.ikappa * idistkd;
        } else {
            if (role == 0) e[D_QQ] += P.fctn * idist;
            de = P.fctn * pow3(idist);
        }
    } else {  // T_SLV
        double dist = sqrt(dot3(v1, v1));
        double cte1 = exp((13.38 - dist) * 0.1875);
        if (role == 0) e[D_SOLV] += P.epsolv * ((1.0 - cte1) * (1.0 - cte1)) - P.epsolv;
        de = 0.375 * P.epsolv * cte1 * (cte1 - 1.0) / dist;
    }
    if (force) {
        double s = role == 0 ? -de : de;
        f.x += s * v1.x; f.y += s * v1.y; f.z += s * v1.z;
    }
}

// An angle or dihedral term incident on site `me` (energy.mono.f90:240-334), from its
// resolved record: x = word (kind | role | list index) | first other site << 32, y = the
// second and third other sites (32 bits each); the term's sites in list order are the
// others with `me` at position `role`.  Force on `me` (added to f) and the term's energy
// (added to e when role == 0 and EN).  The angle's sin(theta) is sqrt((1 - c)(1 + c)) of
// c = cos(theta); the dihedral's sin(phi - phi0) for the force comes from cos and sin of phi
// (the two arguments of the reference's atan2 over their norm) and the stored cos and sin
// of phi0 by the angle-difference identity, so a pass without energies calls acos alone.
// The dihedral energy is formed as the reference forms it, 1 - cos(atan2(..) - phi0): near
// phi0 that difference is rounding, which any other form would not reproduce.
template <bool EN>
__device__ inline void dna_multi(const StepParams &P, double2 rec, int me, bool free_me, V3 &f,
                                 double *e) {
    const uint64_t bx = (uint64_t)__double_as_longlong(rec.x);
    const uint64_t by = (uint64_t)__double_as_longlong(rec.y);
    const uint32_t w = (uint32_t)bx;
    const int kind = w >> 29, role = (w >> 27) & 3, t = w & 0x7FFFFFF;
    const int o[3] = {(int)(bx >> 32), (int)(uint32_t)by, (int)(by >> 32)};
    int sl[4];
    for (int r = 0, q = 0; r < 4; r++) sl[r] = r == role ? me : o[q++];
    const double eps = P.epsnuc;
    if (kind == T_ANGLE) {
        const V3 rb = ld3(P, sl[1]);
        const V3 v1 = sub3(ld3(P, sl[0]), rb), v2 = sub3(ld3(P, sl[2]), rb);
        const double ang0 = P.lval[1][t];
        const double m1 = dot3(v1, v1), m2 = dot3(v2, v2), m3 = dot3(v1, v2);
        const double modval = 1.0 / sqrt(m1 * m2);
        const double c = m3 * modval;
        const double varang = acos(c) - ang0;
        if (EN && role == 0) e[D_ANG] += 700.0 * eps * (varang * varang);
        if (free_me) {
            const double de = 1400.0 * eps * varang * modval / sqrt((1.0 - c) * (1.0 + c));
            const V3 f1 = scl3(de, v1), f2 = scl3(de, v2);
            if (role == 0) { f.x += f2.x; f.y += f2.y; f.z += f2.z; }
            else if (role == 1) { f.x -= f1.x + f2.x; f.y -= f1.y + f2.y; f.z -= f1.z + f2.z; }
            else { f.x += f1.x; f.y += f1.y; f.z += f1.z; }
        }
    } else {
        const V3 r2 = ld3(P, sl[1]), r3 = ld3(P, sl[2]);
        const V3 v1 = sub3(ld3(P, sl[0]), r2), v2 = sub3(r3, r2), v3 = sub3(r3, ld3(P, sl[3]));
        const double2 cs0 = P.dihcs[t];   // cos and sin of the list's phi0
        const V3 n1 = cross3(v1, v2), n2 = cross3(v2, v3);
        const double v22 = dot3(v2, v2);
        const double v2v = sqrt(v22);
        const double yv = v2v * dot3(v1, n2), xv = dot3(n1, n2);
        const double hv = sqrt(xv * xv + yv * yv);
        const double cphi = hv > 0.0 ? xv / hv : 1.0, sphi = hv > 0.0 ? yv / hv : 0.0;
        const double svar = sphi * cs0.x - cphi * cs0.y;   // sin(phi - phi0)
        if constexpr (EN) {
            if (role == 0) e[D_DIHE] += 28.0 * eps * (1.0 - cos(atan2(yv, xv) - P.lval[2][t]));
        }
        if (free_me) {
            const double n1v = dot3(n1, n1), n2v = dot3(n2, n2);
            double v12 = dot3(v1, v2), v23 = dot3(v2, v3);
            const double de = 28.0 * eps * svar * v2v;
            const V3 f1 = scl3(-de / n1v, n1), f4 = scl3(de / n2v, n2);
            const double iv22 = 1.0 / v22;
            v12 = v12 * iv22;
            v23 = v23 * iv22;
            V3 g;
            if (role == 0) g = f1;
            else if (role == 3) g = f4;
            else if (role == 1) g = {-f1.x + v12 * f1.x - v23 * f4.x, -f1.y + v12 * f1.y - v23 * f4.y,
                                     -f1.z + v12 * f1.z - v23 * f4.z};
            else g = {-f4.x - v12 * f1.x + v23 * f4.x, -f4.y - v12 * f1.y + v23 * f4.y,
                      -f4.z - v12 * f1.z + v23 * f4.z};
            f.x += g.x; f.y += g.y; f.z += g.z;
        }
    }
}

// 1/sqrt(x2) to within an ulp: hardware estimate plus two Newton steps.  The wall and
// tail forms of the table lookup and the force factor derive 1/r and 1/r^2 from it.
__device__ __forceinline__ double rsqrt_nr(double x2) {
    double r = __builtin_amdgcn_rsq(x2);
    double e = fma(-x2 * r, r, 1.0);
    r = fma(r * e, fma(0.375, e, 0.5), r);
    e = fma(-x2 * r, r, 1.0);
    r = fma(r * e, 0.5, r);
    return r;
}

// 1/sqrt(x2) from the hardware estimate and one third-order step (the series of
// (1 - e)^(-1/2) to e^2): the step leaves about (5/16) e^3 of the estimate's relative error
// e, so the result is within a few ulp; for force factors, not for the table index.
__device__ __forceinline__ double rsqrt_c3(double x2) {
    const double r = __builtin_amdgcn_rsq(x2);
    const double e = fma(-x2 * r, r, 1.0);
    return fma(r * e, fma(0.375, e, 0.5), r);
}

// Gaussians of particle i for this launch's step: injected rows are addressed by the
// step counter, Philox rows were generated for the current chunk of steps by
// k_philox_fill (row = step_off).
__device__ __forceinline__ const double *noise_row(const StepParams &P, int i) {
    const int64_t row = P.noise_mode == 0 ? P.step_ctr[1] + P.step_ctr[0] + P.step_off
                                          : P.step_off;
    return P.noise + row * P.noise_stride + 3 * (int64_t)i;
}

// Table pair (getyd, polynom.f90:197-230) in tail-first form: every lane forms the
// analytic tail sc3/r^6 + fct/r from rinv = 1/sqrt(d2); only lanes inside the table range
// (d2 <= dm2hi) enter the branch that replaces it with the wall or the tabulated quadratic,
// so far partners cost ~25 operations and near ones pay for the knot gather alone.
// tl = (sc3, tail fct, dm2hi, unused) of the pair type.  C3: 1/r from rsqrt_c3.
template <bool C3 = false>
__device__ __forceinline__ double pair_near_far(const DevPair &pp, const double4 &tl,
                                               const double4 *__restrict__ ep, double ires,
                                               double d2, double &y) {
    const double rinv = C3 ? rsqrt_c3(d2) : rsqrt_nr(d2);
    const double ix2 = rinv * rinv;
    const double c = tl.x * pow3(ix2);
    const double f = tl.y * rinv;
    y = c + f;
    double de = (6.0 * c + f) * ix2;
    if (d2 <= tl.z) {
        if (d2 < pp.dm2lo) {
            const double cw = pp.sc1 * pow6(ix2);
            y = cw + pp.sc2;
            de = 12.0 * cw * ix2;
        } else {
            // the knot index from the correctly rounded root, as getyd forms it
            // (polynom.f90:225-226): 1/sqrt is within an ulp only, and one ulp at a knot
            // selects the neighbouring segment, where the table jumps by ~1e-9
            const double x = sqrt(d2);
            int k = (int)((x - pp.dmi) * ires) + pp.nxi;
            k = min(max(k, pp.nxi), pp.nlast);
            const double4 e = ep[k];
            y = e.x + x * (e.y + x * e.z);
            de = -e.y * rinv - 2.0 * e.z;
        }
    }
    return de;
}

// analytic pair energy as INTERACT writes it (interact.f90:182-204,365-383): the SRPMF
// smoothing reads e*exp(...)-srpy (ENERGY has e*(exp(...)-srpy))
__device__ __forceinline__ double interact_analytic(const DevPair &p, const StepParams &P,
                                                    double dist2) {
    double e = 0.0, idist = 0.0, idist2 = 1.0 / dist2;
    if (p.flags & (PF_CHR | PF_SRPMF)) idist = sqrt(idist2);
    if (p.flags & PF_CHR) e = e + p.fct * idist;
    if (p.flags & PF_LJ) {
        double dist6 = pow3(p.sgp2 * idist2);
        e = e + p.epp4 * dist6 * (dist6 - 1.0);
    }
    if ((p.flags & PF_SRPMF) && dist2 <= P.rth) {
        double dist = 1.0 / idist;
        double es = p.c0 * exp((p.c1 - dist) * p.c2) * cos(p.c3 * PI * (p.c1 - dist)) +
                    p.c4 * pow6(p.c1 * idist);
        if (dist >= P.srpx) es = es * exp(-P.srpk * (dist - P.srpx)) - P.srpy;
        e = e + es;
    }
    return e;
}

// gety, polynom.f90:172-194
__device__ __forceinline__ double gety(const DevPair &p, const double4 *__restrict__ ep,
                                       double ires, double x2) {
    if (x2 < p.dm2lo) {
        double ix2 = 1.0 / x2;
        return p.sc1 * pow6(ix2) + p.sc2;
    } else if (x2 > p.dm2hi) {
        double ix2 = 1.0 / x2;
        if (p.flags & PF_CHR) return p.sc3 * pow3(ix2) + p.fct * sqrt(ix2);
        return p.sc3 * pow3(ix2);
    }
    double x = sqrt(x2);
    int i = (int)((x - p.dmi) * ires) + p.nxi;
    double4 e = ep[i];
    return e.x + x * (e.y + x * e.z);
}


// STATICF0 (staticf0.f90:19-82) and VDWGD0TRLN (vdwgd0trln.f90:19-117) of one particle at
// (xj, yj, zj) with atnam2 type jt2: called by a whole wavefront, lanes 0..7 take the 8
// corners; the energy comes back on every lane.  charged: the particle feels the static
// field with charge chi (an ion, or a P site).  wall: the repulsion corners reach thold8,
// in which case the repulsion energy is 1e10.  vmap: the first node of the particle's
// repulsion map (vdw_map).
__device__ inline double interact_grids(const StepParams &P, const float *phix, bool charged,
                                        double chi, double xj, double yj, double zj, int jt2,
                                        bool &wall, size_t vmap = 0) {
    const int lane = threadIdx.x & (WAVE - 1);
    double es = 0.0, ev = 0.0, ph = 0.0;
    bool okv = false;
    if ((P.flags & QPHIX) && lane < 8) {
        const GridGeom &g = P.gx;
        const bool ok = charged && xj <= g.xbcen + g.tranx && xj >= g.xbcen - g.tranx &&
                        yj <= g.ybcen + g.trany && yj >= g.ybcen - g.trany &&
                        zj <= g.zbcen + g.tranz && zj >= g.zbcen - g.tranz;
        if (ok) {
            double gx_, gy_, gz_;
            int ix, iy, iz;
            grid_cell(g, xj, yj, zj, gx_, gy_, gz_, ix, iy, iz);
            int n1 = ix + (lane >> 2), n2 = iy + ((lane >> 1) & 1), n3 = iz + (lane & 1);
            double ai = 1.0 - fabs(gx_ - n1 * g.dcel) * g.idcel;
            double bi = 1.0 - fabs(gy_ - n2 * g.dcel) * g.idcel;
            double ci = 1.0 - fabs(gz_ - n3 * g.dcel) * g.idcel;
            double fi = ai * bi * ci;
            es = fi * chi * (double)phix[((size_t)n1 * g.ny + n2) * g.nz + n3] * P.celec;
        }
    }
    if ((P.flags & QPHIV) && (P.flags & QBSPLINE)) {
        // VDWGD0SPLN (vdwgd0spln.f90:19-130): a node counts where its weight is nonzero
        const GridGeom &g = P.gv;
        okv = xj <= g.xbcen + g.tranx && xj >= g.xbcen - g.tranx && yj <= g.ybcen + g.trany &&
              yj >= g.ybcen - g.trany && zj <= P.vzmax && zj >= P.vzmin;
        if (okv && lane < 27) {
            double xi, yi, zi;
            int ix, iy, iz;
            spline_cell(g, xj, yj, zj, xi, yi, zi, ix, iy, iz);
            double esvdw = P.svdw;
            if (P.flags & QSVDW) esvdw = esvdw * P.scal[jt2];
            const int k = ix - 1 + lane / 9, l = iy - 1 + (lane / 3) % 3, m = iz - 1 + lane % 3;
            double ai, bi, ci, dd;
            bspline_m3(1.5 - (xi - k * g.dcel) * g.idcel, ai, dd);
            bspline_m3(1.5 - (yi - l * g.dcel) * g.idcel, bi, dd);
            bspline_m3(1.5 - (zi - m * g.dcel) * g.idcel, ci, dd);
            const double fi = ai * bi * ci;
            if (ai != 0.0 && bi != 0.0 && fi != 0.0) {
                ph = (double)(P.phiv + vmap)[((size_t)k * g.ny + l) * g.nz + m];
                ev = fi * esvdw * ph;
            }
        }
    } else if (P.flags & QPHIV) {
        const GridGeom &g = P.gv;
        okv = xj <= g.xbcen + g.tranx && xj >= g.xbcen - g.tranx && yj <= g.ybcen + g.trany &&
              yj >= g.ybcen - g.trany && zj <= P.vzmax && zj >= P.vzmin;
        if (okv && lane < 8) {
            double gx_, gy_, gz_;
            int ix, iy, iz;
            grid_cell(g, xj, yj, zj, gx_, gy_, gz_, ix, iy, iz);
            double esvdw = P.svdw;
            if (P.flags & QSVDW) esvdw = esvdw * P.scal[jt2];
            int n1 = ix + (lane >> 2), n2 = iy + ((lane >> 1) & 1), n3 = iz + (lane & 1);
            double ai = 1.0 - fabs(gx_ - n1 * g.dcel) * g.idcel;
            double bi = 1.0 - fabs(gy_ - n2 * g.dcel) * g.idcel;
            double ci = 1.0 - fabs(gz_ - n3 * g.dcel) * g.idcel;
            double fi = ai * bi * ci;
            ph = (double)(P.phiv + vmap)[((size_t)n1 * g.ny + n2) * g.nz + n3];
            ev = fi * esvdw * ph;
        }
    }
    es = wave_sum(es);
    ev = wave_sum(ev);
    ph = wave_sum(ph);
    double e = es;
    wall = false;
    if (okv) {
        if (ph >= ((P.flags & QBSPLINE) ? P.thold27 : P.thold8)) {
            ev = 1.0e10;
            wall = true;
        }
        e += ev;
    }
    return e;
}
}  // namespace bromoc
