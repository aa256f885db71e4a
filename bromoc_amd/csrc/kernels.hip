// kernels.hip -- gfx950 kernels of the BROMOC force / Brownian-dynamics step.
//
// k_step: one wavefront per target particle (every DNA site and every moving ion).
//   The block's four wavefronts stream the whole particle set through LDS in tiles of 256;
//   each lane evaluates its share of the tile against the wavefront's target (all-pairs, as
//   the reference's ENERGY has no cutoff: energy.mono.f90:117-194,584-687), then walks the
//   target's DNA-term incidence list (energy.mono.f90:196-508) and, on lanes 0..7, the
//   8 trilinear corners of the static-field and repulsion grids (staticf1.f90:45-118,
//   vdwgd1trln.f90:50-161).  A butterfly reduction gives the total force on lane 0, which
//   applies the Ermak-McCammon update of DYNAMICS0 / DYNAMICS0NUC (dynamics0.mono.f90:31-46,
//   dynamics0nuc.f90:31-44) into the other position buffer, so one launch is one BD step.
//   Every force is a gather over partners in a fixed order: results are deterministic and
//   need no atomics; energies are per-particle partials reduced in a fixed tree.
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

// host-side flag values (include/bromoc_gpu.h)
constexpr uint32_t QPAR = 1u << 0, QNUCL = 1u << 1, QNONBOND = 1u << 2, QNOBOND = 1u << 3;
constexpr uint32_t QPHIX = 1u << 4, QPHIV = 1u << 5, QSVDW = 1u << 6, QSOLV = 1u << 7;
constexpr uint32_t QDEBY = 1u << 8, QCOUNTION = 1u << 9, QDNAFREE = 1u << 10;

constexpr double PI = 3.14159265358979323846264338327950288419716939937510;
constexpr double RSMALL = 1E-3;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// integer powers in the order the Fortran compiler expands them (binary powering)
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
// DYNAMICS0/1/2 for ions (dynamics0.mono.f90:31-46, dynamics1.mono.f90:71-86,
// dynamics2.mono.f90:75-85) with countions, DYNAMICS0NUC (dynamics0nuc.f90:31-44) for
// sites; fixcoor unless the site is free of the box (dnafree).
__device__ __forceinline__ double4 bd_update(const StepParams &P, int i, bool is_site,
                                             const double4 &pa, const double F[3],
                                             const double *g) {
    double fac1, fac2, ddz = 0.0;
    int itype = 0;
    if (is_site) {
        fac1 = P.fact0n1;
        fac2 = P.fact0n2;
    } else {
        itype = abs(P.typei[i]);
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
    if (!is_site && P.dmode != 0) zn = zn + ddz;
    if (!is_site && (P.flags & QCOUNTION)) {
        const int row = itype - P.nold - 1;
        for (int c = 0; c < P.cntpts; c++) {
            const double zc = P.zcont[c];
            if (pa.z < zc && zn >= zc) atomicAdd(&P.nfwd[row * P.cntpts + c], 1);
            else if (pa.z >= zc && zn < zc) atomicAdd(&P.nbwd[row * P.cntpts + c], 1);
        }
    }
    if (!is_site || !(P.flags & QDNAFREE)) fixcoor(P, xn, yn, zn);
    return make_double4(xn, yn, zn, pa.w);
}

// Reaction-field parameters of an ion at (x, y, z) of atnam type itype: srfe and reff
// from the 8 corners of gsrfen / sqrfac*greff (rfpar.f90:75-142), with their gradients
// when grad is set.  The corner loop leaves the innermost (z) loop at a zero gsrfen corner,
// as the reference's `exit` does (rfpar.f90:121).  Returns false outside the grid box.
__device__ bool rf_params(const StepParams &P, double x, double y, double z, int itype,
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

// One DNA term incident on site `me` in role `role`: force on `me` (added to f) and the
// term's energy (added to e when role == 0).  energy.mono.f90:205-507.
__device__ void dna_term(const StepParams &P, uint32_t w, bool free_me, V3 &f, double *e) {
    const int kind = w >> 29, role = (w >> 27) & 3, t = w & 0x7FFFFFF;
    const double eps = P.epsnuc;
    const int n = P.nlist[kind];
    const int32_t *L = P.lst[kind];
    if (kind == T_BOND) {
        int a = L[t], b = L[n + t];
        V3 v1 = sub3(ld3(P, b), ld3(P, a));
        double dist = sqrt(dot3(v1, v1));
        double vard = dist - P.lval[0][t];
        double vard2 = vard * vard;
        if (role == 0) e[D_BOND] += eps * vard2 * (1.0 + 100.0 * vard2);
        if (free_me) {
            double de = eps * vard * (2.0 + 400.0 * vard2) / dist;
            V3 f1 = scl3(de, v1);
            double s = role == 0 ? 1.0 : -1.0;
            f.x += s * f1.x; f.y += s * f1.y; f.z += s * f1.z;
        }
    } else if (kind == T_ANGLE) {
        int a = L[t], b = L[n + t], d = L[2 * n + t];
        V3 rb = ld3(P, b);
        V3 v1 = sub3(ld3(P, a), rb), v2 = sub3(ld3(P, d), rb);
        double m1 = dot3(v1, v1), m2 = dot3(v2, v2), m3 = dot3(v1, v2);
        double modval = 1.0 / sqrt(m1 * m2);
        double ang = acos(m3 * modval);
        double varang = ang - P.lval[1][t];
        if (role == 0) e[D_ANG] += 700.0 * eps * (varang * varang);
        if (free_me) {
            double de = 1400.0 * eps * varang * modval / sin(ang);
            V3 f1 = scl3(de, v1), f2 = scl3(de, v2);
            if (role == 0) { f.x += f2.x; f.y += f2.y; f.z += f2.z; }
            else if (role == 1) { f.x -= f1.x + f2.x; f.y -= f1.y + f2.y; f.z -= f1.z + f2.z; }
            else { f.x += f1.x; f.y += f1.y; f.z += f1.z; }
        }
    } else if (kind == T_DIHE) {
        int i1 = L[t], i2 = L[n + t], i3 = L[2 * n + t], i4 = L[3 * n + t];
        V3 r2 = ld3(P, i2), r3 = ld3(P, i3);
        V3 v1 = sub3(ld3(P, i1), r2), v2 = sub3(r3, r2), v3 = sub3(r3, ld3(P, i4));
        V3 n1 = cross3(v1, v2), n2 = cross3(v2, v3);
        double v22 = dot3(v2, v2);
        double v2v = sqrt(v22);
        double av = dot3(v1, n2), bv = dot3(n1, n2);
        double dihe = atan2(v2v * av, bv);
        double vardihe = dihe - P.lval[2][t];
        if (role == 0) e[D_DIHE] += 28.0 * eps * (1.0 - cos(vardihe));
        if (free_me) {
            double n1v = dot3(n1, n1), n2v = dot3(n2, n2);
            double v12 = dot3(v1, v2), v23 = dot3(v2, v3);
            double de = 28.0 * eps * sin(vardihe) * v2v;
            V3 f1 = scl3(-de / n1v, n1), f4 = scl3(de / n2v, n2);
            double iv22 = 1.0 / v22;
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
    } else {
        int a = L[t], b = L[n + t];
        V3 v1 = sub3(ld3(P, b), ld3(P, a));
        double de = 0.0;
        bool force = free_me;
        if (kind == T_STACK) {
            double dist2 = dot3(v1, v1);
            double idist2 = 1.0 / dist2;
            double sg = P.lval[3][t];
            double cte1 = pow3(sg * sg * idist2);
            if (role == 0) e[D_STACK] += 4.0 * eps * cte1 * (cte1 - 1.0);
            de = 24.0 * eps * cte1 * (2.0 * cte1 - 1.0) * idist2;
        } else if (kind == T_BP) {
            int na = P.namsite[a];
            double epsln = (na == 5 || na == 4) ? 2.532 * eps * P.scalepairing
                                                : 2.0 * eps * P.scalepairing;
            double dist2 = dot3(v1, v1);
            double idist2 = 1.0 / dist2;
            double sg = P.lval[4][t];
            double cte1 = sg * sg * idist2;
            if (role == 0) e[D_BP] += epsln * pow5(cte1) * (20.0 * cte1 - 24.0);
            de = 240.0 * epsln * pow5(cte1) * (cte1 - 1.0) * idist2;
        } else if (kind == T_EX) {
            double dist2 = dot3(v1, v1);
            double idist2 = 1.0 / dist2;
            double sgex2 = P.lval[5][t] * P.lval[5][t];
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
            if (P.flags & QDEBY) {
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
}

// 1/sqrt(x2) to within an ulp: hardware estimate plus two Newton steps.  Every branch
// of the table lookup derives r, 1/r and 1/r^2 from it.
__device__ __forceinline__ double rsqrt_nr(double x2) {
    double r = __builtin_amdgcn_rsq(x2);
    double e = fma(-x2 * r, r, 1.0);
    r = fma(r * e, fma(0.375, e, 0.5), r);
    e = fma(-x2 * r, r, 1.0);
    r = fma(r * e, 0.5, r);
    return r;
}

// Table pair (getyd, polynom.f90:197-230) from x2 and rinv = 1/sqrt(x2): wall below dmi,
// analytic tail above the last knot, piecewise quadratic in between.
__device__ __forceinline__ void efpot_eval(const DevPair &p, const double4 *__restrict__ ep,
                                           double ires, double x2, double rinv, double &y,
                                           double &yp) {
    const double ix2 = rinv * rinv;
    if (x2 < p.dm2lo) {
        const double c = p.sc1 * pow6(ix2);
        y = c + p.sc2;
        yp = 12.0 * c * ix2;
    } else if (x2 > p.dm2hi) {
        const double c = p.sc3 * pow3(ix2);
        const double f = p.fct * rinv;     // fct is zero for uncharged pairs
        y = c + f;
        yp = (6.0 * c + f) * ix2;
    } else {
        const double x = x2 * rinv;
        int i = (int)((x - p.dmi) * ires) + p.nxi;
        i = min(max(i, p.nxi), p.nlast);
        const double4 e = ep[i];
        y = e.x + x * (e.y + x * e.z);
        yp = -e.y * rinv - 2.0 * e.z;
    }
}

// Gaussians of particle i for this launch's step: injected rows are addressed by the
// step counter, Philox rows were generated for the current chunk of steps by
// k_philox_fill (row = step_off).
__device__ __forceinline__ const double *noise_row(const StepParams &P, int i) {
    const int64_t row = P.noise_mode == 0 ? P.step_ctr[1] + P.step_ctr[0] + P.step_off
                                          : P.step_off;
    return P.noise + row * P.noise_stride + 3 * (int64_t)i;
}

// One partner j of a target: the table pair of getyd (polynom.f90:197-230) evaluated
// without branches.  Wall, tail and table forms are all formed from rinv = 1/sqrt(d2);
// the knot index is clamped into the pair's table, so the gather is always in bounds and
// is simply not selected outside [dmi, r_n].  Returns the force factor; adds the energy.
__device__ __forceinline__ double pair_table(const DevPair &pp, const double4 *__restrict__ ep,
                                            double ires, double d2, double &y) {
    const double rinv = rsqrt_nr(d2);
    const double ix2 = rinv * rinv;
    const double x = d2 * rinv;
    int k = (int)((x - pp.dmi) * ires) + pp.nxi;
    k = min(max(k, pp.nxi), pp.nlast);
    const double4 e = ep[k];
    const double c3 = pp.sc3 * pow3(ix2);
    const double f = pp.fct * rinv;              // zero for uncharged pairs
    const double c12 = pp.sc1 * pow6(ix2);
    const bool wall = d2 < pp.dm2lo;
    const bool tail = d2 > pp.dm2hi;
    const double yt = e.x + x * (e.y + x * e.z);
    const double ypt = -e.y * rinv - 2.0 * e.z;
    y = wall ? c12 + pp.sc2 : (tail ? c3 + f : yt);
    return wall ? 12.0 * c12 * ix2 : (tail ? (6.0 * c3 + f) * ix2 : ypt);
}

// Table pair (getyd, polynom.f90:197-230) in tail-first form: every lane forms the
// analytic tail sc3/r^6 + fct/r from rinv = 1/sqrt(d2); only lanes inside the table range
// (d2 <= dm2hi) enter the branch that replaces it with the wall or the tabulated quadratic,
// so far partners cost ~25 operations and near ones pay for the knot gather alone.
// tl = (sc3, tail fct, dm2hi, unused) of the pair type.
__device__ __forceinline__ double pair_near_far(const DevPair &pp, const double4 &tl,
                                               const double4 *__restrict__ ep, double ires,
                                               double d2, double &y) {
    const double rinv = rsqrt_nr(d2);
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
            const double x = d2 * rinv;
            int k = (int)((x - pp.dmi) * ires) + pp.nxi;
            k = min(max(k, pp.nxi), pp.nlast);
            const double4 e = ep[k];
            y = e.x + x * (e.y + x * e.z);
            de = -e.y * rinv - 2.0 * e.z;
        }
    }
    return de;
}

// k_step: T target particles per block of 256 threads (targets tbeg.. of the launch).
// Every thread walks the partner set with stride 256, two partners per iteration, and
// evaluates each partner against all T targets.  Per-target sums: butterfly within each
// wavefront, then a fixed-order sum over the 4 wavefronts in LDS.  Wavefront w < T also
// handles target w's 8 grid corners; thread a < T finishes target a (force, energies,
// BD update).  CG-DNA list forces of site targets come from k_dna (fdna).
template <int FEAT, int T>
__global__ __launch_bounds__(BLOCK) void k_step(StepParams P) {
    __shared__ DevPair s_row[T][MAX_LDS_TYPES];   // pair parameters by partner type
    __shared__ double4 s_tail[T][MAX_LDS_TYPES];  // (sc3, tail fct, dm2hi, 0) by partner type
    __shared__ double s_red[WPB][T][7];
    __shared__ double s_dg[T][5];
    __shared__ double s_g[T][3];
    __shared__ int s_warn[T];

    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid >> 6;
    const int ns = P.nsites, nf = P.nfix, ntot = P.ntot;
    const int nt1 = P.nttyp + 1;
    const int wbase = P.tbeg + blockIdx.x * T;

    int ti[T];
    bool tval[T], tsite[T], tfree[T];
    double tx[T], ty[T], tz[T];
#pragma unroll
    for (int a = 0; a < T; a++) {
        const int w = wbase + a;
        tval[a] = w < P.tend;
        const int i = !tval[a] ? 0 : (w < ns ? w : w + nf);
        ti[a] = i;
        tsite[a] = i < ns;
        const double4 pa = P.pos[i];
        tx[a] = pa.x; ty[a] = pa.y; tz[a] = pa.z;
        tfree[a] = !tsite[a] || ((FEAT & F_DNA) && P.stfree[i] != 0);
        // pair rows of this block's targets
        const int row = __double2loint(pa.w) * nt1;
        for (int k = tid; k < nt1; k += BLOCK) {
            const DevPair d = P.rows[row + k];
            s_row[a][k] = d;
            s_tail[a][k] = make_double4(d.sc3, d.fct, d.dm2hi, 0.0);
        }
    }
    __syncthreads();

    // Gaussians of the BD update, fetched by the last wavefront (lanes 0..T-1) before its
    // share of the pair loop
    if constexpr ((FEAT & F_UPDATE) != 0) {
        const int a = tid - (BLOCK - WAVE);
        if (a >= 0 && a < T && tval[a]) {
            const int i = ti[a];
            const double *nr = noise_row(P, i);
            s_g[a][0] = nr[0]; s_g[a][1] = nr[1]; s_g[a][2] = nr[2];
        }
    }

    double fx[T], fy[T], fz[T], et[T], eel[T], evd[T], esr[T];
#pragma unroll
    for (int a = 0; a < T; a++) fx[a] = fy[a] = fz[a] = et[a] = eel[a] = evd[a] = esr[a] = 0.0;

    // ---- pair loop (energy.mono.f90:117-194, 584-687)
    const bool nonbond = (P.flags & QNONBOND) && (P.flags & QPAR);
    if (nonbond) {
        for (int k0 = tid; k0 < ntot; k0 += 2 * BLOCK) {
            // partners in spatial order, two per iteration (two load chains in flight)
            const int k1 = k0 + BLOCK;
            const bool has1 = k1 < ntot;
            const double4 p0 = P.spos[k0];
            const double4 p1 = P.spos[has1 ? k1 : k0];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const double4 pj = u == 0 ? p0 : p1;
                const bool jok = u == 0 || has1;
                const int t2j = __double2loint(pj.w);
                const int j = __double2hiint(pj.w);
#pragma unroll
                for (int a = 0; a < T; a++) {
                    const bool skip = !jok || !tval[a] || j == ti[a] || (tsite[a] && j < ns);
                    const double dx = tx[a] - pj.x, dy = ty[a] - pj.y, dz = tz[a] - pj.z;
                    double d2 = dx * dx + dy * dy + dz * dz;
                    d2 = skip ? 1e30 : d2;
                    // a pair's energy is counted once: by the site of a site-ion pair, by
                    // the higher-index ion of an ion-ion pair
                    const bool ce = !skip && (tsite[a] || (j >= ns && j < ti[a]));
                    double y, de;
                    if (P.all_efpot || (s_row[a][t2j].flags & PF_EFPOT)) {
                        de = pair_near_far(s_row[a][t2j], s_tail[a][t2j], P.ep, P.ires, d2, y);
                        et[a] += ce ? y : 0.0;
                    } else {
                        double q0 = 0.0, q1 = 0.0, q2 = 0.0;
                        de = pair_analytic(s_row[a][t2j], P, d2, q0, q1, q2);
                        eel[a] += ce ? q0 : 0.0;
                        evd[a] += ce ? q1 : 0.0;
                        esr[a] += ce ? q2 : 0.0;
                    }
                    de = (skip || !tfree[a]) ? 0.0 : de;
                    fx[a] += de * dx;
                    fy[a] += de * dy;
                    fz[a] += de * dz;
                }
            }
        }
    }

    // pair sums to LDS (the pair accumulators are dead after this)
#pragma unroll
    for (int a = 0; a < T; a++) {
        double v[7] = {fx[a], fy[a], fz[a], et[a], eel[a], evd[a], esr[a]};
#pragma unroll
        for (int k = 0; k < 7; k++) {
            if (k >= 3 && !P.write_energy) break;
            v[k] = wave_sum(v[k]);
            if (lane == 0) s_red[wv][a][k] = v[k];
        }
    }

    // ---- grids of target `wv` (if any), lanes 0..7 of wavefront wv
    double gfx = 0.0, gfy = 0.0, gfz = 0.0, egx = 0.0, egv = 0.0;
    bool warned = false;
    const bool mine = wv < T;
    int i_me = 0;
    bool val_me = false, site_me = false;
    double xi = 0.0, yi = 0.0, zi = 0.0;
    int t2i = 0;
    if (mine) {
        const int w = wbase + wv;
        val_me = w < P.tend;
        i_me = !val_me ? 0 : (w < ns ? w : w + nf);
        site_me = i_me < ns;
        const double4 pa = P.pos[i_me];
        xi = pa.x; yi = pa.y; zi = pa.z;
        t2i = __double2loint(pa.w);
    }
    if constexpr ((FEAT & F_PHIX) != 0) {
        if (mine && val_me && lane < 8) {
            const GridGeom &g = P.gx;
            double chi;
            bool ok;
            if (site_me) { ok = P.namsite[i_me] == 1; chi = P.cgnuc; }
            else { ok = true; chi = P.cg[abs(P.typei[i_me])]; }
            ok = ok && xi <= g.xbcen + g.tranx && xi >= g.xbcen - g.tranx &&
                 yi <= g.ybcen + g.trany && yi >= g.ybcen - g.trany &&
                 zi <= g.zbcen + g.tranz && zi >= g.zbcen - g.tranz;
            if (ok) {
                double gx_, gy_, gz_;
                int ix, iy, iz;
                grid_cell(g, xi, yi, zi, gx_, gy_, gz_, ix, iy, iz);
                const double one1 = 1.0;
                const int n1 = ix + (lane >> 2), n2 = iy + ((lane >> 1) & 1), n3 = iz + (lane & 1);
                double ai = gx_ - n1 * g.dcel;
                const double aisign = fsign(one1, ai);
                ai = one1 - fabs(ai) * g.idcel;
                double bi = gy_ - n2 * g.dcel;
                const double bisign = fsign(one1, bi);
                bi = one1 - fabs(bi) * g.idcel;
                double ci = gz_ - n3 * g.dcel;
                const double cisign = fsign(one1, ci);
                ci = one1 - fabs(ci) * g.idcel;
                const double fi = ai * bi * ci;
                const double ph = (double)P.phix[((size_t)n1 * g.ny + n2) * g.nz + n3];
                const double prefac = ph * P.celec * chi * g.idcel;
                if (ai < (1.0 - RSMALL) && ai > RSMALL) gfx += aisign * bi * ci * prefac;
                if (bi < (1.0 - RSMALL) && bi > RSMALL) gfy += bisign * ai * ci * prefac;
                if (ci < (1.0 - RSMALL) && ci > RSMALL) gfz += cisign * ai * bi * prefac;
                egx += fi * chi * ph * P.celec;
            }
        }
    }
    if constexpr ((FEAT & F_PHIV) != 0) {
        const GridGeom &g = P.gv;
        const bool ok = mine && val_me && xi <= g.xbcen + g.tranx && xi >= g.xbcen - g.tranx &&
                        yi <= g.ybcen + g.trany && yi >= g.ybcen - g.trany && zi <= P.vzmax &&
                        zi >= P.vzmin;
        double ph = 0.0;
        if (ok && lane < 8) {
            double gx_, gy_, gz_;
            int ix, iy, iz;
            grid_cell(g, xi, yi, zi, gx_, gy_, gz_, ix, iy, iz);
            double esvdw = P.svdw;
            if (P.flags & QSVDW) esvdw = P.svdw * P.scal[t2i];
            ph = tri_corner(g, P.phiv, gx_, gy_, gz_, ix, iy, iz, lane, esvdw, gfx, gfy, gfz, egv);
        }
        double sum = ph;
        sum += __shfl_xor(sum, 1, WAVE);
        sum += __shfl_xor(sum, 2, WAVE);
        sum += __shfl_xor(sum, 4, WAVE);
        warned = ok && sum >= P.thold8;
    }
    if (mine) {
        gfx = wave_sum(gfx);
        gfy = wave_sum(gfy);
        gfz = wave_sum(gfz);
        if (P.write_energy) {
            egx = wave_sum(egx);
            egv = wave_sum(egv);
        }
        if (lane == 0) {
            s_dg[wv][0] = gfx; s_dg[wv][1] = gfy; s_dg[wv][2] = gfz;
            s_dg[wv][3] = egx; s_dg[wv][4] = egv;
            s_warn[wv] = warned ? 1 : 0;
        }
    }
    __syncthreads();
    if (tid >= T) return;

    // ---- thread a finishes target a
    const int a = tid;
    const int w = wbase + a;
    if (w >= P.tend) return;
    const int i = w < ns ? w : w + nf;
    const bool is_site = i < ns;
    double F[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < WPB; q++) acc += s_red[q][a][k];
        F[k] = acc + s_dg[a][k];
    }
    if constexpr ((FEAT & F_DNA) != 0) {
        if (is_site) {
            const double4 fd = P.fdna[i];   // CG-DNA list terms from k_dna
            F[0] += fd.x; F[1] += fd.y; F[2] += fd.z;
        }
    }
    if (P.rfon && !is_site) {   // reaction-field forces from k_rfpair
        const double4 fr = P.frf[i];
        F[0] += fr.x; F[1] += fr.y; F[2] += fr.z;
    }
    const double4 pa = P.pos[i];
    const int t2a = __double2loint(pa.w);
    if (s_warn[a]) atomicAdd(&P.warn[t2a], 1);
    if (P.write_energy) {
        double pe[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < WPB; q++) acc += s_red[q][a][3 + k];
            pe[k] = acc;
        }
        // this kernel owns the pair and grid categories; k_dna writes the DNA ones
        double *ep_ = P.epart + i;
        const size_t cs = P.ecap;
        ep_[E_GSBPA * cs] = s_dg[a][3];
        ep_[E_VDWGD * cs] = s_dg[a][4];
        if (is_site) {
            ep_[E_EFPOTMX * cs] = pe[0]; ep_[E_QQMX * cs] = pe[1];
            ep_[E_VDWMX * cs] = pe[2]; ep_[E_SRPMFMX * cs] = pe[3];
        } else {
            ep_[E_EFPOT * cs] = pe[0]; ep_[E_ELEC * cs] = pe[1];
            ep_[E_VDW * cs] = pe[2]; ep_[E_SRPMF * cs] = pe[3];
        }
    }
    if (P.write_forces) {
        P.fx[i] = F[0];
        P.fy[i] = F[1];
        P.fz[i] = F[2];
    }
    if constexpr ((FEAT & F_UPDATE) != 0) {
        const bool fr = !is_site || ((FEAT & F_DNA) && P.stfree[i] != 0);
        if (!fr) {
            P.poso[i] = pa;
            P.sposo[P.rank[i]] = pa;
            return;
        }
        const double g[3] = {s_g[a][0], s_g[a][1], s_g[a][2]};
        const double4 pn = bd_update(P, i, is_site, pa, F, g);
        P.poso[i] = pn;
        P.sposo[P.rank[i]] = pn;
    }
}

// k_dna: CG-DNA bonded and list terms (energy.mono.f90:196-508), one block per site.
// The block's threads take the site's term incidences in order (they are grouped by term
// kind, so wavefronts rarely diverge); sums go butterfly-then-LDS in a fixed order.  The
// force on the site goes to fdna[i] (zero for a fixed site); the energies of the terms the
// site heads go to epart.
__global__ __launch_bounds__(BLOCK) void k_dna(StepParams P) {
    __shared__ double s_r[WPB][3 + 8];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid >> 6;
    const int i = blockIdx.x;
    const bool free_me = P.stfree[i] != 0;
    V3 f = {0.0, 0.0, 0.0};
    double e[ND];
#pragma unroll
    for (int k = 0; k < ND; k++) e[k] = 0.0;
    const int k0 = P.inc_ptr[i], k1 = P.inc_ptr[i + 1];
    for (int k = k0 + tid; k < k1; k += BLOCK) {
        const uint32_t wd = P.inc[k];
        const int kind = wd >> 29;
        const bool on = kind <= T_DIHE ? (P.flags & QNOBOND) != 0
                                       : ((P.flags & QNONBOND) != 0 &&
                                          (kind != T_SLV || (P.flags & QSOLV) != 0));
        if (on) dna_term(P, wd, free_me, f, e);
    }
    f.x = wave_sum(f.x);
    f.y = wave_sum(f.y);
    f.z = wave_sum(f.z);
    if (P.write_energy) {
#pragma unroll
        for (int k = 0; k < 8; k++) e[k] = wave_sum(e[k]);
    }
    if (lane == 0) {
        s_r[wv][0] = f.x; s_r[wv][1] = f.y; s_r[wv][2] = f.z;
        for (int k = 0; k < 8; k++) s_r[wv][3 + k] = e[k];
    }
    __syncthreads();
    if (tid < 3 + 8) {
        double acc = 0.0;
        for (int q = 0; q < WPB; q++) acc += s_r[q][tid];
        if (tid < 3) {
            reinterpret_cast<double *>(&P.fdna[i])[tid] = acc;
        } else if (P.write_energy) {
            P.epart[(size_t)ND_CAT[tid - 3] * P.ecap + i] = acc;
        }
    }
}

// k_wave: one wavefront per target, four targets per block.  The block streams the
// partner array through LDS in tiles of 256, double-buffered (tile t+1 is fetched from
// global memory while tile t is evaluated); each lane evaluates four partners of a tile
// against its wavefront's target, groups of 64 taking the analytic-tail path when every
// lane is beyond the table range.  The wavefront's butterfly sum is the target's total;
// lane 0 finishes it as in k_step.
template <int FEAT>
__global__ __launch_bounds__(BLOCK) void k_wave(StepParams P) {
    __shared__ double4 s_part[2][BLOCK];
    __shared__ DevPair s_row[WPB][MAX_LDS_TYPES];
    __shared__ double4 s_tail[WPB][MAX_LDS_TYPES];

    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid >> 6;
    const int ns = P.nsites, nf = P.nfix, ntot = P.ntot;
    const int nt1 = P.nttyp + 1;
    const int w = P.tbeg + blockIdx.x * WPB + wv;
    const bool valid = w < P.tend;
    const int i = !valid ? 0 : (w < ns ? w : w + nf);
    const bool is_site = i < ns;
    const double4 pa = P.pos[i];
    const int t2i = __double2loint(pa.w);
    const bool free_me = !is_site || ((FEAT & F_DNA) && P.stfree[i] != 0);
    if (lane < nt1) {
        const DevPair d = P.rows[t2i * nt1 + lane];
        s_row[wv][lane] = d;
        s_tail[wv][lane] = make_double4(d.sc3, d.fct, d.dm2hi, 0.0);
    }
    double4 nxt = tid < ntot ? P.spos[tid] : make_double4(0.0, 0.0, 0.0, 0.0);
    s_part[0][tid] = nxt;
    __syncthreads();

    double fx = 0.0, fy = 0.0, fz = 0.0, et = 0.0, eel = 0.0, evd = 0.0, esr = 0.0;
    const bool nonbond = (P.flags & QNONBOND) && (P.flags & QPAR);
    const int ntiles = (ntot + BLOCK - 1) / BLOCK;
    for (int t = 0; t < ntiles; t++) {
        const int cur = t & 1;
        const int jn = (t + 1) * BLOCK + tid;
        if (t + 1 < ntiles) nxt = jn < ntot ? P.spos[jn] : make_double4(0.0, 0.0, 0.0, 0.0);
        const int kend = min(BLOCK, ntot - t * BLOCK);
        if (valid && nonbond) {
#pragma unroll
            for (int q = 0; q < WPB; q++) {
                const int k = q * WAVE + lane;
                const bool jok = k < kend;
                const double4 pj = s_part[cur][k];
                const int t2j = __double2loint(pj.w);
                const int j = __double2hiint(pj.w);
                const bool skip = !jok || j == i || (is_site && j < ns);
                const double dx = pa.x - pj.x, dy = pa.y - pj.y, dz = pa.z - pj.z;
                double d2 = dx * dx + dy * dy + dz * dz;
                d2 = skip ? 1e30 : d2;
                const bool ce = !skip && (is_site || (j >= ns && j < i));
                const int tj = skip ? 0 : t2j;
                double y, de;
                if (P.all_efpot || (s_row[wv][tj].flags & PF_EFPOT)) {
                    de = pair_near_far(s_row[wv][tj], s_tail[wv][tj], P.ep, P.ires, d2, y);
                    et += ce ? y : 0.0;
                } else {
                    double q0 = 0.0, q1 = 0.0, q2 = 0.0;
                    de = pair_analytic(s_row[wv][tj], P, d2, q0, q1, q2);
                    eel += ce ? q0 : 0.0;
                    evd += ce ? q1 : 0.0;
                    esr += ce ? q2 : 0.0;
                }
                de = (skip || !free_me) ? 0.0 : de;
                fx += de * dx;
                fy += de * dy;
                fz += de * dz;
            }
        }
        if (t + 1 < ntiles) s_part[cur ^ 1][tid] = nxt;
        __syncthreads();
    }
    if (!valid) return;

    // ---- grids (lanes 0..7)
    double egx = 0.0, egv = 0.0;
    bool warned = false;
    if constexpr ((FEAT & F_PHIX) != 0) {
        if (lane < 8) {
            const GridGeom &g = P.gx;
            double chi;
            bool ok;
            if (is_site) { ok = P.namsite[i] == 1; chi = P.cgnuc; }
            else { ok = true; chi = P.cg[abs(P.typei[i])]; }
            ok = ok && pa.x <= g.xbcen + g.tranx && pa.x >= g.xbcen - g.tranx &&
                 pa.y <= g.ybcen + g.trany && pa.y >= g.ybcen - g.trany &&
                 pa.z <= g.zbcen + g.tranz && pa.z >= g.zbcen - g.tranz;
            if (ok) {
                double gx_, gy_, gz_;
                int ix, iy, iz;
                grid_cell(g, pa.x, pa.y, pa.z, gx_, gy_, gz_, ix, iy, iz);
                const double one1 = 1.0;
                const int n1 = ix + (lane >> 2), n2 = iy + ((lane >> 1) & 1), n3 = iz + (lane & 1);
                double ai = gx_ - n1 * g.dcel;
                const double aisign = fsign(one1, ai);
                ai = one1 - fabs(ai) * g.idcel;
                double bi = gy_ - n2 * g.dcel;
                const double bisign = fsign(one1, bi);
                bi = one1 - fabs(bi) * g.idcel;
                double ci = gz_ - n3 * g.dcel;
                const double cisign = fsign(one1, ci);
                ci = one1 - fabs(ci) * g.idcel;
                const double fi = ai * bi * ci;
                const double ph = (double)P.phix[((size_t)n1 * g.ny + n2) * g.nz + n3];
                const double prefac = ph * P.celec * chi * g.idcel;
                if (ai < (1.0 - RSMALL) && ai > RSMALL) fx += aisign * bi * ci * prefac;
                if (bi < (1.0 - RSMALL) && bi > RSMALL) fy += bisign * ai * ci * prefac;
                if (ci < (1.0 - RSMALL) && ci > RSMALL) fz += cisign * ai * bi * prefac;
                egx += fi * chi * ph * P.celec;
            }
        }
    }
    if constexpr ((FEAT & F_PHIV) != 0) {
        const GridGeom &g = P.gv;
        const bool ok = pa.x <= g.xbcen + g.tranx && pa.x >= g.xbcen - g.tranx &&
                        pa.y <= g.ybcen + g.trany && pa.y >= g.ybcen - g.trany && pa.z <= P.vzmax &&
                        pa.z >= P.vzmin;
        double ph = 0.0;
        if (ok && lane < 8) {
            double gx_, gy_, gz_;
            int ix, iy, iz;
            grid_cell(g, pa.x, pa.y, pa.z, gx_, gy_, gz_, ix, iy, iz);
            double esvdw = P.svdw;
            if (P.flags & QSVDW) esvdw = P.svdw * P.scal[t2i];
            ph = tri_corner(g, P.phiv, gx_, gy_, gz_, ix, iy, iz, lane, esvdw, fx, fy, fz, egv);
        }
        double sum = ph;
        sum += __shfl_xor(sum, 1, WAVE);
        sum += __shfl_xor(sum, 2, WAVE);
        sum += __shfl_xor(sum, 4, WAVE);
        warned = ok && sum >= P.thold8;
    }
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    if (P.write_energy) {
        et = wave_sum(et);
        eel = wave_sum(eel);
        evd = wave_sum(evd);
        esr = wave_sum(esr);
        egx = wave_sum(egx);
        egv = wave_sum(egv);
    }
    if (lane != 0) return;
    double F[3] = {fx, fy, fz};
    if constexpr ((FEAT & F_DNA) != 0) {
        if (is_site) {
            const double4 fd = P.fdna[i];
            F[0] += fd.x; F[1] += fd.y; F[2] += fd.z;
        }
    }
    if (P.rfon && !is_site) {
        const double4 fr = P.frf[i];
        F[0] += fr.x; F[1] += fr.y; F[2] += fr.z;
    }
    if (warned) atomicAdd(&P.warn[t2i], 1);
    if (P.write_energy) {
        double *ep_ = P.epart + i;
        const size_t cs = P.ecap;
        ep_[E_GSBPA * cs] = egx;
        ep_[E_VDWGD * cs] = egv;
        if (is_site) {
            ep_[E_EFPOTMX * cs] = et; ep_[E_QQMX * cs] = eel;
            ep_[E_VDWMX * cs] = evd; ep_[E_SRPMFMX * cs] = esr;
        } else {
            ep_[E_EFPOT * cs] = et; ep_[E_ELEC * cs] = eel;
            ep_[E_VDW * cs] = evd; ep_[E_SRPMF * cs] = esr;
        }
    }
    if (P.write_forces) {
        P.fx[i] = F[0];
        P.fy[i] = F[1];
        P.fz[i] = F[2];
    }
    if constexpr ((FEAT & F_UPDATE) != 0) {
        if (!free_me) {
            P.poso[i] = pa;
            P.sposo[P.rank[i]] = pa;
            return;
        }
        const double *g = noise_row(P, i);
        const double4 pn = bd_update(P, i, is_site, pa, F, g);
        P.poso[i] = pn;
        P.sposo[P.rank[i]] = pn;
    }
}

// Philox Gaussians of rows 0..nrows-1 of the current chunk: row r holds step
// ctl[0] + r, counter ctl[2] + ctl[0] + r, key (ctl[3], 0), particle i (BROMOC_NOISE_PHILOX)
__global__ void k_philox_fill(const int64_t *ctl, uint64_t stream, int nrows, int ntot,
                              int64_t stride, double *out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nrows * ntot) return;
    const int r = (int)(k / ntot), i = (int)(k % ntot);
    double g[3];
    philox_gauss3((uint64_t)ctl[3], (uint64_t)(ctl[2] + ctl[0] + r), stream, i, g);
    double *o = out + r * stride + 3 * (int64_t)i;
    o[0] = g[0]; o[1] = g[1]; o[2] = g[2];
}

hipError_t launch_philox_fill(const int64_t *ctl, uint64_t stream, int nrows, int ntot,
                              int64_t stride, double *out, hipStream_t s) {
    const int64_t n = (int64_t)nrows * ntot;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_philox_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ctl,
                       stream, nrows, ntot, stride, out);
    return hipGetLastError();
}

static void dispatch_wave(const StepParams &p, int feat, hipStream_t s) {
    dim3 grid((p.tend - p.tbeg + WPB - 1) / WPB), block(BLOCK);
#define CASE(F) case F: hipLaunchKernelGGL((k_wave<F>), grid, block, 0, s, p); break;
    switch (feat) {
        CASE(0) CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
        CASE(8) CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15)
    }
#undef CASE
}

template <int T>
static void dispatch_step(const StepParams &p, int feat, hipStream_t s) {
    dim3 grid((p.tend - p.tbeg + T - 1) / T), block(BLOCK);
#define CASE(F) case F: hipLaunchKernelGGL((k_step<F, T>), grid, block, 0, s, p); break;
    switch (feat) {
        CASE(0) CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
        CASE(8) CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15)
    }
#undef CASE
}

hipError_t launch_dna(const StepParams &p, hipStream_t s) {
    if ((p.flags & QNUCL) && p.nsites > 0)
        hipLaunchKernelGGL(k_dna, dim3(p.nsites), dim3(BLOCK), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_pairs(const StepParams &p, hipStream_t s) {
    if (p.tend <= p.tbeg) return hipSuccess;
    int feat = 0;
    if (p.flags & QNUCL) feat |= F_DNA;
    if (p.flags & QPHIX) feat |= F_PHIX;
    if (p.flags & QPHIV) feat |= F_PHIV;
    if (p.mode == 1) feat |= F_UPDATE;
    switch (p.tpb) {
        case 1: dispatch_step<1>(p, feat, s); break;
        case 2: dispatch_step<2>(p, feat, s); break;
        case 4: dispatch_step<4>(p, feat, s); break;
        default: dispatch_wave(p, feat, s); break;   // one wavefront per target
    }
    return hipGetLastError();
}

hipError_t launch_step(const StepParams &p, hipStream_t s) {
    hipError_t e = launch_rfpar(p, true, s);
    if (e != hipSuccess) return e;
    e = launch_dna(p, s);
    if (e != hipSuccess) return e;
    return launch_pairs(p, s);
}

// ---- RFPARION (rfpar.f90:20-228)
// k_rfpar: srfe, reff and their gradients for every ion slot from the current positions.
__global__ void k_rfpar(StepParams P, int grad) {
    const int i = P.nsites + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.ntot) return;
    const double4 pa = P.pos[i];
    double srfe, reff, ds[3], dr[3];
    const bool in = rf_params(P, pa.x, pa.y, pa.z, abs(P.typei[i]), grad != 0, srfe, reff, ds, dr);
    P.rfa[i] = make_double4(srfe, reff, in ? 1.0 : 0.0, 0.0);
    if (grad) {
        P.rfd[2 * (size_t)i] = make_double4(ds[0], ds[1], ds[2], 0.0);
        P.rfd[2 * (size_t)i + 1] = make_double4(dr[0], dr[1], dr[2], 0.0);
    }
}

// k_rfpair: reaction-field energy and force of each ion, one wavefront per target.
// The reference visits each ion pair once, from its higher index (rfpar.f90:157-223): a
// pair counts when the lower ion's srfe is nonzero and the higher ion is moving and inside
// the grid box.  Gathered here per target t over every other ion p, in the reference's
// operand order (outer = higher index): aux0 = celec*q_hi*q_lo, energy charged to the
// higher index, and the force on t from the variation of srfe_t, reff_t and |r_t - r_p|.
// A moving ion's force goes to frf (added by the step kernel); a fixed ion, which the
// step kernel does not visit, gets its (unused) reaction-field force in fx directly.
__global__ __launch_bounds__(BLOCK) void k_rfpair(StepParams P) {
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x >> 6;
    const int t = P.nsites + blockIdx.x * WPB + wv;
    if (t >= P.ntot) return;
    const int nmov0 = P.nsites + P.nfix;
    const bool tmov = t >= nmov0;
    const double4 rt = P.rfa[t];
    if (rt.z == 0.0) {   // outside the grid box: no reaction-field term at all
        if (lane == 0) {
            P.frf[t] = make_double4(0.0, 0.0, 0.0, 0.0);
            if (P.write_energy) P.epart[(size_t)E_RFPAR * P.ecap + t] = 0.0;
            if (!tmov && P.write_forces) { P.fx[t] = 0.0; P.fy[t] = 0.0; P.fz[t] = 0.0; }
        }
        return;
    }
    const double4 pt = P.pos[t];
    const double cgt = P.cg[abs(P.typei[t])];
    const double srt = rt.x, rft = rt.y;
    double A = 0.0, B = 0.0, cx = 0.0, cy = 0.0, cz = 0.0, e = 0.0;
    for (int p = P.nsites + lane; p < P.ntot; p += WAVE) {
        if (p == t) continue;
        const double4 rp = P.rfa[p];
        // lower ion's srfe nonzero, higher ion moving and inside the box (t is inside)
        const bool hi_t = p < t;
        if (hi_t ? (rp.x == 0.0 || !tmov) : (srt == 0.0 || rp.z == 0.0 || p < nmov0)) continue;
        const double4 pp = P.pos[p];
        const double cgp = P.cg[abs(P.typei[p])];
        const double s_lo = hi_t ? rp.x : srt, s_hi = hi_t ? srt : rp.x;
        const double r_lo = hi_t ? rp.y : rft, r_hi = hi_t ? rft : rp.y;
        const double x_lo = hi_t ? pp.x : pt.x, x_hi = hi_t ? pt.x : pp.x;
        const double y_lo = hi_t ? pp.y : pt.y, y_hi = hi_t ? pt.y : pp.y;
        const double z_lo = hi_t ? pp.z : pt.z, z_hi = hi_t ? pt.z : pp.z;
        const double aux0 = P.celec * (hi_t ? cgt : cgp) * (hi_t ? cgp : cgt);
        const double dx = x_lo - x_hi, dy = y_lo - y_hi, dz = z_lo - z_hi;
        const double dist2 = dx * dx + dy * dy + dz * dz;
        const double srfeij = s_lo * s_hi, reffij = r_lo * r_hi;
        const double rfdn = 1.0 / sqrt(reffij * reffij + dist2);
        const double rfcf = reffij * rfdn;
        const double aux1 = aux0 * rfcf;
        if (hi_t) e += aux1 * srfeij;
        const double aux2 = aux0 * srfeij * (rfdn * (rfdn * rfdn));
        const double aux3 = -aux2 * (dist2 + r_lo * r_hi * 0.5);
        A += -aux1 * rp.x;             // times d srfe_t
        B += aux3 * rp.y;              // times d reff_t
        const double de = aux2 * reffij;
        cx += de * (pt.x - pp.x);
        cy += de * (pt.y - pp.y);
        cz += de * (pt.z - pp.z);
    }
    A = wave_sum(A); B = wave_sum(B);
    cx = wave_sum(cx); cy = wave_sum(cy); cz = wave_sum(cz);
    if (P.write_energy) e = wave_sum(e);
    if (lane != 0) return;
    const double4 dst = P.rfd[2 * (size_t)t], drt = P.rfd[2 * (size_t)t + 1];
    // self term of a moving ion (rfpar.f90:146-156): aux = celec*q*q*srfe, energy
    // aux*srfe/2, force -aux*dsrfe
    const double aux = tmov ? P.celec * cgt * cgt * srt : 0.0;
    const double a_s = A - aux;
    const double4 f = make_double4(a_s * dst.x + B * drt.x + cx, a_s * dst.y + B * drt.y + cy,
                                   a_s * dst.z + B * drt.z + cz, 0.0);
    if (tmov) {
        P.frf[t] = f;
        if (P.write_energy) P.epart[(size_t)E_RFPAR * P.ecap + t] = 0.5 * aux * srt + e;
    } else if (P.write_forces) {
        P.fx[t] = f.x; P.fy[t] = f.y; P.fz[t] = f.z;
    }
}

hipError_t launch_rfpar(const StepParams &p, bool pairs, hipStream_t s) {
    const int nion = p.ntot - p.nsites;
    if (!p.rfon || nion <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rfpar, dim3((nion + 255) / 256), dim3(256), 0, s, p, pairs ? 1 : 0);
    if (pairs)
        hipLaunchKernelGGL(k_rfpair, dim3((nion + WPB - 1) / WPB), dim3(BLOCK), 0, s, p);
    return hipGetLastError();
}

// ---- energy reduction: one block per category, fixed-order tree
__global__ __launch_bounds__(BLOCK) void k_reduce_energy(const double *epart, int ecap, int n,
                                                         double *out) {
    __shared__ double s[BLOCK];
    const int cat = blockIdx.x, tid = threadIdx.x;
    const double *src = epart + (size_t)cat * ecap;
    double acc = 0.0;
    for (int k = tid; k < n; k += BLOCK) acc += src[k];
    s[tid] = acc;
    __syncthreads();
    for (int h = BLOCK / 2; h > 0; h >>= 1) {
        if (tid < h) s[tid] += s[tid + h];
        __syncthreads();
    }
    if (tid == 0) out[cat] = s[0];
}

hipError_t launch_reduce_energy(const double *epart, int ecap, int n, double *out,
                                hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_energy, dim3(NE), dim3(BLOCK), 0, s, epart, ecap, n, out);
    return hipGetLastError();
}

__global__ void k_advance(int64_t *ctr, int64_t by) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ctr[0] += by;
}

hipError_t launch_advance(int64_t *ctr, int64_t by, hipStream_t s) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s, ctr, by);
    return hipGetLastError();
}

// ---- INTERACT (interact.f90:19-419): energy of one particle against the system
// energy-only DNA term as INTERACT writes it (interact.f90:212-348): angle through
// angbond.f90, dihedral through angdih.f90, base pairing without scalepairing
__device__ double dna_term_interact(const StepParams &P, uint32_t w) {
    const int kind = w >> 29, t = w & 0x7FFFFFF;
    const double eps = P.epsnuc;
    const int n = P.nlist[kind];
    const int32_t *L = P.lst[kind];
    if (kind == T_BOND) {
        V3 d = sub3(ld3(P, L[t]), ld3(P, L[n + t]));
        double dist = sqrt(dot3(d, d));
        double vard = dist - P.lval[0][t];
        return eps * (vard * vard) * (1.0 + 100.0 * (vard * vard));
    } else if (kind == T_ANGLE) {
        V3 rb = ld3(P, L[n + t]);
        V3 v1 = sub3(ld3(P, L[t]), rb), v2 = sub3(ld3(P, L[2 * n + t]), rb);
        double r1 = sqrt(dot3(v1, v1)), r2 = sqrt(dot3(v2, v2));
        double ang = acos(dot3(v1, v2) / (r1 * r2));
        double varang = ang - P.lval[1][t];
        return 700.0 * eps * (varang * varang);
    } else if (kind == T_DIHE) {
        V3 r1 = ld3(P, L[t]), r2 = ld3(P, L[n + t]), r3 = ld3(P, L[2 * n + t]), r4 = ld3(P, L[3 * n + t]);
        V3 b1 = sub3(r2, r1), b2 = sub3(r3, r2), b3 = sub3(r4, r3);
        V3 n1 = cross3(b1, b2), n2 = cross3(b2, b3);
        double rr = sqrt(dot3(b2, b2));
        double sinphi = rr * dot3(b1, n2);
        double dihe = atan2(sinphi, dot3(n1, n2));
        double vardihe = dihe - P.lval[2][t];
        return 28.0 * eps * (1.0 - cos(vardihe));
    }
    int a = L[t], b = L[n + t];
    V3 v = sub3(ld3(P, b), ld3(P, a));
    double dist2 = dot3(v, v);
    if (kind == T_STACK) {
        double sg = P.lval[3][t];
        double cte1 = pow3(sg * sg / dist2);
        return 4.0 * eps * cte1 * (cte1 - 1.0);
    } else if (kind == T_BP) {
        int na = P.namsite[a];
        double epsln = (na == 5 || na == 4) ? 2.532 * eps : 2.0 * eps;
        double sg = P.lval[4][t];
        double cte1 = sg * sg / dist2;
        return epsln * pow5(cte1) * (20.0 * cte1 - 24.0);
    } else if (kind == T_EX) {
        double sgex2 = P.lval[5][t] * P.lval[5][t];
        if (dist2 < sgex2) {
            double cte1 = pow3(sgex2 / dist2);
            return 4.0 * eps * cte1 * (cte1 - 1.0) + eps;
        }
        return 0.0;
    } else if (kind == T_QQ) {
        double dist = sqrt(dist2);
        double idist = 1.0 / dist;
        if (!(P.flags & QDEBY)) return P.fctn * idist;
        return P.fctn * idist * exp(-dist * P.ikappa);
    }
    double dist = sqrt(dist2);
    double cte1 = exp((13.38 - dist) * 0.1875);
    return P.epsolv * ((1.0 - cte1) * (1.0 - cte1)) - P.epsolv;
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

__global__ __launch_bounds__(BLOCK) void k_interact(InteractParams Q) {
    const StepParams &P = Q.sp;
    __shared__ DevPair s_pairs[MAX_LDS_PAIRS];
    __shared__ int32_t s_pidx[MAX_LDS_TYPES * MAX_LDS_TYPES];
    __shared__ double s_red[WPB];
    __shared__ double s_rf[3];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    const int ns = P.nsites, ntot = P.ntot, nt1 = P.nttyp + 1;
    for (int k = tid; k <= P.npair; k += BLOCK) s_pairs[k] = P.pairs[k];
    for (int k = tid; k < nt1 * nt1; k += BLOCK) s_pidx[k] = P.pairidx[k];
    __syncthreads();
    const int j = Q.j;
    const double xj = Q.xj, yj = Q.yj, zj = Q.zj;
    const bool beion = (P.flags & QPAR) && j >= ns;
    const int jt2 = beion ? P.nwtype[Q.jtype] : __double2loint(P.pos[j].w);
    const int prow = jt2 * nt1;
    double e = 0.0;
    // staticf0.f90:19-82 and vdwgd0trln.f90:19-117 on lanes 0..7 of wave 0
    if (tid < WAVE) {
        double es = 0.0, ev = 0.0, ph = 0.0;
        bool okv = false;
        if ((P.flags & QPHIX) && tid < 8) {
            const GridGeom &g = P.gx;
            bool besite = (P.flags & QNUCL) && j < ns && P.namsite[j] == 1;
            bool ok = beion || besite;
            double chi = beion ? P.cg[Q.jtype] : P.cgnuc;
            ok = ok && xj <= g.xbcen + g.tranx && xj >= g.xbcen - g.tranx &&
                 yj <= g.ybcen + g.trany && yj >= g.ybcen - g.trany &&
                 zj <= g.zbcen + g.tranz && zj >= g.zbcen - g.tranz;
            if (ok) {
                double gx_, gy_, gz_;
                int ix, iy, iz;
                grid_cell(g, xj, yj, zj, gx_, gy_, gz_, ix, iy, iz);
                int n1 = ix + (tid >> 2), n2 = iy + ((tid >> 1) & 1), n3 = iz + (tid & 1);
                double ai = 1.0 - fabs(gx_ - n1 * g.dcel) * g.idcel;
                double bi = 1.0 - fabs(gy_ - n2 * g.dcel) * g.idcel;
                double ci = 1.0 - fabs(gz_ - n3 * g.dcel) * g.idcel;
                double fi = ai * bi * ci;
                es = fi * chi * (double)P.phix[((size_t)n1 * g.ny + n2) * g.nz + n3] * P.celec;
            }
        }
        if (P.flags & QPHIV) {
            const GridGeom &g = P.gv;
            okv = xj <= g.xbcen + g.tranx && xj >= g.xbcen - g.tranx && yj <= g.ybcen + g.trany &&
                  yj >= g.ybcen - g.trany && zj <= P.vzmax && zj >= P.vzmin;
            if (okv && tid < 8) {
                double gx_, gy_, gz_;
                int ix, iy, iz;
                grid_cell(g, xj, yj, zj, gx_, gy_, gz_, ix, iy, iz);
                double esvdw = P.svdw;
                if (P.flags & QSVDW) esvdw = esvdw * P.scal[jt2];
                int n1 = ix + (tid >> 2), n2 = iy + ((tid >> 1) & 1), n3 = iz + (tid & 1);
                double ai = 1.0 - fabs(gx_ - n1 * g.dcel) * g.idcel;
                double bi = 1.0 - fabs(gy_ - n2 * g.dcel) * g.idcel;
                double ci = 1.0 - fabs(gz_ - n3 * g.dcel) * g.idcel;
                double fi = ai * bi * ci;
                ph = (double)P.phiv[((size_t)n1 * g.ny + n2) * g.nz + n3];
                ev = fi * esvdw * ph;
            }
        }
        es = wave_sum(es);
        ev = wave_sum(ev);
        ph = wave_sum(ph);
        if (tid == 0) {
            e += es;
            if (okv) {
                if (ph >= P.thold8) {
                    ev = 1.0e10;
                    if (Q.alert) atomicAdd(&P.warn[jt2], 1);
                }
                e += ev;
            }
        }
    }
    const bool nonbond = (P.flags & QNONBOND) != 0;
    if (beion && nonbond) {
        for (int i = ns + tid; i < ntot; i += BLOCK) {
            if (i == j) continue;
            const double4 q = P.pos[i];
            double dx = q.x - xj, dy = q.y - yj, dz = q.z - zj;
            double d2 = dx * dx + dy * dy + dz * dz;
            const DevPair &pp = s_pairs[s_pidx[prow + __double2loint(q.w)]];
            e += (pp.flags & PF_EFPOT) ? gety(pp, P.ep, P.ires, d2) : interact_analytic(pp, P, d2);
        }
    }
    if ((P.flags & QNUCL) && !beion) {
        const int k0 = P.inc_ptr[j], k1 = P.inc_ptr[j + 1];
        for (int k = k0 + tid; k < k1; k += BLOCK) {
            uint32_t wd = P.inc[k];
            int kind = wd >> 29;
            bool on = kind <= T_DIHE ? (P.flags & QNOBOND) != 0
                                     : (nonbond && (kind != T_SLV || (P.flags & QSOLV) != 0));
            if (on) e += dna_term_interact(P, wd);
        }
    }
    if ((P.flags & QPAR) && (P.flags & QNUCL) && nonbond) {
        const int lo = beion ? 0 : ns, hi = beion ? ns : ntot;
        for (int i = lo + tid; i < hi; i += BLOCK) {
            const double4 q = P.pos[i];
            double dx = q.x - xj, dy = q.y - yj, dz = q.z - zj;
            double d2 = dx * dx + dy * dy + dz * dz;
            const DevPair &pp = s_pairs[s_pidx[prow + __double2loint(q.w)]];
            e += (pp.flags & PF_EFPOT) ? gety(pp, P.ep, P.ires, d2) : interact_analytic(pp, P, d2);
        }
    }
    // RFPARIONJ (rfpar.f90:232-370) for an ion: its own srfe/reff at the trial position,
    // the self term and every other ion with nonzero srfe (ions' parameters from k_rfpar).
    // Nothing when the trial position is outside the grid box.  (For a site the reference
    // reads srfe outside its array; this path gives sites no reaction-field term.)
    if (P.rfon && beion) {
        if (tid == 0) {
            double sj, rj, ds[3], dr[3];
            const bool in = rf_params(P, xj, yj, zj, Q.jtype, false, sj, rj, ds, dr);
            s_rf[0] = sj; s_rf[1] = rj; s_rf[2] = in ? 1.0 : 0.0;
        }
        __syncthreads();
        if (s_rf[2] != 0.0) {
            const double sj = s_rf[0], rj = s_rf[1];
            const double tau = P.celec * P.cg[Q.jtype];
            if (tid == 0) e += 0.5 * tau * P.cg[Q.jtype] * (sj * sj);
            for (int i = ns + tid; i < ntot; i += BLOCK) {
                if (i == j) continue;
                const double4 ra = P.rfa[i];
                if (ra.x == 0.0) continue;
                const double4 q = P.pos[i];
                const double dx = q.x - xj, dy = q.y - yj, dz = q.z - zj;
                const double dist2 = dx * dx + dy * dy + dz * dz;
                const double reffij = rj * ra.y;
                e += tau * P.cg[abs(P.typei[i])] * reffij * sj * ra.x /
                     sqrt(reffij * reffij + dist2);
            }
        }
    }
    e = wave_sum(e);
    if (lane == 0) s_red[tid >> 6] = e;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int k = 0; k < WPB; k++) t += s_red[k];
        Q.out[0] = t;
    }
}

hipError_t launch_interact(const InteractParams &p, hipStream_t s) {
    hipLaunchKernelGGL(k_interact, dim3(1), dim3(BLOCK), 0, s, p);
    return hipGetLastError();
}

__global__ void k_philox_raw(const uint64_t *ctr, const uint64_t *key, int n, uint64_t *out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint64_t c[4] = {ctr[4 * k], ctr[4 * k + 1], ctr[4 * k + 2], ctr[4 * k + 3]};
    philox4x64(c, key[2 * k], key[2 * k + 1]);
    for (int q = 0; q < 4; q++) out[4 * k + q] = c[q];
}

hipError_t launch_philox_raw(const uint64_t *ctr, const uint64_t *key, int n, uint64_t *out,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_philox_raw, dim3((n + 255) / 256), dim3(256), 0, s, ctr, key, n, out);
    return hipGetLastError();
}

}  // namespace bromoc
