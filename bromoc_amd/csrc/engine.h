// engine.h -- device-side state shared by the kernels (kernels.hip) and the C ABI
// (api.hip) of the gfx950 BROMOC force / BD path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bromoc_gpu.h"

namespace bromoc {

// Number of per-particle energy slots: the bromoc_energies fields after `ener`.
constexpr int NE = 22;
constexpr int RED_BLOCKS = 32;   // blocks per energy term in the reduction
enum ECat {
    E_ELEC = 0, E_VDW, E_SRPMF, E_SRPMFMX, E_EFPOT, E_EFPOTMX, E_QQMX, E_VDWMX,
    E_BOND, E_ANG, E_DIHE, E_STACK, E_BP, E_EX, E_QQ, E_SOLV,
    E_GSBPA, E_GSBPB, E_VDWGD, E_RFPAR, E_CON, E_MEMB
};

// The value of an order-independent fixed-point sum of W words (fix_add, device.h): the
// first in units of 2^-32 (two's complement), the others in units of 2^-64, 2^-96.
__host__ __device__ inline double fix_value(const unsigned long long *acc, int W = 2) {
    double v = (double)(long long)acc[0] * (1.0 / 4294967296.0);
    double unit = 1.0 / 4294967296.0;
    for (int w = 1; w < W; w++) {
        unit *= 1.0 / 4294967296.0;
        v += (double)acc[w] * unit;
    }
    return v;
}

// pair-type flags
constexpr int PF_EFPOT = 1, PF_CHR = 2, PF_LJ = 4, PF_SRPMF = 8;

// One pair type (efpmod/extramod/grandmod fields), laid out for LDS.
struct DevPair {
    double dmi, dm2lo, dm2hi, sc1, sc2, sc3, fct, epp4;
    double sgp2, c0, c1, c2, c3, c4;
    int32_t nxi, nlast, flags, pad;   // first and last knot of the table
};

struct GridGeom {
    int32_t nx, ny, nz, on;
    double dcel, idcel, xbcen, ybcen, zbcen, tranx, trany, tranz;
};

// DNA term kinds in the per-site incidence list
enum TermKind { T_BOND = 0, T_ANGLE, T_DIHE, T_STACK, T_BP, T_EX, T_QQ, T_SLV };
// the list value a two-site term carries (lval slot of its kind; qq and solvent none)
__host__ __device__ constexpr int pair_lval_slot(int kind) {
    return kind == T_BOND ? 0 : kind == T_STACK ? 3 : kind == T_BP ? 4 : kind == T_EX ? 5 : -1;
}
// incidence word: kind (3 bits) | role (2 bits) | term index (27 bits)
__host__ __device__ inline uint32_t inc_pack(int kind, int role, int term) {
    return (uint32_t(kind) << 29) | (uint32_t(role) << 27) | uint32_t(term);
}

constexpr int MAX_LDS_PAIRS = 64;     // pair types held in LDS
constexpr int MAX_LDS_TYPES = 16;     // atnam2 types (pair index table in LDS)

// Curved systems (grandmod cx, cy, cz, rsphe, rsphe2, iecx, iecy): kind 0 box, 1 sphere,
// 2 elliptic cylinder, with the bounding box lx2m..lz2p.  The same code serves the host
// (MOVE, the host GRAND loop) and the device, in the reference's order of operations.
struct Curved {
    int32_t kind, pad;
    double cx, cy, cz, rsphe, rsphe2, iecx, iecy;
    double lx2m, lx2p, ly2m, ly2p, lz2m, lz2p;
    double sx, sy, sz, thr;   // the device OUTBOX's quadratic form (device.h outbox)
    // OUTBOX (grandmod.f90:74-97); MOVE's test is the same (move.f90:39-46)
    __host__ __device__ __forceinline__ bool outside(double x, double y, double z) const {
#pragma clang fp contract(off)
        if (kind == 1)
            return (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz) > rsphe2;
        if (kind == 2) {
            const double ex = (x - cx) * iecx, ey = (y - cy) * iecy;
            return ex * ex + ey * ey > 1.0 || z < lz2m || z > lz2p;
        }
        return x < lx2m || x > lx2p || y < ly2m || y > ly2p || z < lz2m || z > lz2p;
    }
    // FIXCOOR's curved branches (fixcoor.f90:24-46): outside the sphere a particle is
    // reflected through the surface along its radius; outside the ellipse it is scaled back
    // onto it by 0.999999999999999/sqrt(r2) and its z clamped.  Returns whether it moved.
    __host__ __device__ __forceinline__ bool fix(double &x, double &y, double &z) const {
#pragma clang fp contract(off)
        bool moved = false;
        if (kind == 1) {
            const double r2 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz);
            if (r2 > rsphe2) {
                const double lzrad = 2.0 * rsphe / sqrt(r2) - 1.0;
                x = lzrad * (x - cx) + cx;
                y = lzrad * (y - cy) + cy;
                z = lzrad * (z - cz) + cz;
                moved = true;
            }
        } else if (kind == 2) {
            const double ex = (x - cx) * iecx, ey = (y - cy) * iecy;
            const double r2 = ex * ex + ey * ey;
            if (r2 > 1.0) {
                const double r = 0.999999999999999 / sqrt(r2);
                x = r * (x - cx) + cx;
                y = r * (y - cy) + cy;
                moved = true;
            }
            if (z < lz2m) { z = lz2m; moved = true; }
            else if (z > lz2p) { z = lz2p; moved = true; }
        }
        return moved;
    }
    // INSERT's acceptance (insert.f90:33-40) of a point drawn in buffer slab ib: strictly
    // inside the shell rmin..rmax around the GCMC centre (gx, gy, gz) of the sphere, inside
    // the ellipse, anywhere in the box
    __host__ __device__ __forceinline__ bool insert_ok(double x, double y, double z, double gx, double gy,
                                       double gz, double rmin, double rmax) const {
#pragma clang fp contract(off)
        if (kind == 1) {
            const double r2 = (x - gx) * (x - gx) + (y - gy) * (y - gy) + (z - gz) * (z - gz);
            return r2 > rmin * rmin && r2 < rmax * rmax;
        }
        if (kind == 2) {
            const double ex = (x - gx) * iecx, ey = (y - gy) * iecy;
            return ex * ex + ey * ey < 1.0;
        }
        return true;
    }
};

struct StepParams {
    // sizes
    int32_t ntot, nsites, nfix, ntargets;
    int32_t nttyp, ntype, nold, npair;
    uint32_t flags;
    int32_t mode;            // 0: ENERGY (forces out, no update), 1: ENERGY + DYNAMICS
    int32_t write_energy, write_forces;
    int32_t tpb, all_efpot;  // k_step targets per block (1, 2, 4) or 8: k_wave; all tabulated
    int32_t tbeg, tend;      // target range of this launch (sites 0..nsites-1 come first)
    // positions (x, y, z, atnam2 type in the low word of w) in, updated positions out
    const double4 *pos;
    double4 *poso;
    // the same particles in spatial (partner) order, w = type | index << 32, and the slot
    // of each particle in that order
    const double4 *spos;
    double4 *sposo;
    const int32_t *rank;
    double r2near;           // squared range beyond which every table pair is in its tail
    const int32_t *typei;    // signed atnam type
    const int32_t *t2;       // atnam2 type per particle
    const int8_t *stfree;    // [nsites]
    const int32_t *namsite;  // [nsites]
    double *fx, *fy, *fz;
    double4 *fdna;           // [nsites] CG-DNA list-term force of each site (k_dna)
    double *epart;           // [NE][ecap]
    int32_t ecap, pad0;
    // tables
    const int32_t *pairidx;  // [(nttyp+1)^2]
    const DevPair *pairs;    // [npair+1]
    const DevPair *rows;     // [(nttyp+1)^2] pairs[pairidx[a*(nttyp+1)+b]]
    const double4 *ep;       // knots (ep1, ep2, ep3, 0)
    double ires, celec;
    const double *cg, *fact1a, *fact2a;
    const int32_t *nwtype;
    double rth, srpx, srpk, srpy;
    // CG-DNA
    const int32_t *inc_ptr;  // [nsites+1] each site's term incidences in inc (INTERACT)
    const uint32_t *inc;
    // resolved incidences for the force kernels, the site's angles and dihedrals first
    // (bits of x = word | first other site << 32, bits of y = second | third other site
    // << 32, the other sites in list order), then the two-site terms (bits of x = word |
    // other site << 32 | (base pair of a C/G site a) << 63, y = the term's list value)
    const int32_t *incr_ptr;   // [nsites+1]
    const int32_t *inc_split;  // [nsites] first two-site record of each site in incr
    const double2 *incr;
    double4 *fmul;             // [records of incr] angle / dihedral results of k_dna_multi
    // one-word records of the bulk two-site terms: other site (bits 0-23), term (24-25:
    // excluded volume, P-P, solvent, base pair), role (26), C/G site a of a base pair (27),
    // slot of the list value in bpar[16 * term + slot] (28-31)
    const int32_t *incb_ptr;   // [nsites+1]
    const uint32_t *incb;
    const double *bpar;        // [4 * 16]
    const double2 *dihcs;      // [ndihe] cos and sin of each dihedral's phi0
    const int32_t *lst[8];   // component-major site lists
    const double *lval[6];
    int32_t nlist[8];
    double epsnuc, scalepairing, fctn, kappa, ikappa, epsolv, cgnuc;
    // grids
    GridGeom gx, gv;
    const float *phix;
    const int8_t *phiv;
    double svdw, thold8, vzmin, vzmax, thold27;
    const double *scal;
    int32_t *warn;
    // Brownian dynamics
    double bdmax, fact0n1, fact0n2;
    double lx2m, lx2p, ly2m, ly2p, lz2m, lz2p;   // FIXCOOR's clamp (open along curved axes)
    Curved cv;                                    // OUTBOX; the curved FIXCOOR after a step
    int32_t cntpts, noise_mode;
    const double *zcont;
    int32_t *nfwd, *nbwd;
    const double *noise;           // Gaussian rows [row][3*i]: injected rows, or the Philox
                                   // rows of the current chunk (noise_row)
    int64_t noise_stride;          // doubles per row
    const int64_t *step_ctr;       // device step counter (graph-friendly)
    int64_t step_off;              // this launch's step = *step_ctr + step_off
    int64_t row0;                  // injected row of step counter 0
    uint64_t seed;
    uint64_t stream;
    // ion update: 0 DYNAMICS0 (fact1a/fact2a), 1 DYNAMICS1 (pore switch), 2 DYNAMICS2
    // (spline profile), 3 DYNAMICS3 (DNA proximity: ibeta, diff0, diffcutoff)
    int32_t dmode, nspline;
    double ibeta, diff0, diffcutoff;
    double dt, kbtdt, plength2, pcenter;
    const double *diffusion, *ampl3, *p3;
    const double *xs, *ys, *bs, *cs, *ds;
    // reaction-field parameter grids (RFPARION)
    GridGeom gr;
    const float *gsrfen, *greff;
    double sqrfac;
    int32_t rfon, rfsingle;
    double4 *rfa;            // [cap] (srfe, reff, 1 if inside the grid box, 1 if moving)
    double4 *rfq;            // [cap] (x, y, z, charge) of each ion, for the pair kernel
    double4 *rfd;            // [2*cap] (d srfe / d r, 0), (d reff / d r, 0)
    double4 *frf;            // [cap] reaction-field force on each moving ion (k_rfpair)
    // APFOR: summed applied force of each site (zero where none) [nsites]
    const double4 *apf;
    // CONTRANS: ctn entries (csn 1-based, 0 = centroid), restraint centres [ctn+1]; cen:
    // site position sums of the current positions (all sites, first strand, second
    // strand), written by k_centroid
    int32_t ctn, unsplit, nsites1st, notrans_axes;
    const int32_t *csn;
    const double *kx, *ky, *kz, *contrx, *contry, *contrz;
    double insites;
    double *cen;             // [9]
    // NOTRANS centre
    double notrx, notry, notrz;
    // MEMBRANE (per atnam2 type arrays: ampl1, ampl2, rcylinder [nttyp+1], p1, p2 [2*(nttyp+1)])
    int32_t qpore, pad1;
    double voltage, afact, ceps, zmemb1, zmemb2;
    const double *ampl1, *p1, *ampl2, *p2, *rcylinder;
    // virial pressure: F . x of every target added per step (the positions before the
    // update), or null when the pressure is not gathered [cap]
    double *vacc;
    // MMIJ reaction field, RECTBOX (rect_rf1.f90): its force on each moving ion goes to frf
    int32_t mmon, ntpol, xnpol, ynpol, znpol, npl;   // npl = xnpol + ynpol + znpol
    const int32_t *lstp;     // [ntpol] degrees, x | y << 8 | z << 16
    const double *bnorm;     // [ntpol]
    const double *mmij;      // [ntpol * ntpol]
    double xscal, yscal, zscal, rfscal;
    double *mleg;            // [cap+1][npl] P_0.. of x, y, z scaled (zeros outside the region)
    double *mchg;            // [cap+1] charge of each moving ion inside the region, else 0
    double *mcoef, *mq, *mer;   // [ntpol] coefficients, M q rfscal, per-row energy terms
    // tile mode (tiles.hip): tiles of 64 * ttpl consecutive slots of the spatial order;
    // tpart[((K * ntile + I) * tnc + c) * TT + t]: component c (fx, fy, fz, and with the
    // fused reaction field A, B) of slot t of tile I from the tile pair of I and K, each
    // entry written once per step; tener[cat][item]: pair energies of each tile pair
    double *tpart, *tener;
    int32_t *tflag;          // [titems] 1 where a tile pair wrote its reaction-field rows
    int32_t ntile, ttpl, tnc, titems;
};

// tpb values of tile mode: TPB_TILE + targets per lane (1, 2 or 4)
constexpr int TPB_TILE = 16;
constexpr int TILE_WAVES = 2;      // wavefronts of a k_tiles block (tile pairs sharing tile J)
__host__ __device__ inline bool tile_mode(int tpb) { return tpb > TPB_TILE; }

// pair-energy categories of the tile kernel (tener rows) and the ECat each one feeds
constexpr int NTE = 9;
enum { TE_EFPOT = 0, TE_ELEC, TE_VDW, TE_SRPMF, TE_EFPOTMX, TE_QQMX, TE_VDWMX, TE_SRPMFMX,
       TE_RFPAR };

// spatial partner order (sort.hip)
struct SortGeom {
    double lo[3];
    double inv;        // 1 / cell size
    int32_t n[3];
    int32_t nsites, nfix, pad;   // particle classes: sites, fixed ions, then moving ions
};
struct SortBufs {
    SortGeom geom;
    uint32_t *keys, *keys_out;     // [cap]: keys, then the sorted keys
    int32_t *vals, *order, *rank;  // [cap]: slot of each particle; particle in each slot
    int32_t *hist;                 // [sort_hist_ints(cap)] digit histograms of a pass
    int32_t *cflag;                // [cap] cell numbers (exclusive scan of the key changes)
    int32_t *cell_start;           // [cap+1] first slot of each occupied cell, then n
    int32_t *ncell;                // [1] occupied cells
};
size_t sort_hist_ints(int n);
hipError_t sort_partners(const SortBufs &b, const double4 *pos, int n, double4 *spos0,
                         double4 *spos1, hipStream_t s);

// launchers (kernels.hip)
hipError_t launch_step(const StepParams &p, hipStream_t s);   // launch_dna + launch_pairs
hipError_t launch_dna(const StepParams &p, hipStream_t s);    // k_dna (CG-DNA list terms)
hipError_t launch_pairs(const StepParams &p, hipStream_t s);  // k_step (pairs, grids, update)
// RFPARION: k_rfpar (srfe, reff, and their gradients when pairs or grad) then, with pairs,
// the cross term k_rfpair
hipError_t launch_rfpar(const StepParams &p, bool pairs, hipStream_t s, bool grad = false);
hipError_t launch_notrans(const StepParams &p, hipStream_t s);  // curved FIXCOOR, NOTRANS after a step
hipError_t launch_fixed(const StepParams &p, hipStream_t s);    // terms of the fixed ions
// MMIJ reaction field (mmij.hip): RECT_RF1 for ENERGY (force into frf, energy into slot 0 of
// E_GSBPB), RECT_RF0 for INTERACT (added to *out)
hipError_t launch_mmij(const StepParams &p, hipStream_t s);
hipError_t launch_mmij_interact(const StepParams &p, double xj, double yj, double zj, int jtype,
                                int j, double *out, hipStream_t s);
// out[cat] = sum over particles of epart[cat], plus for the pair categories the sum over
// the tile pairs of tener (tile mode), both in a fixed order; out holds NE * (1 +
// RED_BLOCKS) doubles (the range sums follow the NE results)
hipError_t launch_reduce_energy(const double *epart, int ecap, int n, double *out,
                                hipStream_t s, const double *tener = nullptr, int titems = 0);
// tile mode (tiles.hip): k_tiles over every tile pair (which & 1), then k_tfin over every
// slot (which & 2)
hipError_t launch_tiles(const StepParams &p, hipStream_t s, int which = 3);
// device bytes of the tile buffers for `cap` particles at ttpl targets per lane (fused
// reaction field: rf != 0)
size_t tile_part_doubles(int cap, int ttpl, int rf);
size_t tile_ener_doubles(int cap, int ttpl);
hipError_t launch_advance(int64_t *ctr, int64_t by, hipStream_t s);
hipError_t launch_philox_fill(const int64_t *ctl, uint64_t stream, int nrows, int ntot,
                              int64_t stride, double *out, hipStream_t s);
hipError_t launch_philox_raw(const uint64_t *ctr, const uint64_t *key, int n, uint64_t *out,
                             hipStream_t s);

struct InteractParams {
    StepParams sp;              // system state (positions x,y,z; tables; grids; lists)
    double xj, yj, zj;
    int32_t jtype, j, alert, pad;
    double *out;                // [1]
};
hipError_t launch_interact(const InteractParams &p, hipStream_t s);

// ---- GRAND of one system on the device (grand.hip)
constexpr int GC_BLOCKS = 256;   // workgroups of k_gc_interact (partials per attempt)

// error codes of GcState::err
enum { GCE_OK = 0, GCE_UNIFORMS = 1, GCE_CAPACITY = 2, GCE_STATE = 3 };

struct GcState {
    int64_t k;               // uniforms drawn
    int32_t ntot, err;
    int32_t mode, ib, ip, itype, j;   // trial: mode 0 none, 1 insertion, 2 deletion
    int32_t nacc;            // accepted insertions and deletions
    int32_t natt, pad;       // attempts finished (rows of GcParams::natlog written)
    double x, y, z, ener;
};

struct GcParams {
    StepParams sp;           // system tables and grids; sp.pos is the current buffer
    double4 *pos0, *pos1;    // both position buffers (a change is written to both)
    int32_t *typei, *t2, *ibuffer;
    const double *u;
    int64_t nu;
    int32_t nbuffer, pad0;
    int32_t *nat;
    int32_t *natlog;         // [attempts][nbuffer]: each buffer's count after each attempt
                             // (the host adds them to the caller's prob histogram)
    const double *avnum, *mu, *kb, *lzmin, *lzmax;   // [nbuffer+1]
    const double *rmin, *rmax;                        // [nbuffer+1] (sphere)
    const int32_t *ibfftyp, *ntotat;
    const int8_t *bias;
    double lx, ly, cx, cy, cz, ikbt;
    int64_t icycle;
    int32_t cap, pad;
    GcState *st;
    double *part;            // [GC_BLOCKS]
};

// one attempt: k_gc_pre, (k_gc_rfpar), k_gc_interact, k_gc_post
// a GRAND call of ngcmc attempts (grand.hip)
hipError_t launch_grand(const GcParams &g, int ngcmc, hipStream_t s);

// ---- EFPOT tables on the device (tables.hip)
struct TabEnds;
struct TabJob {
    int32_t ntab, npair, nrows, pad;
    const double *r, *u;        // samples of every table, concatenated
    const int64_t *off;         // [ntab+1] first sample of each table
    const int32_t *nxi;         // [ntab] first knot of each table in ep
    const double *kf;           // [ntab] fct of a charged pair, else 0
    const int32_t *tab_of;      // [npair+1] table of each pair type, -1: not updated
    double *scratch;            // [6 * total samples]
    TabEnds *ends;              // [ntab]
    double4 *ep;                // the context's knots
    DevPair *pairs, *rows;      // the context's pair parameters
    const int32_t *pairidx;     // [nrows]
};
// k_tab_build (one thread per table: SPLINEPOT into ep) then k_tab_rows (the pair rows)
hipError_t launch_table_update(const TabJob &j, hipStream_t s);
// the fixed workspace of the allocation-free table VJP (tables.hip): its device address on
// `device` (current), taken once per device; and its size in doubles
hipError_t vjp_scratch_init(int device, double **out);
size_t vjp_scratch_doubles();
// grad_u = (d table / d u)^T (grad_ep, grad_sc) by the reverse sweep (table_splinepot_vjp),
// one thread; scratch: 6n doubles
hipError_t launch_table_vjp(int n, const double *r, const double *gep, const double *gsc,
                            double *out, double *scratch, hipStream_t s);

// ---- per-cycle analyses (analysis.hip)
struct AnParams {
    const double4 *pos;        // current positions (atnam2 type in the low word of w)
    const int32_t *typei;
    int32_t ntot, nsites, nfix, what;
    int32_t nzmax, igr, npoints, datom;
    int32_t ntype, nold, nttyp, chan;  // chan: 0 no prob2, 1 membrane pore, 2 czmin..czmax
    double zmini, idelz, dr, czmin, czmax, zmemb1, zmemb2;
    const double *rcylinder;   // [nttyp+1] (membrane)
    const int32_t *nwtype;     // [ntype+1]
    unsigned long long *rho, *rdna, *gr, *prob2, *prob3;
    int32_t *irmax;            // [1]
    int32_t *ncnt;             // [2][ntype+1] channel and system counts of the cycle
};
hipError_t launch_analyze(const AnParams &a, hipStream_t s);   // k_hist, then k_prob
// vsum[0] += sum of vacc[0..n) in a fixed order; vacc zeroed
hipError_t launch_vir_fold(double *vacc, int n, double *vsum, hipStream_t s);

// ---- trajectory adjoint (adjoint.hip)
struct AdjParams {
    StepParams sp;
    const double4 *x;       // state x_s [ntot]
    const double4 *lam1;    // lambda_{s+1} [ntot]
    double4 *lam0;          // lambda_s [ntot]
    double4 *f;             // forces F(x_s) [ntot]
    double4 *mu;            // [ntot]
    double4 *dbox;          // [ntot] 1/0 per component: fixcoor left it alone
    const double *xi;       // noise row of step s, [3*ntot]
    const int32_t *group;   // [nt1*nt1] RDF group of an atnam2 type pair (-1: none)
    int32_t ngroup, nbins;
    double r0, dr, sigma, cut;
    const double *w;        // [ngroup*nbins] residual weights dL/dH
    double *hist;           // [ntot][ngroup*nbins] per-target raw histograms
    // gradients as order-independent fixed-point sums of ADJ_W words (fix_add):
    // [nknots][2 (ep2, ep3)][ADJ_W] and [npair+1][2 (sc1, sc3)][ADJ_W]
    unsigned long long *gep, *gsc;
    int32_t obs, pad;
    double zflux;           // terminal call: lambda_z of ion i gains zflux * q_i (flux loss)
    // tangent-linear pass (k_tlm): dx_s in, dx_{s+1} out, the direction of the parameters
    // ([nknots][2] ep2, ep3; [npair+1][2] sc1, sc3) and the derivative of every histogram
    // bin as fixed-point sums [ngroup*nbins][ADJ_W]
    const double4 *dx;
    double4 *dxo;
    const double *vep, *vsc;
    unsigned long long *dh;
};

// words of the adjoint's and the tangent's fixed-point sums (resolution 2^-96 of a sum
// whose sources the host scales to about 1)
constexpr int ADJ_W = 3;
// which: 0 k_rdf, 1 k_adj_force, 2 k_adj_mu, 3 k_adj_pairs, 4 k_tlm
hipError_t launch_adj(int which, const AdjParams &a, hipStream_t s);

// ---- replica batches (batch.hip)
struct BatchParams {
    StepParams sp;              // shared tables, grid geometry, BD constants
    int32_t nrep, persistent;   // replicas; 1: waves take items from queue[0]
    int32_t mode, write_energy; // mode 1: BD step, 0: ENERGY only
    int32_t write_forces, ngcmc;
    const int32_t *order;       // [nrep] replicas by decreasing capacity
    const int32_t *off, *cap, *nfix;   // [nrep] first slot, slots, leading fixed ions
    int32_t *ntot;              // [nrep] particles now (GRAND changes it)
    int32_t *wstart;            // [nrep+1] work-item prefix over `order` (k_batch_prep)
    int32_t *queue;             // [2] work counter, error flag
    const float *const *phix;   // [nrep] static field of each replica
    const double4 *pos;         // [slots] (x, y, z, atnam2 type in the low word)
    double4 *poso;              // the other buffer (BD step output)
    double4 *posw;              // the current buffer, written by GRAND
    int32_t *typei, *ibuffer;   // [slots]
    double *fx, *fy, *fz;       // [slots]
    double *epart;              // [NE][ecap]
    int64_t ecap;
    int32_t *nfwd, *nbwd, *warn;   // per replica, strides below
    int32_t flux_stride, warn_stride;
    int32_t gstride, pad;       // per-replica stride of the buffer arrays (nbuffer max + 1)
    const int32_t *nbuf;        // [nrep] buffers of each replica (0: no GCMC)
    const int32_t *ibfftyp, *ntotat;              // [nrep][gstride]
    const double *avnum, *mu, *kb, *lzmin, *lzmax;   // [nrep][gstride]
    const int8_t *bias;                            // [nrep][gstride]
    const double *gbox;         // [nrep][6] lx, ly, cx, cy, cz, ikbt
    int32_t *nat;               // [nrep][gstride]
    int32_t *prob;              // per replica [nbuffer][cap+1] at probofs[r]
    const int64_t *probofs;
    double *gener;              // [nrep] GRAND's energy accumulator
    int64_t *ctl;               // [8]: steps, counter0, cycles, seed, step counter, cycle,
                                //      cycle0, spare
};
hipError_t launch_batch_prep(const BatchParams &b, int gcmc, hipStream_t s);
hipError_t launch_batch_step(const BatchParams &b, int blocks, int lanes, hipStream_t s);
hipError_t launch_batch_gcmc(const BatchParams &b, hipStream_t s);
hipError_t launch_batch_reduce(const BatchParams &b, double *out, hipStream_t s);

// the shared part of a batch: a single-system context over the shared bromoc_system
// (no particles), whose device tables the batch kernels read (api.hip)
int shared_context(const bromoc_system *s, int device, bromoc_ctx **out, StepParams *base);

}  // namespace bromoc
