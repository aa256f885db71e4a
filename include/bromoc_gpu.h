/*
 * bromoc_gpu.h -- C ABI of the MI355X (gfx950) BROMOC force / Brownian-dynamics path.
 *
 * The reference driver (cg-gcmc-bd, Fortran) keeps all state in module globals and calls
 * argument-less or by-reference routines.  This ABI takes that state explicitly, through
 * plain pointers and sizes, so a thin iso_c_binding module can forward the module arrays
 * and the Fortran driver keeps its routine names and call sites.
 *
 * Entry point                  replaces (reference interface)
 * ---------------------------  -----------------------------------------------------------
 * bromoc_create                one-time state set-up done by shell_simul.f90 (PARTICLE,
 *                              NUCLEOTIDE, EFPOT, GSBP, SYSTEM commands) before SIMUL
 * bromoc_energy                SUBROUTINE ENERGY            energy.mono.f90:19-692
 * bromoc_step                  simul1.f90:363-390: ENERGY + DYNAMICS0 (dynamics0.mono.f90:19-48)
 *                              + DYNAMICS0NUC (dynamics0nuc.f90:19-46), nbd times
 * bromoc_dynamics              DYNAMICS0(ninit,nfinal) + DYNAMICS0NUC(1,nsites) alone
 * bromoc_interact              SUBROUTINE INTERACT          interact.f90:19-419
 * bromoc_grand                 SUBROUTINE GRAND             grand.f90:19-121
 * bromoc_build_table           EFPOT table construction     splinepot.f90:19-96 +
 *                              squarespline polynom.f90:20-48
 * bromoc_discretize            SUBROUTINE DISCRETIZE        discretize.f90:19-120
 * bromoc_update_tables         SPLINEPOT on the device, into a live context
 * bromoc_spline                SUBROUTINE SPLINE            spline.f90:19-98
 * bromoc_metropolis            SUBROUTINE METROPOLIS        metropolis.f90:19-58 + MOVE
 *                              (move.f90:19-50)
 * bromoc_read_map              SUBROUTINE READPHI           readphi.f90:19-349
 * bromoc_read_rfpar            SUBROUTINE READRFPAR         rfpar.f90:373-582
 * bromoc_repwalls              SUBROUTINE REPWALLS          repwalls.f90:19-67
 * bromoc_set_analysis,         the statistics of the cycle loop: rho(z), rDNA(z), g(r),
 * bromoc_analyze,              prob2/prob3 (simul1.f90:477-538) and the virial pressure
 * bromoc_get_analysis          (simul1.f90:365-368)
 *
 * bromoc_step / bromoc_dynamics move the ions with DYNAMICS0, DYNAMICS1
 * (dynamics1.mono.f90:19-90, flag BROMOC_QDIFFUSE), DYNAMICS2 (dynamics2.mono.f90:19-89,
 * BROMOC_QPROFILE) or DYNAMICS3 (dynamics3.mono.f90:19-75, BROMOC_QPROXDIFF), as
 * simul1.f90:369-380 selects; bromoc_step also applies NOTRANS (simul1.f90:382-386,
 * BROMOC_QNOTRANS) after each step.  ENERGY and INTERACT include RFPARION / RFPARIONJ
 * (rfpar.f90:20-370) under BROMOC_QRFPAR, the B-spline repulsion map (vdwgd1spln.f90:19-
 * 170, vdwgd0spln.f90) under BROMOC_QBSPLINE, the membrane (membrane.f90:19-107,
 * interact.f90:58-119) under BROMOC_QMEMB, the MMIJ reaction field (rect_rf1.f90,
 * rect_rf0.f90) under BROMOC_QMMIJ; ENERGY includes APFOR and CONTRANS
 * (energy.mono.f90:509-580) under BROMOC_QAPFOR / BROMOC_QCONTRANS.
 *
 * Conventions
 *  - Particle indices are 0-based (C index i is Fortran index i+1).  Sites occupy
 *    0..nsites-1, fixed ions nsites..nsites+nfix-1, moving ions nsites+nfix..ntot-1,
 *    the same order as grandmod's x,y,z arrays.
 *  - Type tables keep the reference's 1-based type numbering: arrays indexed by type
 *    have length ntype+1 (or nttyp+1) and entry 0 unused.
 *  - Two-index site lists (sitebond, siteangle, ...) are component-major:
 *    list[k*n + t] is site k of term t (0-based site ids).
 *  - Grids are x-major, z fastest: value(ix,iy,iz) = data[(ix*ncly + iy)*nclz + iz],
 *    as in staticf1.f90:91.
 *  - All arithmetic is IEEE binary64.  Grid payloads keep the reference storage
 *    precision (phix float32, phiv int8; gsbpmod.f90:73-74) and are promoted exactly.
 *  - Every entry point returns a status code; nothing aborts the process.
 */
#ifndef BROMOC_GPU_H
#define BROMOC_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BROMOC_ABI_VERSION 5

/* status codes */
#define BROMOC_OK              0
#define BROMOC_E_ARG           1  /* null pointer, bad size or index */
#define BROMOC_E_UNSUPPORTED   2  /* a reference option this path does not implement */
#define BROMOC_E_CAPACITY      3  /* ntot would exceed capacity (reference: ntot>datom, grand.f90:66) */
#define BROMOC_E_HIP           4  /* a HIP runtime call failed */
#define BROMOC_E_NODEVICE      5  /* no gfx950 device visible */
#define BROMOC_E_STATE         6  /* call order (e.g. injected noise not uploaded) */

/* flags: the reference's logical switches that select ENERGY/DYNAMICS branches */
#define BROMOC_QPAR        (1u << 0)   /* ions present (nucleotmod Qpar) */
#define BROMOC_QNUCL       (1u << 1)   /* CG-DNA present (Qnucl) */
#define BROMOC_QNONBOND    (1u << 2)   /* nonbonded terms (Qnonbond) */
#define BROMOC_QNOBOND     (1u << 3)   /* bonded DNA terms (Qnobond) */
#define BROMOC_QPHIX       (1u << 4)   /* static field grid (Qphix) */
#define BROMOC_QPHIV       (1u << 5)   /* repulsion grid, trilinear (Qphiv .and. Qtrln) */
#define BROMOC_QSVDW       (1u << 6)   /* per-type repulsion scaling (Qsvdw) */
#define BROMOC_QSOLV       (1u << 7)   /* solvent-induced S-S term (Qsolv) */
#define BROMOC_QDEBY       (1u << 8)   /* Debye screening of P-P Coulomb (Qdeby) */
#define BROMOC_QCOUNTION   (1u << 9)   /* flux counting across z planes (Qcountion) */
#define BROMOC_QDNAFREE    (1u << 10)  /* DNA sites not confined by fixcoor (Qdnafree) */
#define BROMOC_QDIFFUSE    (1u << 28)  /* DYNAMICS1: pore-switched diffusion (Qdiffuse) */
#define BROMOC_QPROFILE    (1u << 29)  /* DYNAMICS2: spline diffusion profile (Qprofile) */
#define BROMOC_QRFPAR      (1u << 25)  /* reaction-field parameter grids, rfpar.f90 (Qrfpar) */
#define BROMOC_QRFPSIN     (1u << 30)  /* one reaction-field grid for every ion type (Qrfpsin) */
#define BROMOC_QBSPLINE    (1u << 18)  /* repulsion map by 27-point B-spline (Qbspline:
                                        * vdwgd1spln.f90, vdwgd0spln.f90) instead of the
                                        * trilinear stencil; needs QPHIV */
#define BROMOC_QPROXDIFF   (1u << 21)  /* DYNAMICS3: DNA-proximity diffusion (Qproxdiff) */
#define BROMOC_QNOTRANS    (1u << 22)  /* NOTRANS: DNA centroid reset after each step of
                                        * bromoc_step (simul1.f90:382-386) */
#define BROMOC_QAPFOR      (1u << 23)  /* APFOR: constant forces on sites (Qapfor) */
#define BROMOC_QCONTRANS   (1u << 24)  /* CONTRANS: harmonic site/centroid restraints */
#define BROMOC_QMEMB       (1u << 16)  /* analytic membrane (membrane.f90) */
#define BROMOC_QMMIJ       (1u << 17)  /* MMIJ reaction field, RECTBOX (rect_rf1/rect_rf0);
                                        * the SPHERE shape (sphe_rf1) is rejected */
#define BROMOC_QDEBYHYB    (1u << 20)  /* hybrid Debye (Qdebyhyb): P-P pairs screened when
                                        * both sites are outside the box; needs QDNAFREE */
#define BROMOC_QNMCDEN     (1u << 19)  /* per-species phiv maps (Qnmcden): phiv.data holds
                                        * nclx*ncly*nclz bytes per map, the six site maps
                                        * S, P, Ab, Tb, Cb, Gb (with QNUCL) then one per ion
                                        * type nold+1..ntype (with QPAR); needs QPHIV, not
                                        * with QSVDW */
#define BROMOC_QSPHERE     (1u << 26)  /* spherical system of radius rsphe around (cx, cy, cz):
                                        * FIXCOOR reflects inward (fixcoor.f90:24-31), OUTBOX,
                                        * MOVE, INSERT (rmin..rmax shells) and COUNT test the
                                        * sphere (grandmod.f90:79-84, move.f90:39-41,
                                        * insert.f90:35-37, count.f90:52-64) */
#define BROMOC_QECYL       (1u << 27)  /* elliptic-cylinder system: semi-axes 1/iecx, 1/iecy
                                        * around (cx, cy), z in lz2m..lz2p (fixcoor.f90:33-46,
                                        * grandmod.f90:85-90, move.f90:42-43, insert.f90:38-39) */

/* site names (nucleotmod namsite) */
#define BROMOC_SITE_S   0
#define BROMOC_SITE_P   1
#define BROMOC_SITE_AB  2
#define BROMOC_SITE_TB  3
#define BROMOC_SITE_CB  4
#define BROMOC_SITE_GB  5

/* Parameters of one pair type `is` (nindex.f90).  Fields follow efpmod/extramod/grandmod:
 * Qefpot, Qchr, Qlj, Qsrpmfi, fct, epp4, sgp2, c0..c4, dmi, dm2(1:2), sc(1:3). */
typedef struct bromoc_pair {
    int32_t efpot, chr, lj, srpmf;
    int32_t nxi;         /* 0-based index of this pair's first knot in ep */
    int32_t nknots;      /* knots of this pair's table (0 when efpot == 0) */
    double  fct, epp4, sgp2;
    double  c0, c1, c2, c3, c4;
    double  dmi, dm2lo, dm2hi;
    double  sc1, sc2, sc3;
} bromoc_pair;

/* A regular grid as read by readphi.f90:43-49,160-170. */
typedef struct bromoc_grid {
    int32_t     nclx, ncly, nclz;
    int32_t     nmaps;   /* phiv: maps in data -- with NMCDEN exactly 6 (with CG-DNA) plus one
                          * per ion type nold+1..ntype (with PARTICLE), else 0 or 1 (one map);
                          * phix and the reaction-field grids: 0 */
    double      dcel;
    double      xbcen, ybcen, zbcen;
    const void *data;    /* phix: float[ncl^3]; phiv: int8_t[ncl^3] per map */
} bromoc_grid;

/* Complete ENERGY/DYNAMICS state.  Pointers are read during bromoc_create only. */
typedef struct bromoc_system {
    int32_t abi_version;
    int32_t ntot, nsites, nfix;
    int32_t capacity;            /* device slots for particles (>= ntot; GCMC growth) */
    int32_t ntype, nold, nttyp;  /* atnam types, nucleotide types, atnam2 types */
    int32_t npair;               /* pair types 1..npair */
    uint32_t flags;              /* BROMOC_Q* */
    /* per particle */
    const double  *x, *y, *z;    /* [ntot] */
    const int32_t *typei;        /* [ntot] signed atnam type; entries < nsites unused */
    const int32_t *typtyp;       /* [nsites] atnam2 type of each site */
    const int32_t *namsite;      /* [nsites] BROMOC_SITE_* */
    const int8_t  *stfree;       /* [nsites] site moves / receives pair forces */
    /* per type */
    const int32_t *nwtype;       /* [ntype+1] atnam -> atnam2 */
    const double  *cg;           /* [ntype+1] charge */
    const double  *fact1a;       /* [ntype+1] D*dt/kT (simul1.f90:347) */
    const double  *fact2a;       /* [ntype+1] sqrt(2*D*dt) (simul1.f90:348) */
    const int32_t *pairidx;      /* [(nttyp+1)^2] nindex(a,b), 0 where undefined */
    const bromoc_pair *pairs;    /* [npair+1], entry 0 unused */
    int32_t nknots, reserved0;
    const double  *ep;           /* [3*nknots] (ep1,ep2,ep3) per knot, efpmod ep(3,:) */
    double res, ires;            /* table spacing and its inverse (efpmod) */
    double rth, srpx, srpk, srpy;/* SRPMF globals (grandmod) */
    /* CG-DNA (nucleotmod) */
    double cgnuc, epsnuc, scalepairing, fctn, kappa, ikappa, epsolv;
    int32_t nbond, nangle, ndihe, nstack, nbp, nex, nqq, nsolv;
    const int32_t *sitebond;  const double *distbond;   /* [2*nbond], [nbond] */
    const int32_t *siteangle; const double *valangle;   /* [3*nangle] */
    const int32_t *sitedihe;  const double *valdihe;    /* [4*ndihe] */
    const int32_t *sitestack; const double *sgstack;    /* [2*nstack] */
    const int32_t *sitebp;    const double *sgbp;       /* [2*nbp] */
    const int32_t *siteex;    const double *sgex;       /* [2*nex] */
    const int32_t *siteqq;                              /* [2*nqq] */
    const int32_t *siteslv;                             /* [2*nsolv] */
    /* external grids (gsbpmod) */
    bromoc_grid phix;            /* static field, float32 */
    bromoc_grid phiv;            /* repulsion map, int8 */
    double svdw, thold8, vzmin, vzmax;
    const double *scal;          /* [nttyp+1] per-type svdw scale (QSVDW) */
    /* Brownian dynamics (grandmod, simul1.f90:342-349) */
    double bdmax, fact0n1, fact0n2;
    double lx2m, lx2p, ly2m, ly2p, lz2m, lz2p;
    int32_t cntpts, reserved1;
    const double *zcont;         /* [cntpts] flux-counting planes */
    /* position-dependent diffusion of the ions (QDIFFUSE: DYNAMICS1, dynamics1.mono.f90;
     * QPROFILE: DYNAMICS2, dynamics2.mono.f90) */
    double dt, kbtdt;            /* time step and dt/kT (simul1.f90:342) */
    const double *diffusion;     /* [ntype+1] grandmod diffusion(itype) */
    double plength2, pcenter;    /* DIFFUSION command: half pore length, pore centre */
    const double *ampl3, *p3;    /* [ntype+1] lowfraction, switchlength per ion type */
    int32_t nspline, reserved2;  /* PROFILE knots (splinemod) */
    const double *xs, *ys;       /* [nspline] knots and values */
    const double *bs, *cs, *ds;  /* [nspline] spline coefficients b, c, d (spline.f90) */
    /* reaction-field parameter grids (QRFPAR, gsbpmod: nclx3.., gsrfen, greff) */
    bromoc_grid rf;              /* geometry; data unused */
    const float *gsrfen, *greff; /* [ngrid * nclx*ncly*nclz], ngrid = 1 with QRFPSIN, else
                                  * ntype - nold (one grid per ion type, in type order) */
    double sqrfac;               /* scale of greff (gsbpmod sqrfac) */
    /* APFOR (apfmod, energy.mono.f90:510-518): force af(1:3,k) on site sn(k) */
    int32_t afn, ctn;            /* APFOR entries; CONTRANS entries */
    const int32_t *sn;           /* [afn] 0-based site index */
    const double  *af;           /* [3*afn] */
    /* CONTRANS (nucleotmod, energy.mono.f90:520-580): entry k restrains site csn(k)
     * (1-based as the reference stores it; 0 = the centroid of all sites, or with unsplit
     * of each strand) along the axes whose k* is nonzero */
    const int32_t *csn;          /* [ctn] */
    const double  *kx, *ky, *kz; /* [ctn] */
    const double  *contrx, *contry, *contrz;   /* [ctn+1]; entry ctn: the second strand */
    int32_t unsplit, nsites1st;  /* Qunsplit; sites of the first strand */
    /* NOTRANS (nucleotmod notr*): axes bit 0 x, 1 y, 2 z (Qnotrx, Qnotry, Qnotrz) */
    int32_t notrans_axes, reserved3;
    double notrx, notry, notrz;
    /* PROXDIFF (grandmod): 1/beta, diff0, diffcutoff of DYNAMICS3 */
    double ibeta, diff0, diffcutoff;
    double thold27;              /* B-spline wall threshold (vdwgd1spln.f90:151) */
    /* MEMBRANE (grandmod, membrane.f90:19-107): Nernst transmembrane potential and a
     * membrane slab zmemb1..zmemb2 with switched walls, optionally pierced by a cylindrical
     * pore (Qpore); the Debye factor is ikappa above.  Per atnam2 type: ampl1, ampl2,
     * rcylinder [nttyp+1]; p1, p2 [2*(nttyp+1)] as (p, 1/p) pairs */
    int32_t qpore, reserved4;
    double voltage, afact, ceps, zmemb1, zmemb2;
    const double *ampl1, *p1, *ampl2, *p2, *rcylinder;
    /* MMIJ reaction field (QMMIJ, gsbpmod; readmmij.f90, rect_norm.f90, rect_rf1.f90,
     * rect_rf0.f90): shape 1 = RECTBOX, the only one supported (SPHERE, sphe_rf1.f90, is
     * rejected).  ntpol = xnpol*ynpol*znpol basis functions, each a product of Legendre
     * polynomials of degrees lstpx/lstpy/lstpz (0-based degrees, < xnpol etc.; every npol
     * >= 3) in the scaled coordinates xscal*x, ..., normalised by bnorm (rect_norm.f90);
     * the region is |x| <= 1/xscal, |y| <= 1/yscal, |z| <= 1/zscal; mmij [ntpol*ntpol] is
     * the symmetric reaction-field matrix, row-major, both triangles filled (as readmmij.f90
     * leaves it); rfscal scales the energy. */
    int32_t mmij_shape, xnpol, ynpol, znpol, ntpol, reserved5;
    const int32_t *lstpx, *lstpy, *lstpz;   /* [ntpol] */
    const double *bnorm;                     /* [ntpol] */
    const double *mmij;                      /* [ntpol*ntpol] */
    double xscal, yscal, zscal, rfscal;
    /* curved systems (grandmod cx, cy, cz, rsphe, rsphe2, iecx, iecy): the centre, with
     * QSPHERE the radius and its square as the reference holds both, with QECYL the
     * inverse semi-axes.  lx2m..lz2p stay the bounding box (shell_simul.f90:875-895). */
    double cx, cy, cz, rsphe, rsphe2, iecx, iecy;
} bromoc_system;

/* Energy terms as in grandmod/gsbpmod/nucleotmod (energy.mono.f90:46-68,689). */
typedef struct bromoc_energies {
    double ener;
    double eelec, evdw, esrpmf, esrpmfmx, eefpot, eefpotmx, eqqmx, evdwmx;
    double ebond, eang, edihe, estack, ebp, eex, eqq, esolv;
    double egsbpa, egsbpb, evdwgd, erfpar, econ, ememb;
} bromoc_energies;

/* GCMC buffers (grandmod: nbuffer, ibfftyp, avnum, mu, kb, LZmin, LZmax, Qbufferbias,
 * ntotat; box lx, ly, cx, cy, cz; ikbt). */
typedef struct bromoc_gcmc {
    int32_t nbuffer, reserved;
    const int32_t *ibfftyp;      /* [nbuffer+1] ion type of each buffer */
    const double  *avnum, *mu, *kb, *lzmin, *lzmax;   /* [nbuffer+1] */
    const int8_t  *bufferbias;   /* [nbuffer+1] */
    const int32_t *ntotat;       /* [nbuffer+1] */
    double lx, ly, cx, cy, cz, ikbt;
    const double  *rmin, *rmax;  /* [nbuffer+1] spherical shells of the buffers (QSPHERE;
                                  * may be null otherwise) */
} bromoc_gcmc;

typedef struct bromoc_ctx bromoc_ctx;

/* noise modes of bromoc_step / bromoc_dynamics */
#define BROMOC_NOISE_INJECTED 0   /* Gaussians uploaded with bromoc_upload_noise */
#define BROMOC_NOISE_PHILOX   1   /* counter-based Philox4x64-10 + Box-Muller */

/* Create a device context on HIP device `device` from `sys` (copied to HBM). */
int bromoc_create(const bromoc_system *sys, int device, bromoc_ctx **out);
int bromoc_destroy(bromoc_ctx *ctx);
/* Human-readable text of the last error on this context (static storage). */
const char *bromoc_last_error(const bromoc_ctx *ctx);

/* Positions and types: ntot may change (GCMC); typei entries for sites are ignored. */
int bromoc_set_positions(bromoc_ctx *ctx, int32_t ntot, const double *x, const double *y,
                         const double *z, const int32_t *typei);
int bromoc_get_positions(bromoc_ctx *ctx, double *x, double *y, double *z, int32_t *typei,
                         int32_t *ntot);
int bromoc_get_forces(bromoc_ctx *ctx, double *fx, double *fy, double *fz);

/* ENERGY: forces (if want_forces) and energy terms for the current positions. */
int bromoc_energy(bromoc_ctx *ctx, int32_t want_forces, bromoc_energies *out);

/* Injected Gaussian noise for nsteps steps, layout [step][particle i][xyz] over ntot
 * particles (entries of non-moving particles are ignored).  Kept in HBM. */
int bromoc_upload_noise(bromoc_ctx *ctx, int32_t nsteps, const double *noise);

/* nsteps x { ENERGY ; DYNAMICS0(ions) ; DYNAMICS0NUC(sites) }.  With injected noise the
 * step s uses noise row (first_row + s).  With Philox, the Gaussian for (step, i, c) is
 * a function of (seed, counter0 + s, i, c) only.  energies (optional) receives the terms
 * of the last step's ENERGY. */
int bromoc_step(bromoc_ctx *ctx, int32_t nsteps, int32_t noise_mode, int32_t first_row,
                uint64_t seed, uint64_t counter0, bromoc_energies *energies);

/* The spatial order the step kernels use, rebuilt from the current positions: keys[ntot]
 * the sorted keys (the particle's class in bits 30-31 -- 0 CG-DNA site, 1 fixed ion,
 * 2 moving ion -- over the Morton code of its 8 A cell of the box, 10 bits per axis),
 * order[ntot] the particle in each slot (a stable sort: equal keys keep index order),
 * rank[ntot] the slot of each particle, cell_start[ncell+1] the first slot of each occupied
 * cell (then ntot) and *ncell.  Any output may be NULL.  No counterpart in the reference,
 * whose pair loops run over all pairs in index order. */
int bromoc_spatial_order(bromoc_ctx *ctx, uint32_t *keys, int32_t *order, int32_t *rank,
                         int32_t *cell_start, int32_t *ncell);

/* Measurement: nsteps steps as bromoc_step (launched one by one, no graph), with a HIP
 * event pair around every launch of the step kernel on the library's stream; *total_ms is
 * first-to-last, *kernel_ms the mean step-kernel duration.  No counterpart in the reference (it times runs with date_and_time,
 * cputime.f90:19-39). */
int bromoc_time_steps(bromoc_ctx *ctx, int32_t nsteps, int32_t noise_mode, int32_t first_row,
                      uint64_t seed, uint64_t counter0, double *total_ms, double *kernel_ms);

/* Launch shape picked by the library's autotuner for this system: 1, 2 or 4 targets per
 * block of the block-per-target step kernel, or 8 for the wavefront-per-target kernel;
 * runs the tuner if it has not run.  No counterpart in the reference. */
int bromoc_get_tuning(bromoc_ctx *ctx, int32_t *targets_per_block);

/* One DYNAMICS pass with the current forces (as left by bromoc_energy). */
int bromoc_dynamics(bromoc_ctx *ctx, int32_t noise_mode, int32_t row, uint64_t seed,
                    uint64_t counter);

/* Flux counters nforward/nbackward, [ntype-nold][cntpts] row-major (countions.f90). */
int bromoc_get_flux(bromoc_ctx *ctx, int32_t *nforward, int32_t *nbackward);
/* Per-atnam2-type wall warnings warn(1:nttyp) (vdwgd1trln.f90:143-153), [nttyp+1]. */
int bromoc_get_warn(bromoc_ctx *ctx, int32_t *warn);

/* INTERACT: energy of particle j (0-based; j == ntot means an insertion trial) at
 * (xj,yj,zj) with atnam type jtype against the rest of the system. */
int bromoc_interact(bromoc_ctx *ctx, double xj, double yj, double zj, int32_t jtype,
                    int32_t j, int32_t alert, double *dener);

/* GCMC buffers for bromoc_grand. */
int bromoc_set_gcmc(bromoc_ctx *ctx, const bromoc_gcmc *g);
/* Buffer membership ibuffer(1:ntot) of grandmod (0 = none), 0-based particle order. */
int bromoc_set_ibuffer(bromoc_ctx *ctx, const int32_t *ibuffer);
int bromoc_get_ibuffer(bromoc_ctx *ctx, int32_t *ibuffer);
/* GRAND(ngcmc): uniforms[] supplies every rndm()/rndm1() draw in call order; *used
 * returns how many were consumed.  nat[nbuffer+1] is read and updated, prob is
 * [nbuffer][capacity+1] (prob(ib,nat) of grand.f90:114-116), ener is updated. */
int bromoc_grand(bromoc_ctx *ctx, int32_t ngcmc, const double *uniforms, int64_t nuniforms,
                 int64_t *used, int32_t *nat, int32_t *prob, int64_t icycle, double *ener);

/* METROPOLIS(nmcm) (metropolis.f90:31-55 with MOVE, move.f90:32-47; box, sphere or cylinder):
 * nmcm trials, each moving one random moving ion by up to mcmax/2 per axis (mcmax as
 * grandmod holds it: twice the MCMAX keyword), accepted when the INTERACT energy falls or
 * with probability exp(-dE ikbt).  uniforms[] supplies every rndm() in call order.  A trial
 * runs only when its MOVE draws and one more uniform are available: otherwise the call
 * stops there with *done < nmcm (no state changed by the unfinished trial); *used counts
 * the uniforms of the finished trials.  *ener is updated as grandmod ener. */
int bromoc_metropolis(bromoc_ctx *ctx, int32_t nmcm, double mcmax, double ikbt,
                      const double *uniforms, int64_t nuniforms, int64_t *used, int32_t *done,
                      double *ener);

/* ---------------------------------------------------------------- map files
 * Grid files as the reference reads them: Fortran unformatted sequential records with
 * 4-byte length markers (a negative leading marker continues the record in the next
 * subrecord).  Record 1: nclx, ncly, nclz (int32), dcel, xbcen, ybcen, zbcen (real*8);
 * record 2: six real*8 (PHIX/PHIV: epsw, epsp, conc, tmemb, zmemb, epsm; RFPAR: the Born
 * radius and five unused); then the float32 values, x-major with z fastest. */
typedef struct bromoc_map_header {
    int32_t nclx, ncly, nclz, reserved;
    double dcel, xbcen, ybcen, zbcen;
    double rec2[6];
} bromoc_map_header;

#define BROMOC_MAP_PHIX 0   /* static field: float32 values as stored */
#define BROMOC_MAP_PHIV 1   /* repulsion map: int8, 0 where the file value is nonzero, else 1 */

/* READPHI (readphi.f90:19-349, one map per file, no Qnmcden): reads the map at path.  With
 * adjust the map is cropped to the system box[6] = (lx2m, lx2p, ly2m, ly2p, lz2m, lz2p)
 * plus one node beyond its upper faces, as readphi.f90:98-156 does, and hdr describes the
 * cropped map.  data (float* for PHIX, int8_t* for PHIV) receives nclx*ncly*nclz values
 * of hdr; with data == NULL only hdr is filled in.  capacity: elements data holds. */
int bromoc_read_map(const char *path, int32_t kind, int32_t adjust, const double *box,
                    bromoc_map_header *hdr, void *data, int64_t capacity);

/* READRFPAR (rfpar.f90:373-582): nfiles reaction-field parameter files, one per ion type
 * (or one with QRFPSIN), each holding record 1, (radion, 5 unused), gsrfen and greff on the
 * same lattice (files that differ in geometry are rejected, rfpar.f90:455-463).  radion
 * [nfiles]; gsrfen, greff [nfiles * ncel3] of the (optionally cropped) lattice in hdr;
 * with gsrfen == NULL only hdr and radion are filled in. */
int bromoc_read_rfpar(int32_t nfiles, const char *const *paths, int32_t adjust,
                      const double *box, bromoc_map_header *hdr, double *radion,
                      float *gsrfen, float *greff, int64_t capacity);

/* REPWALLS (repwalls.f90:19-67) on device `device`: successive wall layers on a binarised
 * repulsion map in place, phiv[nclx*ncly*nclz] (host memory): layer n+1 marks the nodes of
 * layer n whose 27-node neighbourhood is all layer n, for n = 1..walls, stopping at the
 * first layer with no node. */
int bromoc_repwalls(int32_t nclx, int32_t ncly, int32_t nclz, int8_t *phiv, int32_t walls,
                    int device);

/* ------------------------------------------------------------- per-cycle analyses
 * The statistics simul1.f90 gathers in its cycle loop, accumulated on the device so the
 * host reads them once at the end instead of copying positions every cycle:
 *   BROMOC_AN_RHO   rho(type, z) of the moving ions          simul1.f90:477-485
 *   BROMOC_AN_RDNA  rDNA(site type, z) of the DNA sites      simul1.f90:487-495
 *   BROMOC_AN_GR    g(r) of the moving ions around particle igr, and irmax
 *                                                            simul1.f90:497-509
 *   BROMOC_AN_PROB  prob2 (ions in the channel) and prob3 (ions in the system) number
 *                   distributions per type                   simul1.f90:511-538
 *   BROMOC_AN_PRES  virial pressure of every bromoc_step step simul1.f90:365-368
 * bromoc_analyze adds one cycle's histograms from the current positions (the reference
 * does this after the cycle's BD steps and COUNT); the pressure is added by every step of
 * bromoc_step while BROMOC_AN_PRES is set (bromoc_dynamics, one DYNAMICS call, adds none).
 * Bins follow the reference: iz = int((z - zmini) * idelz) + 1 kept when iz <= nzmax,
 * ir = int(dist / dr) + 1 kept when ir <= npoints, a count n kept when n <= datom.  The
 * reference writes out of its arrays for iz < 1 (a site or ion more than 1/idelz below
 * zmini) or a count above datom; those samples are left out here. */
#define BROMOC_AN_RHO   1
#define BROMOC_AN_RDNA  2
#define BROMOC_AN_GR    4
#define BROMOC_AN_PROB  8
#define BROMOC_AN_PRES  16

typedef struct bromoc_analysis {
    int32_t what;        /* BROMOC_AN_* */
    int32_t nzmax;       /* z bins (simul1.f90:183-188: int(lz*idelz), at most 10000) */
    double  zmini;       /* lower z of bin 1 (lz2m) */
    double  idelz;       /* bins per Angstrom (2.0) */
    int32_t igr;         /* 0-based index of the g(r) centre particle */
    int32_t npoints;     /* g(r) shells (10000) */
    double  dr;          /* shell width (0.1) */
    int32_t datom;       /* largest count of prob2/prob3 (grandmod datom, 23000) */
    int32_t reserved;
    double  czmin, czmax;/* channel of prob2 without a membrane; both 0: no prob2 */
    double  tvol;        /* system volume for the pressure (Angstrom^3) */
    double  temp;        /* K */
    double  ikbt;        /* 1/kBT in (kcal/mol)^-1, shell_simul.f90:865-866 */
} bromoc_analysis;

/* Set the analyses and zero their accumulators (also the pressure). */
int bromoc_set_analysis(bromoc_ctx *ctx, const bromoc_analysis *an);
/* One cycle's histograms (rho, rDNA, g(r), prob2/prob3 as selected) from the current
 * positions; enqueued on the context's stream, no host synchronisation. */
int bromoc_analyze(bromoc_ctx *ctx);
/* The accumulators (any pointer may be NULL): rho[ntype*nzmax] (row t-1: atnam type t),
 * rdna[nttyp*nzmax] (row t-1: atnam2 type t), gr[ntype*npoints], irmax,
 * prob2[ntype*(datom+1)], prob3[ntype*(datom+1)] (row t-1, column n), the reference's
 * pres sum over the steps and the number of steps it holds (the reference prints
 * pres / steps, simul1.f90:656). */
int bromoc_get_analysis(bromoc_ctx *ctx, int64_t *rho, int64_t *rdna, int64_t *gr,
                        int32_t *irmax, int64_t *prob2, int64_t *prob3, double *pres,
                        int64_t *pres_steps);

/* EFPOT table from (r,U) samples at uniform spacing: writes ep[3*n] and pair fields
 * dmi, dm2lo, dm2hi, sc1..sc3 (host-side, splinepot.f90:19-96). */
int bromoc_build_table(int32_t n, const double *r, const double *u, int32_t chr, double fct,
                       double *ep, bromoc_pair *pair);

/* EFPOT table built from a pair type's analytic parameters (discretize.f90:19-120, the
 * EFPOT ... build command, shell_simul.f90:1840-1853,1869-1871): params' chr, fct, lj,
 * epp4, sgp2, srpmf and c0..c4 sampled at r = k*res for k = nnp..mnp (Coulomb fct/r of a
 * charged pair, LJ, SRPMF where r^2 <= rth with the srpx/srpk/srpy smoothing), through
 * squarespline; the wall takes the table's value and the sampled slope at the first knot.
 * Writes ep[3*(mnp-nnp+1)] and out->sc1..sc3; dmi and dm2 are the EFPOT command's
 * (nnp*res and its square, (mnp*res)^2; the reference recomputes nnp = int(dmi*ires)
 * before calling DISCRETIZE, shell_simul.f90:1869, and callers do the same).  Host-side. */
int bromoc_discretize(const bromoc_pair *params, int32_t nnp, int32_t mnp, double res,
                      double rth, double srpx, double srpk, double srpy, double *ep,
                      bromoc_pair *out);

/* Cubic-spline coefficients of the diffusion profile (spline.f90:19-98, the
 * Forsythe-Malcolm-Moler SPLINE): b, c, d of n >= 2 knots x (strictly monotone), y.
 * Host-side, as shell_simul.f90:1968 calls it for the PROFILE command. */
int bromoc_spline(int32_t n, const double *x, const double *y, double *b, double *c, double *d);

/* n raw Philox4x64-10 blocks computed on the device (the generator behind
 * BROMOC_NOISE_PHILOX): out[4k..4k+3] = philox(ctr[4k..4k+3], key[2k..2k+1]).  No
 * counterpart in the reference (its noise is libgfortran rand()). */
int bromoc_philox4x64(int32_t n, const uint64_t *ctr, const uint64_t *key, uint64_t *out);

/* ------------------------------------------------------------- trajectory adjoint
 * Gradient of a loss on the smoothed radial distribution functions of a BD trajectory
 * with respect to the pair tables, by a reverse pass through the steps (no counterpart in
 * the reference, whose potentials are fitted by the separate inverse-Monte-Carlo
 * program).  The trajectory: nsteps steps of ENERGY + DYNAMICS0 (moving ions, every pair
 * tabulated; the static field phix and the trilinear repulsion map phiv, with QSVDW and
 * NMCDEN, allowed; no B-spline map, no CG-DNA) from the current positions with the injected
 * noise rows
 * 0..nsteps-1.  The RDF of group g (a pair of atnam ion types) in bin b is
 *   g[g][b] = V / (nsteps * P_g * 4 pi r_b^2 sigma sqrt(2 pi)) * sum over the steps' states
 *             x_1..x_nsteps and the group's pairs of exp(-(r - r_b)^2 / (2 sigma^2)),
 * terms with |r - r_b| > cut left out, r_b = r0 + (b + 1/2) dr, V the box volume, P_g the
 * number of the group's pairs (moving-moving, or moving-fixed).  The flux is the charge
 * current through the box, J = sum over the moving ions of q_i (z_i(x_nsteps) - z_i(x_0)) /
 * (nsteps dt) in e A / ps (the smooth counterpart of the reference's plane-crossing counts,
 * countions.f90:19-32).  The loss is
 * L = sum_{g,b} (g[g][b] - target[g][b])^2 + wflux (J - jtarget)^2. */
typedef struct bromoc_rdf {
    int32_t ngroup, nbins;
    const int32_t *groups;       /* [2*ngroup] atnam types (a, b) of each group */
    double r0, dr, sigma, cut;
    const double *target;        /* [ngroup*nbins], NULL: zero */
    double wflux, jtarget;       /* weight of the flux term (0: none) and its target */
} bromoc_rdf;

/* Forward pass (storing the positions every checkpoint_every steps) and, when grad_ep is
 * given, the reverse pass recomputing each segment from its checkpoint.  Outputs: *loss,
 * rdf[ngroup*nbins] (optional), grad_ep[3*nknots] = dL/d(ep1, ep2, ep3) of every knot,
 * grad_sc[3*(npair+1)] = dL/d(sc1, sc2, sc3) of every pair type.  Positions end at the
 * trajectory's last state; flux counters are not updated. */
int bromoc_adjoint_rdf(bromoc_ctx *ctx, int32_t nsteps, int32_t checkpoint_every,
                       const bromoc_rdf *rdf, double *loss, double *rdf_out, double *grad_ep,
                       double *grad_sc);

/* Tangent-linear pass of the same trajectory: dloss = <dL/d(ep, sc), v>, the derivative of
 * the loss along the direction v_ep[3*nknots] (per knot: ep1, ep2, ep3) and
 * v_sc[3*(npair+1)] (per pair type: sc1, sc2, sc3), carried forward step by step with the
 * trajectory (the ep1 and sc2 entries do not move the forces and take no part).  With the
 * same checkpoint_every the trajectory is bitwise the one bromoc_adjoint_rdf runs, so
 * <grad, v> of the reverse pass equals dloss to rounding for every v (the dot-product
 * test of the adjoint).  Also returns the loss.  Positions end at the last state. */
int bromoc_tangent_rdf(bromoc_ctx *ctx, int32_t nsteps, int32_t checkpoint_every,
                       const bromoc_rdf *rdf, const double *v_ep, const double *v_sc,
                       double *loss, double *dloss);

/* Chain rule through bromoc_build_table for one pair type's table of n samples at r: the
 * table (ep[3n], sc1..sc3) is affine in the sample values u (the charge term is the
 * offset), so grad_u[k] = sum of grad_ep . d ep / du_k + grad_sc . d(sc1, sc2, sc3) / du_k,
 * computed by one reverse sweep through the construction's stages (O(n)). */
int bromoc_table_vjp(int32_t n, const double *r, const double *grad_ep, const double *grad_sc,
                     double *grad_u);
/* The same product on the context's device (the same reverse sweep, one thread);
 * bit-identical to bromoc_table_vjp. */
int bromoc_table_vjp_device(bromoc_ctx *ctx, int32_t n, const double *r, const double *grad_ep,
                            const double *grad_sc, double *grad_u);
/* The same product, results bit-identical, with no allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (a fitting loop under HIP graph
 * capture, a serving loop).  Inputs, outputs and the sweep's workspace (11n + 3 doubles)
 * are staged in a fixed 16 MiB device array per GPU, so n <= 190649 (BROMOC_E_CAPACITY
 * beyond).  Every
 * call on one GPU shares that array: calls that overlap on one GPU race. */
int bromoc_table_vjp_device_noalloc(bromoc_ctx *ctx, int32_t n, const double *r,
                                    const double *grad_ep, const double *grad_sc,
                                    double *grad_u);
/* Rebuild ntab tables of the context's pair types on its device from new samples
 * (SPLINEPOT, splinepot.f90:19-96, the construction bromoc_build_table runs on the host,
 * bit-identical), in place: table t is pair type is[t] with n[t] samples (its knot count
 * at bromoc_create) at r, u (all tables' samples concatenated); its Coulomb tail keeps
 * the pair's chr and fct.  The potential-fitting loop's update of the knot values, with
 * no new context and no host copy of the knots.  No counterpart in the reference, which
 * builds its tables once per input deck. */
int bromoc_update_tables(bromoc_ctx *ctx, int32_t ntab, const int32_t *is, const int32_t *n,
                         const double *r, const double *u);

/* ------------------------------------------------------------------ replica batches
 * Many independent ion systems (no CG-DNA, no reaction-field grids) that share one
 * bromoc_system's types, pair tables, repulsion map, box, flux planes and BD constants, and
 * differ in their particles, their static field and their GCMC buffers.  Every replica
 * runs the same loop as simul1.f90:352-390 -- GRAND (grand.f90:19-121) every gcmc_every
 * steps, then ENERGY + DYNAMICS per step -- and all replicas advance together: one launch
 * per BD step covers every replica, its waves taking (replica, ion) work items from a
 * device queue ordered by replica size; one workgroup per replica runs its GCMC attempts.
 * No counterpart in the reference (one system per process); results per replica are those
 * of the single-system entry points with Philox stream r (replica r draws its Gaussians
 * with counter (step, ion, r, 0), its GCMC uniforms with counter (cycle, k/4, r, 1): word
 * k%4 of the block, (w >> 11) * 2^-53). */
typedef struct bromoc_replica {
    int32_t ntot, nfix, capacity, reserved;  /* particles now, leading fixed ions, slots */
    const double  *x, *y, *z;                /* [ntot] */
    const int32_t *typei;                    /* [ntot] signed atnam type (ions only) */
    const float   *phix;                     /* own static field on the shared phix
                                              * geometry (QPHIX), [nclx*ncly*nclz] */
    const bromoc_gcmc *gcmc;                 /* own buffers, or NULL for no GCMC */
    const int32_t *ibuffer;                  /* [ntot] buffer of each particle (gcmc) */
    const int32_t *nat;                      /* [nbuffer+1] particles per buffer (gcmc) */
} bromoc_replica;

typedef struct bromoc_batch bromoc_batch;

int bromoc_batch_create(const bromoc_system *shared, int32_t nrep, const bromoc_replica *reps,
                        int device, bromoc_batch **out);
int bromoc_batch_destroy(bromoc_batch *b);
const char *bromoc_batch_last_error(const bromoc_batch *b);
/* nsteps BD steps of every replica with Philox noise (seed, step counters counter0..);
 * with gcmc_every > 0, ngcmc GRAND attempts per replica before every gcmc_every-th step
 * (simul1.f90:353-356 order: GRAND, then the cycle's nbd steps), GCMC cycles numbered
 * from cycle0.  energies (optional, [nrep]) receive each replica's ENERGY terms of the
 * last step. */
int bromoc_batch_run(bromoc_batch *b, int32_t nsteps, uint64_t seed, uint64_t counter0,
                     int32_t gcmc_every, int32_t ngcmc, uint64_t cycle0,
                     bromoc_energies *energies);
/* ENERGY of every replica at its current positions; forces stay on the device. */
int bromoc_batch_energy(bromoc_batch *b, bromoc_energies *energies);
/* Replica r's particles (x, y, z, typei up to capacity entries), ntot, buffer membership,
 * nat[nbuffer+1], GCMC energy accumulator (GRAND's ener), any may be NULL. */
int bromoc_batch_get(bromoc_batch *b, int32_t r, double *x, double *y, double *z,
                     int32_t *typei, int32_t *ntot, int32_t *ibuffer, int32_t *nat,
                     double *gcmc_ener);
int bromoc_batch_get_forces(bromoc_batch *b, int32_t r, double *fx, double *fy, double *fz);
/* flux counters [ntype-nold][cntpts] and prob[nbuffer][capacity+1] of replica r */
int bromoc_batch_get_flux(bromoc_batch *b, int32_t r, int32_t *nforward, int32_t *nbackward);
int bromoc_batch_get_prob(bromoc_batch *b, int32_t r, int32_t *prob);
/* Measurement: nsteps steps launched one by one (no graph, no GCMC) with a HIP event pair
 * around each step kernel; mean step-kernel duration in *kernel_ms. */
int bromoc_batch_time_steps(bromoc_batch *b, int32_t nsteps, uint64_t seed, uint64_t counter0,
                            double *total_ms, double *kernel_ms);
/* Launch shape the batch autotuner chose: workgroups of the persistent step kernel (0: one
 * work item per lane group, no queue) and lanes per target (16, 32 or 64). */
int bromoc_batch_get_tuning(bromoc_batch *b, int32_t *blocks, int32_t *lanes);

#ifdef __cplusplus
}
#endif
#endif /* BROMOC_GPU_H */
