// api.hip -- C ABI of libbromoc_gpu.so (declared in include/bromoc_gpu.h).
//
// Host side of the gfx950 path: validates the reference state handed over by the caller
// (every index a kernel will form is bounds-checked here, before any launch), lays it out
// in HBM, and drives the kernels of kernels.hip.  GRAND's accept/reject loop runs here
// around the INTERACT kernel, in the reference's draw order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bromoc_gpu.h"
#include "engine.h"
#include "tables.h"

using namespace bromoc;

namespace {

constexpr int GRAPH_K = 32;   // BD steps per captured graph (even)

constexpr uint32_t SUPPORTED = BROMOC_QPAR | BROMOC_QNUCL | BROMOC_QNONBOND | BROMOC_QNOBOND |
                               BROMOC_QPHIX | BROMOC_QPHIV | BROMOC_QSVDW | BROMOC_QSOLV |
                               BROMOC_QDEBY | BROMOC_QCOUNTION | BROMOC_QDNAFREE |
                               BROMOC_QDIFFUSE | BROMOC_QPROFILE | BROMOC_QRFPAR |
                               BROMOC_QRFPSIN | BROMOC_QBSPLINE | BROMOC_QPROXDIFF |
                               BROMOC_QNOTRANS | BROMOC_QAPFOR | BROMOC_QCONTRANS |
                               BROMOC_QMEMB | BROMOC_QMMIJ | BROMOC_QDEBYHYB | BROMOC_QNMCDEN |
                               BROMOC_QSPHERE | BROMOC_QECYL;

// The system's shape for OUTBOX, FIXCOOR, MOVE and INSERT (grandmod.f90:57,74-97)
static Curved curved_of(const bromoc_system *s) {
    Curved v{};
    v.kind = (s->flags & BROMOC_QSPHERE) ? 1 : (s->flags & BROMOC_QECYL) ? 2 : 0;
    v.cx = s->cx; v.cy = s->cy; v.cz = s->cz;
    v.rsphe = s->rsphe; v.rsphe2 = s->rsphe2; v.iecx = s->iecx; v.iecy = s->iecy;
    v.lx2m = s->lx2m; v.lx2p = s->lx2p; v.ly2m = s->ly2m; v.ly2p = s->ly2p;
    v.lz2m = s->lz2m; v.lz2p = s->lz2p;
    v.sx = v.kind == 1 ? 1.0 : v.kind == 2 ? s->iecx : 0.0;
    v.sy = v.kind == 1 ? 1.0 : v.kind == 2 ? s->iecy : 0.0;
    v.sz = v.kind == 1 ? 1.0 : 0.0;
    v.thr = v.kind == 1 ? s->rsphe2 : v.kind == 2 ? 1.0 : HUGE_VAL;
    return v;
}

// Repulsion maps the phiv grid holds: one, or with NMCDEN (readphi.f90:176-240) six site
// maps (S, P, Ab, Tb, Cb, Gb) with DNA followed by one per ion type nold+1..ntype.
static int64_t phiv_maps(const bromoc_system *s) {
    if (!(s->flags & BROMOC_QNMCDEN)) return 1;
    return ((s->flags & BROMOC_QNUCL) ? 6 : 0) + ((s->flags & BROMOC_QPAR) ? s->ntype - s->nold : 0);
}

// constamod.f90:21-32 (binary64)
double celec_value() {
    const double pi = 3.14159265358979323846264338327950288419716939937510;
    const double kcalmol = 4184.0 / 6.0221412927E23;
    const double coulomb = 1.60217656535E-19, epermit = 8.854187817620E-12;
    return coulomb * coulomb / (4.0 * pi * epermit * kcalmol * 1E-10);
}

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
};

}  // namespace

struct bromoc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;          // ion targets run here, beside the site work
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::string err;
    bromoc_system sys{};       // scalars (pointer members are not kept)
    int cap = 0, ntot = 0, ns = 0, nf = 0;
    // device state
    double4 *pos[2] = {nullptr, nullptr};   // (x, y, z, atnam2 type | index) ping-pong buffers
    double4 *spos[2] = {nullptr, nullptr};  // the same, in spatial order
    SortBufs sb{};
    bool sort_dirty = true;
    int since_sort = 0;
    double r2near = 0.0;
    int cur = 0;
    int tpb = 0;                             // targets per block chosen by the autotuner
                                             // (TPB_TILE + targets per lane: tile mode)
    double *tpart = nullptr, *tener = nullptr;   // tile-mode partial sums and pair energies
    size_t tpart_n = 0, tener_n = 0;
    int32_t *warn_scratch = nullptr;
    int32_t *typei = nullptr, *t2 = nullptr, *namsite = nullptr, *nwtype = nullptr;
    int8_t *stfree = nullptr;
    double *fx = nullptr, *fy = nullptr, *fz = nullptr;
    double4 *fdna = nullptr;
    double *epart = nullptr, *ered = nullptr;
    int32_t *pairidx = nullptr;
    DevPair *pairs = nullptr, *rows = nullptr;
    double4 *ep = nullptr;
    double *cg = nullptr, *fact1a = nullptr, *fact2a = nullptr, *scal = nullptr, *zcont = nullptr;
    int32_t *inc_ptr = nullptr, *inc_split = nullptr, *incr_ptr = nullptr, *incb_ptr = nullptr;
    uint32_t *inc = nullptr, *incb = nullptr;
    double2 *incr = nullptr;
    double4 *fmul = nullptr;
    double *bpar = nullptr;
    double2 *dihcs = nullptr;
    int32_t *lst[8] = {};
    double *lval[6] = {};
    float *phix = nullptr;
    int8_t *phiv = nullptr;
    int32_t *warn = nullptr, *nfwd = nullptr, *nbwd = nullptr;
    double *noise = nullptr;
    int noise_rows = 0;
    double *pbuf = nullptr;     // Philox Gaussians of the current chunk [GRAPH_K][3*cap]
    int64_t noise_stride = 0;
    int64_t *ctl = nullptr;     // [step, row0, counter0, seed]
    int64_t h_ctl[4] = {0, 0, 0, 0};   // host source of the ctl upload (outlives the copy)
    void *pin = nullptr;               // pinned staging of a call's small copies
    size_t pin_cap = 0;
    double4 h_move{};                  // staging of one particle moved by GRAND
    double *iout = nullptr;
    std::vector<void *> allocs;
    // host mirrors
    std::vector<int32_t> h_typei, h_nwtype, h_typtyp;
    std::vector<double> h_cg;
    std::vector<bromoc_pair> h_pairs;   // pair types 0..npair as created (tables updated)
    std::vector<int32_t> h_pairidx;
    // GCMC
    bool have_gcmc = false;
    bromoc_gcmc g{};
    std::vector<int32_t> ibfftyp, ntotat, ibuffer;
    std::vector<double> avnum, mu, kb, lzmin, lzmax, rmin, rmax;
    std::vector<int8_t> bufferbias;
    // device GRAND: buffer parameters, nat, prob, the buffer of each slot, state, the
    // per-workgroup INTERACT partials and the caller's uniforms
    struct GcDev {
        int nb = -1;
        double *avnum = nullptr, *mu = nullptr, *kb = nullptr, *lzmin = nullptr, *lzmax = nullptr;
        double *rmin = nullptr, *rmax = nullptr;
        int32_t *ibfftyp = nullptr, *ntotat = nullptr, *nat = nullptr;
        int32_t *natlog = nullptr;        // [lcap] buffer counts after each attempt
        int64_t lcap = 0;
        std::vector<int32_t> h_natlog;
        int32_t *ibuffer = nullptr;
        bool ib_current = false;          // ibuffer equals the host's copy
        int8_t *bias = nullptr;
        GcState *st = nullptr;
        double *part = nullptr, *u = nullptr;
        int64_t ucap = 0;
        std::vector<void *> mem;
    } gd;
    StepParams base{};
    // per-cycle analyses: integer histograms (freed on re-set), the per-particle virial
    // sums of the running bromoc_step call, their fold, and the steps folded so far
    bromoc_analysis an{};
    unsigned long long *an_rho = nullptr, *an_rdna = nullptr, *an_gr = nullptr;
    unsigned long long *an_prob2 = nullptr, *an_prob3 = nullptr;
    int32_t *an_irmax = nullptr, *an_ncnt = nullptr;
    double *vacc = nullptr, *vsum = nullptr;
    int64_t pres_steps = 0, pres_ntot_steps = 0;
    // captured K-step graphs, one per starting buffer parity
    struct Graph {
        hipGraphExec_t exec = nullptr;
        int ntot = -1, tpb = -1, noise_mode = -1;
        const double *noise = nullptr;
        const double *vacc = nullptr;
        int64_t stride = -1;
    } graph[2];
};

namespace {

int fail(bromoc_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(c, expr)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail((c), BROMOC_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
hipError_t dalloc(bromoc_ctx *c, T **p, size_t n) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return e;
    c->allocs.push_back(q);
    *p = (T *)q;
    return hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(T));
}

template <typename T>
hipError_t dupload(bromoc_ctx *c, T **p, const T *src, size_t n) {
    hipError_t e = dalloc(c, p, n);
    if (e != hipSuccess || n == 0 || !src) return e;
    return hipMemcpy(*p, src, n * sizeof(T), hipMemcpyHostToDevice);
}

void energies_from(const double *r, bool rfpar, bool mmij, bromoc_energies *out) {
    bromoc_energies E;
    double *f = &E.eelec;
    for (int k = 0; k < NE; k++) f[k] = r[k];
    // energy.mono.f90:83-115,689 (egsbpb is the MMIJ reaction field or erfpar)
    E.ener = E.ememb;
    if (mmij) E.ener = E.ener + E.egsbpb;
    E.ener = E.ener + E.egsbpa;
    E.ener = E.ener + E.evdwgd;
    if (rfpar) {
        E.ener = E.ener + E.erfpar;
        E.egsbpb = E.erfpar;
    }
    E.ener = E.ener + E.eelec + E.evdw + E.esrpmf + E.esrpmfmx + E.ebond + E.eang + E.edihe +
             E.estack + E.ebp + E.eex + E.eqq + E.esolv + E.eqqmx + E.evdwmx + E.eefpot +
             E.eefpotmx + E.econ;
    *out = E;
}

int validate(const bromoc_system *s, std::string &msg) {
    auto bad = [&](const std::string &m) { msg = m; return BROMOC_E_ARG; };
    if (s->abi_version != BROMOC_ABI_VERSION) return bad("abi_version mismatch");
    if (s->flags & ~SUPPORTED) {
        msg = "unsupported reference option in flags";
        return BROMOC_E_UNSUPPORTED;
    }
    if (s->ntot < 0 || s->nsites < 0 || s->nfix < 0 || s->nsites + s->nfix > s->ntot)
        return bad("particle counts");
    if (s->capacity < s->ntot || s->capacity > (1 << 26)) return bad("capacity");
    if ((s->flags & BROMOC_QSPHERE) && (s->flags & BROMOC_QECYL))
        return bad("QSPHERE and QECYL together");
    if ((s->flags & BROMOC_QSPHERE) && !(s->rsphe > 0.0 && s->rsphe2 > 0.0))
        return bad("QSPHERE needs rsphe > 0 and rsphe2");
    if ((s->flags & BROMOC_QECYL) && !(s->iecx > 0.0 && s->iecy > 0.0))
        return bad("QECYL needs iecx, iecy > 0");
    if (!s->x || !s->y || !s->z || !s->typei) return bad("positions/typei missing");
    if (s->ntype < 1 || s->nttyp < 1 || s->npair < 0) return bad("type counts");
    if (s->nttyp + 1 > MAX_LDS_TYPES) return bad("too many atnam2 types for the LDS pair table");
    if (s->npair + 1 > MAX_LDS_PAIRS) return bad("too many pair types for LDS");
    if (!s->nwtype || !s->cg || !s->fact1a || !s->fact2a || !s->pairidx || !s->pairs)
        return bad("type tables missing");
    if (s->nsites > 0 && (!s->typtyp || !s->namsite || !s->stfree)) return bad("site arrays");
    if ((s->flags & BROMOC_QNUCL) == 0 && s->nsites > 0) return bad("sites without QNUCL");
    const int nt1 = s->nttyp + 1;
    for (int i = 0; i < nt1 * nt1; i++)
        if (s->pairidx[i] < 0 || s->pairidx[i] > s->npair) return bad("pairidx out of range");
    for (int t = 1; t <= s->npair; t++) {
        const bromoc_pair &p = s->pairs[t];
        if (!p.efpot) continue;
        if (!s->ep || p.nxi < 0 || p.nknots < 2 || (int64_t)p.nxi + p.nknots > s->nknots)
            return bad("pair table outside ep");
        if (!(s->ires > 0.0) || !(p.dm2lo > 0.0) || !(p.dm2hi >= p.dm2lo)) return bad("table range");
        // largest index the lookup can form (polynom.f90:226): x <= sqrt(dm2hi)
        double xmax = std::sqrt(p.dm2hi);
        long long imax = (long long)((xmax - p.dmi) * s->ires);
        if (imax < 0 || imax >= p.nknots) return bad("table range exceeds its knots");
    }
    std::vector<char> site_t2(nt1, 0), ion_t2(nt1, 0);
    for (int i = 0; i < s->nsites; i++) {
        if (s->typtyp[i] < 1 || s->typtyp[i] > s->nttyp) return bad("typtyp out of range");
        if (s->namsite[i] < 0 || s->namsite[i] > 5) return bad("namsite out of range");
        site_t2[s->typtyp[i]] = 1;
    }
    for (int i = s->nsites; i < s->ntot; i++) {
        int t = std::abs(s->typei[i]);
        if (t < 1 || t > s->ntype) return bad("typei out of range");
        int t2 = s->nwtype[t];
        if (t2 < 1 || t2 > s->nttyp) return bad("nwtype out of range");
        ion_t2[t2] = 1;
    }
    // every ion type may meet every site type and every ion type
    for (int a = 1; a < nt1; a++) {
        if (!ion_t2[a]) continue;
        for (int b = 1; b < nt1; b++)
            if ((ion_t2[b] || site_t2[b]) && s->pairidx[a * nt1 + b] == 0)
                return bad("pairidx has no pair type for a present type pair");
    }
    // ion types reachable later (GCMC inserts any ion type): every ion atnam type must have
    // a pair type with every site type present and with every ion type
    std::vector<char> any_ion_t2(nt1, 0);
    for (int t = s->nold + 1; t <= s->ntype; t++) {
        int t2 = s->nwtype[t];
        if (t2 < 1 || t2 > s->nttyp) return bad("nwtype of an ion type out of range");
        any_ion_t2[t2] = 1;
    }
    for (int a = 1; a < nt1; a++) {
        if (!any_ion_t2[a]) continue;
        for (int b = 1; b < nt1; b++)
            if ((any_ion_t2[b] || site_t2[b]) && s->pairidx[a * nt1 + b] == 0)
                return bad("pairidx incomplete for ion types");
    }
    for (int i = s->nsites; i < s->ntot; i++)
        if (std::abs(s->typei[i]) <= s->nold) return bad("an ion carries a nucleotide type");
    const int32_t *lp[8] = {s->sitebond, s->siteangle, s->sitedihe, s->sitestack,
                            s->sitebp, s->siteex, s->siteqq, s->siteslv};
    const int cnt[8] = {s->nbond, s->nangle, s->ndihe, s->nstack, s->nbp, s->nex, s->nqq, s->nsolv};
    const int ar[8] = {2, 3, 4, 2, 2, 2, 2, 2};
    if (s->flags & BROMOC_QNUCL) {
        for (int k = 0; k < 8; k++) {
            if (cnt[k] < 0 || cnt[k] >= (1 << 27)) return bad("list size");
            if (cnt[k] && !lp[k]) return bad("list pointer missing");
            for (int64_t q = 0; q < (int64_t)cnt[k] * ar[k]; q++)
                if (lp[k][q] < 0 || lp[k][q] >= s->nsites) return bad("site list index out of range");
        }
        const double *vp[6] = {s->distbond, s->valangle, s->valdihe, s->sgstack, s->sgbp, s->sgex};
        for (int k = 0; k < 6; k++)
            if (cnt[k] && !vp[k]) return bad("list values missing");
    }
    const bromoc_grid *gs[2] = {&s->phix, &s->phiv};
    const uint32_t gf[2] = {BROMOC_QPHIX, BROMOC_QPHIV};
    for (int k = 0; k < 2; k++) {
        if (!(s->flags & gf[k])) continue;
        const bromoc_grid &g = *gs[k];
        if (!g.data || g.nclx < 2 || g.ncly < 2 || g.nclz < 2 || !(g.dcel > 0.0))
            return bad("grid geometry");
    }
    if (s->flags & BROMOC_QPHIV) {
        double tz = 0.5 * (s->phiv.nclz - 1) * s->phiv.dcel;
        if (s->vzmax > s->phiv.zbcen + tz || s->vzmin < s->phiv.zbcen - tz)
            return bad("vzmin/vzmax outside the repulsion map");
        if ((s->flags & BROMOC_QSVDW) && !s->scal) return bad("scal missing");
    }
    if (s->flags & BROMOC_QNMCDEN) {
        if (!(s->flags & BROMOC_QPHIV)) return bad("NMCDEN needs QPHIV");
        if (s->flags & BROMOC_QSVDW) return bad("SVDW not compatible with NMCDEN (shell_simul.f90:2875)");
        // shell_simul.f90:2730-2737; a moving ion without PARTICLE would read the map left
        // by the particle before it (vdwgd1trln.f90:61-80)
        if (!(s->flags & (BROMOC_QNUCL | BROMOC_QPAR))) return bad("NMCDEN needs NUCLEOTIDES or PARTICLE");
        if (!(s->flags & BROMOC_QPAR) && s->ntot > s->nsites + s->nfix)
            return bad("NMCDEN maps of moving ions need PARTICLE");
        if (phiv_maps(s) < 1) return bad("NMCDEN map count");
        // the ion maps follow the site maps from type nold+1; without CG-DNA there are no
        // nucleotide types for them to follow
        if (!(s->flags & BROMOC_QNUCL) && s->nold != 0) return bad("NMCDEN with nold != 0 needs NUCLEOTIDES");
        if (s->phiv.nmaps != phiv_maps(s))
            return bad("phiv.nmaps = " + std::to_string(s->phiv.nmaps) + " but NMCDEN reads " +
                       std::to_string(phiv_maps(s)) + " maps");
    } else if ((s->flags & BROMOC_QPHIV) && (s->phiv.nmaps < 0 || s->phiv.nmaps > 1)) {
        return bad("phiv.nmaps > 1 without NMCDEN");
    }
    if ((s->flags & BROMOC_QCOUNTION) && (s->cntpts < 0 || (s->cntpts > 0 && !s->zcont)))
        return bad("flux planes");
    // simul1.f90:338-340
    if ((s->flags & BROMOC_QDIFFUSE) && (s->flags & BROMOC_QPROFILE))
        return bad("DIFFUSION and PROFILE together");
    if (s->flags & (BROMOC_QDIFFUSE | BROMOC_QPROFILE)) {
        if (!s->diffusion || !std::isfinite(s->dt) || !std::isfinite(s->kbtdt))
            return bad("diffusion constants missing");
    }
    if ((s->flags & BROMOC_QDIFFUSE) && (!s->ampl3 || !s->p3)) return bad("DIFFUSION arrays missing");
    if (s->flags & BROMOC_QPROFILE) {
        if (s->nspline < 2 || s->nspline > (1 << 20) || !s->xs || !s->ys || !s->bs || !s->cs || !s->ds)
            return bad("PROFILE spline missing");
        if (!(s->xs[0] != s->xs[s->nspline - 1])) return bad("PROFILE knots have equal ends");
    }
    if (s->flags & BROMOC_QRFPAR) {
        const bromoc_grid &g = s->rf;
        if (g.nclx < 2 || g.ncly < 2 || g.nclz < 2 || !(g.dcel > 0.0)) return bad("RFPAR grid geometry");
        if (!s->gsrfen || !s->greff) return bad("RFPAR grids missing");
        if (!(s->flags & BROMOC_QRFPSIN) && s->ntype - s->nold < 1) return bad("RFPAR ion types");
    }
    if ((s->flags & BROMOC_QBSPLINE) && !(s->flags & BROMOC_QPHIV)) return bad("QBSPLINE needs QPHIV");
    if ((s->flags & BROMOC_QBSPLINE) && (s->phiv.nclx < 3 || s->phiv.ncly < 3 || s->phiv.nclz < 3))
        return bad("B-spline repulsion map needs 3 nodes per axis");
    // shell_simul.f90:1903 (PROXDIFF needs ions and DNA)
    if (s->flags & BROMOC_QPROXDIFF) {
        if (!(s->flags & BROMOC_QPAR) || !(s->flags & BROMOC_QNUCL) || s->nsites < 1)
            return bad("PROXDIFF needs ions and DNA");
        if (!s->diffusion || !std::isfinite(s->dt) || !std::isfinite(s->kbtdt) ||
            !std::isfinite(s->ibeta) || !std::isfinite(s->diff0))
            return bad("PROXDIFF constants missing");
    }
    const uint32_t dnaopts = BROMOC_QNOTRANS | BROMOC_QAPFOR | BROMOC_QCONTRANS;
    if ((s->flags & dnaopts) && (!(s->flags & BROMOC_QNUCL) || s->nsites < 1))
        return bad("NOTRANS / APFOR / CONTRANS need DNA sites");
    if (s->flags & BROMOC_QAPFOR) {
        if (s->afn < 0 || (s->afn > 0 && (!s->sn || !s->af))) return bad("APFOR arrays");
        for (int k = 0; k < s->afn; k++)
            if (s->sn[k] < 0 || s->sn[k] >= s->nsites) return bad("APFOR site out of range");
    }
    if (s->flags & BROMOC_QCONTRANS) {
        if (s->ctn < 0 || (s->ctn > 0 && (!s->csn || !s->kx || !s->ky || !s->kz)) ||
            !s->contrx || !s->contry || !s->contrz)
            return bad("CONTRANS arrays");
        for (int k = 0; k < s->ctn; k++)
            if (s->csn[k] < 0 || s->csn[k] > s->nsites) return bad("CONTRANS site out of range");
        if (s->unsplit && (s->nsites1st < 0 || s->nsites1st > s->nsites))
            return bad("CONTRANS strand split");
    }
    if ((s->flags & BROMOC_QNOTRANS) && (s->notrans_axes & ~7)) return bad("NOTRANS axes");
    if ((s->flags & BROMOC_QMEMB) &&
        (!s->ampl1 || !s->p1 || !s->ampl2 || !s->p2 || !s->rcylinder || !(s->zmemb2 >= s->zmemb1)))
        return bad("MEMBRANE parameters");
    // hybrid Debye (shell_simul.f90:820-846): not with plain Debye, only for free DNA
    if ((s->flags & BROMOC_QDEBYHYB) &&
        ((s->flags & BROMOC_QDEBY) || !(s->flags & BROMOC_QDNAFREE) || !(s->flags & BROMOC_QNUCL) ||
         !(s->ikappa > 0.0)))
        return bad("DEBYHYB needs DNAFREE, no DEBY and a Debye length");
    if (s->flags & BROMOC_QMMIJ) {
        if (s->mmij_shape == 2) {
            msg = "the SPHERE reaction field (sphe_rf1.f90) is not supported";
            return BROMOC_E_UNSUPPORTED;
        }
        if (s->mmij_shape != 1) return bad("MMIJ shape");
        if (s->flags & BROMOC_QRFPAR) return bad("MMIJ and RFPAR together (shell_simul.f90:2869)");
        // rect_rf1.f90:48-57 fills three polynomials per axis whatever npol is
        if (s->xnpol < 3 || s->ynpol < 3 || s->znpol < 3 || s->xnpol > 64 || s->ynpol > 64 ||
            s->znpol > 64 || s->ntpol < 1 || s->ntpol > 32768)
            return bad("MMIJ polynomial counts");
        if (!s->lstpx || !s->lstpy || !s->lstpz || !s->bnorm || !s->mmij) return bad("MMIJ arrays missing");
        if (!(s->xscal > 0.0) || !(s->yscal > 0.0) || !(s->zscal > 0.0) || !std::isfinite(s->rfscal))
            return bad("MMIJ region or rfscal");
        for (int k = 0; k < s->ntpol; k++)
            if (s->lstpx[k] < 0 || s->lstpx[k] >= s->xnpol || s->lstpy[k] < 0 ||
                s->lstpy[k] >= s->ynpol || s->lstpz[k] < 0 || s->lstpz[k] >= s->znpol)
                return bad("MMIJ degree out of range");
    }
    return BROMOC_OK;
}

// the atnam2 type rides in the low word of the fourth component of each position, the
// particle's own index in the high word
double pack_type(int32_t t2, int32_t i) {
    uint64_t bits = (uint64_t)(uint32_t)t2 | ((uint64_t)(uint32_t)i << 32);
    double d;
    std::memcpy(&d, &bits, sizeof(d));
    return d;
}

GridGeom geom_of(const bromoc_grid &g, bool on) {
    GridGeom r{};
    r.nx = g.nclx; r.ny = g.ncly; r.nz = g.nclz; r.on = on;
    r.dcel = g.dcel;
    r.idcel = 1.0 / g.dcel;
    r.xbcen = g.xbcen; r.ybcen = g.ybcen; r.zbcen = g.zbcen;
    // readphi.f90:49-51
    r.tranx = 0.5 * (g.nclx - 1) * g.dcel;
    r.trany = 0.5 * (g.ncly - 1) * g.dcel;
    r.tranz = 0.5 * (g.nclz - 1) * g.dcel;
    return r;
}

StepParams params_for(bromoc_ctx *c) {
    StepParams p = c->base;
    p.ntot = c->ntot;
    p.ntargets = c->ns + (c->ntot - c->ns - c->nf);
    p.tbeg = 0;
    p.tend = p.ntargets;
    p.pos = c->pos[c->cur];
    p.poso = c->pos[c->cur ^ 1];
    p.spos = c->spos[c->cur];
    p.sposo = c->spos[c->cur ^ 1];
    p.tpb = c->tpb > 0 ? c->tpb : 4;
    if (tile_mode(p.tpb)) {
        const int tt = 64 * (p.tpb - TPB_TILE);
        p.ttpl = p.tpb - TPB_TILE;
        p.ntile = (p.ntot + tt - 1) / tt;
        p.titems = p.ntile * (p.ntile + 1) / 2;
        p.tnc = p.rfon ? 5 : 3;
        p.tpart = c->tpart;
        p.tener = c->tener;
        // one word per tile pair after the energies: its reaction-field sums were written
        p.tflag = reinterpret_cast<int32_t *>(c->tener + (size_t)NTE * p.titems);
    }
    return p;
}

// Tile-mode buffers for the context's capacity at ttpl targets per lane (kept; grown when
// a larger shape needs more).  Fails with BROMOC_E_CAPACITY when they would pass
// TILE_MAX_BYTES.
constexpr size_t TILE_MAX_BYTES = size_t(3) << 30;
int ensure_tiles(bromoc_ctx *c, int ttpl) {
    const size_t np = tile_part_doubles(c->cap, ttpl, c->base.rfon);
    const size_t ne = tile_ener_doubles(c->cap, ttpl);
    if (8 * (np + ne) > TILE_MAX_BYTES)
        return fail(c, BROMOC_E_CAPACITY, "tile buffers beyond their limit");
    if (np > c->tpart_n) {
        if (c->tpart) (void)hipFree(c->tpart);
        c->tpart = nullptr;
        c->tpart_n = 0;
        HIPCHK(c, hipMalloc(&c->tpart, 8 * np));
        c->tpart_n = np;
    }
    if (ne > c->tener_n) {
        if (c->tener) (void)hipFree(c->tener);
        c->tener = nullptr;
        c->tener_n = 0;
        HIPCHK(c, hipMalloc(&c->tener, 8 * ne));
        c->tener_n = ne;
    }
    return BROMOC_OK;
}

// tile mode is not offered where the step kernel does more than the tile kernels: the
// nearest-site records of DYNAMICS3 with DNA
bool tiles_allowed(const bromoc_ctx *c) {
    return !(c->base.dmode == 3 && c->ns > 0);
}

// Re-sort the partner array when positions were replaced or after GRAPH_K steps.
int ensure_sorted(bromoc_ctx *c, bool force = false) {
    if (!force && !c->sort_dirty && c->since_sort < GRAPH_K) return BROMOC_OK;
    HIPCHK(c, sort_partners(c->sb, c->pos[c->cur], c->ntot, c->spos[c->cur], c->spos[c->cur ^ 1],
                            c->stream));
    c->sort_dirty = false;
    c->since_sort = 0;
    return BROMOC_OK;
}

// Pinned host staging for a call's small copies (the step loop's control words, GRAND's
// uniforms, counts and state): from pageable memory each such copy is a slow staged
// transfer that leaves the GPU idle between calls.  Every call synchronises before it
// returns, so the next call may reuse or regrow the buffer.
static unsigned char *pinned(bromoc_ctx *c, size_t bytes) {
    if (bytes > c->pin_cap) {
        if (c->pin) (void)hipHostFree(c->pin);
        c->pin = nullptr;
        c->pin_cap = 0;
        if (hipHostMalloc(&c->pin, bytes) != hipSuccess) return nullptr;
        c->pin_cap = bytes;
    }
    return (unsigned char *)c->pin;
}

int sync_err(bromoc_ctx *c) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, BROMOC_E_HIP, std::string("kernel: ") + hipGetErrorString(e));
    return BROMOC_OK;
}

int reduce_energies(bromoc_ctx *c, bromoc_energies *out) {
    const StepParams p = params_for(c);
    const bool tiles = tile_mode(p.tpb);
    HIPCHK(c, launch_reduce_energy(c->epart, c->cap, c->ntot, c->ered, c->stream,
                                   tiles ? c->tener : nullptr, tiles ? p.titems : 0));
    double r[NE];
    HIPCHK(c, hipMemcpyAsync(r, c->ered, sizeof(r), hipMemcpyDeviceToHost, c->stream));
    int st = sync_err(c);
    if (st) return st;
    energies_from(r, (c->sys.flags & BROMOC_QRFPAR) != 0, (c->sys.flags & BROMOC_QMMIJ) != 0, out);
    return BROMOC_OK;
}

}  // namespace

int bromoc::shared_context(const bromoc_system *s, int device, bromoc_ctx **out,
                           StepParams *base) {
    int st = bromoc_create(s, device, out);
    if (st) return st;
    *base = (*out)->base;
    return BROMOC_OK;
}

extern "C" {

const char *bromoc_last_error(const bromoc_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int bromoc_destroy(bromoc_ctx *c) {
    if (!c) return BROMOC_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    for (auto &g : c->graph)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (void *p : c->allocs) (void)hipFree(p);
    for (void *p : c->gd.mem) (void)hipFree(p);
    if (c->pin) (void)hipHostFree(c->pin);
    if (c->gd.u) (void)hipFree(c->gd.u);
    if (c->gd.natlog) (void)hipFree(c->gd.natlog);
    for (void *p : {(void *)c->tpart, (void *)c->tener, (void *)c->an_rho, (void *)c->an_rdna, (void *)c->an_gr, (void *)c->an_prob2,
                    (void *)c->an_prob3})
        if (p) (void)hipFree(p);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return BROMOC_OK;
}

int bromoc_create(const bromoc_system *s, int device, bromoc_ctx **out) {
    if (!s || !out) return BROMOC_E_ARG;
    *out = nullptr;
    std::string msg;
    int st = validate(s, msg);
    if (st) {
        fprintf(stderr, "bromoc_create: %s\n", msg.c_str());
        return st;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev)
        return BROMOC_E_NODEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return BROMOC_E_NODEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        fprintf(stderr, "bromoc_create: device %d is %s, this library is built for gfx950\n",
                device, prop.gcnArchName);
        return BROMOC_E_NODEVICE;
    }
    bromoc_ctx *c = new bromoc_ctx();
    c->device = device;
    c->sys = *s;
    c->cap = s->capacity;
    c->ntot = s->ntot;
    c->ns = s->nsites;
    c->nf = s->nfix;
    auto cleanup = [&](int code) { c->err = c->err.empty() ? "create failed" : c->err; bromoc_destroy(c); return code; };
    if (hipSetDevice(device) != hipSuccess) return cleanup(BROMOC_E_HIP);
    // the device address of the allocation-free entry point's workspace on this device
    if (vjp_scratch_init(device, nullptr) != hipSuccess) return cleanup(BROMOC_E_HIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return cleanup(BROMOC_E_HIP);
    if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess) return cleanup(BROMOC_E_HIP);
    if (hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess) return cleanup(BROMOC_E_HIP);
    if (hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) return cleanup(BROMOC_E_HIP);
    const int cap = c->cap, ns = c->ns, nt1 = s->nttyp + 1;
#define ACHK(expr) do { if ((expr) != hipSuccess) return cleanup(BROMOC_E_HIP); } while (0)
    c->h_typei.assign(s->typei, s->typei + s->ntot);
    c->h_typei.resize(cap, 0);
    c->h_nwtype.assign(s->nwtype, s->nwtype + s->ntype + 1);
    c->h_cg.assign(s->cg, s->cg + s->ntype + 1);
    if (ns) c->h_typtyp.assign(s->typtyp, s->typtyp + ns);
    std::vector<int32_t> t2(cap, 0);
    for (int i = 0; i < s->ntot; i++)
        t2[i] = i < ns ? s->typtyp[i] : s->nwtype[std::abs(s->typei[i])];
    {
        std::vector<double4> p4(cap, make_double4(0.0, 0.0, 0.0, 0.0));
        for (int i = 0; i < s->ntot; i++) p4[i] = make_double4(s->x[i], s->y[i], s->z[i], pack_type(t2[i], i));
        for (int b = 0; b < 2; b++) ACHK(dupload(c, &c->pos[b], p4.data(), cap));
    }
    for (int b = 0; b < 2; b++) ACHK(dalloc(c, &c->spos[b], cap));
    ACHK(dalloc(c, &c->sb.keys, cap));
    ACHK(dalloc(c, &c->sb.keys_out, cap));
    ACHK(dalloc(c, &c->sb.vals, cap));
    ACHK(dalloc(c, &c->sb.order, cap));
    ACHK(dalloc(c, &c->sb.rank, cap));
    ACHK(dalloc(c, &c->sb.hist, sort_hist_ints(cap)));
    ACHK(dalloc(c, &c->sb.cflag, cap));
    ACHK(dalloc(c, &c->sb.cell_start, (size_t)cap + 1));
    ACHK(dalloc(c, &c->sb.ncell, 1));
    {
        // Morton cells of 8 A over the box (10 bits per axis)
        SortGeom &g = c->sb.geom;
        const double lo[3] = {s->lx2m, s->ly2m, s->lz2m}, hi[3] = {s->lx2p, s->ly2p, s->lz2p};
        const double cell = 8.0;
        g.inv = 1.0 / cell;
        g.nsites = ns;
        g.nfix = s->nfix;
        for (int k = 0; k < 3; k++) {
            g.lo[k] = lo[k];
            const double ext = hi[k] - lo[k];
            int n = ext > 0.0 ? (int)std::ceil(ext / cell) : 1;
            g.n[k] = std::min(std::max(n, 1), 1024);
        }
    }
    ACHK(dupload(c, &c->typei, c->h_typei.data(), cap));
    ACHK(dupload(c, &c->t2, t2.data(), cap));
    ACHK(dupload(c, &c->stfree, ns ? s->stfree : (const int8_t *)nullptr, ns));
    ACHK(dupload(c, &c->namsite, ns ? s->namsite : (const int32_t *)nullptr, ns));
    ACHK(dupload(c, &c->nwtype, s->nwtype, s->ntype + 1));
    ACHK(dupload(c, &c->cg, s->cg, s->ntype + 1));
    ACHK(dupload(c, &c->fact1a, s->fact1a, s->ntype + 1));
    ACHK(dupload(c, &c->fact2a, s->fact2a, s->ntype + 1));
    ACHK(dalloc(c, &c->fx, cap));
    ACHK(dalloc(c, &c->fy, cap));
    ACHK(dalloc(c, &c->fz, cap));
    ACHK(dalloc(c, &c->fdna, std::max(ns, 1)));
    ACHK(dalloc(c, &c->epart, (size_t)NE * cap));
    ACHK(dalloc(c, &c->ered, (size_t)NE * (1 + RED_BLOCKS)));
    ACHK(dupload(c, &c->pairidx, s->pairidx, (size_t)nt1 * nt1));
    c->h_pairs.assign(s->pairs, s->pairs + s->npair + 1);
    c->h_pairidx.assign(s->pairidx, s->pairidx + (size_t)nt1 * nt1);
    std::vector<DevPair> dp(s->npair + 1);
    for (int t = 0; t <= s->npair; t++) {
        const bromoc_pair &p = s->pairs[t];
        DevPair &d = dp[t];
        std::memset(&d, 0, sizeof(d));
        if (t == 0) continue;
        d.dmi = p.dmi; d.dm2lo = p.dm2lo; d.dm2hi = p.dm2hi;
        d.sc1 = p.sc1; d.sc2 = p.sc2; d.sc3 = p.sc3; d.fct = p.fct;
        d.epp4 = p.epp4; d.sgp2 = p.sgp2;
        d.c0 = p.c0; d.c1 = p.c1; d.c2 = p.c2; d.c3 = p.c3; d.c4 = p.c4;
        d.nxi = p.nxi;
        d.nlast = p.efpot ? p.nxi + p.nknots - 1 : p.nxi;
        if (!p.chr) d.fct = 0.0;   // the table tail adds fct/r for charged pairs only
        d.flags = (p.efpot ? PF_EFPOT : 0) | (p.chr ? PF_CHR : 0) | (p.lj ? PF_LJ : 0) |
                  (p.srpmf ? PF_SRPMF : 0);
    }
    ACHK(dupload(c, &c->pairs, dp.data(), dp.size()));
    {
        std::vector<DevPair> rows((size_t)nt1 * nt1);
        for (int k = 0; k < nt1 * nt1; k++) rows[k] = dp[s->pairidx[k]];
        ACHK(dupload(c, &c->rows, rows.data(), rows.size()));
    }
    // two zero knots of padding: a lookup that rounds one knot past a table's end (a 1-ulp
    // difference in sqrt at x = r_n) stays inside the allocation
    std::vector<double4> ep4(std::max(s->nknots, 0) + 2, make_double4(0.0, 0.0, 0.0, 0.0));
    for (int k = 0; k < s->nknots; k++)
        ep4[k] = make_double4(s->ep[3 * k], s->ep[3 * k + 1], s->ep[3 * k + 2], 0.0);
    ACHK(dupload(c, &c->ep, ep4.data(), ep4.size()));
    // CG-DNA: component-major lists and the per-site incidence (CSR) of every term
    const int32_t *lp[8] = {s->sitebond, s->siteangle, s->sitedihe, s->sitestack,
                            s->sitebp, s->siteex, s->siteqq, s->siteslv};
    const int cnt[8] = {s->nbond, s->nangle, s->ndihe, s->nstack, s->nbp, s->nex, s->nqq, s->nsolv};
    const int ar[8] = {2, 3, 4, 2, 2, 2, 2, 2};
    const double *vp[6] = {s->distbond, s->valangle, s->valdihe, s->sgstack, s->sgbp, s->sgex};
    std::vector<int32_t> deg(ns + 1, 0);
    bool dna = (s->flags & BROMOC_QNUCL) && ns > 0;
    for (int k = 0; k < 8; k++) {
        int n = dna ? cnt[k] : 0;
        ACHK(dupload(c, &c->lst[k], n ? lp[k] : (const int32_t *)nullptr, (size_t)n * ar[k]));
        if (k < 6) ACHK(dupload(c, &c->lval[k], n ? vp[k] : (const double *)nullptr, n));
        for (int64_t q = 0; q < (int64_t)n * ar[k]; q++) deg[lp[k][q] + 1]++;
    }
    for (int i = 0; i < ns; i++) deg[i + 1] += deg[i];
    // each site's terms (INTERACT's list) in the order angle, dihedral, bond, stacking, base
    // pair, excluded volume, P-P, solvent
    static const int order[8] = {T_ANGLE, T_DIHE, T_BOND, T_STACK, T_BP, T_EX, T_QQ, T_SLV};
    std::vector<uint32_t> inc(std::max<int64_t>(deg[ns], 1));
    {
        std::vector<int32_t> fill(deg.begin(), deg.end());
        for (int ko = 0; ko < 8 && dna; ko++)
            for (int r = 0, k = order[ko]; r < ar[k]; r++)
                for (int t = 0; t < cnt[k]; t++) inc[fill[lp[k][(int64_t)r * cnt[k] + t]]++] = inc_pack(k, r, t);
    }
    // Resolved records for the force kernels.  The two-site terms that come in bulk
    // (excluded volume, P-P, solvent, base pairs) take one word each when the sites fit in
    // 24 bits and the list's values in a table of 16: the other site, the term, the site's
    // role, whether site a of a base pair is a C or G base, and the value's table slot.
    // The others take 16 bytes: the word (kind, role, list index), the other site, that
    // flag and the list value; a site's angle and dihedral records come first.
    static const int bulk_kind[4] = {T_EX, T_QQ, T_SLV, T_BP};
    bool bulk[8] = {};
    std::vector<double> bpar(4 * 16, 0.0);
    for (int q = 0; q < 4 && dna && ns <= (1 << 24); q++) {
        const int k = bulk_kind[q], sl = pair_lval_slot(k);
        bool ok = true;
        std::vector<double> vals;
        for (int t = 0; t < cnt[k] && ok && sl >= 0; t++) {
            const double v = vp[sl][t];
            bool seen = false;
            for (size_t m = 0; m < vals.size() && !seen; m++)
                seen = std::memcmp(&vals[m], &v, sizeof(v)) == 0;
            if (!seen) {
                ok = vals.size() < 16;
                vals.push_back(v);
            }
        }
        if (!ok) continue;
        bulk[k] = true;
        for (size_t m = 0; m < vals.size(); m++) bpar[16 * q + m] = vals[m];
    }
    std::vector<int32_t> rptr(ns + 1, 0), bptr(ns + 1, 0), split(std::max(ns, 1), 0);
    for (int k = 0; k < 8 && dna; k++)
        for (int64_t q = 0; q < (int64_t)cnt[k] * ar[k]; q++) (bulk[k] ? bptr : rptr)[lp[k][q] + 1]++;
    for (int i = 0; i < ns; i++) {
        rptr[i + 1] += rptr[i];
        bptr[i + 1] += bptr[i];
    }
    std::vector<double2> incr(std::max(rptr[ns], 1), make_double2(0.0, 0.0));
    std::vector<uint32_t> incb(std::max(bptr[ns], 1), 0u);
    {
        std::vector<int32_t> rfill(rptr.begin(), rptr.end()), bfill(bptr.begin(), bptr.end());
        for (int i = 0; i < ns; i++) split[i] = rptr[i];
        for (int k = T_ANGLE; k <= T_DIHE && dna; k++)
            for (int64_t q = 0; q < (int64_t)cnt[k] * ar[k]; q++) split[lp[k][q]]++;
        for (int ko = 0; ko < 8 && dna; ko++)
            for (int r = 0, k = order[ko]; r < ar[k]; r++)
                for (int t = 0; t < cnt[k]; t++) {
                    const int site = lp[k][(int64_t)r * cnt[k] + t];
                    const int sl = pair_lval_slot(k);
                    const double par = sl >= 0 ? vp[sl][t] : 0.0;
                    const int other = ar[k] == 2 ? lp[k][(int64_t)(1 - r) * cnt[k] + t] : 0;
                    if (ar[k] > 2) {   // angle, dihedral: the other sites in list order
                        uint32_t os[3] = {0u, 0u, 0u};
                        for (int rr = 0, q = 0; rr < ar[k]; rr++)
                            if (rr != r) os[q++] = (uint32_t)lp[k][(int64_t)rr * cnt[k] + t];
                        const int at = rfill[site]++;
                        const uint64_t bx = inc_pack(k, r, t) | (uint64_t)os[0] << 32;
                        const uint64_t by = (uint64_t)os[1] | (uint64_t)os[2] << 32;
                        std::memcpy(&incr[at].x, &bx, sizeof(bx));
                        std::memcpy(&incr[at].y, &by, sizeof(by));
                        continue;
                    }
                    const int na = k == T_BP ? s->namsite[lp[k][t]] : 0;
                    const bool wide = na == 4 || na == 5;
                    if (bulk[k]) {
                        const int q = k == T_EX ? 0 : k == T_QQ ? 1 : k == T_SLV ? 2 : 3;
                        int slot = 0;
                        while (sl >= 0 && std::memcmp(&bpar[16 * q + slot], &par, sizeof(par)) != 0) slot++;
                        incb[bfill[site]++] = (uint32_t)other | (uint32_t)q << 24 | (uint32_t)r << 26 |
                                              (uint32_t)(wide ? 1 : 0) << 27 | (uint32_t)slot << 28;
                        continue;
                    }
                    const int at = rfill[site]++;
                    uint64_t bits = inc_pack(k, r, t);
                    if (ar[k] == 2) {
                        bits |= (uint64_t)(uint32_t)other << 32;
                        if (wide) bits |= (uint64_t)1 << 63;
                    }
                    std::memcpy(&incr[at].x, &bits, sizeof(bits));
                    incr[at].y = par;
                }
    }
    ACHK(dupload(c, &c->inc_ptr, deg.data(), deg.size()));
    ACHK(dupload(c, &c->inc, inc.data(), inc.size()));
    ACHK(dupload(c, &c->incr_ptr, rptr.data(), rptr.size()));
    ACHK(dupload(c, &c->inc_split, split.data(), split.size()));
    ACHK(dupload(c, &c->incr, incr.data(), incr.size()));
    ACHK(dalloc(c, &c->fmul, incr.size()));
    ACHK(dupload(c, &c->incb_ptr, bptr.data(), bptr.size()));
    ACHK(dupload(c, &c->incb, incb.data(), incb.size()));
    ACHK(dupload(c, &c->bpar, bpar.data(), bpar.size()));
    {
        std::vector<double2> cs(std::max(dna ? s->ndihe : 0, 1), make_double2(1.0, 0.0));
        for (int t = 0; dna && t < s->ndihe; t++)
            cs[t] = make_double2(std::cos(s->valdihe[t]), std::sin(s->valdihe[t]));
        ACHK(dupload(c, &c->dihcs, cs.data(), cs.size()));
    }
    if (s->flags & BROMOC_QPHIX)
        ACHK(dupload(c, &c->phix, (const float *)s->phix.data,
                     (size_t)s->phix.nclx * s->phix.ncly * s->phix.nclz));
    if (s->flags & BROMOC_QPHIV)
        ACHK(dupload(c, &c->phiv, (const int8_t *)s->phiv.data,
                     (size_t)phiv_maps(s) * s->phiv.nclx * s->phiv.ncly * s->phiv.nclz));
    std::vector<double> scal(nt1, 1.0);
    if ((s->flags & BROMOC_QSVDW) && s->scal) scal.assign(s->scal, s->scal + nt1);
    ACHK(dupload(c, &c->scal, scal.data(), nt1));
    ACHK(dalloc(c, &c->warn, nt1));
    ACHK(dalloc(c, &c->warn_scratch, nt1));
    const int nion_t = std::max(1, s->ntype - s->nold), np = std::max(1, s->cntpts);
    ACHK(dalloc(c, &c->nfwd, (size_t)nion_t * np));
    ACHK(dalloc(c, &c->nbwd, (size_t)nion_t * np));
    ACHK(dupload(c, &c->zcont, s->cntpts > 0 ? s->zcont : (const double *)nullptr,
                 std::max(0, s->cntpts)));
    ACHK(dalloc(c, &c->ctl, 4));
    ACHK(dalloc(c, &c->pbuf, (size_t)GRAPH_K * 3 * cap));
    ACHK(dalloc(c, &c->iout, 2));
    // position-dependent diffusion (DYNAMICS1 / DYNAMICS2)
    const bool dyn12 = (s->flags & (BROMOC_QDIFFUSE | BROMOC_QPROFILE)) != 0;
    const int nsp = (s->flags & BROMOC_QPROFILE) ? s->nspline : 0;
    double *d_diff = nullptr, *d_ampl3 = nullptr, *d_p3 = nullptr;
    double *d_sp[5] = {};
    ACHK(dupload(c, &d_diff, dyn12 ? s->diffusion : (const double *)nullptr, dyn12 ? s->ntype + 1 : 0));
    const bool sw = (s->flags & BROMOC_QDIFFUSE) != 0;
    ACHK(dupload(c, &d_ampl3, sw ? s->ampl3 : (const double *)nullptr, sw ? s->ntype + 1 : 0));
    ACHK(dupload(c, &d_p3, sw ? s->p3 : (const double *)nullptr, sw ? s->ntype + 1 : 0));
    const double *sp_src[5] = {s->xs, s->ys, s->bs, s->cs, s->ds};
    for (int k = 0; k < 5; k++) ACHK(dupload(c, &d_sp[k], nsp ? sp_src[k] : (const double *)nullptr, nsp));
    // reaction-field parameter grids and per-ion parameters
    const bool rfon = (s->flags & BROMOC_QRFPAR) != 0;
    float *d_gs = nullptr, *d_gr = nullptr;
    {
        const size_t ncel3 = rfon ? (size_t)s->rf.nclx * s->rf.ncly * s->rf.nclz : 0;
        const size_t ng = (s->flags & BROMOC_QRFPSIN) ? 1 : (size_t)std::max(1, s->ntype - s->nold);
        ACHK(dupload(c, &d_gs, rfon ? s->gsrfen : (const float *)nullptr, ncel3 * ng));
        ACHK(dupload(c, &d_gr, rfon ? s->greff : (const float *)nullptr, ncel3 * ng));
    }
    double4 *d_rfa = nullptr, *d_rfd = nullptr, *d_frf = nullptr, *d_rfq = nullptr;
    ACHK(dalloc(c, &d_rfa, rfon ? cap : 0));
    ACHK(dalloc(c, &d_rfq, rfon ? cap : 0));
    ACHK(dalloc(c, &d_rfd, rfon ? 2 * (size_t)cap : 0));
    const bool mmon = (s->flags & BROMOC_QMMIJ) != 0;
    ACHK(dalloc(c, &d_frf, (rfon || mmon) ? cap : 0));
    // MMIJ reaction field: degrees packed per basis function, norms, matrix, scratch
    int32_t *d_lstp = nullptr;
    double *d_bnorm = nullptr, *d_mmij = nullptr, *d_mleg = nullptr, *d_mchg = nullptr;
    double *d_mcoef = nullptr, *d_mq = nullptr, *d_mer = nullptr;
    const int ntpol = mmon ? s->ntpol : 0;
    const int npl = mmon ? s->xnpol + s->ynpol + s->znpol : 0;
    {
        std::vector<int32_t> lp(std::max(ntpol, 1), 0);
        for (int k = 0; k < ntpol; k++) lp[k] = s->lstpx[k] | (s->lstpy[k] << 8) | (s->lstpz[k] << 16);
        ACHK(dupload(c, &d_lstp, mmon ? lp.data() : (const int32_t *)nullptr, ntpol));
        ACHK(dupload(c, &d_bnorm, mmon ? s->bnorm : (const double *)nullptr, ntpol));
        ACHK(dupload(c, &d_mmij, mmon ? s->mmij : (const double *)nullptr, (size_t)ntpol * ntpol));
        ACHK(dalloc(c, &d_mleg, mmon ? ((size_t)cap + 1) * npl : 0));
        ACHK(dalloc(c, &d_mchg, mmon ? (size_t)cap + 1 : 0));
        ACHK(dalloc(c, &d_mcoef, ntpol));
        ACHK(dalloc(c, &d_mq, ntpol));
        ACHK(dalloc(c, &d_mer, ntpol));
    }
    // DYNAMICS3 needs the diffusion constants even without DYNAMICS1/2
    if (!dyn12 && (s->flags & BROMOC_QPROXDIFF)) ACHK(dupload(c, &d_diff, s->diffusion, s->ntype + 1));
    // APFOR: the applied force of each site summed over its entries in entry order
    double4 *d_apf = nullptr;
    {
        std::vector<double4> apf(std::max(ns, 1), make_double4(0.0, 0.0, 0.0, 0.0));
        if (s->flags & BROMOC_QAPFOR)
            for (int k = 0; k < s->afn; k++) {
                double4 &a = apf[s->sn[k]];
                a.x = a.x + s->af[3 * k]; a.y = a.y + s->af[3 * k + 1]; a.z = a.z + s->af[3 * k + 2];
            }
        ACHK(dupload(c, &d_apf, apf.data(), apf.size()));
    }
    // CONTRANS
    const bool con = (s->flags & BROMOC_QCONTRANS) != 0;
    const int ctn = con ? s->ctn : 0;
    int32_t *d_csn = nullptr;
    double *d_k[3] = {}, *d_c[3] = {}, *d_cen = nullptr;
    ACHK(dupload(c, &d_csn, ctn ? s->csn : (const int32_t *)nullptr, ctn));
    const double *ksrc[3] = {s->kx, s->ky, s->kz}, *csrc[3] = {s->contrx, s->contry, s->contrz};
    for (int a = 0; a < 3; a++) {
        ACHK(dupload(c, &d_k[a], ctn ? ksrc[a] : (const double *)nullptr, ctn));
        ACHK(dupload(c, &d_c[a], con ? csrc[a] : (const double *)nullptr, con ? ctn + 1 : 0));
    }
    ACHK(dalloc(c, &d_cen, 9));
    // MEMBRANE per atnam2 type
    const bool memb = (s->flags & BROMOC_QMEMB) != 0;
    double *d_m[5] = {};
    {
        const double *src[5] = {s->ampl1, s->p1, s->ampl2, s->p2, s->rcylinder};
        const size_t len[5] = {(size_t)nt1, 2 * (size_t)nt1, (size_t)nt1, 2 * (size_t)nt1, (size_t)nt1};
        for (int k = 0; k < 5; k++)
            ACHK(dupload(c, &d_m[k], memb ? src[k] : (const double *)nullptr, memb ? len[k] : 0));
    }
#undef ACHK

    StepParams &p = c->base;
    p.ntot = s->ntot; p.nsites = ns; p.nfix = s->nfix;
    p.nttyp = s->nttyp; p.ntype = s->ntype; p.nold = s->nold; p.npair = s->npair;
    p.flags = s->flags;
    p.typei = c->typei; p.t2 = c->t2; p.stfree = c->stfree; p.namsite = c->namsite;
    p.fx = c->fx; p.fy = c->fy; p.fz = c->fz;
    p.fdna = c->fdna;
    p.epart = c->epart; p.ecap = cap;
    p.pairidx = c->pairidx; p.pairs = c->pairs; p.rows = c->rows; p.ep = c->ep;
    p.ires = s->ires; p.celec = celec_value();
    p.cg = c->cg; p.fact1a = c->fact1a; p.fact2a = c->fact2a; p.nwtype = c->nwtype;
    p.rth = s->rth; p.srpx = s->srpx; p.srpk = s->srpk; p.srpy = s->srpy;
    p.inc_ptr = c->inc_ptr; p.inc_split = c->inc_split; p.inc = c->inc; p.incr = c->incr;
    p.incr_ptr = c->incr_ptr; p.incb_ptr = c->incb_ptr; p.incb = c->incb; p.bpar = c->bpar;
    p.fmul = c->fmul; p.dihcs = c->dihcs;
    for (int k = 0; k < 8; k++) { p.lst[k] = c->lst[k]; p.nlist[k] = dna ? cnt[k] : 0; }
    for (int k = 0; k < 6; k++) p.lval[k] = c->lval[k];
    p.epsnuc = s->epsnuc; p.scalepairing = s->scalepairing; p.fctn = s->fctn;
    p.kappa = s->kappa; p.ikappa = s->ikappa; p.epsolv = s->epsolv; p.cgnuc = s->cgnuc;
    p.gx = geom_of(s->phix, (s->flags & BROMOC_QPHIX) != 0);
    p.gv = geom_of(s->phiv, (s->flags & BROMOC_QPHIV) != 0);
    p.phix = c->phix; p.phiv = c->phiv;
    p.svdw = s->svdw; p.thold8 = s->thold8; p.vzmin = s->vzmin; p.vzmax = s->vzmax;
    p.scal = c->scal; p.warn = c->warn;
    p.bdmax = s->bdmax; p.fact0n1 = s->fact0n1; p.fact0n2 = s->fact0n2;
    p.lx2m = s->lx2m; p.lx2p = s->lx2p; p.ly2m = s->ly2m; p.ly2p = s->ly2p;
    p.lz2m = s->lz2m; p.lz2p = s->lz2p;
    p.cv = curved_of(s);
    if (p.cv.kind != 0) {   // the step kernels clamp the box axes only; k_curved does the rest
        p.lx2m = p.ly2m = -HUGE_VAL;
        p.lx2p = p.ly2p = HUGE_VAL;
        if (p.cv.kind == 1) {
            p.lz2m = -HUGE_VAL;
            p.lz2p = HUGE_VAL;
        }
    }
    p.cntpts = (s->flags & BROMOC_QCOUNTION) ? s->cntpts : 0;
    p.zcont = c->zcont; p.nfwd = c->nfwd; p.nbwd = c->nbwd;
    p.step_ctr = c->ctl;
    p.stream = 0;
    // simul1.f90:370-379: DIFFUSION, else PROFILE, else PROXDIFF
    p.dmode = (s->flags & BROMOC_QDIFFUSE) ? 1 : (s->flags & BROMOC_QPROFILE) ? 2
            : (s->flags & BROMOC_QPROXDIFF) ? 3 : 0;
    p.ibeta = s->ibeta; p.diff0 = s->diff0; p.diffcutoff = s->diffcutoff;
    p.thold27 = s->thold27;
    p.apf = d_apf;
    p.ctn = ctn; p.csn = d_csn;
    p.kx = d_k[0]; p.ky = d_k[1]; p.kz = d_k[2];
    p.contrx = d_c[0]; p.contry = d_c[1]; p.contrz = d_c[2];
    p.unsplit = s->unsplit ? 1 : 0; p.nsites1st = s->nsites1st;
    p.insites = ns > 0 ? 1.0 / ns : 0.0;   // insites = 1.0/nsites (shell_simul.f90:701)
    p.cen = d_cen;
    p.notrans_axes = (s->flags & BROMOC_QNOTRANS) ? s->notrans_axes : 0;
    p.notrx = s->notrx; p.notry = s->notry; p.notrz = s->notrz;
    p.qpore = s->qpore ? 1 : 0;
    p.voltage = s->voltage; p.afact = s->afact; p.ceps = s->ceps;
    p.zmemb1 = s->zmemb1; p.zmemb2 = s->zmemb2;
    p.ampl1 = d_m[0]; p.p1 = d_m[1]; p.ampl2 = d_m[2]; p.p2 = d_m[3]; p.rcylinder = d_m[4];
    p.nspline = nsp;
    p.dt = s->dt; p.kbtdt = s->kbtdt; p.plength2 = s->plength2; p.pcenter = s->pcenter;
    p.diffusion = d_diff; p.ampl3 = d_ampl3; p.p3 = d_p3;
    p.xs = d_sp[0]; p.ys = d_sp[1]; p.bs = d_sp[2]; p.cs = d_sp[3]; p.ds = d_sp[4];
    p.gr = geom_of(s->rf, rfon);
    p.gsrfen = d_gs; p.greff = d_gr; p.sqrfac = s->sqrfac;
    p.rfon = rfon ? 1 : 0;
    p.rfsingle = (s->flags & BROMOC_QRFPSIN) ? 1 : 0;
    p.rfa = d_rfa; p.rfd = d_rfd; p.frf = d_frf; p.rfq = d_rfq;
    p.mmon = mmon ? 1 : 0;
    p.ntpol = ntpol; p.npl = npl;
    p.xnpol = mmon ? s->xnpol : 0; p.ynpol = mmon ? s->ynpol : 0; p.znpol = mmon ? s->znpol : 0;
    p.lstp = d_lstp; p.bnorm = d_bnorm; p.mmij = d_mmij;
    p.xscal = s->xscal; p.yscal = s->yscal; p.zscal = s->zscal; p.rfscal = s->rfscal;
    p.mleg = d_mleg; p.mchg = d_mchg; p.mcoef = d_mcoef; p.mq = d_mq; p.mer = d_mer;
    p.all_efpot = 1;
    c->r2near = 0.0;
    for (int k = 0; k < nt1 * nt1; k++) {
        const int is = s->pairidx[k];
        if (is > 0 && !s->pairs[is].efpot) p.all_efpot = 0;
        if (is > 0 && s->pairs[is].efpot) c->r2near = std::max(c->r2near, s->pairs[is].dm2hi);
    }
    // the analytic-tail fast path needs every pair tabulated
    if (!p.all_efpot) c->r2near = HUGE_VAL;
    p.r2near = c->r2near;
    p.rank = c->sb.rank;
    *out = c;
    return BROMOC_OK;
}

int bromoc_set_positions(bromoc_ctx *c, int32_t ntot, const double *x, const double *y,
                         const double *z, const int32_t *typei) {
    if (!c || !x || !y || !z) return BROMOC_E_ARG;
    if (ntot < c->ns + c->nf || ntot > c->cap) return fail(c, BROMOC_E_ARG, "ntot out of range");
    (void)hipSetDevice(c->device);
    c->gd.ib_current = false;   // the particle count may change: upload ibuffer again
    std::vector<int32_t> t2(ntot);
    if (typei) {
        for (int i = c->ns; i < ntot; i++) {
            int t = std::abs(typei[i]);
            if (t <= c->sys.nold || t > c->sys.ntype) return fail(c, BROMOC_E_ARG, "typei out of range");
            int tt = c->h_nwtype[t];
            if (tt < 1 || tt > c->sys.nttyp) return fail(c, BROMOC_E_ARG, "nwtype out of range");
        }
        for (int i = c->ns; i < ntot; i++) c->h_typei[i] = typei[i];
    } else if (ntot > c->ntot) {
        return fail(c, BROMOC_E_ARG, "typei required when ntot grows");
    }
    for (int i = 0; i < ntot; i++)
        t2[i] = i < c->ns ? c->h_typtyp[i] : c->h_nwtype[std::abs(c->h_typei[i])];
    std::vector<double4> p4(ntot);
    for (int i = 0; i < ntot; i++) p4[i] = make_double4(x[i], y[i], z[i], pack_type(t2[i], i));
    for (int b = 0; b < 2; b++)
        HIPCHK(c, hipMemcpyAsync(c->pos[b], p4.data(), sizeof(double4) * ntot, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->typei, c->h_typei.data(), sizeof(int32_t) * ntot, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->t2, t2.data(), sizeof(int32_t) * ntot, hipMemcpyHostToDevice, c->stream));
    int st = sync_err(c);
    c->ntot = ntot;
    c->sort_dirty = true;
    return st;
}

int bromoc_get_positions(bromoc_ctx *c, double *x, double *y, double *z, int32_t *typei,
                         int32_t *ntot) {
    if (!c) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    const size_t n = c->ntot;
    int st = BROMOC_OK;
    if (x || y || z) {
        std::vector<double4> p4(n);
        HIPCHK(c, hipMemcpyAsync(p4.data(), c->pos[c->cur], sizeof(double4) * n, hipMemcpyDeviceToHost, c->stream));
        st = sync_err(c);
        for (size_t i = 0; i < n; i++) {
            if (x) x[i] = p4[i].x;
            if (y) y[i] = p4[i].y;
            if (z) z[i] = p4[i].z;
        }
    }
    if (typei) std::memcpy(typei, c->h_typei.data(), sizeof(int32_t) * n);
    if (ntot) *ntot = c->ntot;
    return st;
}

int bromoc_get_forces(bromoc_ctx *c, double *fx, double *fy, double *fz) {
    if (!c || !fx || !fy || !fz) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    const size_t n = c->ntot;
    HIPCHK(c, hipMemcpyAsync(fx, c->fx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(fy, c->fy, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(fz, c->fz, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    return sync_err(c);
}

int bromoc_energy(bromoc_ctx *c, int32_t want_forces, bromoc_energies *out) {
    if (!c) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    int sst = ensure_sorted(c);
    if (sst) return sst;
    StepParams p = params_for(c);
    p.mode = 0;
    p.write_energy = 1;
    p.write_forces = want_forces ? 1 : 0;
    HIPCHK(c, launch_step(p, c->stream));
    bromoc_energies E;
    int st = reduce_energies(c, &E);
    if (st) return st;
    if (out) *out = E;
    return BROMOC_OK;
}

int bromoc_upload_noise(bromoc_ctx *c, int32_t nsteps, const double *noise) {
    if (!c || nsteps <= 0 || !noise) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    const int64_t stride = 3 * (int64_t)c->ntot;
    const size_t n = (size_t)nsteps * stride;
    if (c->noise && (size_t)c->noise_rows * c->noise_stride < n) {
        (void)hipFree(c->noise);
        c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void *)c->noise));
        c->noise = nullptr;
    }
    if (!c->noise) HIPCHK(c, dalloc(c, &c->noise, n));
    HIPCHK(c, hipMemcpy(c->noise, noise, n * sizeof(double), hipMemcpyHostToDevice));
    c->noise_rows = nsteps;
    c->noise_stride = stride;
    return BROMOC_OK;
}

// One BD step.  Site targets (with the CG-DNA list kernel before them) and ion targets
// are independent within a step -- each reads the old positions and writes its own new
// ones -- so the ion launch runs on a second stream, forked from and joined back into the
// main one (also inside captured graphs).  timing, when given, brackets the ion launch.
static hipError_t enqueue_step(bromoc_ctx *c, const StepParams &p, hipEvent_t t0 = nullptr,
                               hipEvent_t t1 = nullptr) {
    const int ns = p.nsites, nt = p.ntargets;
    hipError_t e;
    if ((e = launch_rfpar(p, !tile_mode(p.tpb), c->stream, true)) != hipSuccess) return e;
    if ((e = launch_mmij(p, c->stream)) != hipSuccess) return e;
    if ((e = launch_fixed(p, c->stream)) != hipSuccess) return e;
    static const bool fork = [] {
        const char *v = std::getenv("BROMOC_FORK");
        return v && v[0] == '1';
    }();
    if (tile_mode(p.tpb)) {
        // timing brackets the pair sweep (k_tiles), the dominant kernel of the step
        if ((e = launch_dna(p, c->stream)) != hipSuccess) return e;
        if (t0 && (e = hipEventRecord(t0, c->stream)) != hipSuccess) return e;
        if ((e = launch_tiles(p, c->stream, 1)) != hipSuccess) return e;
        if (t1 && (e = hipEventRecord(t1, c->stream)) != hipSuccess) return e;
        if ((e = launch_tiles(p, c->stream, 2)) != hipSuccess) return e;
        return launch_notrans(p, c->stream);
    }
    if (!fork || ns == 0 || nt == ns) {
        if ((e = launch_dna(p, c->stream)) != hipSuccess) return e;
        if (t0 && (e = hipEventRecord(t0, c->stream)) != hipSuccess) return e;
        if ((e = launch_pairs(p, c->stream)) != hipSuccess) return e;
        if (t1 && (e = hipEventRecord(t1, c->stream)) != hipSuccess) return e;
        return launch_notrans(p, c->stream);
    }
    if ((e = hipEventRecord(c->ev_fork, c->stream)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(c->stream2, c->ev_fork, 0)) != hipSuccess) return e;
    StepParams q = p;
    q.tbeg = ns;
    q.tend = nt;
    if (t0 && (e = hipEventRecord(t0, c->stream2)) != hipSuccess) return e;
    if ((e = launch_pairs(q, c->stream2)) != hipSuccess) return e;
    if (t1 && (e = hipEventRecord(t1, c->stream2)) != hipSuccess) return e;
    q.tbeg = 0;
    q.tend = ns;
    if ((e = launch_dna(q, c->stream)) != hipSuccess) return e;
    if ((e = launch_pairs(q, c->stream)) != hipSuccess) return e;
    if ((e = hipEventRecord(c->ev_join, c->stream2)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(c->stream, c->ev_join, 0)) != hipSuccess) return e;
    return launch_notrans(p, c->stream);
}

// Autotuner: the pair kernel's launch shape for this system -- k_step with 1, 2 or 4
// targets per block, k_wave (8), or tile mode with 1, 2 or 4 targets per lane (17, 18,
// 20) -- chosen by timing ENERGY-only launches (warnings go to a scratch counter) on the
// ctx's stream: the reaction-field kernels the shape needs, the CG-DNA kernel and its pair
// kernels.  The environment variable BROMOC_TPB forces a value.
static int autotune(bromoc_ctx *c) {
    if (c->tpb > 0) return BROMOC_OK;
    if (const char *env = std::getenv("BROMOC_TPB")) {
        int v = std::atoi(env);
        if (v == 1 || v == 2 || v == 4 || v == 8) { c->tpb = v; return BROMOC_OK; }
        if ((v == TPB_TILE + 1 || v == TPB_TILE + 2 || v == TPB_TILE + 4) && tiles_allowed(c)) {
            int st = ensure_tiles(c, v - TPB_TILE);
            if (st) return st;
            c->tpb = v;
            return BROMOC_OK;
        }
    }
    int sst = ensure_sorted(c);
    if (sst) return sst;
    std::vector<int> cand = {1, 2, 4, 8};
    if (tiles_allowed(c)) {
        for (int t : {1, 2, 4}) {
            const size_t bytes = 8 * (tile_part_doubles(c->cap, t, c->base.rfon) +
                                      tile_ener_doubles(c->cap, t));
            // a tile must hold enough slots for the shape to be worth timing, and the
            // partial sums must stay within their limit
            if (bytes <= TILE_MAX_BYTES && c->ntot >= 64 * t * 4 && ensure_tiles(c, t) == BROMOC_OK)
                cand.push_back(TPB_TILE + t);
        }
    }
    const int nc = (int)cand.size();
    std::vector<double> best(nc, 1e300);
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    // three interleaved rounds of five launches per shape; each shape keeps its best round
    // (shapes within a few per cent of each other are told apart reliably this way)
    for (int round = 0; round < 3; round++) {
        for (int q = 0; q < nc; q++) {
            c->tpb = cand[q];
            StepParams p = params_for(c);
            c->tpb = 0;
            p.mode = 0;
            p.write_energy = 0;
            p.write_forces = 0;
            p.warn = c->warn_scratch;
            // one whole ENERGY (the DNA, reaction-field and fixed-ion terms the pair kernel
            // reads), then the shape's reaction-field, CG-DNA and pair kernels are timed (the
            // one-target shape computes the CG-DNA terms in its pair kernel)
            HIPCHK(c, launch_step(p, c->stream));
            HIPCHK(c, hipEventRecord(e0, c->stream));
            for (int r = 0; r < 5; r++) {
                HIPCHK(c, launch_rfpar(p, !tile_mode(p.tpb), c->stream, true));
                HIPCHK(c, launch_dna(p, c->stream));
                HIPCHK(c, launch_pairs(p, c->stream));
            }
            HIPCHK(c, hipEventRecord(e1, c->stream));
            int st = sync_err(c);
            if (st) return st;
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
            best[q] = std::min(best[q], (double)ms);
        }
    }
    int pick = 4;
    double bt = 1e300;
    for (int q = 0; q < nc; q++)
        if (best[q] < bt) { bt = best[q]; pick = cand[q]; }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->tpb = pick;
    return BROMOC_OK;
}

int bromoc_get_tuning(bromoc_ctx *c, int32_t *targets_per_block) {
    if (!c || !targets_per_block) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    int st = autotune(c);
    if (st) return st;
    *targets_per_block = c->tpb;
    return BROMOC_OK;
}

// The BD loop.  Steps run in chunks of at most GRAPH_K: with Philox noise a chunk starts
// with one launch that draws the Gaussians of all its steps (k_philox_fill), then its
// steps follow; the device counter ctl[0] holds the chunk's first step.  Whole chunks
// replay a captured graph (when graphs are on and no events are requested); the rest,
// including the last step (which writes forces and energies), is launched directly.
// ev, when given, receives an event pair around every step's pair kernel.
static int run_steps(bromoc_ctx *c, int32_t nsteps, int32_t noise_mode, int32_t first_row,
                     uint64_t seed, uint64_t counter0, bool want_energy,
                     std::vector<hipEvent_t> *ev = nullptr, bool notrans = true) {
    if (noise_mode != BROMOC_NOISE_INJECTED && noise_mode != BROMOC_NOISE_PHILOX)
        return fail(c, BROMOC_E_ARG, "noise_mode");
    if (noise_mode == BROMOC_NOISE_INJECTED) {
        if (!c->noise) return fail(c, BROMOC_E_STATE, "no injected noise uploaded");
        if (first_row < 0 || (int64_t)first_row + nsteps > c->noise_rows)
            return fail(c, BROMOC_E_STATE, "noise rows out of range");
        if (3 * (int64_t)c->ntot > c->noise_stride)
            return fail(c, BROMOC_E_STATE, "ntot grew beyond the uploaded noise rows");
    }
    int st0 = autotune(c);
    if (st0) return st0;
    c->h_ctl[0] = 0; c->h_ctl[1] = first_row; c->h_ctl[2] = (int64_t)counter0; c->h_ctl[3] = (int64_t)seed;
    int64_t *hc = (int64_t *)pinned(c, 4096);
    if (!hc) return fail(c, BROMOC_E_HIP, "pinned staging");
    std::memcpy(hc, c->h_ctl, sizeof(c->h_ctl));
    HIPCHK(c, hipMemcpyAsync(c->ctl, hc, sizeof(c->h_ctl), hipMemcpyHostToDevice, c->stream));
    const bool philox = noise_mode == BROMOC_NOISE_PHILOX;
    StepParams p = params_for(c);
    p.mode = 1;
    if (!notrans) p.flags &= ~BROMOC_QNOTRANS;   // NOTRANS belongs to the step loop
    p.noise_mode = noise_mode;
    p.noise = philox ? c->pbuf : c->noise;
    p.noise_stride = philox ? 3 * (int64_t)c->cap : c->noise_stride;
    // the virial pressure is a statistic of the step loop (simul1.f90:365-368)
    p.vacc = (notrans && (c->an.what & BROMOC_AN_PRES)) ? c->vacc : nullptr;
    static const bool graphs_on = [] {
        const char *e = std::getenv("BROMOC_GRAPH");
        return !(e && e[0] == '0');
    }();
    int done = 0;
    while (graphs_on && !ev && nsteps - done > GRAPH_K) {
        int sst = ensure_sorted(c, true);
        if (sst) return sst;
        bromoc_ctx::Graph &g = c->graph[c->cur];
        if (!g.exec || g.ntot != c->ntot || g.tpb != c->tpb || g.noise_mode != noise_mode ||
            g.noise != p.noise || g.stride != p.noise_stride || g.vacc != p.vacc) {
            if (g.exec) {
                (void)hipGraphExecDestroy(g.exec);
                g.exec = nullptr;
            }
            hipGraph_t graph = nullptr;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeGlobal));
            StepParams q = p;
            q.write_energy = 0;
            q.write_forces = 0;
            hipError_t le = hipSuccess;
            if (philox)
                le = launch_philox_fill(c->ctl, q.stream, GRAPH_K, c->ntot, q.noise_stride,
                                        c->pbuf, c->stream);
            for (int k = 0; k < GRAPH_K && le == hipSuccess; k++) {
                q.pos = c->pos[c->cur ^ (k & 1)];
                q.poso = c->pos[c->cur ^ (k & 1) ^ 1];
                q.spos = c->spos[c->cur ^ (k & 1)];
                q.sposo = c->spos[c->cur ^ (k & 1) ^ 1];
                q.step_off = k;
                le = enqueue_step(c, q);
            }
            if (le == hipSuccess) le = launch_advance(c->ctl, GRAPH_K, c->stream);
            hipError_t ce = hipStreamEndCapture(c->stream, &graph);
            if (le != hipSuccess || ce != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return fail(c, BROMOC_E_HIP, std::string("graph capture: ") +
                                                 hipGetErrorString(le != hipSuccess ? le : ce));
            }
            hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) return fail(c, BROMOC_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
            g.ntot = c->ntot; g.tpb = c->tpb; g.noise_mode = noise_mode;
            g.noise = p.noise; g.stride = p.noise_stride; g.vacc = p.vacc;
        }
        HIPCHK(c, hipGraphLaunch(g.exec, c->stream));
        done += GRAPH_K;   // K is even: the buffer parity is unchanged
        c->since_sort += GRAPH_K;
    }
    while (done < nsteps) {
        const int n = std::min(GRAPH_K, nsteps - done);
        if (philox)
            HIPCHK(c, launch_philox_fill(c->ctl, p.stream, n, c->ntot, p.noise_stride, c->pbuf, c->stream));
        for (int r = 0; r < n; r++) {
            const int s = done + r;
            const bool last = s == nsteps - 1;
            int sst = ensure_sorted(c);
            if (sst) return sst;
            p.pos = c->pos[c->cur];
            p.poso = c->pos[c->cur ^ 1];
            p.spos = c->spos[c->cur];
            p.sposo = c->spos[c->cur ^ 1];
            p.step_off = r;
            p.write_energy = (last && want_energy) ? 1 : 0;
            p.write_forces = last ? 1 : 0;
            if (ev) HIPCHK(c, enqueue_step(c, p, (*ev)[2 * s], (*ev)[2 * s + 1]));
            else HIPCHK(c, enqueue_step(c, p));
            c->cur ^= 1;
            c->since_sort++;
        }
        done += n;
        if (done < nsteps) HIPCHK(c, launch_advance(c->ctl, n, c->stream));
    }
    if (p.vacc) {   // fold this call's virial sums before GRAND can renumber particles
        HIPCHK(c, launch_vir_fold(c->vacc, c->ntot, c->vsum, c->stream));
        c->pres_steps += nsteps;
        c->pres_ntot_steps += (int64_t)nsteps * c->ntot;
    }
    return BROMOC_OK;
}

int bromoc_step(bromoc_ctx *c, int32_t nsteps, int32_t noise_mode, int32_t first_row,
                uint64_t seed, uint64_t counter0, bromoc_energies *energies) {
    if (!c || nsteps < 0) return BROMOC_E_ARG;
    if (nsteps == 0) return BROMOC_OK;
    (void)hipSetDevice(c->device);
    int st = run_steps(c, nsteps, noise_mode, first_row, seed, counter0, energies != nullptr);
    if (st) return st;
    if (energies) return reduce_energies(c, energies);
    return sync_err(c);
}

// nsteps BD steps with a HIP event pair around every launch of the pair kernel (k_step,
// k_wave or, in tile mode, k_tiles), on the stream the kernels run on: *total_ms is
// first-to-last, *kernel_ms the mean pair-kernel duration.  A separate k_dna launch (tile
// mode, several targets per block) falls outside the pair; the one-target k_step's own
// CG-DNA terms fall inside it.
int bromoc_time_steps(bromoc_ctx *c, int32_t nsteps, int32_t noise_mode, int32_t first_row,
                      uint64_t seed, uint64_t counter0, double *total_ms, double *kernel_ms) {
    if (!c || nsteps <= 0 || !total_ms || !kernel_ms) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    std::vector<hipEvent_t> ev(2 * (size_t)nsteps, nullptr);
    int st = BROMOC_OK;
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) st = fail(c, BROMOC_E_HIP, "hipEventCreate");
    if (!st) st = run_steps(c, nsteps, noise_mode, first_row, seed, counter0, false, &ev);
    if (!st) st = sync_err(c);
    if (!st) {
        double sum = 0.0;
        for (int s = 0; s < nsteps; s++) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]) != hipSuccess) {
                st = fail(c, BROMOC_E_HIP, "hipEventElapsedTime");
                break;
            }
            sum += ms;
        }
        float tot = 0.f;
        if (!st && hipEventElapsedTime(&tot, ev[0], ev[2 * nsteps - 1]) != hipSuccess)
            st = fail(c, BROMOC_E_HIP, "hipEventElapsedTime");
        *total_ms = tot;
        *kernel_ms = sum / nsteps;
    }
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    return st;
}

int bromoc_dynamics(bromoc_ctx *c, int32_t noise_mode, int32_t row, uint64_t seed,
                    uint64_t counter) {
    if (!c) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    // The forces of the current positions are recomputed in the same fixed order as the
    // preceding ENERGY, so the update sees bit-identical forces.
    int st = run_steps(c, 1, noise_mode, row, seed, counter, false, nullptr, false);
    if (st) return st;
    return sync_err(c);
}

int bromoc_spatial_order(bromoc_ctx *c, uint32_t *keys, int32_t *order, int32_t *rank,
                         int32_t *cell_start, int32_t *ncell) {
    if (!c) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    int st = ensure_sorted(c, true);
    if (st) return st;
    const int n = c->ntot;
    int32_t nc = 0;
    HIPCHK(c, hipMemcpyAsync(&nc, c->sb.ncell, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    if (keys) HIPCHK(c, hipMemcpyAsync(keys, c->sb.keys_out, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
    if (order) HIPCHK(c, hipMemcpyAsync(order, c->sb.order, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
    if (rank) HIPCHK(c, hipMemcpyAsync(rank, c->sb.rank, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
    st = sync_err(c);
    if (st) return st;
    if (cell_start) {
        HIPCHK(c, hipMemcpyAsync(cell_start, c->sb.cell_start, sizeof(int32_t) * ((size_t)nc + 1),
                                 hipMemcpyDeviceToHost, c->stream));
        st = sync_err(c);
        if (st) return st;
    }
    if (ncell) *ncell = nc;
    return BROMOC_OK;
}

int bromoc_get_flux(bromoc_ctx *c, int32_t *nforward, int32_t *nbackward) {
    if (!c || !nforward || !nbackward) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    const size_t n = (size_t)std::max(0, c->sys.ntype - c->sys.nold) * std::max(0, c->sys.cntpts);
    if (n == 0) return BROMOC_OK;
    HIPCHK(c, hipMemcpyAsync(nforward, c->nfwd, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(nbackward, c->nbwd, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    return sync_err(c);
}

int bromoc_get_warn(bromoc_ctx *c, int32_t *warn) {
    if (!c || !warn) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipMemcpyAsync(warn, c->warn, (c->sys.nttyp + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    return sync_err(c);
}

// ---- per-cycle analyses (simul1.f90:365-368, 477-538)

int bromoc_set_analysis(bromoc_ctx *c, const bromoc_analysis *an) {
    if (!c || !an) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    const int w = an->what;
    if (w & ~(BROMOC_AN_RHO | BROMOC_AN_RDNA | BROMOC_AN_GR | BROMOC_AN_PROB | BROMOC_AN_PRES))
        return fail(c, BROMOC_E_ARG, "analysis: unknown bits in what");
    if ((w & (BROMOC_AN_RHO | BROMOC_AN_RDNA)) &&
        (an->nzmax < 1 || !(an->idelz > 0.0) || !std::isfinite(an->idelz) || !std::isfinite(an->zmini)))
        return fail(c, BROMOC_E_ARG, "analysis: nzmax, idelz or zmini");
    if ((w & BROMOC_AN_GR) && (an->npoints < 1 || !(an->dr > 0.0) || an->igr < 0 || an->igr >= c->ntot))
        return fail(c, BROMOC_E_ARG, "analysis: npoints, dr or igr");
    if ((w & BROMOC_AN_PROB) && (an->datom < 0 || an->czmin > an->czmax))
        return fail(c, BROMOC_E_ARG, "analysis: datom or czmin > czmax");
    if ((w & BROMOC_AN_PRES) && (!(an->tvol > 0.0) || !std::isfinite(an->ikbt)))
        return fail(c, BROMOC_E_ARG, "analysis: tvol or ikbt");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unsigned long long **bufs[5] = {&c->an_rho, &c->an_rdna, &c->an_gr, &c->an_prob2, &c->an_prob3};
    for (auto b : bufs)
        if (*b) {
            (void)hipFree(*b);
            *b = nullptr;
        }
    const size_t nt = (size_t)c->sys.ntype, n2 = (size_t)c->sys.nttyp;
    size_t sz[5] = {0, 0, 0, 0, 0};
    if (w & BROMOC_AN_RHO) sz[0] = nt * an->nzmax;
    if (w & BROMOC_AN_RDNA) sz[1] = n2 * an->nzmax;
    if (w & BROMOC_AN_GR) sz[2] = nt * an->npoints;
    if (w & BROMOC_AN_PROB) sz[3] = sz[4] = nt * ((size_t)an->datom + 1);
    for (int k = 0; k < 5; k++) {
        if (!sz[k]) continue;
        HIPCHK(c, hipMalloc((void **)bufs[k], sz[k] * sizeof(unsigned long long)));
        HIPCHK(c, hipMemsetAsync(*bufs[k], 0, sz[k] * sizeof(unsigned long long), c->stream));
    }
    if (!c->an_irmax) HIPCHK(c, dalloc(c, &c->an_irmax, 1));
    if (!c->an_ncnt) HIPCHK(c, dalloc(c, &c->an_ncnt, 2 * (nt + 1)));
    if (!c->vacc) HIPCHK(c, dalloc(c, &c->vacc, c->cap));
    if (!c->vsum) HIPCHK(c, dalloc(c, &c->vsum, 1));
    HIPCHK(c, hipMemsetAsync(c->an_irmax, 0, sizeof(int32_t), c->stream));
    HIPCHK(c, hipMemsetAsync(c->an_ncnt, 0, 2 * (nt + 1) * sizeof(int32_t), c->stream));
    HIPCHK(c, hipMemsetAsync(c->vacc, 0, (size_t)c->cap * sizeof(double), c->stream));
    HIPCHK(c, hipMemsetAsync(c->vsum, 0, sizeof(double), c->stream));
    c->an = *an;
    c->pres_steps = c->pres_ntot_steps = 0;
    return sync_err(c);
}

int bromoc_analyze(bromoc_ctx *c) {
    if (!c) return BROMOC_E_ARG;
    const int w = c->an.what & (BROMOC_AN_RHO | BROMOC_AN_RDNA | BROMOC_AN_GR | BROMOC_AN_PROB);
    if (!w) return BROMOC_OK;
    (void)hipSetDevice(c->device);
    if ((w & BROMOC_AN_GR) && c->an.igr >= c->ntot)
        return fail(c, BROMOC_E_STATE, "analysis: igr beyond the current particles");
    AnParams a{};
    a.pos = c->pos[c->cur];
    a.typei = c->typei;
    a.ntot = c->ntot; a.nsites = c->ns; a.nfix = c->nf; a.what = w;
    a.nzmax = c->an.nzmax; a.igr = c->an.igr; a.npoints = c->an.npoints; a.datom = c->an.datom;
    a.ntype = c->sys.ntype; a.nold = c->sys.nold; a.nttyp = c->sys.nttyp;
    // prob2 counts a channel only with a membrane or a czmin..czmax range (simul1.f90:514-516)
    if (c->sys.flags & BROMOC_QMEMB) a.chan = 1;
    else a.chan = (c->an.czmin == 0.0 && c->an.czmax == 0.0) ? 0 : 2;
    a.zmini = c->an.zmini; a.idelz = c->an.idelz; a.dr = c->an.dr;
    a.czmin = c->an.czmin; a.czmax = c->an.czmax;
    a.zmemb1 = c->base.zmemb1; a.zmemb2 = c->base.zmemb2;
    a.rcylinder = c->base.rcylinder; a.nwtype = c->base.nwtype;
    a.rho = c->an_rho; a.rdna = c->an_rdna; a.gr = c->an_gr;
    a.prob2 = c->an_prob2; a.prob3 = c->an_prob3;
    a.irmax = c->an_irmax; a.ncnt = c->an_ncnt;
    HIPCHK(c, launch_analyze(a, c->stream));
    return BROMOC_OK;
}

int bromoc_get_analysis(bromoc_ctx *c, int64_t *rho, int64_t *rdna, int64_t *gr,
                        int32_t *irmax, int64_t *prob2, int64_t *prob3, double *pres,
                        int64_t *pres_steps) {
    if (!c) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    const size_t nt = (size_t)c->sys.ntype, n2 = (size_t)c->sys.nttyp;
    struct { int64_t *dst; unsigned long long *src; size_t n; } cp[5] = {
        {rho, c->an_rho, nt * c->an.nzmax},
        {rdna, c->an_rdna, n2 * c->an.nzmax},
        {gr, c->an_gr, nt * c->an.npoints},
        {prob2, c->an_prob2, nt * ((size_t)c->an.datom + 1)},
        {prob3, c->an_prob3, nt * ((size_t)c->an.datom + 1)}};
    for (auto &q : cp) {
        if (!q.dst) continue;
        if (!q.src) return fail(c, BROMOC_E_STATE, "analysis: that statistic is not gathered");
        HIPCHK(c, hipMemcpyAsync(q.dst, q.src, q.n * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    }
    int32_t irm = 0;
    double vs = 0.0;
    if (irmax && c->an_irmax) HIPCHK(c, hipMemcpyAsync(&irm, c->an_irmax, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    if (pres && c->vsum) HIPCHK(c, hipMemcpyAsync(&vs, c->vsum, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    int st = sync_err(c);
    if (st) return st;
    if (irmax) *irmax = irm;
    if (pres) {
        // sum over the steps of (ntot - vir/3 * ikbt) * itvol * kba3bar * temp
        // (simul1.f90:367, constamod.f90:36), the ntot terms and the virials summed first
        const double i3 = 1.0 / 3.0, kba3bar = 1.380648813E-23 * 1e25;
        const double itvol = 1.0 / c->an.tvol;
        *pres = ((double)c->pres_ntot_steps - vs * i3 * c->an.ikbt) * itvol * kba3bar * c->an.temp;
    }
    if (pres_steps) *pres_steps = c->pres_steps;
    return BROMOC_OK;
}

static int interact_async(bromoc_ctx *c, double xj, double yj, double zj, int32_t jtype, int32_t j,
                          int32_t alert, int slot = 0) {
    const bool beion = (c->sys.flags & BROMOC_QPAR) && j >= c->ns;
    if (j < 0 || j > c->ntot) return fail(c, BROMOC_E_ARG, "j out of range");
    if (beion) {
        // an ion carries an ion type (nold < jtype <= ntype, shell_simul.f90:530); the
        // reaction-field grids are indexed by it
        if (jtype <= c->sys.nold || jtype > c->sys.ntype)
            return fail(c, BROMOC_E_ARG, "jtype is not an ion type");
        int t2 = c->h_nwtype[jtype];
        if (t2 < 1 || t2 > c->sys.nttyp) return fail(c, BROMOC_E_ARG, "nwtype(jtype) out of range");
    } else if (j >= c->ns) {
        return fail(c, BROMOC_E_ARG, "j is not a site and QPAR is off");
    }
    InteractParams q{};
    q.sp = params_for(c);
    q.xj = xj; q.yj = yj; q.zj = zj;
    q.jtype = jtype; q.j = j; q.alert = alert;
    q.out = c->iout + slot;
    HIPCHK(c, launch_rfpar(q.sp, false, c->stream));
    HIPCHK(c, launch_interact(q, c->stream));
    HIPCHK(c, launch_mmij_interact(q.sp, xj, yj, zj, jtype, j, q.out, c->stream));
    return BROMOC_OK;
}

int bromoc_interact(bromoc_ctx *c, double xj, double yj, double zj, int32_t jtype, int32_t j,
                    int32_t alert, double *dener) {
    if (!c || !dener) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    int st = interact_async(c, xj, yj, zj, jtype, j, alert);
    if (st) return st;
    HIPCHK(c, hipMemcpyAsync(dener, c->iout, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return sync_err(c);
}

int bromoc_set_gcmc(bromoc_ctx *c, const bromoc_gcmc *g) {
    if (!c || !g || g->nbuffer < 0) return BROMOC_E_ARG;
    const int nb = g->nbuffer;
    if (nb > 0 && (!g->ibfftyp || !g->avnum || !g->mu || !g->kb || !g->lzmin || !g->lzmax ||
                   !g->bufferbias || !g->ntotat))
        return fail(c, BROMOC_E_ARG, "gcmc arrays missing");
    const bool sphere = (c->sys.flags & BROMOC_QSPHERE) != 0;
    if (nb > 0 && sphere && (!g->rmin || !g->rmax))
        return fail(c, BROMOC_E_ARG, "gcmc rmin/rmax missing for a spherical system");
    for (int b = 1; b <= nb; b++)
        if (g->ibfftyp[b] <= c->sys.nold || g->ibfftyp[b] > c->sys.ntype)
            return fail(c, BROMOC_E_ARG, "buffer ion type out of range");
    c->g = *g;
    c->ibfftyp.assign(g->ibfftyp, g->ibfftyp + nb + 1);
    c->ntotat.assign(g->ntotat, g->ntotat + nb + 1);
    c->avnum.assign(g->avnum, g->avnum + nb + 1);
    c->mu.assign(g->mu, g->mu + nb + 1);
    c->kb.assign(g->kb, g->kb + nb + 1);
    c->lzmin.assign(g->lzmin, g->lzmin + nb + 1);
    c->lzmax.assign(g->lzmax, g->lzmax + nb + 1);
    if (sphere && nb > 0) {
        c->rmin.assign(g->rmin, g->rmin + nb + 1);
        c->rmax.assign(g->rmax, g->rmax + nb + 1);
    } else {
        c->rmin.assign(nb + 1, 0.0);
        c->rmax.assign(nb + 1, 0.0);
    }
    c->g.rmin = c->g.rmax = nullptr;   // the caller's arrays are not kept
    c->bufferbias.assign(g->bufferbias, g->bufferbias + nb + 1);
    c->ibuffer.assign(c->cap + 1, 0);
    c->gd.ib_current = false;
    // device copies for the device GRAND (arrays of a previous call are freed here)
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (void *p : c->gd.mem) (void)hipFree(p);
    c->gd.mem.clear();
    auto dev = [&](auto **ptr, size_t n, const void *src) -> hipError_t {
        *ptr = nullptr;
        hipError_t e = hipMalloc((void **)ptr, sizeof(**ptr) * std::max<size_t>(n, 1));
        if (e != hipSuccess) return e;
        c->gd.mem.push_back((void *)*ptr);
        if (src) return hipMemcpy(*ptr, src, sizeof(**ptr) * n, hipMemcpyHostToDevice);
        return hipMemset(*ptr, 0, sizeof(**ptr) * std::max<size_t>(n, 1));
    };
    const size_t n1 = nb + 1;
    HIPCHK(c, dev(&c->gd.avnum, n1, nb > 0 ? g->avnum : nullptr));
    HIPCHK(c, dev(&c->gd.mu, n1, nb > 0 ? g->mu : nullptr));
    HIPCHK(c, dev(&c->gd.kb, n1, nb > 0 ? g->kb : nullptr));
    HIPCHK(c, dev(&c->gd.lzmin, n1, nb > 0 ? g->lzmin : nullptr));
    HIPCHK(c, dev(&c->gd.lzmax, n1, nb > 0 ? g->lzmax : nullptr));
    HIPCHK(c, dev(&c->gd.rmin, n1, c->rmin.data()));
    HIPCHK(c, dev(&c->gd.rmax, n1, c->rmax.data()));
    HIPCHK(c, dev(&c->gd.ibfftyp, n1, nb > 0 ? g->ibfftyp : nullptr));
    HIPCHK(c, dev(&c->gd.ntotat, n1, nb > 0 ? g->ntotat : nullptr));
    HIPCHK(c, dev(&c->gd.bias, n1, nb > 0 ? g->bufferbias : nullptr));
    HIPCHK(c, dev(&c->gd.nat, n1, nullptr));
    HIPCHK(c, dev(&c->gd.ibuffer, (size_t)c->cap + 1, nullptr));
    HIPCHK(c, dev(&c->gd.st, 1, nullptr));
    HIPCHK(c, dev(&c->gd.part, GC_BLOCKS, nullptr));
    c->gd.nb = nb;
    c->have_gcmc = true;
    return BROMOC_OK;
}

int bromoc_set_ibuffer(bromoc_ctx *c, const int32_t *ib) {
    if (!c || !ib || !c->have_gcmc) return BROMOC_E_ARG;
    for (int i = 0; i < c->ntot; i++)
        if (ib[i] < 0 || ib[i] > c->g.nbuffer) return fail(c, BROMOC_E_ARG, "ibuffer out of range");
    std::copy(ib, ib + c->ntot, c->ibuffer.begin());
    c->gd.ib_current = false;
    return BROMOC_OK;
}

int bromoc_get_ibuffer(bromoc_ctx *c, int32_t *ib) {
    if (!c || !ib || !c->have_gcmc) return BROMOC_E_ARG;
    std::copy(c->ibuffer.begin(), c->ibuffer.begin() + c->ntot, ib);
    return BROMOC_OK;
}

// GRAND on the device (grand.hip): the caller's uniforms, nat and prob go up, ngcmc
// attempts run back to back on the stream, and the state comes back in one copy.
static int grand_device(bromoc_ctx *c, int32_t ngcmc, const double *u, int64_t nu,
                        int64_t *used, int32_t *nat, int32_t *prob, int64_t icycle,
                        double *ener) {
    const bromoc_gcmc &g = c->g;
    const int nb = g.nbuffer;
    const int pstride = c->cap + 1;
    auto &D = c->gd;
    if (nu > D.ucap) {
        if (D.u) (void)hipFree(D.u);
        D.u = nullptr;
        D.ucap = 0;
        HIPCHK(c, hipMalloc(&D.u, sizeof(double) * nu));
        D.ucap = nu;
    }
    const int64_t nlog = (int64_t)std::max(ngcmc, 1) * std::max(nb, 1);
    if (nlog > D.lcap) {
        if (D.natlog) (void)hipFree(D.natlog);
        D.natlog = nullptr;
        D.lcap = 0;
        HIPCHK(c, hipMalloc(&D.natlog, sizeof(int32_t) * nlog));
        D.lcap = nlog;
    }
    const int n0 = c->ntot;
    // pinned staging: uniforms, nat and the state up; the state, nat and the log back
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_nat = al(sizeof(double) * nu), o_st = o_nat + al(sizeof(int32_t) * (nb + 1));
    const size_t o_st2 = o_st + al(sizeof(GcState)), o_nat2 = o_st2 + al(sizeof(GcState));
    const size_t o_log = o_nat2 + al(sizeof(int32_t) * (nb + 1));
    unsigned char *pin = pinned(c, o_log + sizeof(int32_t) * (size_t)nlog);
    if (!pin) return fail(c, BROMOC_E_HIP, "pinned staging");
    if (nu > 0) {
        std::memcpy(pin, u, sizeof(double) * nu);
        HIPCHK(c, hipMemcpyAsync(D.u, pin, sizeof(double) * nu, hipMemcpyHostToDevice, c->stream));
    }
    std::memcpy(pin + o_nat, nat, sizeof(int32_t) * (nb + 1));
    HIPCHK(c, hipMemcpyAsync(D.nat, pin + o_nat, sizeof(int32_t) * (nb + 1), hipMemcpyHostToDevice, c->stream));
    // the device's ibuffer is the host's unless the host changed it since the last call
    if (!D.ib_current)
        HIPCHK(c, hipMemcpyAsync(D.ibuffer, c->ibuffer.data(), sizeof(int32_t) * n0, hipMemcpyHostToDevice, c->stream));
    GcState h{};
    h.k = 0; h.ntot = n0; h.err = GCE_OK; h.ener = *ener;
    std::memcpy(pin + o_st, &h, sizeof(h));
    HIPCHK(c, hipMemcpyAsync(D.st, pin + o_st, sizeof(h), hipMemcpyHostToDevice, c->stream));
    GcParams G{};
    G.sp = params_for(c);
    G.pos0 = c->pos[0]; G.pos1 = c->pos[1];
    G.typei = c->typei; G.t2 = c->t2; G.ibuffer = D.ibuffer;
    G.u = D.u; G.nu = nu;
    G.nbuffer = nb;
    G.nat = D.nat; G.natlog = D.natlog;
    G.avnum = D.avnum; G.mu = D.mu; G.kb = D.kb; G.lzmin = D.lzmin; G.lzmax = D.lzmax;
    G.rmin = D.rmin; G.rmax = D.rmax;
    G.ibfftyp = D.ibfftyp; G.ntotat = D.ntotat; G.bias = D.bias;
    G.lx = g.lx; G.ly = g.ly; G.cx = g.cx; G.cy = g.cy; G.cz = g.cz; G.ikbt = g.ikbt;
    G.icycle = icycle;
    G.cap = c->cap;
    G.st = D.st; G.part = D.part;
    HIPCHK(c, launch_grand(G, ngcmc, c->stream));
    HIPCHK(c, hipMemcpyAsync(pin + o_st2, D.st, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    int st = sync_err(c);
    if (st) return st;
    std::memcpy(&h, pin + o_st2, sizeof(h));
    const int n1 = h.ntot;
    HIPCHK(c, hipMemcpyAsync(pin + o_nat2, D.nat, sizeof(int32_t) * (nb + 1), hipMemcpyDeviceToHost, c->stream));
    D.h_natlog.resize((size_t)nlog);
    if (nb > 0 && h.natt > 0)
        HIPCHK(c, hipMemcpyAsync(pin + o_log, D.natlog, sizeof(int32_t) * (size_t)h.natt * nb,
                                 hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->ibuffer.data(), D.ibuffer, sizeof(int32_t) * n1, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->h_typei.data(), c->typei, sizeof(int32_t) * n1, hipMemcpyDeviceToHost, c->stream));
    st = sync_err(c);
    if (st) return st;
    std::memcpy(nat, pin + o_nat2, sizeof(int32_t) * (nb + 1));
    if (nb > 0 && h.natt > 0)
        std::memcpy(D.h_natlog.data(), pin + o_log, sizeof(int32_t) * (size_t)h.natt * nb);
    D.ib_current = true;
    // prob(b, count) += 1 for every logged count, attempt by attempt
    for (int a = 0; a < h.natt; a++)
        for (int b = 1; b <= nb; b++)
            prob[(size_t)(b - 1) * pstride + D.h_natlog[(size_t)a * nb + (b - 1)]] += 1;
    if (h.nacc > 0) c->sort_dirty = true;
    c->ntot = n1;
    *ener = h.ener;
    if (used) *used = h.k;
    switch (h.err) {
        case GCE_UNIFORMS: return fail(c, BROMOC_E_ARG, "uniform stream exhausted");
        case GCE_CAPACITY: return fail(c, BROMOC_E_CAPACITY, "ntot exceeds capacity");
        case GCE_STATE: return fail(c, BROMOC_E_STATE, "buffer bookkeeping inconsistent");
        default: return BROMOC_OK;
    }
}

// GRAND, grand.f90:40-118 (INSERT insert.f90:29-41 rectangular; FIND find.f90:26-32).  On
// the device (grand_device) except for MMIJ systems, whose INTERACT term (rect_rf0.f90)
// runs as its own kernel sequence: their attempts go through the host loop below.
int bromoc_grand(bromoc_ctx *c, int32_t ngcmc, const double *u, int64_t nu, int64_t *used,
                 int32_t *nat, int32_t *prob, int64_t icycle, double *ener) {
    if (!c || !u || !nat || !prob || !ener || ngcmc < 0 || icycle <= 0) return BROMOC_E_ARG;
    if (!c->have_gcmc) return fail(c, BROMOC_E_STATE, "bromoc_set_gcmc first");
    (void)hipSetDevice(c->device);
    static const bool host_grand = [] {
        const char *v = std::getenv("BROMOC_HOST_GRAND");
        return v && v[0] == '1';
    }();
    if (!c->base.mmon && !host_grand)
        return grand_device(c, ngcmc, u, nu, used, nat, prob, icycle, ener);
    c->gd.ib_current = false;   // the host loop below edits ibuffer
    const bromoc_gcmc &g = c->g;
    const int nbuffer = g.nbuffer;
    const double nbufferr = (double)nbuffer;
    const int pstride = c->cap + 1;
    const Curved cv = curved_of(&c->sys);
    int64_t k = 0;
    auto rndm = [&](double &v) {
        if (k >= nu) return false;
        v = u[k++];
        return true;
    };
#define RNDM(v) do { if (!rndm(v)) return fail(c, BROMOC_E_ARG, "uniform stream exhausted"); } while (0)
    double dener = 0.0;
    for (int ig = 0; ig < ngcmc; ig++) {
        double r0;
        RNDM(r0);
        if (r0 < 0.5) {
            double r1;
            RNDM(r1);
            const int ib = (int)(nbufferr * r1) + 1;
            if (ib <= nbuffer && c->avnum[ib] > 0.0) {
                const int itype = c->ibfftyp[ib];
                double xn, yn, zn;
                do {   // INSERT (insert.f90:29-41)
                    double a, b, d;
                    RNDM(a); RNDM(b); RNDM(d);
                    xn = g.lx * (a - 0.5) + g.cx;
                    yn = g.ly * (b - 0.5) + g.cy;
                    zn = c->lzmin[ib] + (c->lzmax[ib] - c->lzmin[ib]) * d;
                } while (!cv.insert_ok(xn, yn, zn, g.cx, g.cy, g.cz, c->rmin[ib], c->rmax[ib]));
                int st = bromoc_interact(c, xn, yn, zn, itype, c->ntot, 0, &dener);
                if (st) return st;
                double rate;
                if (c->bufferbias[ib]) {
                    rate = (c->avnum[ib] + c->kb[ib] * (c->avnum[ib] - (double)(c->ntotat[ib] / icycle)));
                    rate = rate * std::exp(-(dener - c->mu[ib]) * g.ikbt) / (double)(nat[ib] + 1);
                } else {
                    rate = (c->avnum[ib] / (nat[ib] + 1)) * std::exp(-(dener - c->mu[ib]) * g.ikbt);
                }
                rate = rate / (1.0 + rate);
                double r2;
                RNDM(r2);
                if (r2 <= rate) {
                    if (c->ntot + 1 > c->cap) return fail(c, BROMOC_E_CAPACITY, "ntot exceeds capacity");
                    *ener = *ener + dener;
                    nat[ib] = nat[ib] + 1;
                    const int n = c->ntot;
                    const int32_t ti = zn < g.cz ? -itype : itype;
                    const int32_t t2 = c->h_nwtype[itype];
                    c->h_move = make_double4(xn, yn, zn, pack_type(t2, n));
                    for (int bb = 0; bb < 2; bb++)
                        HIPCHK(c, hipMemcpyAsync(c->pos[bb] + n, &c->h_move, sizeof(double4), hipMemcpyHostToDevice, c->stream));
                    HIPCHK(c, hipMemcpyAsync(c->typei + n, &ti, sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
                    HIPCHK(c, hipMemcpyAsync(c->t2 + n, &t2, sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
                    int st2 = sync_err(c);
                    if (st2) return st2;
                    c->h_typei[n] = ti;
                    c->ibuffer[n] = ib;
                    c->ntot = n + 1;
                    c->sort_dirty = true;
                }
            }
        } else {
            double r1, r2;
            RNDM(r1);
            const int ib = (int)(nbufferr * r1) + 1;
            RNDM(r2);
            const double natib = ib <= nbuffer ? (double)nat[ib] : 0.0;
            const int ip = (int)(natib * r2) + 1;
            if (ib <= nbuffer && ip <= nat[ib]) {
                const int itype = c->ibfftyp[ib];
                int iat = -1, ipx = 0;
                for (int a = c->nf + c->ns; a < c->ntot; a++)
                    if (c->ibuffer[a] == ib && ++ipx == ip) { iat = a; break; }
                if (iat < 0) return fail(c, BROMOC_E_STATE, "buffer bookkeeping inconsistent");
                double4 pa;
                HIPCHK(c, hipMemcpyAsync(&pa, c->pos[c->cur] + iat, sizeof(double4), hipMemcpyDeviceToHost, c->stream));
                int st = sync_err(c);
                if (st) return st;
                st = bromoc_interact(c, pa.x, pa.y, pa.z, itype, iat, 1, &dener);
                if (st) return st;
                double rate;
                if (c->bufferbias[ib]) {
                    rate = (c->avnum[ib] + c->kb[ib] * (c->avnum[ib] - (double)(c->ntotat[ib] / icycle)));
                    rate = rate * std::exp(-(dener - c->mu[ib]) * g.ikbt) / (double)nat[ib];
                } else {
                    rate = (c->avnum[ib] / nat[ib]) * std::exp(-(dener - c->mu[ib]) * g.ikbt);
                }
                rate = 1.0 / (1.0 + rate);
                double r3;
                RNDM(r3);
                if (r3 <= rate) {
                    *ener = *ener - dener;
                    const int last = c->ntot - 1;
                    // grand.f90:102-109: the last particle takes the deleted one's slot
                    HIPCHK(c, hipMemcpyAsync(&c->h_move, c->pos[c->cur] + last, sizeof(double4), hipMemcpyDeviceToHost, c->stream));
                    int st1 = sync_err(c);
                    if (st1) return st1;
                    c->h_move.w = pack_type(c->h_nwtype[std::abs(c->h_typei[last])], iat);
                    for (int bb = 0; bb < 2; bb++)
                        HIPCHK(c, hipMemcpyAsync(c->pos[bb] + iat, &c->h_move, sizeof(double4), hipMemcpyHostToDevice, c->stream));
                    HIPCHK(c, hipMemcpyAsync(c->typei + iat, c->typei + last, sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
                    HIPCHK(c, hipMemcpyAsync(c->t2 + iat, c->t2 + last, sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
                    int st2 = sync_err(c);
                    if (st2) return st2;
                    c->h_typei[iat] = c->h_typei[last];
                    c->ibuffer[iat] = c->ibuffer[last];
                    nat[ib] = nat[ib] - 1;
                    c->ntot = last;
                    c->sort_dirty = true;
                }
            }
        }
        for (int b = 1; b <= nbuffer; b++) {
            if (nat[b] < 0 || nat[b] > c->cap) return fail(c, BROMOC_E_STATE, "nat out of range");
            prob[(size_t)(b - 1) * pstride + nat[b]] += 1;
        }
    }
#undef RNDM
    if (used) *used = k;
    return BROMOC_OK;
}

// METROPOLIS, metropolis.f90:31-55, with MOVE (move.f90:32-47, rectangular box).  The
// positions live on the device; a host copy taken at entry serves MOVE's draws and is kept
// in step with every accepted move.  Per trial the two INTERACT calls (old position with
// Qalert, trial position without) run back to back and are read in one copy.
int bromoc_metropolis(bromoc_ctx *c, int32_t nmcm, double mcmax, double ikbt,
                      const double *u, int64_t nu, int64_t *used, int32_t *done, double *ener) {
    if (!c || (!u && nu > 0) || nu < 0 || !ener || nmcm < 0) return BROMOC_E_ARG;
    if (used) *used = 0;
    if (done) *done = 0;
    if (nmcm == 0) return BROMOC_OK;
    if (!(c->sys.flags & BROMOC_QPAR)) return fail(c, BROMOC_E_STATE, "METROPOLIS needs ions");
    const int first = c->ns + c->nf, nmove = c->ntot - first;
    if (nmove <= 0) return fail(c, BROMOC_E_STATE, "METROPOLIS: no moving ions");
    (void)hipSetDevice(c->device);
    std::vector<double4> hp(c->ntot);
    HIPCHK(c, hipMemcpyAsync(hp.data(), c->pos[c->cur], sizeof(double4) * c->ntot,
                             hipMemcpyDeviceToHost, c->stream));
    int st = sync_err(c);
    if (st) return st;
    const Curved cv = curved_of(&c->sys);
    const double natom2 = (double)nmove;   // natom - nsites (move.f90:33)
    int64_t k = 0;
    int32_t ndone = 0;
    for (int im = 0; im < nmcm; im++) {
        int64_t kk = k;
        int iat = -1;
        double xn = 0.0, yn = 0.0, zn = 0.0;
        bool out = true;
        while (out) {
            if (kk + 4 > nu) break;
            if (!(u[kk] >= 0.0 && u[kk] < 1.0)) return fail(c, BROMOC_E_ARG, "uniform outside [0,1)");
            iat = first + (int)(natom2 * u[kk]);
            const double4 pa = hp[iat];
            xn = pa.x + mcmax * (u[kk + 1] - 0.5);
            yn = pa.y + mcmax * (u[kk + 2] - 0.5);
            zn = pa.z + mcmax * (u[kk + 3] - 0.5);
            kk += 4;
            out = cv.outside(xn, yn, zn);
        }
        if (out || kk + 1 > nu) break;   // the trial could not finish
        const int jt = std::abs(c->h_typei[iat]);
        const double4 pa = hp[iat];
        if ((st = interact_async(c, pa.x, pa.y, pa.z, jt, iat, 1, 0))) return st;
        if ((st = interact_async(c, xn, yn, zn, jt, iat, 0, 1))) return st;
        double eo[2];
        HIPCHK(c, hipMemcpyAsync(eo, c->iout, sizeof(eo), hipMemcpyDeviceToHost, c->stream));
        if ((st = sync_err(c))) return st;
        const double eold = eo[0], enew = eo[1];
        bool acc = enew <= eold;
        if (!acc) {
            const double bltz = std::exp(-(enew - eold) * ikbt);
            acc = u[kk++] < bltz;
        }
        if (acc) {
            hp[iat] = make_double4(xn, yn, zn, pa.w);
            for (int b = 0; b < 2; b++)
                HIPCHK(c, hipMemcpyAsync(c->pos[b] + iat, &hp[iat], sizeof(double4), hipMemcpyHostToDevice, c->stream));
            *ener = *ener + (enew - eold);
            c->sort_dirty = true;
        }
        k = kk;
        ndone = im + 1;
    }
    if ((st = sync_err(c))) return st;
    if (used) *used = k;
    if (done) *done = ndone;
    return BROMOC_OK;
}

// EFPOT tables (splinepot.f90:19-96, discretize.f90:19-120): the shared construction in
// tables.h, on the host here and on the device in tables.hip.
int bromoc_build_table(int32_t n, const double *r, const double *u, int32_t chr, double fct,
                       double *ep, bromoc_pair *pair) {
    if (n < 2 || !r || !u || !ep || !pair) return BROMOC_E_ARG;
    if (r[0] == 0.0 || r[n - 1] == 0.0) return BROMOC_E_ARG;
    for (int k = 1; k < n; k++)
        if (!(r[k] > r[k - 1])) return BROMOC_E_ARG;
    std::vector<double> scratch(3 * (size_t)n);
    TabEnds o{};
    table_splinepot(n, r, u, chr ? fct : 0.0, ep, o, scratch.data());
    pair->dmi = o.dmi; pair->dm2lo = o.dm2lo; pair->dm2hi = o.dm2hi;
    pair->sc1 = o.sc1; pair->sc2 = o.sc2; pair->sc3 = o.sc3;
    pair->nknots = n;
    pair->efpot = 1;
    pair->chr = chr;
    pair->fct = fct;
    return BROMOC_OK;
}

static TabPair tab_pair_of(const bromoc_pair &p) {
    TabPair t{};
    t.chr = p.chr; t.lj = p.lj; t.srpmf = p.srpmf;
    t.fct = p.fct; t.epp4 = p.epp4; t.sgp2 = p.sgp2;
    t.c0 = p.c0; t.c1 = p.c1; t.c2 = p.c2; t.c3 = p.c3; t.c4 = p.c4;
    return t;
}

int bromoc_discretize(const bromoc_pair *params, int32_t nnp, int32_t mnp, double res, double rth,
                      double srpx, double srpk, double srpy, double *ep, bromoc_pair *out) {
    if (!params || !ep || !out || nnp < 1 || mnp - nnp + 1 < 3 || !(res > 0.0)) return BROMOC_E_ARG;
    const int n = mnp - nnp + 1;
    std::vector<double> scratch(5 * (size_t)n);
    TabEnds o{};
    table_discretize(tab_pair_of(*params), nnp, mnp, res, rth, srpx, srpk, srpy, ep, o,
                     scratch.data());
    out->sc1 = o.sc1; out->sc2 = o.sc2; out->sc3 = o.sc3;
    return BROMOC_OK;
}

// SPLINE (spline.f90:19-98): the Forsythe-Malcolm-Moler cubic interpolating spline with
// end third derivatives from divided differences; a tridiagonal solve by forward
// elimination and back substitution, in the reference's operation order.
int bromoc_spline(int32_t n, const double *x, const double *y, double *b, double *c, double *d) {
    if (n < 2 || !x || !y || !b || !c || !d) return BROMOC_E_ARG;
    const int nm1 = n - 1;
    // 1-based views
    const double *X = x - 1, *Y = y - 1;
    double *B = b - 1, *Cc = c - 1, *D = d - 1;
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
            Cc[1] = Cc[1] * (D[1] * D[1]) / (X[4] - X[1]);
            Cc[n] = -Cc[n] * (D[n - 1] * D[n - 1]) / (X[n] - X[n - 3]);
        }
        for (int i = 2; i <= n; i++) {
            const double t = D[i - 1] / B[i - 1];
            B[i] = B[i] - t * D[i - 1];
            Cc[i] = Cc[i] - t * Cc[i - 1];
        }
        Cc[n] = Cc[n] / B[n];
        for (int ib = 1; ib <= nm1; ib++) {
            const int i = n - ib;
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

int bromoc_update_tables(bromoc_ctx *c, int32_t ntab, const int32_t *is, const int32_t *n,
                         const double *r, const double *u) {
    if (!c || ntab < 1 || !is || !n || !r || !u) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    const int np = c->sys.npair;
    std::vector<int32_t> tab_of(np + 1, -1), nxi(ntab), nk(ntab);
    std::vector<int64_t> off(ntab + 1, 0);
    std::vector<double> kf(ntab);
    for (int t = 0; t < ntab; t++) {
        if (is[t] < 1 || is[t] > np) return fail(c, BROMOC_E_ARG, "update_tables: pair type");
        const bromoc_pair &p = c->h_pairs[is[t]];
        if (!p.efpot || n[t] != p.nknots || n[t] < 2)
            return fail(c, BROMOC_E_ARG, "update_tables: not a table of that many knots");
        if (tab_of[is[t]] >= 0) return fail(c, BROMOC_E_ARG, "update_tables: pair type twice");
        tab_of[is[t]] = t;
        nxi[t] = p.nxi;
        nk[t] = n[t];
        kf[t] = p.chr ? p.fct : 0.0;
        off[t + 1] = off[t] + n[t];
    }
    const int64_t tot = off[ntab];
    for (int t = 0; t < ntab; t++) {
        const double *rt = r + off[t];
        if (rt[0] == 0.0) return fail(c, BROMOC_E_ARG, "update_tables: r");
        for (int k = 1; k < n[t]; k++)
            if (!(rt[k] > rt[k - 1])) return fail(c, BROMOC_E_ARG, "update_tables: r not increasing");
    }
    // staging: r, u, per-table metadata, scratch and the results
    TabJob jb{};
    jb.ntab = ntab;
    std::vector<void *> tmp;
    auto up = [&](void **dst, const void *src, size_t bytes) -> hipError_t {
        hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 8));
        if (e != hipSuccess) return e;
        tmp.push_back(*dst);
        return src ? hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, c->stream) : hipSuccess;
    };
    hipError_t e = hipSuccess;
    if (!e) e = up((void **)&jb.r, r, sizeof(double) * tot);
    if (!e) e = up((void **)&jb.u, u, sizeof(double) * tot);
    if (!e) e = up((void **)&jb.off, off.data(), sizeof(int64_t) * (ntab + 1));
    if (!e) e = up((void **)&jb.nxi, nxi.data(), sizeof(int32_t) * ntab);
    if (!e) e = up((void **)&jb.kf, kf.data(), sizeof(double) * ntab);
    if (!e) e = up((void **)&jb.tab_of, tab_of.data(), sizeof(int32_t) * (np + 1));
    if (!e) e = up((void **)&jb.scratch, nullptr, sizeof(double) * 6 * tot);
    if (!e) e = up((void **)&jb.ends, nullptr, sizeof(TabEnds) * ntab);
    if (!e) {
        jb.ep = c->ep;
        jb.pairs = c->pairs;
        jb.rows = c->rows;
        jb.pairidx = c->pairidx;
        jb.nrows = (c->sys.nttyp + 1) * (c->sys.nttyp + 1);
        jb.npair = np;
        e = launch_table_update(jb, c->stream);
    }
    int st = e ? fail(c, BROMOC_E_HIP, std::string("update_tables: ") + hipGetErrorString(e)) : sync_err(c);
    for (void *q : tmp) (void)hipFree(q);
    if (st) return st;
    // the host mirror and the squared range past which every table pair is in its tail
    c->r2near = 0.0;
    for (int t = 0; t < ntab; t++) {
        const double rn = r[off[t] + n[t] - 1], r1 = r[off[t]];
        bromoc_pair &p = c->h_pairs[is[t]];
        p.dmi = r1; p.dm2lo = r1 * r1; p.dm2hi = rn * rn;
    }
    bool all = true;
    for (int32_t k : c->h_pairidx) {
        if (k <= 0) continue;
        if (!c->h_pairs[k].efpot) all = false;
        else c->r2near = std::max(c->r2near, c->h_pairs[k].dm2hi);
    }
    if (!all) c->r2near = HUGE_VAL;
    c->base.r2near = c->r2near;
    return BROMOC_OK;
}

int bromoc_table_vjp_device(bromoc_ctx *c, int32_t n, const double *r, const double *grad_ep,
                            const double *grad_sc, double *grad_u) {
    if (!c || n < 2 || !r || !grad_ep || !grad_sc || !grad_u) return BROMOC_E_ARG;
    if (r[0] == 0.0) return BROMOC_E_ARG;
    for (int k = 1; k < n; k++)
        if (!(r[k] > r[k - 1])) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    double *d_r = nullptr, *d_gep = nullptr, *d_gsc = nullptr, *d_out = nullptr, *d_scr = nullptr;
    hipError_t e = hipMalloc((void **)&d_r, sizeof(double) * n);
    if (!e) e = hipMalloc((void **)&d_gep, sizeof(double) * 3 * n);
    if (!e) e = hipMalloc((void **)&d_gsc, sizeof(double) * 3);
    if (!e) e = hipMalloc((void **)&d_out, sizeof(double) * n);
    if (!e) e = hipMalloc((void **)&d_scr, sizeof(double) * 6 * n);
    if (!e) e = hipMemcpyAsync(d_r, r, sizeof(double) * n, hipMemcpyHostToDevice, c->stream);
    if (!e) e = hipMemcpyAsync(d_gep, grad_ep, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream);
    if (!e) e = hipMemcpyAsync(d_gsc, grad_sc, sizeof(double) * 3, hipMemcpyHostToDevice, c->stream);
    if (!e) e = launch_table_vjp(n, d_r, d_gep, d_gsc, d_out, d_scr, c->stream);
    if (!e) e = hipMemcpyAsync(grad_u, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
    int st = e ? fail(c, BROMOC_E_HIP, std::string("table_vjp_device: ") + hipGetErrorString(e)) : sync_err(c);
    for (double *q : {d_r, d_gep, d_gsc, d_out, d_scr})
        if (q) (void)hipFree(q);
    return st;
}

// The same product with no allocation: inputs, outputs and the sweep's workspace live in
// the device's fixed `scratch` array (tables.hip), whose address was taken at bromoc_create.
// Layout: r [n], grad_ep [3n], grad_sc [3], grad_u [n], workspace [6n]: 11n + 3 doubles.
int bromoc_table_vjp_device_noalloc(bromoc_ctx *c, int32_t n, const double *r,
                                    const double *grad_ep, const double *grad_sc,
                                    double *grad_u) {
    if (!c || n < 2 || !r || !grad_ep || !grad_sc || !grad_u) return BROMOC_E_ARG;
    if (r[0] == 0.0) return BROMOC_E_ARG;
    for (int k = 1; k < n; k++)
        if (!(r[k] > r[k - 1])) return BROMOC_E_ARG;
    const size_t cap = vjp_scratch_doubles();
    const size_t need = 11 * (size_t)n + 3;
    if (need > cap)
        return fail(c, BROMOC_E_CAPACITY,
                    "table_vjp_device_noalloc: n = " + std::to_string(n) + " needs " +
                        std::to_string(need) + " doubles of workspace; " + std::to_string(cap) +
                        " exist (n <= " + std::to_string((cap - 3) / 11) + ")");
    (void)hipSetDevice(c->device);
    double *base = nullptr;
    HIPCHK(c, vjp_scratch_init(c->device, &base));
    double *d_r = base, *d_gep = d_r + n, *d_gsc = d_gep + 3 * (size_t)n, *d_out = d_gsc + 3;
    double *d_scr = d_out + n;
    HIPCHK(c, hipMemcpyAsync(d_r, r, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_gep, grad_ep, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_gsc, grad_sc, sizeof(double) * 3, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_table_vjp(n, d_r, d_gep, d_gsc, d_out, d_scr, c->stream));
    HIPCHK(c, hipMemcpyAsync(grad_u, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    return sync_err(c);
}

int bromoc_table_vjp(int32_t n, const double *r, const double *grad_ep, const double *grad_sc,
                     double *grad_u) {
    if (n < 2 || !r || !grad_ep || !grad_sc || !grad_u) return BROMOC_E_ARG;
    if (r[0] == 0.0 || r[n - 1] == 0.0) return BROMOC_E_ARG;
    for (int k = 1; k < n; k++)
        if (!(r[k] > r[k - 1])) return BROMOC_E_ARG;
    std::vector<double> scratch(6 * (size_t)n);
    table_splinepot_vjp(n, r, grad_ep, grad_sc, grad_u, scratch.data());
    return BROMOC_OK;
}

// ---- trajectory adjoint of an RDF loss (kernels in adjoint.hip)
namespace {

// What the adjoint and the tangent pass share: the checks, the RDF groups and each group's
// pair count, scratch buffers, the forward steps and the loss.
struct AdjRun {
    bromoc_ctx *c;
    int nsteps, K, nt, nt1, nf0, ng, nb, nh, nk, nseg;
    const bromoc_rdf *rdf;
    std::vector<int32_t> group;
    std::vector<double> npairs;
    std::vector<void *> tmp;
    int32_t *d_group = nullptr;
    double4 *ckpt = nullptr;
    double *hist = nullptr;
    AdjParams A{};

    ~AdjRun() { for (void *q : tmp) (void)hipFree(q); }

    hipError_t alloc(void **q, size_t bytes) {
        hipError_t e = hipMalloc(q, std::max<size_t>(bytes, 8));
        if (e == hipSuccess) {
            tmp.push_back(*q);
            e = hipMemsetAsync(*q, 0, std::max<size_t>(bytes, 8), c->stream);
        }
        return e;
    }

    int setup(bromoc_ctx *c_, int32_t nsteps_, int32_t K_, const bromoc_rdf *rdf_) {
        c = c_; nsteps = nsteps_; K = K_; rdf = rdf_;
        if (nsteps < 1 || K < 1 || !rdf || rdf->ngroup < 1 || rdf->nbins < 1 ||
            rdf->ngroup * (int64_t)rdf->nbins > 4096 || !rdf->groups || !(rdf->dr > 0.0) ||
            !(rdf->sigma > 0.0) || !(rdf->cut >= 0.0) || !(rdf->wflux >= 0.0))
            return BROMOC_E_ARG;
        const bromoc_system &S = c->sys;
        const uint32_t allowed = BROMOC_QPAR | BROMOC_QNONBOND | BROMOC_QNOBOND | BROMOC_QCOUNTION |
                                 BROMOC_QPHIX | BROMOC_QPHIV | BROMOC_QSVDW | BROMOC_QNMCDEN;
        if ((S.flags & ~allowed) || c->ns != 0 || !c->base.all_efpot)
            return fail(c, BROMOC_E_UNSUPPORTED,
                        "adjoint: moving ions with tabulated pairs and the trilinear grids only "
                        "(no B-spline map, CG-DNA, RFPAR, diffusion profiles)");
        if (rdf->wflux > 0.0 && !(c->base.dt > 0.0))
            return fail(c, BROMOC_E_ARG, "adjoint: the flux term needs the time step dt");
        nt = c->ntot; nt1 = S.nttyp + 1; nf0 = c->nf;
        if (nt1 > MAX_LDS_TYPES) return fail(c, BROMOC_E_UNSUPPORTED, "adjoint: too many types");
        if (!c->noise || c->noise_rows < nsteps || c->noise_stride < 3 * (int64_t)nt)
            return fail(c, BROMOC_E_STATE, "adjoint: injected noise rows 0..nsteps-1 required");
        ng = rdf->ngroup; nb = rdf->nbins; nh = ng * nb;
        group.assign((size_t)nt1 * nt1, -1);
        std::vector<int64_t> nmov(S.ntype + 1, 0), nfix(S.ntype + 1, 0);
        for (int i = 0; i < nt; i++) (i < nf0 ? nfix : nmov)[std::abs(c->h_typei[i])]++;
        npairs.assign(ng, 0.0);
        for (int g = 0; g < ng; g++) {
            const int a = rdf->groups[2 * g], b = rdf->groups[2 * g + 1];
            if (a <= S.nold || a > S.ntype || b <= S.nold || b > S.ntype)
                return fail(c, BROMOC_E_ARG, "adjoint: RDF group type is not an ion type");
            const int ta = c->h_nwtype[a], tb = c->h_nwtype[b];
            if (group[ta * nt1 + tb] >= 0) return fail(c, BROMOC_E_ARG, "adjoint: RDF groups overlap");
            group[ta * nt1 + tb] = group[tb * nt1 + ta] = g;
            double pc = a == b ? 0.5 * (double)nmov[a] * (double)(nmov[a] - 1)
                               : (double)nmov[a] * (double)nmov[b];
            pc += (double)nmov[a] * (double)nfix[b] + (a != b ? (double)nmov[b] * (double)nfix[a] : 0.0);
            npairs[g] = pc;
        }
        int st = autotune(c);
        if (st) return st;
        nseg = (nsteps + K - 1) / K;
        nk = S.nknots + 2;
        const size_t v4 = sizeof(double4) * (size_t)nt;
        HIPCHK(c, alloc((void **)&ckpt, v4 * nseg));
        HIPCHK(c, alloc((void **)&hist, sizeof(double) * (size_t)nt * nh));
        HIPCHK(c, alloc((void **)&d_group, sizeof(int32_t) * group.size()));
        HIPCHK(c, hipMemcpyAsync(d_group, group.data(), sizeof(int32_t) * group.size(),
                                 hipMemcpyHostToDevice, c->stream));
        // injected noise row s of step s: ctl = {0, 0, ...} and step_off = s
        c->h_ctl[0] = 0; c->h_ctl[1] = 0; c->h_ctl[2] = 0; c->h_ctl[3] = 0;
        HIPCHK(c, hipMemcpyAsync(c->ctl, c->h_ctl, sizeof(c->h_ctl), hipMemcpyHostToDevice, c->stream));
        A.sp = params_for(c);
        A.group = d_group;
        A.ngroup = ng; A.nbins = nb;
        A.r0 = rdf->r0; A.dr = rdf->dr; A.sigma = rdf->sigma; A.cut = rdf->cut;
        A.hist = hist;
        return BROMOC_OK;
    }

    StepParams step_params(int s) {
        StepParams p = params_for(c);
        // one fixed launch shape (two targets per workgroup), not the autotuner's timed
        // pick: every pass, and every context of the same system, replays the trajectory
        // bitwise
        p.tpb = 2;
        p.mode = 1;
        p.write_energy = 0;
        p.write_forces = 0;
        p.noise_mode = BROMOC_NOISE_INJECTED;
        p.noise = c->noise;
        p.noise_stride = c->noise_stride;
        p.step_off = s;
        p.cntpts = 0;          // the trajectory is replayed: no flux counting
        return p;
    }

    // step s of the forward pass, from the current positions (the partner order is rebuilt
    // at every checkpoint, so a segment recomputed from its checkpoint is bitwise the same)
    int forward_step(int s) {
        if (s % K == 0) {
            HIPCHK(c, hipMemcpyAsync(ckpt + (size_t)(s / K) * nt, c->pos[c->cur],
                                     sizeof(double4) * (size_t)nt, hipMemcpyDeviceToDevice, c->stream));
            int sst = ensure_sorted(c, true);
            if (sst) return sst;
        }
        HIPCHK(c, launch_step(step_params(s), c->stream));
        c->cur ^= 1;
        A.x = c->pos[c->cur];
        HIPCHK(c, launch_adj(0, A, c->stream));
        return BROMOC_OK;
    }

    // the loss of the forward pass (positions at their last state), its per-bin weights
    // dL/dH, and the flux term's coefficient dL/dJ / (nsteps dt)
    int loss(double *L, double *rdf_out, std::vector<double> &w, double &zcoef) {
        std::vector<double> H((size_t)nt * nh);
        std::vector<double4> x0(nt), xT(nt);
        HIPCHK(c, hipMemcpyAsync(H.data(), hist, sizeof(double) * H.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(x0.data(), ckpt, sizeof(double4) * nt, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(xT.data(), c->pos[c->cur], sizeof(double4) * nt, hipMemcpyDeviceToHost, c->stream));
        int st = sync_err(c);
        if (st) return st;
        const bromoc_system &S = c->sys;
        const double PI_ = 3.14159265358979323846264338327950288419716939937510;
        const double V = (S.lx2p - S.lx2m) * (S.ly2p - S.ly2m) * (S.lz2p - S.lz2m);
        std::vector<double> g(nh, 0.0);
        w.assign(nh, 0.0);
        double l = 0.0;
        for (int k = 0; k < nh; k++) {
            double h = 0.0;
            for (int i = nf0; i < nt; i++) h += H[(size_t)i * nh + k];
            const int gi = k / nb, b = k % nb;
            const double rb = rdf->r0 + (b + 0.5) * rdf->dr;
            const double cn = npairs[gi] > 0.0
                ? V / ((double)nsteps * npairs[gi] * 4.0 * PI_ * rb * rb * rdf->sigma * std::sqrt(2.0 * PI_))
                : 0.0;
            g[k] = cn * h;
            const double res = g[k] - (rdf->target ? rdf->target[k] : 0.0);
            l += res * res;
            w[k] = 2.0 * res * cn;
        }
        zcoef = 0.0;
        if (rdf->wflux > 0.0) {
            double q = 0.0;
            for (int i = nf0; i < nt; i++) q += c->h_cg[std::abs(c->h_typei[i])] * (xT[i].z - x0[i].z);
            const double tdt = (double)nsteps * c->base.dt;
            const double J = q / tdt;
            l += rdf->wflux * (J - rdf->jtarget) * (J - rdf->jtarget);
            zcoef = 2.0 * rdf->wflux * (J - rdf->jtarget) / tdt;
        }
        *L = l;
        if (rdf_out) std::memcpy(rdf_out, g.data(), sizeof(double) * nh);
        return BROMOC_OK;
    }
};

}  // namespace

int bromoc_adjoint_rdf(bromoc_ctx *c, int32_t nsteps, int32_t K, const bromoc_rdf *rdf,
                       double *loss, double *rdf_out, double *grad_ep, double *grad_sc) {
    if (!c || !loss) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    const bromoc_system &S = c->sys;
    AdjRun R;
    int st = R.setup(c, nsteps, K, rdf);
    if (st) return st;
    const int nt = R.nt;
    const size_t v4 = sizeof(double4) * (size_t)nt;
    double4 *traj, *xS, *F, *mu, *dbox, *lam[2];
    double *d_w;
    unsigned long long *gep, *gsc;
    HIPCHK(c, R.alloc((void **)&traj, v4 * K));
    HIPCHK(c, R.alloc((void **)&xS, v4));
    HIPCHK(c, R.alloc((void **)&F, v4));
    HIPCHK(c, R.alloc((void **)&mu, v4));
    HIPCHK(c, R.alloc((void **)&dbox, v4));
    HIPCHK(c, R.alloc((void **)&lam[0], v4));
    HIPCHK(c, R.alloc((void **)&lam[1], v4));
    HIPCHK(c, R.alloc((void **)&d_w, sizeof(double) * R.nh));
    HIPCHK(c, R.alloc((void **)&gep, sizeof(unsigned long long) * 2 * ADJ_W * (size_t)R.nk));
    HIPCHK(c, R.alloc((void **)&gsc, sizeof(unsigned long long) * 2 * ADJ_W * (size_t)(S.npair + 1)));
    AdjParams &A = R.A;
    A.w = d_w; A.gep = gep; A.gsc = gsc;
    A.f = F; A.mu = mu; A.dbox = dbox;
    // ---- forward: checkpoints every K steps, smoothed RDF of every state x_1..x_nsteps
    for (int s = 0; s < nsteps; s++)
        if ((st = R.forward_step(s))) return st;
    HIPCHK(c, hipMemcpyAsync(xS, c->pos[c->cur], v4, hipMemcpyDeviceToDevice, c->stream));
    std::vector<double> w;
    double zcoef;
    if ((st = R.loss(loss, rdf_out, w, zcoef))) return st;
    if (grad_ep || grad_sc) {
        // the reverse pass is linear in its sources (the bin weights and the flux
        // coefficient): they go in scaled by a power of two that brings the largest to
        // about 1, and the gradients come out scaled back, both exactly, so the fixed-point
        // sums keep their relative precision whatever the loss's size
        double big = std::fabs(zcoef);
        for (double v : w) big = std::max(big, std::fabs(v));
        const int sc = big > 0.0 ? -std::ilogb(big) : 0;
        for (double &v : w) v = std::ldexp(v, sc);
        HIPCHK(c, hipMemcpyAsync(d_w, w.data(), sizeof(double) * R.nh, hipMemcpyHostToDevice, c->stream));
        st = sync_err(c);       // w is the source of that copy
        if (st) return st;
        // lambda_nsteps = dl/dx at the last state (mu = 0: no step after it), with the flux
        // term's share
        A.x = xS; A.lam1 = lam[0]; A.lam0 = lam[1]; A.obs = 1; A.zflux = std::ldexp(zcoef, sc);
        HIPCHK(c, launch_adj(3, A, c->stream));
        A.zflux = 0.0;
        int lc = 1;
        for (int seg = R.nseg - 1; seg >= 0; seg--) {
            const int s0 = seg * K, s1 = std::min(s0 + K, nsteps);
            HIPCHK(c, hipMemcpyAsync(c->pos[c->cur], R.ckpt + (size_t)seg * nt, v4, hipMemcpyDeviceToDevice, c->stream));
            int sst = ensure_sorted(c, true);
            if (sst) return sst;
            for (int s = s0; s < s1; s++) {
                HIPCHK(c, hipMemcpyAsync(traj + (size_t)(s - s0) * nt, c->pos[c->cur], v4, hipMemcpyDeviceToDevice, c->stream));
                HIPCHK(c, launch_step(R.step_params(s), c->stream));
                c->cur ^= 1;
            }
            for (int s = s1 - 1; s >= s0; s--) {
                A.x = traj + (size_t)(s - s0) * nt;
                A.xi = c->noise + (size_t)s * c->noise_stride;
                A.lam1 = lam[lc];
                A.lam0 = lam[lc ^ 1];
                A.obs = s >= 1 ? 1 : 0;
                HIPCHK(c, launch_adj(1, A, c->stream));
                HIPCHK(c, launch_adj(2, A, c->stream));
                HIPCHK(c, launch_adj(3, A, c->stream));
                lc ^= 1;
            }
        }
        std::vector<unsigned long long> hg(2 * ADJ_W * (size_t)R.nk), hs(2 * ADJ_W * (size_t)(S.npair + 1));
        HIPCHK(c, hipMemcpyAsync(hg.data(), gep, sizeof(unsigned long long) * hg.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(hs.data(), gsc, sizeof(unsigned long long) * hs.size(), hipMemcpyDeviceToHost, c->stream));
        // the last state back in place
        HIPCHK(c, hipMemcpyAsync(c->pos[c->cur], xS, v4, hipMemcpyDeviceToDevice, c->stream));
        st = sync_err(c);
        if (st) return st;
        auto val = [&](const std::vector<unsigned long long> &h, size_t e) {
            return std::ldexp(fix_value(&h[ADJ_W * e], ADJ_W), -sc);
        };
        if (grad_ep)
            for (int k = 0; k < S.nknots; k++) {
                grad_ep[3 * k] = 0.0;
                grad_ep[3 * k + 1] = val(hg, 2 * (size_t)k);
                grad_ep[3 * k + 2] = val(hg, 2 * (size_t)k + 1);
            }
        if (grad_sc)
            for (int t = 0; t <= S.npair; t++) {
                grad_sc[3 * t] = val(hs, 2 * (size_t)t);
                grad_sc[3 * t + 1] = 0.0;
                grad_sc[3 * t + 2] = val(hs, 2 * (size_t)t + 1);
            }
    }
    c->sort_dirty = true;
    return sync_err(c);
}

int bromoc_tangent_rdf(bromoc_ctx *c, int32_t nsteps, int32_t K, const bromoc_rdf *rdf,
                       const double *v_ep, const double *v_sc, double *loss, double *dloss) {
    if (!c || !loss || !dloss || !v_ep || !v_sc) return BROMOC_E_ARG;
    (void)hipSetDevice(c->device);
    const bromoc_system &S = c->sys;
    AdjRun R;
    int st = R.setup(c, nsteps, K, rdf);
    if (st) return st;
    const int nt = R.nt;
    const size_t v4 = sizeof(double4) * (size_t)nt;
    double4 *F, *mu, *dbox, *ones, *dxs[2], *junk;
    double *vep, *vsc;
    unsigned long long *dh;
    int vscale = 0;
    HIPCHK(c, R.alloc((void **)&F, v4));
    HIPCHK(c, R.alloc((void **)&mu, v4));
    HIPCHK(c, R.alloc((void **)&dbox, v4));
    HIPCHK(c, R.alloc((void **)&ones, v4));
    HIPCHK(c, R.alloc((void **)&dxs[0], v4));
    HIPCHK(c, R.alloc((void **)&dxs[1], v4));
    HIPCHK(c, R.alloc((void **)&junk, v4));
    HIPCHK(c, R.alloc((void **)&vep, sizeof(double) * 2 * (size_t)R.nk));
    HIPCHK(c, R.alloc((void **)&vsc, sizeof(double) * 2 * (size_t)(S.npair + 1)));
    HIPCHK(c, R.alloc((void **)&dh, sizeof(unsigned long long) * ADJ_W * (size_t)R.nh));
    {
        std::vector<double4> h1(nt, make_double4(1.0, 1.0, 1.0, 0.0));
        // the tangent is linear in v: v goes in scaled by the power of two that brings its
        // largest entry to about 1 (exact), dloss comes out scaled back
        double big = 0.0;
        for (int k = 0; k < S.nknots; k++) big = std::max({big, std::fabs(v_ep[3 * k + 1]), std::fabs(v_ep[3 * k + 2])});
        for (int t = 0; t <= S.npair; t++) big = std::max({big, std::fabs(v_sc[3 * t]), std::fabs(v_sc[3 * t + 2])});
        vscale = big > 0.0 ? -std::ilogb(big) : 0;
        std::vector<double> he(2 * (size_t)R.nk, 0.0), hs(2 * (size_t)(S.npair + 1), 0.0);
        for (int k = 0; k < S.nknots; k++) {
            he[2 * k] = std::ldexp(v_ep[3 * k + 1], vscale);
            he[2 * k + 1] = std::ldexp(v_ep[3 * k + 2], vscale);
        }
        for (int t = 0; t <= S.npair; t++) {
            hs[2 * t] = std::ldexp(v_sc[3 * t], vscale);
            hs[2 * t + 1] = std::ldexp(v_sc[3 * t + 2], vscale);
        }
        HIPCHK(c, hipMemcpyAsync(ones, h1.data(), v4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(vep, he.data(), sizeof(double) * he.size(), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(vsc, hs.data(), sizeof(double) * hs.size(), hipMemcpyHostToDevice, c->stream));
        st = sync_err(c);     // the host staging vectors go out of scope here
        if (st) return st;
    }
    AdjParams &A = R.A;
    A.f = F; A.mu = mu; A.dbox = dbox;
    A.vep = vep; A.vsc = vsc; A.dh = dh;
    A.lam1 = ones;            // k_adj_mu then gives mu = f1 D_bd D_box
    int dc = 0;               // dxs[dc] holds dx of the current state (zero at x_0)
    for (int s = 0; s < nsteps; s++) {
        // the tangent of step s at x_s, before the step moves the positions
        if (s % K == 0) {      // the checkpoint's sort, as the forward pass makes it
            int sst = ensure_sorted(c, true);
            if (sst) return sst;
        }
        A.x = c->pos[c->cur];
        A.xi = c->noise + (size_t)s * c->noise_stride;
        A.dx = dxs[dc];
        A.dxo = dxs[dc ^ 1];
        A.obs = s >= 1 ? 1 : 0;
        HIPCHK(c, launch_adj(1, A, c->stream));
        HIPCHK(c, launch_adj(2, A, c->stream));
        HIPCHK(c, launch_adj(4, A, c->stream));
        dc ^= 1;
        if ((st = R.forward_step(s))) return st;
    }
    // the histogram derivative at the last state
    A.x = c->pos[c->cur];
    A.dx = dxs[dc];
    A.dxo = junk;
    A.obs = 1;
    A.xi = c->noise + (size_t)(nsteps - 1) * c->noise_stride;
    HIPCHK(c, launch_adj(1, A, c->stream));
    HIPCHK(c, launch_adj(2, A, c->stream));
    HIPCHK(c, launch_adj(4, A, c->stream));
    std::vector<double> w;
    double zcoef;
    if ((st = R.loss(loss, nullptr, w, zcoef))) return st;
    std::vector<unsigned long long> hdh(ADJ_W * (size_t)R.nh);
    std::vector<double4> hdx(nt);
    HIPCHK(c, hipMemcpyAsync(hdh.data(), dh, sizeof(unsigned long long) * hdh.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(hdx.data(), dxs[dc], v4, hipMemcpyDeviceToHost, c->stream));
    st = sync_err(c);
    if (st) return st;
    double d = 0.0;
    for (int k = 0; k < R.nh; k++) d += w[k] * fix_value(&hdh[ADJ_W * (size_t)k], ADJ_W);
    if (zcoef != 0.0)
        for (int i = R.nf0; i < nt; i++) d += zcoef * c->h_cg[std::abs(c->h_typei[i])] * hdx[i].z;
    *dloss = std::ldexp(d, -vscale);
    c->sort_dirty = true;
    return BROMOC_OK;
}

// raw Philox4x64-10 blocks on the device (the generator behind BROMOC_NOISE_PHILOX)
int bromoc_philox4x64(int32_t n, const uint64_t *ctr, const uint64_t *key, uint64_t *out) {
    if (n < 0 || (n > 0 && (!ctr || !key || !out))) return BROMOC_E_ARG;
    if (n == 0) return BROMOC_OK;
    uint64_t *d = nullptr;
    if (hipMalloc(&d, sizeof(uint64_t) * 10 * n) != hipSuccess) return BROMOC_E_HIP;
    (void)hipMemcpy(d, ctr, sizeof(uint64_t) * 4 * n, hipMemcpyHostToDevice);
    (void)hipMemcpy(d + 4 * n, key, sizeof(uint64_t) * 2 * n, hipMemcpyHostToDevice);
    hipError_t e = launch_philox_raw(d, d + 4 * n, n, d + 6 * n, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, d + 6 * n, sizeof(uint64_t) * 4 * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e == hipSuccess ? BROMOC_OK : BROMOC_E_HIP;
}

}  // extern "C"
