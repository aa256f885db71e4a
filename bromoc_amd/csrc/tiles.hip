// tiles.hip -- tile mode of the BD step: every unordered particle pair evaluated once.
//
// The reference's ENERGY visits each ion-ion pair once (energy.mono.f90:117-194) and
// applies Newton's third law; its reaction-field cross term does the same for each ion
// pair (rfpar.f90:177-223), and the ion-DNA loop each site-ion pair (:584-687).  k_step
// gathers instead, every target over all partners, which costs every pair twice.  Tile
// mode keeps the reference's once-per-pair work and stays deterministic:
//
//  * the particles in spatial order (spos) are cut into tiles of TT = 64 * TPL slots;
//  * k_tiles: one wavefront per tile pair (I, J), I <= J, two per workgroup on the same
//    column J, which share tile J's partner data in LDS.  Lane l holds the TPL targets
//    l, l + 64, .. of tile I in registers; tile J's partners sit in LDS.  At step k every
//    lane takes a different partner (a rotation, so no two lanes touch one partner at
//    once), evaluates it against its TPL targets, adds the targets' shares in registers
//    and the partner's share into the partner's LDS accumulator.  Each pair's reaction-
//    field cross term (when the system has one) is evaluated in the same sweep.  At the
//    end the wavefront writes tile I's sums to tpart[J][I] and tile J's to tpart[I][J]
//    (a diagonal tile pair writes their sum once): every entry of tpart is written exactly
//    once per step, and the pair energies go to tener[cat][item];
//  * the sweep a wavefront runs depends on what its two tiles hold (the sort puts CG-DNA
//    sites, fixed ions and moving ions in separate runs, so most tiles hold one class):
//    two tiles of moving ions off the diagonal need no per-pair masks, and their
//    reaction-field term, when every ion of both tiles is either inside the grid with a
//    nonzero srfe or outside it with none, reduces to per-ion weights folded into the
//    charges (or vanishes, when one tile has no ion inside); a tile of sites against a tile
//    of moving ions is all site-ion pairs with no reaction field; tiles of sites and fixed
//    ions only have no pair to evaluate.  Everything else takes the masked sweep;
//  * k_tfin: eight lanes per slot sum the slot's ntile partial rows in a fixed order,
//    gather the grid corners and finish the particle as k_step does (CG-DNA list forces,
//    reaction-field gradients and self term, MEMBRANE, energies, BD update).
#include <type_traits>

#include "device.h"

namespace bromoc {

constexpr int F_RFT = 64;   // reaction-field cross term fused into the pair sweep
// slot flags (bits above the atnam2 type, which takes the low 16): a particle (VALID), a
// site, a fixed ion, receiving pair forces (RECV), moving and inside the reaction-field grid
// (HI), a moving ion (ION); for the reaction field, a moving ion with HI and srfe != 0
// (RFON) and one with exactly one of the two (RFODD)
constexpr int SL_VALID = 1, SL_SITE = 2, SL_FIXED = 4, SL_RECV = 8, SL_HI = 16, SL_ION = 32,
              SL_RFON = 64, SL_RFODD = 128;

size_t tile_part_doubles(int cap, int ttpl, int rf) {
    const size_t tt = 64 * (size_t)ttpl, n = (cap + tt - 1) / tt;
    return n * n * (rf ? 5 : 3) * tt;
}
size_t tile_ener_doubles(int cap, int ttpl) {
    const size_t tt = 64 * (size_t)ttpl, n = (cap + tt - 1) / tt;
    return (size_t)(NTE + 1) * (n * (n + 1) / 2);   // energies, then the tflag words
}

// The particle in slot s of the spatial order: position, and a word of atnam2 type and
// flags (the pair kernels' masks); with the reaction field, srfe, reff and the charge.
struct SlotInfo {
    double x, y, z, sr, rr, q;
    int tf, idx;
};

template <int FEAT>
__device__ __forceinline__ SlotInfo slot_info(const StepParams &P, int s) {
    SlotInfo o;
    const bool ok = s < P.ntot;
    const double4 pa = P.spos[ok ? s : 0];
    const int idx = __double2hiint(pa.w), t2 = __double2loint(pa.w);
    const bool site = idx < P.nsites, fixed = !site && idx < P.nsites + P.nfix;
    bool recv = !site && !fixed;
    if constexpr ((FEAT & F_DNA) != 0) {
        if (site) recv = P.stfree[idx] != 0;
    }
    const bool ion = ok && !site && !fixed;
    int fl = (ok ? SL_VALID : 0) | (site ? SL_SITE : 0) | (fixed ? SL_FIXED : 0) |
             (recv ? SL_RECV : 0) | (ion ? SL_ION : 0);
    o.x = pa.x; o.y = pa.y; o.z = pa.z;
    o.sr = 0.0; o.rr = 0.0; o.q = 0.0;
    if constexpr ((FEAT & F_RFT) != 0) {
        if (!site) {
            // (srfe, reff, inside the grid box, moving) from k_rfpar; charge in rfq.w
            const double4 ra = P.rfa[idx];
            o.sr = ra.x; o.rr = ra.y;
            o.q = P.rfq[idx].w;
            const bool hi = ra.z != 0.0 && ra.w != 0.0;
            if (hi) fl |= SL_HI;
            if (ion && hi && ra.x != 0.0) fl |= SL_RFON;
            if (ion && hi != (ra.x != 0.0)) fl |= SL_RFODD;
        }
    }
    o.tf = t2 | (fl << 16);
    o.idx = idx;
    return o;
}

// the low 8 bits of v AND-ed (OR-ed) over the wavefront
__device__ __forceinline__ int wave_and8(int v) {
    int r = 0;
#pragma unroll
    for (int b = 0; b < 8; b++)
        if (__ballot(((v >> b) & 1) == 0) == 0ull) r |= 1 << b;
    return r;
}
__device__ __forceinline__ int wave_or8(int v) {
    int r = 0;
#pragma unroll
    for (int b = 0; b < 8; b++)
        if (__ballot((v >> b) & 1) != 0ull) r |= 1 << b;
    return r;
}

// Blocks of k_tiles: column J of the tile pairs (I, J), I <= J, in pairs of rows -- block
// c of column J holds the items (2c, J) and (2c + 1, J), one per wavefront, so the two
// wavefronts share tile J's partner data in LDS.  Column j has j / 2 + 1 blocks; the
// columns before J have tile_blocks(J) = m (m + 1) for J = 2m and (m + 1)^2 for J = 2m + 1.
__host__ __device__ __forceinline__ int tile_blocks(int J) {
    const int m = J >> 1;
    return (J & 1) ? (m + 1) * (m + 1) : m * (m + 1);
}
__device__ __forceinline__ void block_tiles(int b, int &J, int &c) {
    int j = (int)(2.0 * sqrt((double)b));
    while (j > 0 && tile_blocks(j) > b) j--;
    while (tile_blocks(j + 1) <= b) j++;
    J = j;
    c = b - tile_blocks(j);
}
// LDS row stride of the type-pair tables: even, so that the rows (t, p) and (p, t), which a
// wavefront reads at once, fall in different banks
__host__ __device__ __forceinline__ int tail_stride(int nt1) { return nt1 + (nt1 & 1); }

template <int FEAT, int TPL>
__global__ __launch_bounds__(TILE_WAVES * WAVE) void k_tiles(StepParams P) {
    constexpr int TT = WAVE * TPL;
    constexpr bool RF = (FEAT & F_RFT) != 0;
    constexpr int NR = RF ? TT : 1;
    // tile J's partner data, shared by the block's wavefronts
    __shared__ double s_x[TT], s_y[TT], s_z[TT];
    __shared__ int s_tf[TT], s_ix[RF ? TT : 1];
    __shared__ double s_sr[NR], s_rr[NR], s_q[NR];
    // each wavefront's partner sums
    __shared__ double s_afx[TILE_WAVES][TT], s_afy[TILE_WAVES][TT], s_afz[TILE_WAVES][TT];
    __shared__ double s_afa[TILE_WAVES][NR], s_afb[TILE_WAVES][NR];
    // by type pair (row stride tail_stride): (sc3, tail fct, dm2hi, flags); the near branch
    // reads the rest of the pair row from global memory
    extern __shared__ double4 s_tail[];

    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid >> 6;
    const int b = blockIdx.x;
    const int n = P.ntile, nt1 = P.nttyp + 1, S = tail_stride(nt1);
    if (b >= tile_blocks(n)) return;
    int J, cp;
    block_tiles(b, J, cp);
    const int I = TILE_WAVES * cp + wv;
    const bool active = I <= J;
    const bool diag = I == J;
    const int item = J * (J + 1) / 2 + I;   // the tile pair's index: (I, J) by columns
    for (int k = tid; k < nt1 * nt1; k += TILE_WAVES * WAVE) {
        const DevPair &d = P.rows[k];
        const int l = (k / nt1) * S + k % nt1;
        s_tail[l] = make_double4(d.sc3, d.fct, d.dm2hi, (double)d.flags);
    }
    double *const s_fx = s_afx[wv], *const s_fy = s_afy[wv], *const s_fz = s_afz[wv];
    double *const s_fa = s_afa[wv], *const s_fb = s_afb[wv];
    // partners: tile J into LDS; this wavefront's sums cleared
    for (int p = tid; p < TT; p += TILE_WAVES * WAVE) {
        const SlotInfo o = slot_info<FEAT>(P, J * TT + p);
        s_x[p] = o.x; s_y[p] = o.y; s_z[p] = o.z; s_tf[p] = o.tf;
        if constexpr (RF) { s_ix[p] = o.idx; s_sr[p] = o.sr; s_rr[p] = o.rr; s_q[p] = o.q; }
    }
#pragma unroll
    for (int u = 0; u < TPL; u++) {
        const int p = lane + WAVE * u;
        s_fx[p] = 0.0; s_fy[p] = 0.0; s_fz[p] = 0.0;
        if constexpr (RF) { s_fa[p] = 0.0; s_fb[p] = 0.0; }
    }
    // targets: tile I in registers
    double tx[TPL], ty[TPL], tz[TPL], tsr[TPL], trr[TPL], tq[TPL];
    int ttf[TPL], tix[TPL];
    double ax_[TPL], ay_[TPL], az_[TPL], aa_[TPL], ab_[TPL];
#pragma unroll
    for (int u = 0; u < TPL; u++) {
        const SlotInfo o = slot_info<FEAT>(P, (active ? I : J) * TT + lane + WAVE * u);
        tx[u] = o.x; ty[u] = o.y; tz[u] = o.z; ttf[u] = o.tf; tix[u] = o.idx;
        tsr[u] = o.sr; trr[u] = o.rr; tq[u] = o.q;
        ax_[u] = ay_[u] = az_[u] = aa_[u] = ab_[u] = 0.0;
    }
    __syncthreads();
    // the block's last barrier: a wavefront past the column's diagonal has no tile pair
    if (!active) return;

    double en[NTE];
#pragma unroll
    for (int c = 0; c < NTE; c++) en[c] = 0.0;

    auto sweep = [&](auto en_on) {
        constexpr bool EN = decltype(en_on)::value;
        for (int k = 0; k < TT; k++) {
            // rotation: at step k the lanes hold 64 different partners, and over the TT
            // steps every lane meets every partner once
            const int p = ((lane + k / TPL) & (WAVE - 1)) + WAVE * (k % TPL);
            const double px = s_x[p], py = s_y[p], pz = s_z[p];
            const int ptf = s_tf[p];
            const int t2p = ptf & 0xFFFF, pfl = ptf >> 16;
            double psr = 0.0, prr = 0.0, pq = 0.0;
            int pix = 0;
            if constexpr (RF) { psr = s_sr[p]; prr = s_rr[p]; pq = s_q[p]; pix = s_ix[p]; }
            double gx = 0.0, gy = 0.0, gz = 0.0;    // pair force on the partner (negated)
            double rx = 0.0, ry = 0.0, rz = 0.0, ga = 0.0, gb = 0.0;   // reaction field
#pragma unroll
            for (int u = 0; u < TPL; u++) {
                const int tfl = ttf[u] >> 16, t2t = ttf[u] & 0xFFFF;
                const int both = tfl & pfl;
                bool skip = !(both & SL_VALID) || (both & (SL_SITE | SL_FIXED)) ||
                            (diag && p <= lane + WAVE * u);
                const double dx = tx[u] - px, dy = ty[u] - py, dz = tz[u] - pz;
                double d2 = dx * dx + dy * dy + dz * dz;
                d2 = skip ? 1e30 : d2;
                const int row = t2t * nt1 + t2p, lrow = t2t * S + t2p;
                const double4 tl = s_tail[lrow];
                const bool mx = ((tfl | pfl) & SL_SITE) != 0;
                double y, de;
                if constexpr ((FEAT & F_TAB) != 0) {
                    de = pair_near_far<false>(P.rows[row], tl, P.ep, P.ires, d2, y);
                    if constexpr (EN) {
                        if (!skip) { if (mx) en[TE_EFPOTMX] += y; else en[TE_EFPOT] += y; }
                    }
                } else if ((int)tl.w & PF_EFPOT) {
                    de = pair_near_far<false>(P.rows[row], tl, P.ep, P.ires, d2, y);
                    if constexpr (EN) {
                        if (!skip) { if (mx) en[TE_EFPOTMX] += y; else en[TE_EFPOT] += y; }
                    }
                } else {
                    double q0 = 0.0, q1 = 0.0, q2 = 0.0;
                    de = pair_analytic(P.rows[row], P, d2, q0, q1, q2);
                    if constexpr (EN) {
                        if (!skip) {
                            en[mx ? TE_QQMX : TE_ELEC] += q0;
                            en[mx ? TE_VDWMX : TE_VDW] += q1;
                            en[mx ? TE_SRPMFMX : TE_SRPMF] += q2;
                        }
                    }
                }
                de = skip ? 0.0 : de;
                const double fxp = de * dx, fyp = de * dy, fzp = de * dz;
                if (tfl & SL_RECV) { ax_[u] += fxp; ay_[u] += fyp; az_[u] += fzp; }
                gx += fxp; gy += fyp; gz += fzp;
                if constexpr (RF) {
                    // rfpar.f90:177-223 for the ion pair (lo, hi) by index: counted when
                    // srfe(lo) is nonzero and hi is moving and inside the grid box
                    const bool thi = tix[u] > pix;
                    const bool on = !skip && !mx &&
                                    (thi ? (psr != 0.0 && (tfl & SL_HI)) : (tsr[u] != 0.0 && (pfl & SL_HI)));
                    const double m = on ? 1.0 : 0.0;
                    const double aux0 = P.celec * tq[u] * pq;
                    const double srfeij = tsr[u] * psr, reffij = trr[u] * prr;
                    const double rfdn = rsqrt_nr(reffij * reffij + (on ? d2 : 1.0));
                    const double rfcf = reffij * rfdn;
                    const double aux1 = aux0 * rfcf;
                    const double aux2 = aux0 * srfeij * (rfdn * (rfdn * rfdn));
                    const double aux3 = -aux2 * ((on ? d2 : 1.0) + reffij * 0.5);
                    const double der = m * (aux2 * reffij);
                    if constexpr (EN) en[TE_RFPAR] = fma(m, aux1 * srfeij, en[TE_RFPAR]);
                    aa_[u] = fma(m, -aux1 * psr, aa_[u]);
                    ab_[u] = fma(m, aux3 * prr, ab_[u]);
                    ga = fma(m, -aux1 * tsr[u], ga);
                    gb = fma(m, aux3 * trr[u], gb);
                    ax_[u] = fma(der, dx, ax_[u]);
                    ay_[u] = fma(der, dy, ay_[u]);
                    az_[u] = fma(der, dz, az_[u]);
                    rx = fma(der, dx, rx);
                    ry = fma(der, dy, ry);
                    rz = fma(der, dz, rz);
                }
            }
            const double pw = (pfl & SL_RECV) ? 1.0 : 0.0;
            s_fx[p] -= pw * gx + rx;
            s_fy[p] -= pw * gy + ry;
            s_fz[p] -= pw * gz + rz;
            if constexpr (RF) {
                s_fa[p] += ga;
                s_fb[p] += gb;
            }
            // LDS operations of one wavefront complete in issue order; this keeps the
            // compiler from moving the next step's partner reads above these updates
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
        }
    };
    // Mask-free sweep over tile pairs of one class each (see the file comment).  RFW: the
    // reaction-field term with the ion weights in the charges (a pair counts when both
    // ions are RFON; with no RFODD ion that is the reference's rule, rfpar.f90:177-223, in
    // either index order); MX: sites against moving ions, all pairs mixed, forces kept
    // only for receiving particles.  The radial reaction-field force joins the pair force.
    auto sweep_fast = [&](auto en_on, auto rf_on, auto mx_on) {
        constexpr bool EN = decltype(en_on)::value;
        constexpr bool RFW = decltype(rf_on)::value;
        constexpr bool MX = decltype(mx_on)::value;
        double tcq[TPL];
#pragma unroll
        for (int u = 0; u < TPL; u++)
            tcq[u] = ((ttf[u] >> 16) & SL_RFON) ? P.celec * tq[u] : 0.0;
        for (int k = 0; k < TT; k++) {
            const int p = ((lane + k / TPL) & (WAVE - 1)) + WAVE * (k % TPL);
            const double px = s_x[p], py = s_y[p], pz = s_z[p];
            const int ptf = s_tf[p];
            const int t2p = ptf & 0xFFFF, pfl = ptf >> 16;
            double psr = 0.0, prr = 0.0, pq = 0.0;
            if constexpr (RFW) {
                psr = s_sr[p];
                prr = s_rr[p];
                pq = (pfl & SL_RFON) ? s_q[p] : 0.0;
            }
            double gx = 0.0, gy = 0.0, gz = 0.0, ga = 0.0, gb = 0.0;
#pragma unroll
            for (int u = 0; u < TPL; u++) {
                const int t2t = ttf[u] & 0xFFFF;
                const double dx = tx[u] - px, dy = ty[u] - py, dz = tz[u] - pz;
                const double d2 = dx * dx + dy * dy + dz * dz;
                const int row = t2t * nt1 + t2p, lrow = t2t * S + t2p;
                const double4 tl = s_tail[lrow];
                double y, de;
                if constexpr ((FEAT & F_TAB) != 0) {
                    de = pair_near_far<true>(P.rows[row], tl, P.ep, P.ires, d2, y);
                    if constexpr (EN) en[MX ? TE_EFPOTMX : TE_EFPOT] += y;
                } else if ((int)tl.w & PF_EFPOT) {
                    de = pair_near_far<true>(P.rows[row], tl, P.ep, P.ires, d2, y);
                    if constexpr (EN) en[MX ? TE_EFPOTMX : TE_EFPOT] += y;
                } else {
                    double q0 = 0.0, q1 = 0.0, q2 = 0.0;
                    de = pair_analytic(P.rows[row], P, d2, q0, q1, q2);
                    if constexpr (EN) {
                        en[MX ? TE_QQMX : TE_ELEC] += q0;
                        en[MX ? TE_VDWMX : TE_VDW] += q1;
                        en[MX ? TE_SRPMFMX : TE_SRPMF] += q2;
                    }
                }
                if constexpr (RFW) {
                    const double srfeij = tsr[u] * psr, reffij = trr[u] * prr;
                    const double rfdn = rsqrt_c3(fma(reffij, reffij, d2));
                    const double aux0 = tcq[u] * pq;
                    const double aux1 = aux0 * (reffij * rfdn);
                    const double aux2 = aux0 * srfeij * (rfdn * (rfdn * rfdn));
                    const double aux3 = -aux2 * (d2 + reffij * 0.5);
                    if constexpr (EN) en[TE_RFPAR] = fma(aux1, srfeij, en[TE_RFPAR]);
                    aa_[u] = fma(-aux1, psr, aa_[u]);
                    ab_[u] = fma(aux3, prr, ab_[u]);
                    ga = fma(-aux1, tsr[u], ga);
                    gb = fma(aux3, trr[u], gb);
                    de = fma(aux2, reffij, de);
                }
                ax_[u] = fma(de, dx, ax_[u]);
                ay_[u] = fma(de, dy, ay_[u]);
                az_[u] = fma(de, dz, az_[u]);
                gx = fma(de, dx, gx);
                gy = fma(de, dy, gy);
                gz = fma(de, dz, gz);
            }
            if constexpr (MX) {
                const double pw = (pfl & SL_RECV) ? 1.0 : 0.0;
                s_fx[p] -= pw * gx;
                s_fy[p] -= pw * gy;
                s_fz[p] -= pw * gz;
            } else {
                s_fx[p] -= gx;
                s_fy[p] -= gy;
                s_fz[p] -= gz;
            }
            if constexpr (RFW) {
                s_fa[p] += ga;
                s_fb[p] += gb;
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
        }
        if constexpr (MX) {
#pragma unroll
            for (int u = 0; u < TPL; u++)
                if (!((ttf[u] >> 16) & SL_RECV)) ax_[u] = ay_[u] = az_[u] = 0.0;
        }
    };

    // the classes of the two tiles, uniform over the wavefront
    int andI = ~0, orI = 0, andJ = ~0, orJ = 0;
#pragma unroll
    for (int u = 0; u < TPL; u++) {
        const int fi = ttf[u] >> 16, fj = s_tf[lane + WAVE * u] >> 16;
        andI &= fi; orI |= fi; andJ &= fj; orJ |= fj;
    }
    andI = wave_and8(andI); orI = wave_or8(orI);
    andJ = wave_and8(andJ); orJ = wave_or8(orJ);
    const bool ionI = andI & SL_ION, ionJ = andJ & SL_ION;
    const bool siteI = (andI & (SL_VALID | SL_SITE)) == (SL_VALID | SL_SITE);
    const bool siteJ = (andJ & (SL_VALID | SL_SITE)) == (SL_VALID | SL_SITE);
    const bool wr = P.write_energy != 0;
    bool rfsums = false;   // this tile pair's reaction-field sums (A, B) may be nonzero
    if (!(orI & SL_ION) && !(orJ & SL_ION)) {
        // sites, fixed ions and empty slots only: no pair of these is evaluated
    } else if (!diag && ionI && ionJ && !(RF && ((orI | orJ) & SL_RFODD))) {
        if (RF && (orI & SL_RFON) && (orJ & SL_RFON)) {
            rfsums = true;
            if (wr) sweep_fast(std::true_type{}, std::bool_constant<RF>{}, std::false_type{});
            else sweep_fast(std::false_type{}, std::bool_constant<RF>{}, std::false_type{});
        } else {
            if (wr) sweep_fast(std::true_type{}, std::false_type{}, std::false_type{});
            else sweep_fast(std::false_type{}, std::false_type{}, std::false_type{});
        }
    } else if (!diag && ((ionI && siteJ) || (siteI && ionJ))) {
        if (wr) sweep_fast(std::true_type{}, std::false_type{}, std::true_type{});
        else sweep_fast(std::false_type{}, std::false_type{}, std::true_type{});
    } else {
        rfsums = RF;
        if (wr) sweep(std::true_type{});
        else sweep(std::false_type{});
    }
    // the wavefront reads its own sums back: its LDS operations complete in order, and the
    // fence keeps the compiler from moving the reads above the sweep's last updates
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_wave_barrier();

    // a tile pair without reaction-field sums writes the three force components only and
    // says so in tflag[item], which k_tfin reads before it reads the A and B rows
    const int NC = RF ? 5 : 3, nc = rfsums ? NC : 3;
    double *const part = P.tpart;
    if constexpr (RF) {
        if (lane == 0) P.tflag[item] = rfsums ? 1 : 0;
    }
#pragma unroll
    for (int u = 0; u < TPL; u++) {
        const int t = lane + WAVE * u;
        double v[5] = {ax_[u], ay_[u], az_[u], aa_[u], ab_[u]};
        if (diag) {
            v[0] += s_fx[t]; v[1] += s_fy[t]; v[2] += s_fz[t];
            if constexpr (RF) { v[3] += s_fa[t]; v[4] += s_fb[t]; }
        }
        double *dst = part + ((size_t)(J * n + I) * NC) * TT + t;
#pragma unroll
        for (int c = 0; c < NC; c++)
            if (c < nc) dst[(size_t)c * TT] = v[c];
        if (!diag) {
            const double w[5] = {s_fx[t], s_fy[t], s_fz[t], RF ? s_fa[t] : 0.0, RF ? s_fb[t] : 0.0};
            double *dp = part + ((size_t)(I * n + J) * NC) * TT + t;
#pragma unroll
            for (int c = 0; c < NC; c++)
                if (c < nc) dp[(size_t)c * TT] = w[c];
        }
    }
    if (P.write_energy) {
#pragma unroll
        for (int c = 0; c < NTE; c++) {
            const double v = wave_sum(en[c]);
            if (lane == 0) P.tener[(size_t)c * P.titems + item] = v;
        }
    }
}

// sum over the 8 lanes of a group (lanes 8g..8g+7); every lane gets the same bits
__device__ __forceinline__ double group8_sum(double v) {
    v += __shfl_xor(v, 1, WAVE);
    v += __shfl_xor(v, 2, WAVE);
    v += __shfl_xor(v, 4, WAVE);
    return v;
}

// k_tfin: slot s of the spatial order, eight lanes per slot (32 slots per block).  The
// lanes split the slot's ntile partial rows, then its grid corners (lane c takes corner c
// of the trilinear stencils, c, c + 8, .. of the 27-node B-spline); lane 0 of the group
// finishes the particle.  Sites and moving ions are finished as by k_step; a fixed ion
// only receives its reaction-field force (rfpar.f90:177-223 applies it to both ions of a
// pair), written when forces are asked for.
template <int FEAT>
__global__ __launch_bounds__(BLOCK) void k_tfin(StepParams P) {
    constexpr bool RF = (FEAT & F_RFT) != 0;
    const int tid = threadIdx.x, sub = tid & 7;
    const int s = blockIdx.x * (BLOCK / 8) + (tid >> 3);
    const bool sv = s < P.ntot;
    const int ss = sv ? s : 0;
    const double4 pa = P.spos[ss];
    const int i = __double2hiint(pa.w), t2 = __double2loint(pa.w);
    const int ns = P.nsites;
    const bool is_site = i < ns, fixed = !is_site && i < ns + P.nfix;
    const int TT = WAVE * P.ttpl, n = P.ntile, NC = RF ? 5 : 3;
    const int I = ss / TT, t = ss - I * TT;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // rows K = sub, sub + 8, .. added in that order, KU of them per iteration so that their
    // loads (the flags, then the rows) are in flight together
    constexpr int KU = 8;
    for (int K0 = sub; K0 < n; K0 += 8 * KU) {
        bool ok[KU], ab[KU];
        const double *src[KU];
#pragma unroll
        for (int u = 0; u < KU; u++) {
            const int K = K0 + 8 * u;
            ok[u] = K < n;
            const int Kc = ok[u] ? K : I;
            src[u] = P.tpart + ((size_t)(Kc * n + I) * NC) * TT + t;
            ab[u] = false;
            // the A and B rows only where the tile pair (min, max) wrote them
            if constexpr (RF) {
                const int lo = min(I, Kc), hi = max(I, Kc);
                const int32_t fl = P.tflag[hi * (hi + 1) / 2 + lo];   // a valid item even when !ok
                ab[u] = ok[u] & (fl != 0);
            }
        }
        double r[KU][5];
#pragma unroll
        for (int u = 0; u < KU; u++) {
#pragma unroll
            for (int c = 0; c < 3; c++) r[u][c] = src[u][(size_t)c * TT];
            r[u][3] = r[u][4] = 0.0;
            if constexpr (RF) {
                if (ab[u]) { r[u][3] = src[u][3 * (size_t)TT]; r[u][4] = src[u][4 * (size_t)TT]; }
            }
        }
#pragma unroll
        for (int u = 0; u < KU; u++) {
            if (!ok[u]) continue;
#pragma unroll
            for (int c = 0; c < NC; c++)
                if (c < 3 || ab[u]) acc[c] += r[u][c];
        }
    }
#pragma unroll
    for (int c = 0; c < NC; c++) acc[c] = group8_sum(acc[c]);

    // grids of a target (a site or a moving ion): staticf1.f90:45-118, vdwgd1trln.f90:
    // 50-161 / vdwgd1spln.f90:50-170, corners split over the group's lanes
    double gfx = 0.0, gfy = 0.0, gfz = 0.0, egx = 0.0, egv = 0.0, ph = 0.0;
    bool vin = false;
    const bool target = sv && !fixed;
    if constexpr ((FEAT & F_PHIX) != 0) {
        const GridGeom &g = P.gx;
        double chi;
        bool ok;
        if (is_site) { ok = P.namsite[i] == 1; chi = P.cgnuc; }
        else { ok = true; chi = P.cg[abs(P.typei[i])]; }
        ok = ok && target && pa.x <= g.xbcen + g.tranx && pa.x >= g.xbcen - g.tranx &&
             pa.y <= g.ybcen + g.trany && pa.y >= g.ybcen - g.trany &&
             pa.z <= g.zbcen + g.tranz && pa.z >= g.zbcen - g.tranz;
        if (ok) {
            double xi, yi, zi;
            int ix, iy, iz;
            grid_cell(g, pa.x, pa.y, pa.z, xi, yi, zi, ix, iy, iz);
            double e = 0.0;
            tri_corner(g, P.phix, xi, yi, zi, ix, iy, iz, sub, P.celec * chi, gfx, gfy, gfz, e);
            egx += e;
        }
    }
    if constexpr ((FEAT & F_PHIV) != 0) {
        const GridGeom &g = P.gv;
        const int8_t *phiv = P.phiv + vdw_map_of(P, i);
        double esvdw = P.svdw;
        if (P.flags & QSVDW) esvdw = P.svdw * P.scal[t2];
        if (P.flags & QBSPLINE) {
            vin = target && pa.x <= g.xbcen + g.tranx - g.dcel && pa.x >= g.xbcen - g.tranx + g.dcel &&
                  pa.y <= g.ybcen + g.trany - g.dcel && pa.y >= g.ybcen - g.trany + g.dcel &&
                  pa.z <= P.vzmax && pa.z >= P.vzmin;
            if (vin) {
                double xi, yi, zi;
                int ix, iy, iz;
                spline_cell(g, pa.x, pa.y, pa.z, xi, yi, zi, ix, iy, iz);
                for (int c = sub; c < 27; c += 8) {
                    const int k = ix - 1 + c / 9, l = iy - 1 + (c / 3) % 3, m = iz - 1 + c % 3;
                    double ai, dai, bi, dbi, ci, dci;
                    bspline_m3(1.5 - (xi - k * g.dcel) * g.idcel, ai, dai);
                    bspline_m3(1.5 - (yi - l * g.dcel) * g.idcel, bi, dbi);
                    bspline_m3(1.5 - (zi - m * g.dcel) * g.idcel, ci, dci);
                    if ((ai != 0.0 || dai != 0.0) && (bi != 0.0 || dbi != 0.0) &&
                        (ci != 0.0 || dci != 0.0)) {
                        const double v = (double)phiv[((size_t)k * g.ny + l) * g.nz + m];
                        ph += v;
                        const double phivsv = v * esvdw;
                        const double prefac = phivsv * g.idcel;
                        gfx += dai * bi * ci * prefac;
                        gfy += ai * dbi * ci * prefac;
                        gfz += ai * bi * dci * prefac;
                        egv += ai * bi * ci * phivsv;
                    }
                }
            }
        } else {
            vin = target && pa.x <= g.xbcen + g.tranx && pa.x >= g.xbcen - g.tranx &&
                  pa.y <= g.ybcen + g.trany && pa.y >= g.ybcen - g.trany && pa.z <= P.vzmax &&
                  pa.z >= P.vzmin;
            if (vin) {
                double xi, yi, zi;
                int ix, iy, iz;
                grid_cell(g, pa.x, pa.y, pa.z, xi, yi, zi, ix, iy, iz);
                ph = tri_corner(g, phiv, xi, yi, zi, ix, iy, iz, sub, esvdw, gfx, gfy, gfz, egv);
            }
        }
    }
    gfx = group8_sum(gfx);
    gfy = group8_sum(gfy);
    gfz = group8_sum(gfz);
    if (P.write_energy) {
        egx = group8_sum(egx);
        egv = group8_sum(egv);
    }
    ph = group8_sum(ph);   // map values are integers: exact in any order
    if (sub != 0 || !sv) return;

    if (fixed) {
        if constexpr (RF) {
            if (P.write_forces) {
                double f[3] = {0.0, 0.0, 0.0};
                if (P.rfa[i].z != 0.0) {
                    const double4 dst = P.rfd[2 * (size_t)i], drt = P.rfd[2 * (size_t)i + 1];
                    f[0] = acc[0] + acc[3] * dst.x + acc[4] * drt.x;
                    f[1] = acc[1] + acc[3] * dst.y + acc[4] * drt.y;
                    f[2] = acc[2] + acc[3] * dst.z + acc[4] * drt.z;
                }
                P.fx[i] = f[0]; P.fy[i] = f[1]; P.fz[i] = f[2];
            }
        }
        return;
    }
    double F[3] = {acc[0] + gfx, acc[1] + gfy, acc[2] + gfz};
    double erf_self = 0.0;
    if constexpr (RF) {
        if (!is_site) {
            // the self term of a moving ion (rfpar.f90:146-156) and the variation of its
            // srfe and reff in every pair (A, B summed over the tile pairs)
            const double4 ra = P.rfa[i];
            if (ra.z != 0.0) {
                const double4 dst = P.rfd[2 * (size_t)i], drt = P.rfd[2 * (size_t)i + 1];
                const double cgt = P.cg[abs(P.typei[i])];
                const double aux = P.celec * cgt * cgt * ra.x;
                const double a_s = acc[3] - aux;
                F[0] += a_s * dst.x + acc[4] * drt.x;
                F[1] += a_s * dst.y + acc[4] * drt.y;
                F[2] += a_s * dst.z + acc[4] * drt.z;
                erf_self = 0.5 * aux * ra.x;
            }
        }
    }
    if constexpr ((FEAT & F_DNA) != 0) {
        if (is_site) {
            const double4 fd = P.fdna[i];   // CG-DNA list terms from k_dna
            F[0] += fd.x; F[1] += fd.y; F[2] += fd.z;
        }
    }
    if (P.mmon && !is_site) {   // MMIJ reaction field (k_mm_force)
        const double4 fr = P.frf[i];
        F[0] += fr.x; F[1] += fr.y; F[2] += fr.z;
    }
    const bool warned = vin && ph >= ((P.flags & QBSPLINE) ? P.thold27 : P.thold8);
    if (warned) atomicAdd(&P.warn[t2], 1);
    if (P.flags & QMEMB) {   // MEMBRANE, membrane.f90:36-104
        const double chg = is_site ? (P.namsite[i] == 1 ? P.cgnuc : 0.0) : P.cg[abs(P.typei[i])];
        bool mw;
        const double em = membrane_energy(P, pa.x, pa.y, pa.z, chg, t2, F, mw);
        if (mw) atomicAdd(&P.warn[t2], 1);
        if (P.write_energy) P.epart[(size_t)E_MEMB * P.ecap + i] = em;
    }
    if (P.write_energy) {
        // the pair categories are summed per tile pair (tener); the grid and reaction-field
        // self terms per particle
        double *ep_ = P.epart + i;
        const size_t cs = P.ecap;
        ep_[E_GSBPA * cs] = egx;
        ep_[E_VDWGD * cs] = egv;
        if (is_site) {
            ep_[E_EFPOTMX * cs] = 0.0; ep_[E_QQMX * cs] = 0.0;
            ep_[E_VDWMX * cs] = 0.0; ep_[E_SRPMFMX * cs] = 0.0;
        } else {
            ep_[E_EFPOT * cs] = 0.0; ep_[E_ELEC * cs] = 0.0;
            ep_[E_VDW * cs] = 0.0; ep_[E_SRPMF * cs] = 0.0;
            if constexpr (RF) ep_[E_RFPAR * cs] = erf_self;
        }
    }
    if (P.write_forces) {
        P.fx[i] = F[0];
        P.fy[i] = F[1];
        P.fz[i] = F[2];
    }
    if constexpr ((FEAT & F_UPDATE) != 0) {
        if (P.vacc) P.vacc[i] += F[0] * pa.x + F[1] * pa.y + F[2] * pa.z;
        const bool fr = !is_site || ((FEAT & F_DNA) && P.stfree[i] != 0);
        if (!fr) {
            P.poso[i] = pa;
            P.sposo[s] = pa;
            return;
        }
        const double4 pn = bd_update(P, i, is_site, pa, F, noise_row(P, i));
        P.poso[i] = pn;
        P.sposo[s] = pn;
    }
}

template <int FEAT>
static void dispatch_tiles(const StepParams &p, hipStream_t s, int which) {
    const int nt1 = p.nttyp + 1;
    const size_t lds = sizeof(double4) * nt1 * tail_stride(nt1);
    const int nblk = tile_blocks(p.ntile);
    if (which & 1) switch (p.ttpl) {
        case 1: hipLaunchKernelGGL((k_tiles<FEAT & (F_DNA | F_TAB | F_RFT), 1>), dim3(nblk), dim3(TILE_WAVES * WAVE), lds, s, p); break;
        case 2: hipLaunchKernelGGL((k_tiles<FEAT & (F_DNA | F_TAB | F_RFT), 2>), dim3(nblk), dim3(TILE_WAVES * WAVE), lds, s, p); break;
        default: hipLaunchKernelGGL((k_tiles<FEAT & (F_DNA | F_TAB | F_RFT), 4>), dim3(nblk), dim3(TILE_WAVES * WAVE), lds, s, p); break;
    }
    if (which & 2)
        hipLaunchKernelGGL((k_tfin<FEAT & ~F_TAB>), dim3((p.ntot + BLOCK / 8 - 1) / (BLOCK / 8)),
                           dim3(BLOCK), 0, s, p);
}

hipError_t launch_tiles(const StepParams &p, hipStream_t s, int which) {
    if (p.ntot <= 0) return hipSuccess;
    int feat = 0;
    if (p.flags & QNUCL) feat |= F_DNA;
    if (p.flags & QPHIX) feat |= F_PHIX;
    if (p.flags & QPHIV) feat |= F_PHIV;
    if (p.mode == 1) feat |= F_UPDATE;
    if (p.all_efpot) feat |= F_TAB;
    if (p.rfon) feat |= F_RFT;
#define CASE(F) case F: dispatch_tiles<F>(p, s, which); break;
    switch (feat) {
        CASE(0) CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
        CASE(8) CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15)
        CASE(16) CASE(17) CASE(18) CASE(19) CASE(20) CASE(21) CASE(22) CASE(23)
        CASE(24) CASE(25) CASE(26) CASE(27) CASE(28) CASE(29) CASE(30) CASE(31)
        CASE(64) CASE(65) CASE(66) CASE(67) CASE(68) CASE(69) CASE(70) CASE(71)
        CASE(72) CASE(73) CASE(74) CASE(75) CASE(76) CASE(77) CASE(78) CASE(79)
        CASE(80) CASE(81) CASE(82) CASE(83) CASE(84) CASE(85) CASE(86) CASE(87)
        CASE(88) CASE(89) CASE(90) CASE(91) CASE(92) CASE(93) CASE(94) CASE(95)
    }
#undef CASE
    return hipGetLastError();
}

}  // namespace bromoc
