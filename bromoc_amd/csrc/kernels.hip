// kernels.hip -- gfx950 kernels of the BROMOC force / Brownian-dynamics step of one system.
//
// One BD step is k_rfpar/k_rfpair (reaction-field parameters, when present), the CG-DNA
// list terms of each site (energy.mono.f90:196-508; in the one-target k_step's own site
// blocks when the launch writes no energies, else k_dna_multi + k_dna) and k_step (or
// k_wave): every target particle -- each free DNA site and each moving ion -- gathers its
// pair forces over all other particles (the reference's ENERGY has no cutoff:
// energy.mono.f90:117-194,584-687), its 8 trilinear corners of the static-field and
// repulsion grids (staticf1.f90:45-118, vdwgd1trln.f90:50-161), then applies the
// Ermak-McCammon update (DYNAMICS0/1/2, DYNAMICS0NUC) into the other position buffer.
// Every force is a gather over partners in a fixed order: results are deterministic and
// need no atomics; energies are per-particle partials reduced in a fixed tree.
#include <type_traits>

#include "device.h"

namespace bromoc {

// The nearest-site record dids(1:5,i) of ion i (energy.mono.f90:583,603-620): the block's
// per-wavefront minima of (dist - dmi, site) combined; the reference starts the record
// only at site 1 (j.eq.1) and later sites replace it only when strictly smaller, so when
// site 1 is outside its table range the record stays zero unless some site gives a
// negative dist - dmi.
template <int T, int NW>
__device__ inline void prox_record(const StepParams &P, int i, const double4 &pa, int a,
                                   const double (*s_pxd)[T], const int (*s_pxj)[T],
                                   const DevPair *row, double dids[5]) {
    double d = INFINITY;
    int jj = 0x7fffffff;
    for (int q = 0; q < NW; q++) {
        const double od = s_pxd[q][a];
        const int oj = s_pxj[q][a];
        if (od < d || (od == d && oj < jj)) { d = od; jj = oj; }
    }
    if (jj == 0x7fffffff) return;
    const double4 p0 = P.pos[0];
    const double e0x = pa.x - p0.x, e0y = pa.y - p0.y, e0z = pa.z - p0.z;
    const double d20 = e0x * e0x + e0y * e0y + e0z * e0z;
    const DevPair &r0 = row[__double2loint(p0.w)];
    const bool first_in = (r0.flags & PF_EFPOT) && !(d20 < r0.dm2lo) &&
                          !(d20 > r0.dm2hi);
    if (!first_in && !(d < 0.0)) return;
    const double4 pj = P.pos[jj];
    const double ex = pa.x - pj.x, ey = pa.y - pj.y, ez = pa.z - pj.z;
    dids[0] = d;
    dids[1] = sqrt(ex * ex + ey * ey + ez * ez);
    dids[2] = ex;
    dids[3] = ey;
    dids[4] = ez;
}

// CONTRANS force on site i along axis ax (energy.mono.f90:520-580), added to f entry by
// entry: a restrained site is pulled to its centre, every site by a restrained centroid
// (of all sites, or of its strand with unsplit); no stfree gate, as in the reference.
__device__ inline double contrans_force(const StepParams &P, int i, int ax, double f) {
    const double *kk = ax == 0 ? P.kx : (ax == 1 ? P.ky : P.kz);
    const double *cc = ax == 0 ? P.contrx : (ax == 1 ? P.contry : P.contrz);
    const double xi = ax == 0 ? P.pos[i].x : (ax == 1 ? P.pos[i].y : P.pos[i].z);
    for (int q = 0; q < P.ctn; q++) {
        const double k = kk[q];
        if (k == 0.0) continue;
        const int j = P.csn[q];
        if (j == 0) {
            if (P.unsplit) {
                if (i < P.nsites1st) {
                    const double con = P.cen[3 + ax] * P.insites * 2.0 - cc[q];
                    f = f - k * con * P.insites * 2.0;
                } else {
                    const double con = P.cen[6 + ax] * P.insites * 2.0 - cc[P.ctn];
                    f = f - k * con * P.insites * 2.0;
                }
            } else {
                const double con = P.cen[ax] * P.insites - cc[q];
                f = f - k * con * P.insites;
            }
        } else if (j - 1 == i) {
            f = f - k * (xi - cc[q]);
        }
    }
    return f;
}

// CG-DNA list terms of site i (energy.mono.f90:196-508): this thread's share -- records
// tid, tid + NB, ... of the site's incidence lists -- of the force on the site and of the
// energies of the terms it heads (e[8], slots D_BOND..D_SOLV; formed as the reference forms
// them with EN, left incomplete without).  The site's angle and
// dihedral records come first; its other records, from inc_split on and then the one-word
// bulk records, are taken DU at a time: the records, then the other sites' positions, then
// the terms, so DU loads of each kind are in flight together.
// records per thread in flight when k_step computes a site's terms
constexpr int DNA_DU = 2;

// The angle and dihedral terms (energy.mono.f90:240-334) of the sites' incidences, one
// wavefront per site, lane by lane over the site's records: the force on the site and the
// term's energy (for the record of its first site) into fmul at the record's slot, for
// k_dna, whose own registers then stay few enough for a full complement of wavefronts.
__global__ __launch_bounds__(WAVE) void k_dna_multi(StepParams P) {
    const int i = blockIdx.x, lane = threadIdx.x;
    const bool free_me = P.stfree[i] != 0;
    const int k0 = P.incr_ptr[i], ks = P.inc_split[i];
    for (int k = k0 + lane; k < ks; k += WAVE) {
        V3 f = {0.0, 0.0, 0.0};
        double e[8];
#pragma unroll
        for (int q = 0; q < 8; q++) e[q] = 0.0;
        if (P.write_energy) dna_multi<true>(P, P.incr[k], i, free_me, f, e);
        else dna_multi<false>(P, P.incr[k], i, free_me, f, e);
        P.fmul[k] = make_double4(f.x, f.y, f.z, e[D_ANG] + e[D_DIHE]);
    }
}

template <int NB, int DU, bool MULTI, bool EN>
__device__ inline void dna_site_terms(const StepParams &P, int i, int tid, V3 &f, double e[8]) {
#pragma unroll
    for (int k = 0; k < 8; k++) e[k] = 0.0;
    const bool free_me = P.stfree[i] != 0;
    const int k0 = P.incr_ptr[i], ks = P.inc_split[i], k1 = P.incr_ptr[i + 1];
    const int b0 = P.incb_ptr[i], b1 = P.incb_ptr[i + 1];
    const V3 rme = ld3(P, i);
    if (P.flags & QNOBOND) {
        for (int k = k0 + tid; k < ks; k += NB) {
            if constexpr (MULTI) {
                dna_multi<EN>(P, P.incr[k], i, free_me, f, e);
            } else {   // evaluated by k_dna_multi
                const uint32_t w = (uint32_t)__double_as_longlong(P.incr[k].x);
                const double4 m = P.fmul[k];
                f.x += m.x; f.y += m.y; f.z += m.z;
                e[(w >> 29) == T_ANGLE ? D_ANG : D_DIHE] += m.w;
            }
        }
    }
    for (int kb = ks + tid; kb < k1; kb += DU * NB) {
        double2 rec[DU];
        V3 ro[DU];
#pragma unroll
        for (int u = 0; u < DU; u++) {
            // resolved incidence: one load gives the term, the other site and the list value
            const int k = kb + u * NB;
            rec[u] = P.incr[k < k1 ? k : ks];
        }
#pragma unroll
        for (int u = 0; u < DU; u++)
            ro[u] = ld3(P, (int)(((uint64_t)__double_as_longlong(rec[u].x) >> 32) & 0x7FFFFFFFu));
#pragma unroll
        for (int u = 0; u < DU; u++) {
            if (kb + u * NB >= k1) break;
            const uint64_t bits = (uint64_t)__double_as_longlong(rec[u].x);
            const uint32_t wd = (uint32_t)bits;
            const int kind = wd >> 29;
            const bool on = kind == T_BOND ? (P.flags & QNOBOND) != 0
                                           : ((P.flags & QNONBOND) != 0 &&
                                              (kind != T_SLV || (P.flags & QSOLV) != 0));
            if (!on) continue;
            const int role = (wd >> 27) & 3;
            const V3 v1 = role == 0 ? sub3(ro[u], rme) : sub3(rme, ro[u]);   // r_b - r_a
            const bool deby = kind == T_QQ && qq_debye(P, rme, ro[u]);
            dna_pair(P, kind, role, v1, rec[u].y, (bits >> 63) != 0, deby, free_me, f, e);
        }
    }
    // the bulk two-site terms, one word per record
    if (!(P.flags & QNONBOND)) return;
    for (int kb = b0 + tid; kb < b1; kb += DU * NB) {
        uint32_t rec[DU];
        V3 ro[DU];
#pragma unroll
        for (int u = 0; u < DU; u++) {
            const int k = kb + u * NB;
            rec[u] = P.incb[k < b1 ? k : b0];
        }
#pragma unroll
        for (int u = 0; u < DU; u++) ro[u] = ld3(P, (int)(rec[u] & 0xFFFFFFu));
#pragma unroll
        for (int u = 0; u < DU; u++) {
            if (kb + u * NB >= b1) break;
            const uint32_t w = rec[u];
            const int q = (w >> 24) & 3;
            const int kind = q == 0 ? T_EX : q == 1 ? T_QQ : q == 2 ? T_SLV : T_BP;
            if (kind == T_SLV && !(P.flags & QSOLV)) continue;
            const int role = (w >> 26) & 1;
            const V3 v1 = role == 0 ? sub3(ro[u], rme) : sub3(rme, ro[u]);   // r_b - r_a
            const bool deby = kind == T_QQ && qq_debye(P, rme, ro[u]);
            dna_pair(P, kind, role, v1, P.bpar[16 * q + (w >> 28)], ((w >> 27) & 1) != 0, deby,
                     free_me, f, e);
        }
    }
}

// The wavefront sums of dna_site_terms' outputs into s_r[wv][0..10] (force, then, with EN,
// energies).
template <int NW, bool EN>
__device__ inline void dna_block_sums(const StepParams &P, int lane, int wv, V3 f,
                                      double e[8], double (*s_r)[3 + 8]) {
    f.x = wave_sum(f.x);
    f.y = wave_sum(f.y);
    f.z = wave_sum(f.z);
    if (EN && P.write_energy) {
#pragma unroll
        for (int k = 0; k < 8; k++) e[k] = wave_sum(e[k]);
    }
    if (lane == 0) {
        s_r[wv][0] = f.x; s_r[wv][1] = f.y; s_r[wv][2] = f.z;
        if constexpr (EN) {
            for (int k = 0; k < 8; k++) s_r[wv][3 + k] = e[k];
        }
    }
}

// Entry c of site i's CG-DNA result from the block's wavefront sums, in wavefront order: a
// force component (c < 3, with APFOR and CONTRANS added) or the energy of slot c - 3.
template <int NW>
__device__ inline double dna_site_finish(const StepParams &P, int i, int c,
                                         const double (*s_r)[3 + 8]) {
    double acc = 0.0;
    for (int q = 0; q < NW; q++) acc += s_r[q][c];
    if (c < 3) {
        if ((P.flags & QAPFOR) && P.stfree[i] != 0) acc += reinterpret_cast<const double *>(&P.apf[i])[c];
        if (P.flags & QCONTRANS) acc = contrans_force(P, i, c, acc);
    }
    return acc;
}

// k_step: T target particles per block of NW wavefronts (targets tbeg.. of the launch).
// Every thread walks the partner set with stride NW*64, two partners per iteration, and
// evaluates each partner against all T targets.  Per-target sums: butterfly within each
// wavefront, then a fixed-order sum over the NW wavefronts in LDS.  Wavefront w < T also
// handles target w's 8 grid corners; thread a < T finishes target a (force, energies,
// BD update).  In a launch that writes no energies (the step loop) a one-target block
// computes its site target's CG-DNA list terms itself (dna_site_terms, before the pair
// loop); otherwise the site targets read them from k_dna (fdna).
template <int FEAT, int T, int NW>
__global__ __launch_bounds__(NW * WAVE) __attribute__((amdgpu_waves_per_eu(T == 1 ? 5 : 1)))
void k_step(StepParams P) {
    constexpr int NB = NW * WAVE;   // threads of the block
    __shared__ DevPair s_row[T][MAX_LDS_TYPES];   // pair parameters by partner type
    __shared__ double4 s_tail[T][MAX_LDS_TYPES];  // (sc3, tail fct, dm2hi, 0) by partner type
    __shared__ double s_red[NW][T][7];
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
        for (int k = tid; k < nt1; k += NB) {
            const DevPair d = P.rows[row + k];
            s_row[a][k] = d;
            s_tail[a][k] = make_double4(d.sc3, d.fct, d.dm2hi, 0.0);
        }
    }
    __syncthreads();

    // CG-DNA list terms of a site target (the block-uniform branch of a one-target block),
    // in the launches that write no energies; the others read k_dna's (fdna)
    constexpr bool DNA_HERE = (FEAT & F_DNA) != 0 && T == 1;
    const bool dna_here = DNA_HERE && !P.write_energy;
    __shared__ double s_dn[DNA_HERE ? NW : 1][3 + 8];
    if constexpr (DNA_HERE) {
        if (dna_here && tval[0] && tsite[0]) {
            V3 f = {0.0, 0.0, 0.0};
            double e[8];
            dna_site_terms<NB, DNA_DU, true, false>(P, ti[0], tid, f, e);
            dna_block_sums<NW, false>(P, lane, wv, f, e, s_dn);
        }
    }

    // Gaussians of the BD update, fetched by the last wavefront (lanes 0..T-1) before its
    // share of the pair loop
    if constexpr ((FEAT & F_UPDATE) != 0) {
        const int a = tid - (NB - WAVE);
        if (a >= 0 && a < T && tval[a]) {
            const int i = ti[a];
            const double *nr = noise_row(P, i);
            s_g[a][0] = nr[0]; s_g[a][1] = nr[1]; s_g[a][2] = nr[2];
        }
    }

    double fx[T], fy[T], fz[T], et[T], eel[T], evd[T], esr[T];
#pragma unroll
    for (int a = 0; a < T; a++) fx[a] = fy[a] = fz[a] = et[a] = eel[a] = evd[a] = esr[a] = 0.0;
    // PROXDIFF: per target ion, the tabulated site partner within its table range of least
    // dist - dmi, ties to the lower site index (energy.mono.f90:603-620 in site order)
    double pxd[T];
    int pxj[T];
#pragma unroll
    for (int a = 0; a < T; a++) { pxd[a] = INFINITY; pxj[a] = 0x7fffffff; }

    // ---- pair loop (energy.mono.f90:117-194, 584-687), compiled twice: with the energy
    // sums, and without them for the steps that write no energies
    const bool nonbond = (P.flags & QNONBOND) && (P.flags & QPAR);
    auto pair_loop = [&](auto en) {
        constexpr bool EN = decltype(en)::value;
        // partners per iteration: two (with the next iteration's partners loaded during
        // this one's, measured best for every shape; three or four only added registers)
        constexpr int PU = 2;
        // the next iteration's partners are loaded while this one's are evaluated
        double4 nx[PU];
#pragma unroll
        for (int u = 0; u < PU; u++) nx[u] = P.spos[min(tid + u * NB, ntot - 1)];
        for (int k0 = tid; k0 < ntot; k0 += PU * NB) {
            // partners in spatial order, PU per iteration
            double4 pp[PU];
            bool pok[PU];
#pragma unroll
            for (int u = 0; u < PU; u++) {
                pp[u] = nx[u];
                pok[u] = k0 + u * NB < ntot;
                nx[u] = P.spos[min(k0 + (PU + u) * NB, ntot - 1)];
            }
#pragma unroll
            for (int u = 0; u < PU; u++) {
                const double4 pj = pp[u];
                const bool jok = pok[u];
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
                    if constexpr ((FEAT & F_TAB) != 0) {
                        de = pair_near_far(s_row[a][t2j], s_tail[a][t2j], P.ep, P.ires, d2, y);
                        if constexpr (EN) et[a] += ce ? y : 0.0;
                    } else if (s_row[a][t2j].flags & PF_EFPOT) {
                        de = pair_near_far(s_row[a][t2j], s_tail[a][t2j], P.ep, P.ires, d2, y);
                        if constexpr (EN) et[a] += ce ? y : 0.0;
                    } else {
                        double q0 = 0.0, q1 = 0.0, q2 = 0.0;
                        de = pair_analytic(s_row[a][t2j], P, d2, q0, q1, q2);
                        if constexpr (EN) {
                            eel[a] += ce ? q0 : 0.0;
                            evd[a] += ce ? q1 : 0.0;
                            esr[a] += ce ? q2 : 0.0;
                        }
                    }
                    if constexpr ((FEAT & F_PROX) != 0) {
                        const DevPair &pr = s_row[a][t2j];
                        if (!skip && !tsite[a] && j < ns && ((FEAT & F_TAB) || (pr.flags & PF_EFPOT)) &&
                            !(d2 < pr.dm2lo) && !(d2 > pr.dm2hi)) {
                            const double dd = sqrt(d2) - pr.dmi;
                            if (dd < pxd[a] || (dd == pxd[a] && j < pxj[a])) { pxd[a] = dd; pxj[a] = j; }
                        }
                    }
                    de = (skip || !tfree[a]) ? 0.0 : de;
                    fx[a] += de * dx;
                    fy[a] += de * dy;
                    fz[a] += de * dz;
                }
            }
        }
    };
    if (nonbond) {
        if (P.write_energy) pair_loop(std::true_type{});
        else pair_loop(std::false_type{});
    }
    __shared__ double s_pxd[(FEAT & F_PROX) ? NW : 1][T];
    __shared__ int s_pxj[(FEAT & F_PROX) ? NW : 1][T];
    if constexpr ((FEAT & F_PROX) != 0) {
#pragma unroll
        for (int a = 0; a < T; a++) {
            double d = pxd[a];
            int jj = pxj[a];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double od = __shfl_xor(d, off, WAVE);
                const int oj = __shfl_xor(jj, off, WAVE);
                if (od < d || (od == d && oj < jj)) { d = od; jj = oj; }
            }
            if (lane == 0) { s_pxd[wv][a] = d; s_pxj[wv][a] = jj; }
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
        if (mine && val_me) warned = vdw_grid(P, xi, yi, zi, t2i, gfx, gfy, gfz, egv, vdw_map_of(P, i_me));
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
        for (int q = 0; q < NW; q++) acc += s_red[q][a][k];
        F[k] = acc + s_dg[a][k];
    }
    if constexpr ((FEAT & F_DNA) != 0) {
        if (is_site && dna_here) {
#pragma unroll
            for (int k = 0; k < 3; k++) F[k] += dna_site_finish<NW>(P, i, k, s_dn);
        } else if (is_site) {
            const double4 fd = P.fdna[i];   // CG-DNA list terms from k_dna
            F[0] += fd.x; F[1] += fd.y; F[2] += fd.z;
        }
    }
    if ((P.rfon || P.mmon) && !is_site) {   // reaction-field forces (k_rfpair or k_mm_force)
        const double4 fr = P.frf[i];
        F[0] += fr.x; F[1] += fr.y; F[2] += fr.z;
    }
    const double4 pa = P.pos[i];
    const int t2a = __double2loint(pa.w);
    if (s_warn[a]) atomicAdd(&P.warn[t2a], 1);
    if (P.flags & QMEMB) {   // MEMBRANE, membrane.f90:36-104
        const double chg = is_site ? (P.namsite[i] == 1 ? P.cgnuc : 0.0) : P.cg[abs(P.typei[i])];
        bool mw;
        const double em = membrane_energy(P, pa.x, pa.y, pa.z, chg, t2a, F, mw);
        if (mw) atomicAdd(&P.warn[t2a], 1);
        if (P.write_energy) P.epart[(size_t)E_MEMB * P.ecap + i] = em;
    }
    if (P.write_energy) {
        double pe[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < NW; q++) acc += s_red[q][a][3 + k];
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
        // virial of the step, F . x at the positions the forces were taken at
        // (simul1.f90:365-367); fixed ions carry no force and add nothing
        if (P.vacc) P.vacc[i] += F[0] * pa.x + F[1] * pa.y + F[2] * pa.z;
        const bool fr = !is_site || ((FEAT & F_DNA) && P.stfree[i] != 0);
        if (!fr) {
            P.poso[i] = pa;
            P.sposo[P.rank[i]] = pa;
            return;
        }
        const double g[3] = {s_g[a][0], s_g[a][1], s_g[a][2]};
        double dids[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if constexpr ((FEAT & F_PROX) != 0) {
            if (!is_site) prox_record<T, NW>(P, i, pa, a, s_pxd, s_pxj, s_row[a], dids);
        }
        const double4 pn = bd_update(P, i, is_site, pa, F, g, dids);
        P.poso[i] = pn;
        P.sposo[P.rank[i]] = pn;
    }
}

// k_centroid: site position sums of the current positions -- all sites, the first strand
// (sites 0..nsites1st-1) and the second -- in a fixed-order tree, then the CONTRANS energy
// econ in the reference's entry and axis order (energy.mono.f90:520-580), written to the
// first site's slot of its energy category.
__global__ __launch_bounds__(BLOCK) void k_centroid(StepParams P) {
    __shared__ double s[9][BLOCK];
    const int tid = threadIdx.x;
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = tid; i < P.nsites; i += BLOCK) {
        const double4 q = P.pos[i];
        const int h = i < P.nsites1st ? 3 : 6;
        a[0] += q.x; a[1] += q.y; a[2] += q.z;
        a[h] += q.x; a[h + 1] += q.y; a[h + 2] += q.z;
    }
    for (int k = 0; k < 9; k++) s[k][tid] = a[k];
    __syncthreads();
    for (int h = BLOCK / 2; h > 0; h >>= 1) {
        if (tid < h)
            for (int k = 0; k < 9; k++) s[k][tid] += s[k][tid + h];
        __syncthreads();
    }
    if (tid < 9) P.cen[tid] = s[tid][0];
    if (tid != 0 || !P.write_energy) return;
    double econ = 0.0;
    for (int q = 0; q < P.ctn; q++) {
        const int j = P.csn[q];
        for (int ax = 0; ax < 3; ax++) {
            const double k = ax == 0 ? P.kx[q] : (ax == 1 ? P.ky[q] : P.kz[q]);
            if (k == 0.0) continue;
            const double *cc = ax == 0 ? P.contrx : (ax == 1 ? P.contry : P.contrz);
            if (j == 0) {
                if (P.unsplit) {
                    double con = s[3 + ax][0] * P.insites * 2.0 - cc[q];
                    econ = econ + 0.5 * k * (con * con);
                    con = s[6 + ax][0] * P.insites * 2.0 - cc[P.ctn];
                    econ = econ + 0.5 * k * (con * con);
                } else {
                    const double con = s[ax][0] * P.insites - cc[q];
                    econ = econ + 0.5 * k * (con * con);
                }
            } else {
                const double4 pj = P.pos[j - 1];
                const double d = (ax == 0 ? pj.x : (ax == 1 ? pj.y : pj.z)) - cc[q];
                econ = econ + 0.5 * k * (d * d);
            }
        }
    }
    P.epart[(size_t)E_CON * P.ecap] = econ;
}

// k_notrans: NOTRANS after the step (simul1.f90:382-386): every site moves so that the
// sites' centroid sits at notr* on the selected axes, x = x - sum(x)/nsites + notrx, on the
// new positions and their spatial-order copies.
__global__ __launch_bounds__(BLOCK) void k_notrans(StepParams P) {
    __shared__ double s[3][BLOCK];
    const int tid = threadIdx.x;
    double a[3] = {0.0, 0.0, 0.0};
    for (int i = tid; i < P.nsites; i += BLOCK) {
        const double4 q = P.poso[i];
        a[0] += q.x; a[1] += q.y; a[2] += q.z;
    }
    for (int k = 0; k < 3; k++) s[k][tid] = a[k];
    __syncthreads();
    for (int h = BLOCK / 2; h > 0; h >>= 1) {
        if (tid < h)
            for (int k = 0; k < 3; k++) s[k][tid] += s[k][tid + h];
        __syncthreads();
    }
    const double sx = s[0][0], sy = s[1][0], sz = s[2][0];
    for (int i = tid; i < P.nsites; i += BLOCK) {
        double4 q = P.poso[i];
        if (P.notrans_axes & 1) q.x = q.x - sx * P.insites + P.notrx;
        if (P.notrans_axes & 2) q.y = q.y - sy * P.insites + P.notry;
        if (P.notrans_axes & 4) q.z = q.z - sz * P.insites + P.notrz;
        P.poso[i] = q;
        P.sposo[P.rank[i]] = q;
    }
}

// k_fixed: the terms of the fixed ions that no step kernel visits (they are never moved):
// their MEMBRANE energy and wall warnings (membrane.f90 loops over every particle).
__global__ void k_fixed(StepParams P) {
    const int i = P.nsites + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nsites + P.nfix) return;
    const double4 pa = P.pos[i];
    const int t2 = __double2loint(pa.w);
    double F[3] = {0.0, 0.0, 0.0};
    bool mw;
    const double em = membrane_energy(P, pa.x, pa.y, pa.z, P.cg[abs(P.typei[i])], t2, F, mw);
    if (mw) atomicAdd(&P.warn[t2], 1);
    if (P.write_energy) P.epart[(size_t)E_MEMB * P.ecap + i] = em;
}

hipError_t launch_fixed(const StepParams &p, hipStream_t s) {
    if ((p.flags & QMEMB) && p.nfix > 0)
        hipLaunchKernelGGL(k_fixed, dim3((p.nfix + 255) / 256), dim3(256), 0, s, p);
    return hipGetLastError();
}

// FIXCOOR of a curved system (fixcoor.f90:24-46) on the positions a step wrote: every moving
// ion, and every free site unless the sites are free of the system (dynamics0.mono.f90:45,
// dynamics0nuc.f90:42).  The step kernels clamp only the box axes (z of the cylinder), so
// this completes FIXCOOR; it reads and writes only the new positions, as the step does.
__global__ void k_curved(StepParams P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.ntot) return;
    if (i < P.nsites ? (P.stfree[i] == 0 || (P.flags & QDNAFREE) != 0) : i < P.nsites + P.nfix)
        return;
    double4 q = P.poso[i];
    if (!P.cv.fix(q.x, q.y, q.z)) return;
    P.poso[i] = q;
    P.sposo[P.rank[i]] = q;
}

hipError_t launch_notrans(const StepParams &p, hipStream_t s) {
    if (p.cv.kind != 0 && p.mode == 1 && p.ntot > 0)
        hipLaunchKernelGGL(k_curved, dim3((p.ntot + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, p);
    if ((p.flags & QNOTRANS) && p.nsites > 0)
        hipLaunchKernelGGL(k_notrans, dim3(1), dim3(BLOCK), 0, s, p);
    return hipGetLastError();
}

// k_dna: CG-DNA bonded and list terms (energy.mono.f90:196-508), one block per site, for
// the pair kernels that do not compute them in their site targets' blocks; the angles and
// dihedrals come from k_dna_multi (fmul).  Sums go butterfly-then-LDS in a fixed order.  The force on the site goes to fdna[i] (zero for a
// fixed site); the energies of the terms the site heads go to epart.
template <int NW>
__global__ __launch_bounds__(NW * WAVE) void k_dna(StepParams P) {
    constexpr int NB = NW * WAVE;   // threads of the block
    __shared__ double s_r[NW][3 + 8];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid >> 6;
    const int i = blockIdx.x;
    V3 f = {0.0, 0.0, 0.0};
    double e[8];
    dna_site_terms<NB, 4, false, true>(P, i, tid, f, e);
    dna_block_sums<NW, true>(P, lane, wv, f, e, s_r);
    __syncthreads();
    if (tid < 3 + 8) {
        const double acc = dna_site_finish<NW>(P, i, tid, s_r);
        if (tid < 3) reinterpret_cast<double *>(&P.fdna[i])[tid] = acc;
        else if (P.write_energy) P.epart[(size_t)ND_CAT[tid - 3] * P.ecap + i] = acc;
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
                if constexpr ((FEAT & F_TAB) != 0) {
                    de = pair_near_far(s_row[wv][tj], s_tail[wv][tj], P.ep, P.ires, d2, y);
                    et += ce ? y : 0.0;
                } else if (s_row[wv][tj].flags & PF_EFPOT) {
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
    if constexpr ((FEAT & F_PHIV) != 0)
        warned = vdw_grid(P, pa.x, pa.y, pa.z, t2i, fx, fy, fz, egv, vdw_map_of(P, i));
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
    if ((P.rfon || P.mmon) && !is_site) {
        const double4 fr = P.frf[i];
        F[0] += fr.x; F[1] += fr.y; F[2] += fr.z;
    }
    if (warned) atomicAdd(&P.warn[t2i], 1);
    if (P.flags & QMEMB) {   // MEMBRANE, membrane.f90:36-104
        const double chg = is_site ? (P.namsite[i] == 1 ? P.cgnuc : 0.0) : P.cg[abs(P.typei[i])];
        bool mw;
        const double em = membrane_energy(P, pa.x, pa.y, pa.z, chg, t2i, F, mw);
        if (mw) atomicAdd(&P.warn[t2i], 1);
        if (P.write_energy) P.epart[(size_t)E_MEMB * P.ecap + i] = em;
    }
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
        // virial of the step, F . x at the positions the forces were taken at
        // (simul1.f90:365-367); fixed ions carry no force and add nothing
        if (P.vacc) P.vacc[i] += F[0] * pa.x + F[1] * pa.y + F[2] * pa.z;
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

#define CASES16(B)                                                                         \
    CASE(B + 0) CASE(B + 1) CASE(B + 2) CASE(B + 3) CASE(B + 4) CASE(B + 5) CASE(B + 6)     \
    CASE(B + 7) CASE(B + 8) CASE(B + 9) CASE(B + 10) CASE(B + 11) CASE(B + 12)              \
    CASE(B + 13) CASE(B + 14) CASE(B + 15)

static void dispatch_wave(const StepParams &p, int feat, hipStream_t s) {
    dim3 grid((p.tend - p.tbeg + WPB - 1) / WPB), block(BLOCK);
#define CASE(F) case F: hipLaunchKernelGGL((k_wave<F>), grid, block, 0, s, p); break;
    switch (feat) { CASES16(0) CASES16(16) }
#undef CASE
}

template <int T, int NW>
static void dispatch_step(const StepParams &p, int feat, hipStream_t s) {
    dim3 grid((p.tend - p.tbeg + T - 1) / T), block(NW * WAVE);
#define CASE(F) case F: hipLaunchKernelGGL((k_step<F, T, NW>), grid, block, 0, s, p); break;
    switch (feat) { CASES16(0) CASES16(16) }
#undef CASE
}

// BD steps of DYNAMICS3 (DNA and ions present, update on): one target per block
static void dispatch_prox(const StepParams &p, int feat, hipStream_t s) {
    dim3 grid(p.tend - p.tbeg), block(BLOCK);
#define CASE(F) case F: hipLaunchKernelGGL((k_step<F | F_PROX, 1, WPB>), grid, block, 0, s, p); break;
    switch (feat) {
        CASE(9) CASE(11) CASE(13) CASE(15) CASE(25) CASE(27) CASE(29) CASE(31)
    }
#undef CASE
}

// Whether the pair kernel launch_pairs picks computes the CG-DNA list terms of its site
// targets itself: the one-target k_step (tpb 1, and every DYNAMICS3 step) in a launch that
// writes no energies.
static bool dna_in_pairs(const StepParams &p) {
    if (tile_mode(p.tpb) || p.write_energy) return false;
    if (p.dmode == 3 && p.mode == 1) return true;
    return p.tpb == 1;
}

hipError_t launch_dna(const StepParams &p, hipStream_t s) {
    if (!(p.flags & QNUCL) || p.nsites <= 0) return hipSuccess;
    if (p.flags & QCONTRANS) hipLaunchKernelGGL(k_centroid, dim3(1), dim3(BLOCK), 0, s, p);
    if (!dna_in_pairs(p)) {
        if (p.flags & QNOBOND) hipLaunchKernelGGL(k_dna_multi, dim3(p.nsites), dim3(WAVE), 0, s, p);
        hipLaunchKernelGGL(k_dna<WPB>, dim3(p.nsites), dim3(BLOCK), 0, s, p);
    }
    return hipGetLastError();
}

hipError_t launch_pairs(const StepParams &p, hipStream_t s) {
    if (tile_mode(p.tpb)) return launch_tiles(p, s);
    if (p.tend <= p.tbeg) return hipSuccess;
    int feat = 0;
    if (p.flags & QNUCL) feat |= F_DNA;
    if (p.flags & QPHIX) feat |= F_PHIX;
    if (p.flags & QPHIV) feat |= F_PHIV;
    if (p.mode == 1) feat |= F_UPDATE;
    if (p.all_efpot) feat |= F_TAB;
    if (p.dmode == 3 && (feat & F_UPDATE) && (feat & F_DNA)) {
        dispatch_prox(p, feat, s);
        return hipGetLastError();
    }
    switch (p.tpb) {
        // one target: two wavefronts per block, so a small system's blocks all fit in one
        // round of waves on the 256 CUs (measured faster than four at 2300 targets)
        case 1: dispatch_step<1, 2>(p, feat, s); break;
        case 2: dispatch_step<2, WPB>(p, feat, s); break;
        case 4: dispatch_step<4, WPB>(p, feat, s); break;
        default: dispatch_wave(p, feat, s); break;   // one wavefront per target
    }
    return hipGetLastError();
}

hipError_t launch_step(const StepParams &p, hipStream_t s) {
    hipError_t e = launch_rfpar(p, !tile_mode(p.tpb), s, true);
    if (e != hipSuccess) return e;
    if ((e = launch_mmij(p, s)) != hipSuccess) return e;
    if ((e = launch_fixed(p, s)) != hipSuccess) return e;
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
    // (srfe, reff, inside the grid box, moving); the pair kernel's partner record
    P.rfa[i] = make_double4(srfe, reff, in ? 1.0 : 0.0, i >= P.nsites + P.nfix ? 1.0 : 0.0);
    P.rfq[i] = make_double4(pa.x, pa.y, pa.z, P.cg[abs(P.typei[i])]);
    if (grad) {
        P.rfd[2 * (size_t)i] = make_double4(ds[0], ds[1], ds[2], 0.0);
        P.rfd[2 * (size_t)i + 1] = make_double4(dr[0], dr[1], dr[2], 0.0);
    }
}

// k_rfpair: reaction-field energy and force of each ion.  The reference visits each ion
// pair once, from its higher index (rfpar.f90:157-223): a pair counts when the lower ion's
// srfe is nonzero and the higher ion is moving and inside the grid box.  Gathered here per
// target t over every other ion p, in the reference's operand order (outer = higher
// index): aux0 = celec*q_hi*q_lo, energy charged to the higher index, and the force on t
// from the variation of srfe_t, reff_t and |r_t - r_p|.  A moving ion's force goes to frf
// (added by the step kernel); a fixed ion, which the step kernel does not visit, gets its
// (unused) reaction-field force in fx directly.
// Layout: each wavefront takes RF_TPW targets, the block's RF_TPW * 4 targets share tiles
// of 256 partner records (position + charge, srfe + reff + inside + moving) staged through
// LDS, double-buffered; each lane evaluates 4 partners of a tile against its targets.
constexpr int RF_TPW = 2;

__global__ __launch_bounds__(BLOCK) void k_rfpair(StepParams P) {
    __shared__ double4 s_q[2][BLOCK], s_a[2][BLOCK];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid >> 6;
    const int ns = P.nsites, ntot = P.ntot, nmov0 = P.nsites + P.nfix;
    int t[RF_TPW];
    bool tval[RF_TPW], tmov[RF_TPW], tin[RF_TPW];
    double4 pt[RF_TPW], rt[RF_TPW];
    double A[RF_TPW], B[RF_TPW], cx[RF_TPW], cy[RF_TPW], cz[RF_TPW], e[RF_TPW];
#pragma unroll
    for (int a = 0; a < RF_TPW; a++) {
        t[a] = ns + (blockIdx.x * WPB + wv) * RF_TPW + a;
        tval[a] = t[a] < ntot;
        const int tt = tval[a] ? t[a] : ns;
        pt[a] = P.rfq[tt];
        rt[a] = P.rfa[tt];
        tmov[a] = tt >= nmov0;
        tin[a] = tval[a] && rt[a].z != 0.0;
        A[a] = B[a] = cx[a] = cy[a] = cz[a] = e[a] = 0.0;
    }
    const int nion = ntot - ns;
    const int ntiles = (nion + BLOCK - 1) / BLOCK;
    double4 nq = make_double4(0.0, 0.0, 0.0, 0.0), na = nq;
    if (tid < nion) { nq = P.rfq[ns + tid]; na = P.rfa[ns + tid]; }
    s_q[0][tid] = nq;
    s_a[0][tid] = na;
    __syncthreads();
    for (int k = 0; k < ntiles; k++) {
        const int cur = k & 1;
        const int jn = (k + 1) * BLOCK + tid;
        if (k + 1 < ntiles) {
            nq = make_double4(0.0, 0.0, 0.0, 0.0);
            na = nq;
            if (jn < nion) { nq = P.rfq[ns + jn]; na = P.rfa[ns + jn]; }
        }
#pragma unroll 1
        for (int u = 0; u < WPB; u++) {
            const int idx = u * WAVE + lane;
            const int p = ns + k * BLOCK + idx;
            const double4 pp = s_q[cur][idx], rp = s_a[cur][idx];
#pragma unroll
            for (int a = 0; a < RF_TPW; a++) {
                const bool hi_t = p < t[a];
                // rp.z: partner inside the grid box; rp.w: partner moving
                const bool on = tin[a] && p < ntot && p != t[a] &&
                                (hi_t ? (rp.x != 0.0 && tmov[a])
                                      : (rt[a].x != 0.0 && rp.z != 0.0 && rp.w != 0.0));
                const double q_hi = hi_t ? pt[a].w : pp.w, q_lo = hi_t ? pp.w : pt[a].w;
                // masks as factors: every term below is finite (a masked pair's square
                // distance is lifted by 1), so m * term is exact 0 or the term
                const double m = on ? 1.0 : 0.0, mh = (on && hi_t) ? 1.0 : 0.0;
                const double aux0 = P.celec * q_hi * q_lo;
                const double dx = pt[a].x - pp.x, dy = pt[a].y - pp.y, dz = pt[a].z - pp.z;
                const double dist2 = dx * dx + dy * dy + dz * dz;
                const double srfeij = rt[a].x * rp.x, reffij = rt[a].y * rp.y;
                const double rfdn = rsqrt_nr(reffij * reffij + dist2 + (1.0 - m));
                const double rfcf = reffij * rfdn;
                const double aux1 = aux0 * rfcf;
                const double aux2 = aux0 * srfeij * (rfdn * (rfdn * rfdn));
                const double aux3 = -aux2 * (dist2 + reffij * 0.5);
                const double de = m * (aux2 * reffij);
                e[a] = fma(mh, aux1 * srfeij, e[a]);
                A[a] = fma(m, -aux1 * rp.x, A[a]);    // times d srfe_t
                B[a] = fma(m, aux3 * rp.y, B[a]);     // times d reff_t
                cx[a] = fma(de, dx, cx[a]);
                cy[a] = fma(de, dy, cy[a]);
                cz[a] = fma(de, dz, cz[a]);
            }
        }
        if (k + 1 < ntiles) {
            s_q[cur ^ 1][tid] = nq;
            s_a[cur ^ 1][tid] = na;
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < RF_TPW; a++) {
        A[a] = wave_sum(A[a]); B[a] = wave_sum(B[a]);
        cx[a] = wave_sum(cx[a]); cy[a] = wave_sum(cy[a]); cz[a] = wave_sum(cz[a]);
        if (P.write_energy) e[a] = wave_sum(e[a]);
    }
    if (lane != 0) return;
#pragma unroll
    for (int a = 0; a < RF_TPW; a++) {
        if (!tval[a]) continue;
        const int ta = t[a];
        if (!tin[a]) {   // outside the grid box: no reaction-field term at all
            P.frf[ta] = make_double4(0.0, 0.0, 0.0, 0.0);
            if (P.write_energy) P.epart[(size_t)E_RFPAR * P.ecap + ta] = 0.0;
            if (!tmov[a] && P.write_forces) { P.fx[ta] = 0.0; P.fy[ta] = 0.0; P.fz[ta] = 0.0; }
            continue;
        }
        const double4 dst = P.rfd[2 * (size_t)ta], drt = P.rfd[2 * (size_t)ta + 1];
        const double srt = rt[a].x, cgt = pt[a].w;
        // self term of a moving ion (rfpar.f90:146-156): aux = celec*q*q*srfe, energy
        // aux*srfe/2, force -aux*dsrfe
        const double aux = tmov[a] ? P.celec * cgt * cgt * srt : 0.0;
        const double a_s = A[a] - aux;
        const double4 f = make_double4(a_s * dst.x + B[a] * drt.x + cx[a],
                                       a_s * dst.y + B[a] * drt.y + cy[a],
                                       a_s * dst.z + B[a] * drt.z + cz[a], 0.0);
        if (tmov[a]) {
            P.frf[ta] = f;
            if (P.write_energy) P.epart[(size_t)E_RFPAR * P.ecap + ta] = 0.5 * aux * srt + e[a];
        } else if (P.write_forces) {
            P.fx[ta] = f.x; P.fy[ta] = f.y; P.fz[ta] = f.z;
        }
    }
}

hipError_t launch_rfpar(const StepParams &p, bool pairs, hipStream_t s, bool grad) {
    const int nion = p.ntot - p.nsites;
    if (!p.rfon || nion <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rfpar, dim3((nion + 255) / 256), dim3(256), 0, s, p,
                       (pairs || grad) ? 1 : 0);
    if (pairs)
        hipLaunchKernelGGL(k_rfpair, dim3((nion + WPB * RF_TPW - 1) / (WPB * RF_TPW)),
                           dim3(BLOCK), 0, s, p);
    return hipGetLastError();
}

// ---- energy reduction: one block per category, fixed-order tree
__device__ __forceinline__ int tile_cat_of(int cat) {
    switch (cat) {
        case E_EFPOT: return TE_EFPOT;
        case E_ELEC: return TE_ELEC;
        case E_VDW: return TE_VDW;
        case E_SRPMF: return TE_SRPMF;
        case E_EFPOTMX: return TE_EFPOTMX;
        case E_QQMX: return TE_QQMX;
        case E_VDWMX: return TE_VDWMX;
        case E_SRPMFMX: return TE_SRPMFMX;
        case E_RFPAR: return TE_RFPAR;
        default: return -1;
    }
}

// Energy terms: category cat sums its per-particle partials followed by its tile pairs'
// energies (tile mode), as one sequence cut into RED_BLOCKS contiguous ranges; block b
// sums range b (four interleaved accumulators per thread, so loads overlap, then a tree),
// and k_reduce_finish adds the RED_BLOCKS range sums in order.  Fixed shapes throughout:
// the same inputs give the same bits.
__global__ __launch_bounds__(BLOCK) void k_reduce_energy(const double *epart, int ecap, int n,
                                                         double *part, const double *tener,
                                                         int titems) {
    __shared__ double s[BLOCK];
    const int cat = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
    const double *src = epart + (size_t)cat * ecap;
    const int tc = tener ? tile_cat_of(cat) : -1;
    const double *tsrc = tc >= 0 ? tener + (size_t)tc * titems : nullptr;
    const int m = n + (tc >= 0 ? titems : 0);
    const int chunk = (m + RED_BLOCKS - 1) / RED_BLOCKS;
    const int lo = b * chunk, hi = min(lo + chunk, m);
    auto val = [&](int k) { return k < n ? src[k] : tsrc[k - n]; };
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int k = lo + tid;
    for (; k + 3 * BLOCK < hi; k += 4 * BLOCK) {
        a0 += val(k);
        a1 += val(k + BLOCK);
        a2 += val(k + 2 * BLOCK);
        a3 += val(k + 3 * BLOCK);
    }
    for (; k < hi; k += BLOCK) a0 += val(k);
    s[tid] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    for (int h = BLOCK / 2; h > 0; h >>= 1) {
        if (tid < h) s[tid] += s[tid + h];
        __syncthreads();
    }
    if (tid == 0) part[(size_t)cat * RED_BLOCKS + b] = s[0];
}

__global__ void k_reduce_finish(const double *part, double *out) {
    const int cat = threadIdx.x;
    if (cat >= NE) return;
    double acc = 0.0;
    for (int b = 0; b < RED_BLOCKS; b++) acc += part[(size_t)cat * RED_BLOCKS + b];
    out[cat] = acc;
}

hipError_t launch_reduce_energy(const double *epart, int ecap, int n, double *out,
                                hipStream_t s, const double *tener, int titems) {
    double *part = out + NE;   // the caller's buffer holds NE * (1 + RED_BLOCKS) doubles
    hipLaunchKernelGGL(k_reduce_energy, dim3(NE, RED_BLOCKS), dim3(BLOCK), 0, s, epart, ecap, n,
                       part, tener, titems);
    hipLaunchKernelGGL(k_reduce_finish, dim3(1), dim3(64), 0, s, (const double *)part, out);
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
    const V3 ra = ld3(P, a), rb = ld3(P, b);
    V3 v = sub3(rb, ra);
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
        if (!qq_debye(P, ra, rb)) return P.fctn * idist;   // interact.f90:323-331
        return P.fctn * idist * exp(-dist * P.ikappa);
    }
    double dist = sqrt(dist2);
    double cte1 = exp((13.38 - dist) * 0.1875);
    return P.epsolv * ((1.0 - cte1) * (1.0 - cte1)) - P.epsolv;
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
    // MEMBRANE part of INTERACT (interact.f90:58-119)
    if ((P.flags & QMEMB) && tid == 0) {
        const bool site_p = (P.flags & QNUCL) && j < ns && P.namsite[j] == 1;
        const double chg = beion ? P.cg[Q.jtype] : (site_p ? P.cgnuc : 0.0);
        bool mw;
        e += membrane_interact(P, xj, yj, zj, chg, jt2, mw);
        if (mw && Q.alert) atomicAdd(&P.warn[jt2], 1);
    }
    // staticf0.f90:19-82 and vdwgd0trln.f90:19-117 on lanes 0..7 of wave 0
    if (tid < WAVE) {
        const bool besite = (P.flags & QNUCL) && j < ns && P.namsite[j] == 1;
        const double chi = beion ? P.cg[Q.jtype] : P.cgnuc;
        bool wall = false;
        const double eg = interact_grids(P, P.phix, beion || besite, chi, xj, yj, zj, jt2, wall,
                                         vdw_map(P, j, Q.jtype));
        if (tid == 0) {
            e += eg;
            if (wall && Q.alert) atomicAdd(&P.warn[jt2], 1);
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
