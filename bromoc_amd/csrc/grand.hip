// grand.hip -- GRAND of one system on the device (grand.f90:40-118 with INSERT,
// insert.f90:29-41, and FIND, find.f90:26-32).
//
// The attempts of a GRAND call are sequential -- each one's acceptance changes the
// particle set the next one sees -- but each attempt's INTERACT is an O(N) sum.  A call
// launches, back to back on the stream with no host round trip:
//   k_gc_rfpar     (reaction field only, once per call) srfe, reff of every ion at the
//                  current positions; an accepted attempt updates the one slot it changes;
//   k_gc_pre       one workgroup: the first trial, drawn from the uniform stream in the
//                  reference's order; for a deletion FIND, the ip-th ion of the buffer
//                  (ballot counts of 64 consecutive slots per wavefront, by groups; the
//                  wavefront that holds it searches one group);
// then per attempt
//   k_gc_interact  G workgroups: INTERACT of the trial ion (interact.f90:19-419 for an ion:
//                  MEMBRANE, grids, ion and site pairs, RFPARIONJ), one partial sum per
//                  workgroup;
//   k_gc_post      one workgroup: the partials summed in a fixed order, the acceptance
//                  test (grand.f90:50-59, 88-97), insertion at the end or deletion by
//                  swap-with-last (grand.f90:60-76, 98-110), the prob histogram, and then
//                  the next attempt's trial (k_gc_pre's work) in the same launch.
// The state (uniform cursor, particle count, energy, error) lives in device memory; the
// host reads it back once per call.  A failed attempt (uniform stream exhausted, capacity,
// inconsistent buffers) sets the error and every later kernel of the call does nothing.
#include "device.h"

namespace bromoc {

constexpr int GC_THREADS = 1024;              // threads of k_gc_pre and k_gc_post
constexpr int GC_WAVES = GC_THREADS / WAVE;
constexpr int FIND_GROUP = 512;   // slots per group of FIND's counts (at least)
constexpr int FIND_MAXG = 128;    // groups per wavefront

// The next trial (grand.f90:42-50, 77-84) and, for a deletion, FIND (find.f90:26-32): the
// ip-th slot, in slot order, of the moving ions whose buffer is ib.  Called by a whole
// workgroup.
__device__ inline void gc_pre(const GcParams &G) {
    // the trial position and acceptance arithmetic as written, unfused: a fused multiply-add
    // rounds differently from the reference's separate operations
#pragma clang fp contract(off)
    __shared__ int s_cnt[GC_WAVES];
    __shared__ int s_grp[GC_WAVES][FIND_MAXG];
    __shared__ int s_found;
    GcState &S = *G.st;
    if (S.err) return;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid >> 6;
    if (tid == 0) {
        int64_t k = S.k;
        int mode = 0, ib = 0, ip = 0, itype = 0;
        bool ok = true;
        auto rndm = [&](double &v) {
            if (k >= G.nu) { ok = false; v = 0.0; return; }
            v = G.u[k++];
        };
        double r0, r1;
        rndm(r0);
        const double nbufferr = (double)G.nbuffer;
        if (ok && r0 < 0.5) {
            rndm(r1);
            ib = (int)(nbufferr * r1) + 1;
            if (ok && ib <= G.nbuffer && G.avnum[ib] > 0.0) {
                mode = 1;
                itype = G.ibfftyp[ib];
                // INSERT (insert.f90:29-41): three draws, again until the point is in the
                // buffer's region (at most as many tries as the uniforms allow)
                const Curved cv = G.sp.cv;
                const double rlo = G.rmin[ib], rhi = G.rmax[ib];
                double x = 0.0, y = 0.0, z = 0.0;
                for (;;) {
                    if (k + 3 > G.nu) { ok = false; k = G.nu; break; }
                    x = G.lx * (G.u[k] - 0.5) + G.cx;
                    y = G.ly * (G.u[k + 1] - 0.5) + G.cy;
                    z = G.lzmin[ib] + (G.lzmax[ib] - G.lzmin[ib]) * G.u[k + 2];
                    k += 3;
                    if (cv.insert_ok(x, y, z, G.cx, G.cy, G.cz, rlo, rhi)) break;
                }
                S.x = x;
                S.y = y;
                S.z = z;
            }
        } else if (ok) {
            double r2;
            rndm(r1);
            ib = (int)(nbufferr * r1) + 1;
            rndm(r2);
            const double natib = ib <= G.nbuffer ? (double)G.nat[ib] : 0.0;
            ip = (int)(natib * r2) + 1;
            if (ok && ib <= G.nbuffer && ip <= G.nat[ib]) {
                mode = 2;
                itype = G.ibfftyp[ib];
            }
        }
        if (!ok) { S.err = GCE_UNIFORMS; mode = 0; }
        S.k = k;
        S.mode = mode; S.ib = ib; S.ip = ip; S.itype = itype;
        S.j = mode == 1 ? S.ntot : -1;
        s_found = -1;
    }
    __syncthreads();
    if (S.mode != 2) return;
    // Wavefront w takes the contiguous w-th part of the moving slots and counts the buffer's
    // ions in it by ballots over 64 consecutive slots, eight loads in flight, keeping the
    // count of every group of FIND_GROUP slots; the wavefront whose part holds the ip-th
    // walks its group counts to the group, then the group's chunks to the slot.
    const int ib = S.ib, ip = S.ip, ntot = S.ntot;
    const int first = G.sp.nsites + G.sp.nfix, cnt = ntot > first ? ntot - first : 0;
    int gsz = FIND_GROUP;   // slots per group: a multiple of 512, at most FIND_MAXG groups
    while ((cnt + GC_WAVES * gsz - 1) / (GC_WAVES * gsz) > FIND_MAXG) gsz *= 2;
    const int ngw = (cnt + GC_WAVES * gsz - 1) / (GC_WAVES * gsz);   // groups per wavefront
    const int lo = first + min(wv * ngw * gsz, cnt), hi = first + min((wv + 1) * ngw * gsz, cnt);
    int c = 0;
    for (int g = 0; g < ngw; g++) {
        int cg = 0;
        const int g0 = lo + g * gsz, g1 = min(g0 + gsz, hi);
        for (int base = g0; base < g1; base += 8 * WAVE) {
            bool m[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int a = base + u * WAVE + lane;
                m[u] = a < g1 && G.ibuffer[a] == ib;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) cg += __popcll(__ballot(m[u]));
        }
        if (lane == 0) s_grp[wv][g] = cg;
        c += cg;
    }
    if (lane == 0) s_cnt[wv] = c;
    __syncthreads();
    int before = 0;
    for (int q = 0; q < wv; q++) before += s_cnt[q];
    if (ip > before && ip <= before + c) {   // the same for the whole wavefront
        int need = ip - before, g = 0;
        while (need > s_grp[wv][g]) need -= s_grp[wv][g++];
        const int g0 = lo + g * gsz, g1 = min(g0 + gsz, hi);
        for (int base = g0; base < g1; base += WAVE) {
            const int a = base + lane;
            const uint64_t m = __ballot(a < g1 && G.ibuffer[a] == ib);
            const int k = __popcll(m);
            if (need <= k) {   // the need-th of this chunk's matches, in lane order
                if (a < g1 && G.ibuffer[a] == ib && __popcll(m & ((1ull << lane) - 1ull)) == need - 1)
                    s_found = a;
                break;
            }
            need -= k;
        }
    }
    __syncthreads();
    if (tid == 0) {
        const int iat = s_found;
        if (iat < 0) {
            S.err = GCE_STATE;
            S.mode = 0;
        } else {
            const double4 q = G.sp.pos[iat];
            S.x = q.x; S.y = q.y; S.z = q.z;
            S.j = iat;
        }
    }
}

__global__ __launch_bounds__(GC_THREADS) void k_gc_pre(GcParams G) { gc_pre(G); }

// srfe, reff (inside, moving) of every ion at the current positions, as k_rfpar without
// gradients, over the device particle count
__global__ void k_gc_rfpar(GcParams G) {
    const GcState &S = *G.st;
    if (S.err) return;
    const StepParams &P = G.sp;
    const int i = P.nsites + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.ntot) return;
    const double4 pa = P.pos[i];
    double srfe, reff, ds[3], dr[3];
    const bool in = rf_params(P, pa.x, pa.y, pa.z, abs(G.typei[i]), false, srfe, reff, ds, dr);
    P.rfa[i] = make_double4(srfe, reff, in ? 1.0 : 0.0, i >= P.nsites + P.nfix ? 1.0 : 0.0);
}

// INTERACT of the trial ion (interact.f90:49-416 with Beion): workgroup 0 adds the
// one-particle terms (MEMBRANE, static field and repulsion grids, the reaction-field self
// term); every workgroup sums its share of the ion partners (i != j) and site partners.
__global__ __launch_bounds__(BLOCK) void k_gc_interact(GcParams G) {
    const GcState &S = *G.st;
    if (S.err || S.mode == 0) return;
    const StepParams &P = G.sp;
    __shared__ double s_red[WPB];
    __shared__ double s_rf[3];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    const int ns = P.nsites, ntot = S.ntot, nt1 = P.nttyp + 1;
    const int j = S.j, jtype = S.itype;
    const bool alert = S.mode == 2;
    const double xj = S.x, yj = S.y, zj = S.z;
    const int jt2 = P.nwtype[jtype];
    const int gtid = blockIdx.x * BLOCK + tid, gstride = gridDim.x * BLOCK;
    double e = 0.0;
    if (blockIdx.x == 0) {
        if ((P.flags & QMEMB) && tid == 0) {
            bool mw;
            e += membrane_interact(P, xj, yj, zj, P.cg[jtype], jt2, mw);
            if (mw && alert) atomicAdd(&P.warn[jt2], 1);
        }
        if (tid < WAVE) {
            bool wall = false;
            const double eg = interact_grids(P, P.phix, true, P.cg[jtype], xj, yj, zj, jt2, wall,
                                             vdw_map(P, j, jtype));
            if (tid == 0) {
                e += eg;
                if (wall && alert) atomicAdd(&P.warn[jt2], 1);
            }
        }
    }
    if (P.flags & QNONBOND) {
        for (int i = ns + gtid; i < ntot; i += gstride) {
            if (i == j) continue;
            const double4 q = P.pos[i];
            const double dx = q.x - xj, dy = q.y - yj, dz = q.z - zj;
            const double d2 = dx * dx + dy * dy + dz * dz;
            const DevPair &pp = P.rows[jt2 * nt1 + __double2loint(q.w)];
            e += (pp.flags & PF_EFPOT) ? gety(pp, P.ep, P.ires, d2) : interact_analytic(pp, P, d2);
        }
        if ((P.flags & QPAR) && (P.flags & QNUCL)) {
            for (int i = gtid; i < ns; i += gstride) {
                const double4 q = P.pos[i];
                const double dx = q.x - xj, dy = q.y - yj, dz = q.z - zj;
                const double d2 = dx * dx + dy * dy + dz * dz;
                const DevPair &pp = P.rows[jt2 * nt1 + __double2loint(q.w)];
                e += (pp.flags & PF_EFPOT) ? gety(pp, P.ep, P.ires, d2) : interact_analytic(pp, P, d2);
            }
        }
    }
    if (P.rfon) {
        // RFPARIONJ (rfpar.f90:232-370): the trial's srfe / reff, its self term and every
        // other ion with nonzero srfe; nothing when the trial is outside the grid box
        if (tid == 0) {
            double sj, rj, ds[3], dr[3];
            const bool in = rf_params(P, xj, yj, zj, jtype, false, sj, rj, ds, dr);
            s_rf[0] = sj; s_rf[1] = rj; s_rf[2] = in ? 1.0 : 0.0;
        }
        __syncthreads();
        if (s_rf[2] != 0.0) {
            const double sj = s_rf[0], rj = s_rf[1];
            const double tau = P.celec * P.cg[jtype];
            if (blockIdx.x == 0 && tid == 0) e += 0.5 * tau * P.cg[jtype] * (sj * sj);
            for (int i = ns + gtid; i < ntot; i += gstride) {
                if (i == j) continue;
                const double4 ra = P.rfa[i];
                if (ra.x == 0.0) continue;
                const double4 q = P.pos[i];
                const double dx = q.x - xj, dy = q.y - yj, dz = q.z - zj;
                const double dist2 = dx * dx + dy * dy + dz * dz;
                const double reffij = rj * ra.y;
                e += tau * P.cg[abs(G.typei[i])] * reffij * sj * ra.x / sqrt(reffij * reffij + dist2);
            }
        }
    }
    e = wave_sum(e);
    if (lane == 0) s_red[tid >> 6] = e;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int k = 0; k < WPB; k++) t += s_red[k];
        G.part[blockIdx.x] = t;
    }
}

// The attempt's acceptance (grand.f90:50-116).  Called by a whole workgroup.
__device__ inline void gc_post(const GcParams &G) {
#pragma clang fp contract(off)
    __shared__ double s[GC_THREADS];
    GcState &S = *G.st;
    const int tid = threadIdx.x;
    if (S.err) return;
    const int mode = S.mode;
    double dener = 0.0;
    if (mode != 0) {
        double a = 0.0;
        for (int b = tid; b < GC_BLOCKS; b += GC_THREADS) a += G.part[b];
        s[tid] = a;
        __syncthreads();
        for (int h = GC_THREADS / 2; h > 0; h >>= 1) {
            if (tid < h) s[tid] += s[tid + h];
            __syncthreads();
        }
        dener = s[0];
    }
    if (tid != 0) return;
    const int ib = S.ib;
    if (mode != 0) {
        // acceptance (grand.f90:50-59, 88-97), uniform drawn after the trial's INTERACT
        if (S.k >= G.nu) { S.err = GCE_UNIFORMS; return; }
        const double r = G.u[S.k++];
        double rate;
        if (mode == 1) {
            if (G.bias[ib]) {
                rate = (G.avnum[ib] + G.kb[ib] * (G.avnum[ib] - (double)(G.ntotat[ib] / G.icycle)));
                rate = rate * exp(-(dener - G.mu[ib]) * G.ikbt) / (double)(G.nat[ib] + 1);
            } else {
                rate = (G.avnum[ib] / (G.nat[ib] + 1)) * exp(-(dener - G.mu[ib]) * G.ikbt);
            }
            rate = rate / (1.0 + rate);
            if (r <= rate) {
                const int n = S.ntot;
                if (n + 1 > G.cap) { S.err = GCE_CAPACITY; return; }
                S.ener = S.ener + dener;
                G.nat[ib] = G.nat[ib] + 1;
                const int itype = S.itype;
                const int32_t ti = S.z < G.cz ? -itype : itype;
                const int32_t t2 = G.sp.nwtype[itype];
                double w;
                {
                    const uint64_t bits = (uint64_t)(uint32_t)t2 | ((uint64_t)(uint32_t)n << 32);
                    w = __longlong_as_double((long long)bits);
                }
                const double4 pn = make_double4(S.x, S.y, S.z, w);
                G.pos0[n] = pn;
                G.pos1[n] = pn;
                G.typei[n] = ti;
                G.t2[n] = t2;
                G.ibuffer[n] = ib;
                if (G.sp.rfon) {   // the new ion's srfe, reff, as k_gc_rfpar computes them
                    double sr, rr, ds[3], dr[3];
                    const bool in = rf_params(G.sp, S.x, S.y, S.z, itype, false, sr, rr, ds, dr);
                    G.sp.rfa[n] = make_double4(sr, rr, in ? 1.0 : 0.0, 1.0);
                }
                S.ntot = n + 1;
                S.nacc++;
            }
        } else {
            if (G.bias[ib]) {
                rate = (G.avnum[ib] + G.kb[ib] * (G.avnum[ib] - (double)(G.ntotat[ib] / G.icycle)));
                rate = rate * exp(-(dener - G.mu[ib]) * G.ikbt) / (double)G.nat[ib];
            } else {
                rate = (G.avnum[ib] / G.nat[ib]) * exp(-(dener - G.mu[ib]) * G.ikbt);
            }
            rate = 1.0 / (1.0 + rate);
            if (r <= rate) {
                S.ener = S.ener - dener;
                const int iat = S.j, last = S.ntot - 1;
                // grand.f90:102-109: the last particle takes the deleted one's slot
                double4 q = G.sp.pos[last];
                const uint64_t bits = (uint64_t)(uint32_t)G.t2[last] | ((uint64_t)(uint32_t)iat << 32);
                q.w = __longlong_as_double((long long)bits);
                G.pos0[iat] = q;
                G.pos1[iat] = q;
                G.typei[iat] = G.typei[last];
                G.t2[iat] = G.t2[last];
                G.ibuffer[iat] = G.ibuffer[last];
                if (G.sp.rfon) G.sp.rfa[iat] = G.sp.rfa[last];
                G.nat[ib] = G.nat[ib] - 1;
                S.ntot = last;
                S.nacc++;
            }
        }
    }
    // prob histogram (grand.f90:114-116), after every attempt: each buffer's count is
    // logged, and the host adds one to prob(b, count) for each entry
    const int a = S.natt;
    for (int b = 1; b <= G.nbuffer; b++) {
        const int nb = G.nat[b];
        if (nb < 0 || nb > G.cap) { S.err = GCE_STATE; return; }
        G.natlog[(size_t)a * G.nbuffer + (b - 1)] = nb;
    }
    S.natt = a + 1;
}

// one attempt's acceptance, then (next) the following attempt's trial
__global__ __launch_bounds__(GC_THREADS) void k_gc_post(GcParams G, int next) {
    gc_post(G);
    __syncthreads();
    if (next) gc_pre(G);
}

hipError_t launch_grand(const GcParams &g, int ngcmc, hipStream_t s) {
    if (ngcmc <= 0) return hipSuccess;
    if (g.sp.rfon) {
        const int nion = g.cap - g.sp.nsites;
        if (nion > 0) hipLaunchKernelGGL(k_gc_rfpar, dim3((nion + 255) / 256), dim3(256), 0, s, g);
    }
    hipLaunchKernelGGL(k_gc_pre, dim3(1), dim3(GC_THREADS), 0, s, g);
    for (int a = 0; a < ngcmc; a++) {
        hipLaunchKernelGGL(k_gc_interact, dim3(GC_BLOCKS), dim3(BLOCK), 0, s, g);
        hipLaunchKernelGGL(k_gc_post, dim3(1), dim3(GC_THREADS), 0, s, g, a + 1 < ngcmc ? 1 : 0);
    }
    return hipGetLastError();
}

}  // namespace bromoc
