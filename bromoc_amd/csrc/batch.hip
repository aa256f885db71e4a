// batch.hip -- replica batches: many independent ion systems advanced together.
//
// Every replica owns a contiguous range of particle slots [off, off + cap) in shared
// arrays; the pair tables, the repulsion map and the BD constants are shared, each replica
// has its own static field and GCMC buffers.  One BD step of all replicas is one launch of
// k_batch_step: its wavefronts take (replica, moving ion) work items -- ordered by replica
// size, largest first -- from a device queue (or one item each in the non-persistent
// shape), gather the ion's pair forces over its own replica's particles (all pairs, as
// energy.mono.f90:117-194), its grid corners, and apply DYNAMICS0/1/2.  k_batch_gcmc runs
// GRAND (grand.f90:40-118) for every replica at once, one workgroup per replica, with
// INTERACT energies (interact.f90:19-419, ions only) reduced across the workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "device.h"

namespace bromoc {

// ---- work-item prefix (one block of 1024 threads) and step / cycle counters
constexpr int PREP_THREADS = 1024;
constexpr int BATCH_LDS_WSTART = 4096;   // work-item prefix held in LDS up to this size

__global__ __launch_bounds__(PREP_THREADS) void k_batch_prep(BatchParams B, int gcmc) {
    const int tid = threadIdx.x;
    if (gcmc) {
        if (tid == 0) {
            B.ctl[5] = B.ctl[6] + B.ctl[2];
            B.ctl[2] += 1;
        }
        return;
    }
    __shared__ int s_wave[PREP_THREADS / WAVE];
    __shared__ int s_carry;
    if (tid == 0) {
        B.ctl[4] = B.ctl[1] + B.ctl[0];
        B.ctl[0] += 1;
        B.queue[0] = 0;
        s_carry = 0;
    }
    __syncthreads();
    const int lane = tid & (WAVE - 1), wv = tid / WAVE;
    for (int base = 0; base < B.nrep; base += PREP_THREADS) {
        const int k = base + tid;
        int v = 0;
        if (k < B.nrep) {
            const int r = B.order[k];
            v = max(0, B.ntot[r] - B.nfix[r]);
        }
        // inclusive scan within the wavefront
        int x = v;
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            const int y = __shfl_up(x, d, WAVE);
            if (lane >= d) x += y;
        }
        if (lane == WAVE - 1) s_wave[wv] = x;
        __syncthreads();
        if (tid == 0) {
            int acc = 0;
            for (int q = 0; q < PREP_THREADS / WAVE; q++) {
                const int t = s_wave[q];
                s_wave[q] = acc;
                acc += t;
            }
        }
        __syncthreads();
        const int excl = s_carry + s_wave[wv] + x - v;
        if (k < B.nrep) B.wstart[k] = excl;
        __syncthreads();
        if (tid == PREP_THREADS - 1) s_carry = excl + v;
        __syncthreads();
    }
    if (tid == 0) B.wstart[B.nrep] = s_carry;
}

hipError_t launch_batch_prep(const BatchParams &b, int gcmc, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_prep, dim3(1), dim3(PREP_THREADS), 0, s, b, gcmc);
    return hipGetLastError();
}

// ---- one BD step (or ENERGY) of every replica
// G lanes per target: each wavefront holds 64/G targets, every lane group of G lanes takes
// its own work item (a moving ion of some replica), strides over that replica's particles
// and reduces its sums within the group.  Small replicas want small groups (more targets
// in flight, the per-item latencies overlap); the autotuner picks G.
// feature bits of the k_batch_step specialisations
constexpr int BF_PHIX = 1;       // static field grids
constexpr int BF_PHIV = 2;       // repulsion map
constexpr int BF_ANALYTIC = 4;   // some pair type is analytic (else every pair is tabulated)

template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// component c of philox_gauss3 (the same operations, so the same bits)
__device__ __forceinline__ double philox_gauss1(uint64_t seed, uint64_t step, uint64_t stream,
                                                int64_t i, int c) {
    uint64_t w[4] = {step, (uint64_t)i, stream, 0};
    philox4x64(w, seed, 0);
    const double inv53 = 1.0 / 9007199254740992.0;
    const double twopi = 2.0 * PI;
    if (c < 2) {
        const double u1 = (double)((w[0] >> 11) + 1) * inv53;
        const double u2 = (double)(w[1] >> 11) * inv53;
        const double r1 = sqrt(-2.0 * log(u1));
        return c == 0 ? r1 * cos(twopi * u2) : r1 * sin(twopi * u2);
    }
    const double u3 = (double)((w[2] >> 11) + 1) * inv53;
    const double u4 = (double)(w[3] >> 11) * inv53;
    const double r2 = sqrt(-2.0 * log(u3));
    return r2 * cos(twopi * u4);
}

template <int FEAT, int G>
__global__ __launch_bounds__(BLOCK) void k_batch_step(BatchParams B) {
    static_assert(G >= 8 && G <= WAVE && (WAVE % G) == 0, "lane group");
    const StepParams &P = B.sp;
    // every (target type, partner type) row of the shared tables, and the work-item
    // prefix when it fits
    extern __shared__ double4 s_dyn[];
    const int nt1 = P.nttyp + 1, nrow = nt1 * nt1;
    DevPair *s_rows = reinterpret_cast<DevPair *>(s_dyn);
    double4 *s_tail = s_dyn + (size_t)nrow * (sizeof(DevPair) / sizeof(double4));
    int32_t *s_ws = reinterpret_cast<int32_t *>(s_tail + nrow);
    const bool ws_lds = B.nrep + 1 <= BATCH_LDS_WSTART;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid >> 6;
    const int sub = lane / G, sl = lane % G;
    for (int k = tid; k < nrow; k += BLOCK) {
        const DevPair d = P.rows[k];
        s_rows[k] = d;
        s_tail[k] = make_double4(d.sc3, d.fct, d.dm2hi, 0.0);
    }
    if (ws_lds)
        for (int k = tid; k <= B.nrep; k += BLOCK) s_ws[k] = B.wstart[k];
    __syncthreads();
    const int32_t *ws = ws_lds ? s_ws : B.wstart;
    const int W = ws[B.nrep];
    const uint64_t step = (uint64_t)B.ctl[4], seed = (uint64_t)B.ctl[3];
    const bool nonbond = (P.flags & QNONBOND) && (P.flags & QPAR);
    int next = (blockIdx.x * WPB + wv) * (WAVE / G) + sub;
    for (;;) {
        int w;
        if (B.persistent) {
            int v = 0;
            if (sl == 0) v = atomicAdd(&B.queue[0], 1);
            w = __shfl(v, sub * G, WAVE);
        } else {
            w = next;
            next = INT_MAX;
        }
        if (w >= W) break;
        // the replica of item w: the last k with wstart[k] <= w
        int lo = 0, hi = B.nrep;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (ws[mid] <= w) lo = mid;
            else hi = mid;
        }
        const int r = B.order[lo];
        const int off = B.off[r], n = B.ntot[r];
        const int t = B.nfix[r] + (w - ws[lo]);
        const double4 pa = B.pos[off + t];
        const int t2 = __double2loint(pa.w);
        const int itype = abs(B.typei[off + t]);
        const DevPair *rowp = s_rows + t2 * nt1;
        const double4 *tailp = s_tail + t2 * nt1;
        double fx = 0.0, fy = 0.0, fz = 0.0, et = 0.0, eel = 0.0, evd = 0.0, esr = 0.0;
        if (nonbond) {
            for (int p = sl; p < n; p += G) {
                if (p == t) continue;
                const double4 q = B.pos[off + p];
                const int tj = __double2loint(q.w);
                const double dx = pa.x - q.x, dy = pa.y - q.y, dz = pa.z - q.z;
                const double d2 = dx * dx + dy * dy + dz * dz;
                // an ion pair's energy is counted by its higher-index ion (energy.mono.f90:117)
                const bool ce = p < t;
                double y, de;
                if ((FEAT & BF_ANALYTIC) == 0 || (rowp[tj].flags & PF_EFPOT)) {
                    de = pair_near_far(rowp[tj], tailp[tj], P.ep, P.ires, d2, y);
                    et += ce ? y : 0.0;
                } else {
                    double q0 = 0.0, q1 = 0.0, q2 = 0.0;
                    de = pair_analytic(rowp[tj], P, d2, q0, q1, q2);
                    eel += ce ? q0 : 0.0;
                    evd += ce ? q1 : 0.0;
                    esr += ce ? q2 : 0.0;
                }
                fx += de * dx;
                fy += de * dy;
                fz += de * dz;
            }
        }
        // grids: lanes 0..7 of the group take the corners (staticf1.f90:45-118,
        // vdwgd1trln.f90:50-161)
        double egx = 0.0, egv = 0.0;
        bool warned = false;
        if constexpr ((FEAT & BF_PHIX) != 0) {
            const GridGeom &g = P.gx;
            const bool ok = pa.x <= g.xbcen + g.tranx && pa.x >= g.xbcen - g.tranx &&
                            pa.y <= g.ybcen + g.trany && pa.y >= g.ybcen - g.trany &&
                            pa.z <= g.zbcen + g.tranz && pa.z >= g.zbcen - g.tranz;
            if (ok && sl < 8) {
                const float *phix = B.phix[r];
                double gx_, gy_, gz_;
                int ix, iy, iz;
                grid_cell(g, pa.x, pa.y, pa.z, gx_, gy_, gz_, ix, iy, iz);
                const double chi = P.cg[itype];
                const double one1 = 1.0;
                const int n1 = ix + (sl >> 2), n2 = iy + ((sl >> 1) & 1), n3 = iz + (sl & 1);
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
                const double ph = (double)phix[((size_t)n1 * g.ny + n2) * g.nz + n3];
                const double prefac = ph * P.celec * chi * g.idcel;
                if (ai < (1.0 - RSMALL) && ai > RSMALL) fx += aisign * bi * ci * prefac;
                if (bi < (1.0 - RSMALL) && bi > RSMALL) fy += bisign * ai * ci * prefac;
                if (ci < (1.0 - RSMALL) && ci > RSMALL) fz += cisign * ai * bi * prefac;
                egx += fi * chi * ph * P.celec;
            }
        }
        if constexpr ((FEAT & BF_PHIV) != 0) {
            const GridGeom &g = P.gv;
            const bool ok = pa.x <= g.xbcen + g.tranx && pa.x >= g.xbcen - g.tranx &&
                            pa.y <= g.ybcen + g.trany && pa.y >= g.ybcen - g.trany &&
                            pa.z <= P.vzmax && pa.z >= P.vzmin;
            double ph = 0.0;
            if (ok && sl < 8) {
                double gx_, gy_, gz_;
                int ix, iy, iz;
                grid_cell(g, pa.x, pa.y, pa.z, gx_, gy_, gz_, ix, iy, iz);
                double esvdw = P.svdw;
                if (P.flags & QSVDW) esvdw = P.svdw * P.scal[t2];
                ph = tri_corner(g, P.phiv, gx_, gy_, gz_, ix, iy, iz, sl, esvdw, fx, fy, fz, egv);
            }
            double sum = ph;
            sum += __shfl_xor(sum, 1, WAVE);
            sum += __shfl_xor(sum, 2, WAVE);
            sum += __shfl_xor(sum, 4, WAVE);
            warned = ok && sum >= P.thold8;
        }
        // Gaussians of the update: component sl on lanes 0..2 of the group
        double gk = 0.0;
        if (B.mode == 1 && sl < 3) gk = philox_gauss1(seed, step, (uint64_t)r, t, sl);
        fx = group_sum<G>(fx);
        fy = group_sum<G>(fy);
        fz = group_sum<G>(fz);
        if (B.write_energy) {
            et = group_sum<G>(et);
            eel = group_sum<G>(eel);
            evd = group_sum<G>(evd);
            esr = group_sum<G>(esr);
            egx = group_sum<G>(egx);
            egv = group_sum<G>(egv);
        }
        const double g1 = __shfl(gk, sub * G + 1, WAVE), g2 = __shfl(gk, sub * G + 2, WAVE);
        if (sl == 0) {
            const int64_t i = off + t;
            if (warned) atomicAdd(&B.warn[(size_t)r * B.warn_stride + t2], 1);
            if (B.write_energy) {
                double *ep_ = B.epart + i;
                const int64_t cs = B.ecap;
                ep_[E_GSBPA * cs] = egx;
                ep_[E_VDWGD * cs] = egv;
                ep_[E_EFPOT * cs] = et;
                ep_[E_ELEC * cs] = eel;
                ep_[E_VDW * cs] = evd;
                ep_[E_SRPMF * cs] = esr;
            }
            if (B.write_forces) {
                B.fx[i] = fx;
                B.fy[i] = fy;
                B.fz[i] = fz;
            }
            if (B.mode == 1) {
                const double g[3] = {gk, g1, g2};
                const double F[3] = {fx, fy, fz};
                B.poso[i] = bd_move(P, false, itype, pa, F, g,
                                    B.nfwd + (size_t)r * B.flux_stride,
                                    B.nbwd + (size_t)r * B.flux_stride);
            }
        }
    }
}

template <int FEAT>
static void launch_batch_feat(const BatchParams &q, int lanes, dim3 grid, size_t lds,
                              hipStream_t s) {
    if (lanes == 16)
        hipLaunchKernelGGL((k_batch_step<FEAT, 16>), grid, dim3(BLOCK), lds, s, q);
    else if (lanes == 32)
        hipLaunchKernelGGL((k_batch_step<FEAT, 32>), grid, dim3(BLOCK), lds, s, q);
    else
        hipLaunchKernelGGL((k_batch_step<FEAT, 64>), grid, dim3(BLOCK), lds, s, q);
}

hipError_t launch_batch_step(const BatchParams &b, int blocks, int lanes, hipStream_t s) {
    int feat = 0;
    if (b.sp.flags & QPHIX) feat |= BF_PHIX;
    if (b.sp.flags & QPHIV) feat |= BF_PHIV;
    if (!b.sp.all_efpot) feat |= BF_ANALYTIC;
    if (lanes != 16 && lanes != 32) lanes = 64;
    BatchParams q = b;
    q.persistent = blocks > 0 ? 1 : 0;
    const int per_block = WPB * (WAVE / lanes);
    dim3 grid(blocks > 0 ? blocks : std::max(1, (b.sp.ntargets + per_block - 1) / per_block));
    const int nt1 = b.sp.nttyp + 1;
    size_t lds = (size_t)nt1 * nt1 * (sizeof(DevPair) + sizeof(double4));
    if (b.nrep + 1 <= BATCH_LDS_WSTART) lds += sizeof(int32_t) * (b.nrep + 1);
    switch (feat) {
        case 0: launch_batch_feat<0>(q, lanes, grid, lds, s); break;
        case 1: launch_batch_feat<1>(q, lanes, grid, lds, s); break;
        case 2: launch_batch_feat<2>(q, lanes, grid, lds, s); break;
        case 3: launch_batch_feat<3>(q, lanes, grid, lds, s); break;
        case 4: launch_batch_feat<4>(q, lanes, grid, lds, s); break;
        case 5: launch_batch_feat<5>(q, lanes, grid, lds, s); break;
        case 6: launch_batch_feat<6>(q, lanes, grid, lds, s); break;
        default: launch_batch_feat<7>(q, lanes, grid, lds, s); break;
    }
    return hipGetLastError();
}

// ---- GRAND of every replica: one workgroup per replica, attempts in order
// uniform k of replica r in GCMC cycle c: word k%4 of Philox(c, k/4, r, 1; seed, 0)
__device__ __forceinline__ double gcmc_uniform(uint64_t seed, uint64_t cycle, int r, uint64_t k) {
    uint64_t c[4] = {cycle, k >> 2, (uint64_t)r, 1};
    philox4x64(c, seed, 0);
    return (double)(c[k & 3] >> 11) * (1.0 / 9007199254740992.0);
}

// INTERACT of an ion of atnam type jtype at (xj, yj, zj) in replica r, as particle j (j ==
// ntot: an insertion trial): grids, then every other ion (interact.f90:155-204).  All
// threads of the block call it; the sum is returned to thread 0.
__device__ double batch_interact(const BatchParams &B, int r, int off, int ntot, double xj,
                                 double yj, double zj, int jtype, int j, bool alert,
                                 double *s_red) {
    const StepParams &P = B.sp;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    const int nt1 = P.nttyp + 1;
    const int jt2 = P.nwtype[jtype];
    double e = 0.0;
    if (tid < WAVE) {
        bool wall = false;
        const double eg = interact_grids(P, B.phix[r], true, P.cg[jtype], xj, yj, zj, jt2, wall);
        if (tid == 0) {
            e += eg;
            if (wall && alert) atomicAdd(&B.warn[(size_t)r * B.warn_stride + jt2], 1);
        }
    }
    if (P.flags & QNONBOND) {
        for (int i = tid; i < ntot; i += BLOCK) {
            if (i == j) continue;
            const double4 q = B.posw[off + i];
            const double dx = q.x - xj, dy = q.y - yj, dz = q.z - zj;
            const double d2 = dx * dx + dy * dy + dz * dz;
            const DevPair &pp = P.rows[jt2 * nt1 + __double2loint(q.w)];
            e += (pp.flags & PF_EFPOT) ? gety(pp, P.ep, P.ires, d2) : interact_analytic(pp, P, d2);
        }
    }
    e = wave_sum(e);
    if (lane == 0) s_red[tid >> 6] = e;
    __syncthreads();
    double t = 0.0;
    if (tid == 0)
        for (int k = 0; k < WPB; k++) t += s_red[k];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(BLOCK) void k_batch_gcmc(BatchParams B) {
    const StepParams &P = B.sp;
    const int r = blockIdx.x;
    const int nb = B.nbuf[r];
    if (nb <= 0 || B.ngcmc <= 0) return;
    __shared__ double s_red[WPB];
    __shared__ double s_d[4];
    __shared__ int s_i[8];
    __shared__ int s_cnt[WPB];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid >> 6;
    const int off = B.off[r], cap = B.cap[r], nf = B.nfix[r];
    const size_t gb = (size_t)r * B.gstride;
    const int32_t *ibfftyp = B.ibfftyp + gb, *ntotat = B.ntotat + gb;
    const double *avnum = B.avnum + gb, *mu = B.mu + gb, *kb = B.kb + gb;
    const double *lzmin = B.lzmin + gb, *lzmax = B.lzmax + gb;
    const int8_t *bias = B.bias + gb;
    int32_t *nat = B.nat + gb;
    int32_t *prob = B.prob + B.probofs[r];
    const double *bx = B.gbox + 6 * (size_t)r;
    const double lx = bx[0], ly = bx[1], cx = bx[2], cy = bx[3], cz = bx[4], ikbt = bx[5];
    const uint64_t seed = (uint64_t)B.ctl[3], cycle = (uint64_t)B.ctl[5];
    const int64_t icycle = (int64_t)cycle + 1;
    const double nbufferr = (double)nb;
    int ntot = B.ntot[r];
    uint64_t k = 0;       // uniforms drawn (thread 0)
    double ener = B.gener[r];
    for (int ig = 0; ig < B.ngcmc; ig++) {
        // ---- the trial (grand.f90:42-50, 77-84), drawn by thread 0
        if (tid == 0) {
            int mode = 0, ib = 0, ip = 0, itype = 0;
            double xn = 0.0, yn = 0.0, zn = 0.0;
            if (gcmc_uniform(seed, cycle, r, k++) < 0.5) {
                ib = (int)(nbufferr * gcmc_uniform(seed, cycle, r, k++)) + 1;
                if (ib <= nb && avnum[ib] > 0.0) {
                    mode = 1;
                    itype = ibfftyp[ib];
                    xn = lx * (gcmc_uniform(seed, cycle, r, k++) - 0.5) + cx;
                    yn = ly * (gcmc_uniform(seed, cycle, r, k++) - 0.5) + cy;
                    zn = lzmin[ib] + (lzmax[ib] - lzmin[ib]) * gcmc_uniform(seed, cycle, r, k++);
                }
            } else {
                ib = (int)(nbufferr * gcmc_uniform(seed, cycle, r, k++)) + 1;
                const double natib = ib <= nb ? (double)nat[ib] : 0.0;
                ip = (int)(natib * gcmc_uniform(seed, cycle, r, k++)) + 1;
                if (ib <= nb && ip <= nat[ib]) {
                    mode = 2;
                    itype = ibfftyp[ib];
                }
            }
            s_i[0] = mode; s_i[1] = ib; s_i[2] = ip; s_i[3] = itype; s_i[4] = -1;
            s_d[0] = xn; s_d[1] = yn; s_d[2] = zn;
        }
        __syncthreads();
        const int mode = s_i[0], ib = s_i[1], ip = s_i[2], itype = s_i[3];
        if (mode == 2) {
            // FIND (find.f90:26-32): the ip-th moving ion of buffer ib, in slot order
            int seen = 0;
            for (int c0 = nf; c0 < ntot && seen < ip; c0 += BLOCK) {
                const int a = c0 + tid;
                const bool hit = a < ntot && B.ibuffer[off + a] == ib;
                const uint64_t m = __ballot(hit);
                if (lane == 0) s_cnt[wv] = __popcll(m);
                __syncthreads();
                int before = seen;
                for (int q = 0; q < wv; q++) before += s_cnt[q];
                const int rank = before + __popcll(m & ((1ull << lane) - 1ull)) + 1;
                if (hit && rank == ip) s_i[4] = a;
                int tot = 0;
                for (int q = 0; q < WPB; q++) tot += s_cnt[q];
                __syncthreads();
                seen += tot;
            }
            __syncthreads();
            if (tid == 0 && s_i[4] >= 0) {
                const double4 q = B.posw[off + s_i[4]];
                s_d[0] = q.x; s_d[1] = q.y; s_d[2] = q.z;
            }
            __syncthreads();
            if (s_i[4] < 0) {   // nat out of step with ibuffer: stop this replica's GRAND
                if (tid == 0) atomicOr(&B.queue[1], 2);
                break;
            }
        }
        if (mode != 0) {
            const int j = mode == 1 ? ntot : s_i[4];
            const double dener = batch_interact(B, r, off, ntot, s_d[0], s_d[1], s_d[2], itype, j,
                                                mode == 2, s_red);
            if (tid == 0) {
                double rate;
                const double natv = mode == 1 ? (double)(nat[ib] + 1) : (double)nat[ib];
                if (bias[ib]) {
                    rate = (avnum[ib] + kb[ib] * (avnum[ib] - (double)(ntotat[ib] / icycle)));
                    rate = rate * exp(-(dener - mu[ib]) * ikbt) / natv;
                } else {
                    rate = mode == 1 ? (avnum[ib] / (nat[ib] + 1)) * exp(-(dener - mu[ib]) * ikbt)
                                     : (avnum[ib] / nat[ib]) * exp(-(dener - mu[ib]) * ikbt);
                }
                rate = mode == 1 ? rate / (1.0 + rate) : 1.0 / (1.0 + rate);
                int nt_new = ntot;
                if (gcmc_uniform(seed, cycle, r, k++) <= rate) {
                    if (mode == 1) {
                        if (ntot + 1 > cap) {
                            atomicOr(&B.queue[1], 1);
                            nt_new = -1;
                        } else {
                            ener = ener + dener;
                            nat[ib] = nat[ib] + 1;
                            const int a = ntot;
                            const int t2 = P.nwtype[itype];
                            B.posw[off + a] = make_double4(s_d[0], s_d[1], s_d[2],
                                                           __hiloint2double(0, t2));
                            B.typei[off + a] = s_d[2] < cz ? -itype : itype;
                            B.ibuffer[off + a] = ib;
                            nt_new = ntot + 1;
                        }
                    } else {
                        ener = ener - dener;
                        const int a = s_i[4], last = ntot - 1;
                        B.posw[off + a] = B.posw[off + last];
                        B.typei[off + a] = B.typei[off + last];
                        B.ibuffer[off + a] = B.ibuffer[off + last];
                        nat[ib] = nat[ib] - 1;
                        nt_new = ntot - 1;
                    }
                }
                s_i[5] = nt_new;
            }
            __syncthreads();
            if (s_i[5] < 0) break;   // capacity exceeded
            ntot = s_i[5];
        }
        if (tid == 0)
            for (int b = 1; b <= nb; b++) prob[(size_t)(b - 1) * (cap + 1) + nat[b]] += 1;
        __syncthreads();
    }
    if (tid == 0) {
        B.ntot[r] = ntot;
        B.gener[r] = ener;
    }
}

hipError_t launch_batch_gcmc(const BatchParams &b, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_gcmc, dim3(b.nrep), dim3(BLOCK), 0, s, b);
    return hipGetLastError();
}

// ---- energies of every replica: one block per replica, fixed-order tree per category
__global__ __launch_bounds__(BLOCK) void k_batch_reduce(BatchParams B, double *out) {
    __shared__ double s[BLOCK];
    const int r = blockIdx.x, tid = threadIdx.x;
    const int off = B.off[r], n = B.ntot[r];
    for (int cat = 0; cat < NE; cat++) {
        const double *src = B.epart + (size_t)cat * B.ecap + off;
        double acc = 0.0;
        for (int k = tid; k < n; k += BLOCK) acc += src[k];
        s[tid] = acc;
        __syncthreads();
        for (int h = BLOCK / 2; h > 0; h >>= 1) {
            if (tid < h) s[tid] += s[tid + h];
            __syncthreads();
        }
        if (tid == 0) out[(size_t)r * NE + cat] = s[0];
        __syncthreads();
    }
}

hipError_t launch_batch_reduce(const BatchParams &b, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_reduce, dim3(b.nrep), dim3(BLOCK), 0, s, b, out);
    return hipGetLastError();
}

}  // namespace bromoc

// ======================================================================= host side
using namespace bromoc;

namespace {

constexpr int BATCH_GRAPH_K = 32;   // steps per captured graph without GCMC

template <typename T>
hipError_t balloc(std::vector<void *> &al, T **p, size_t n) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return e;
    al.push_back(q);
    *p = (T *)q;
    return hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(T));
}

template <typename T>
hipError_t bupload(std::vector<void *> &al, T **p, const T *src, size_t n) {
    hipError_t e = balloc(al, p, n);
    if (e != hipSuccess || !src || n == 0) return e;
    return hipMemcpy(*p, src, n * sizeof(T), hipMemcpyHostToDevice);
}

double pack_t2(int32_t t2) {
    uint64_t bits = (uint64_t)(uint32_t)t2;
    double d;
    std::memcpy(&d, &bits, sizeof(d));
    return d;
}

}  // namespace

struct bromoc_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bromoc_ctx *shared = nullptr;
    bromoc_system sys{};
    int nrep = 0, gstride = 1, flux_stride = 1, warn_stride = 1;
    int64_t total = 0;
    std::vector<int32_t> h_off, h_cap, h_nfix, h_nbuf;
    std::vector<int64_t> h_probofs;
    std::vector<void *> allocs;
    BatchParams base{};
    double4 *pos[2] = {nullptr, nullptr};
    int cur = 0;
    int32_t *warn_scratch = nullptr;
    double *ered = nullptr;
    int blocks = -1;   // autotuned launch shape (0: one item per lane group, no queue)
    int lanes = 64;    // lanes per target
    int64_t h_ctl[8] = {};
    struct Graph {
        hipGraphExec_t exec = nullptr;
        int len = -1, gcmc = -1, ngcmc = -1, blocks = -1, lanes = -1;
    } graph[2];
};

namespace {

int bfail(bromoc_batch *b, int code, const std::string &m) {
    if (b) b->err = m;
    return code;
}

#define BCHK(b, expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return bfail((b), BROMOC_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

BatchParams params_now(bromoc_batch *b) {
    BatchParams p = b->base;
    p.pos = b->pos[b->cur];
    p.poso = b->pos[b->cur ^ 1];
    p.posw = b->pos[b->cur];
    return p;
}

// energy.mono.f90:83-110,689 for an ion-only system (no membrane, no reaction fields)
void batch_energies(const double *r, bromoc_energies *out) {
    bromoc_energies E;
    std::memset(&E, 0, sizeof(E));
    double *f = &E.eelec;
    for (int k = 0; k < NE; k++) f[k] = r[k];
    E.ener = E.ememb;
    E.ener = E.ener + E.egsbpa;
    E.ener = E.ener + E.evdwgd;
    E.ener = E.ener + E.eelec + E.evdw + E.esrpmf + E.esrpmfmx + E.ebond + E.eang + E.edihe +
             E.estack + E.ebp + E.eex + E.eqq + E.esolv + E.eqqmx + E.evdwmx + E.eefpot +
             E.eefpotmx + E.econ;
    *out = E;
}

hipError_t enqueue_batch_step(bromoc_batch *b, BatchParams p, hipStream_t s) {
    hipError_t e = launch_batch_prep(p, 0, s);
    if (e != hipSuccess) return e;
    return launch_batch_step(p, b->blocks, b->lanes, s);
}

hipError_t enqueue_batch_gcmc(const BatchParams &p, hipStream_t s) {
    hipError_t e = launch_batch_prep(p, 1, s);
    if (e != hipSuccess) return e;
    return launch_batch_gcmc(p, s);
}

int batch_sync(bromoc_batch *b) {
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return bfail(b, BROMOC_E_HIP, std::string("kernel: ") + hipGetErrorString(e));
    return BROMOC_OK;
}

// Autotuner: 16, 32 or 64 lanes per target, and a persistent queue with 2 or 8
// workgroups per CU or one work item per lane group; timed on ENERGY-only launches whose
// warnings go to a scratch counter.  BROMOC_BATCH_BLOCKS and BROMOC_BATCH_LANES force a
// shape (blocks 0: one item per lane group).
int batch_autotune(bromoc_batch *b) {
    if (b->blocks >= 0) return BROMOC_OK;
    const char *eb = std::getenv("BROMOC_BATCH_BLOCKS"), *el = std::getenv("BROMOC_BATCH_LANES");
    if (eb || el) {
        b->blocks = eb ? std::max(0, std::atoi(eb)) : 0;
        b->lanes = el ? std::atoi(el) : 64;
        if (b->lanes != 16 && b->lanes != 32) b->lanes = 64;
        return BROMOC_OK;
    }
    hipDeviceProp_t prop;
    BCHK(b, hipGetDeviceProperties(&prop, b->device));
    const int ncu = prop.multiProcessorCount;
    const int cb[3] = {0, 2 * ncu, 8 * ncu}, cl[3] = {16, 32, 64};
    hipEvent_t e0, e1;
    BCHK(b, hipEventCreate(&e0));
    BCHK(b, hipEventCreate(&e1));
    double best = 1e300;
    int pick_b = 0, pick_l = 64;
    for (int l : cl) {
        for (int c : cb) {
            b->blocks = c;
            b->lanes = l;
            BatchParams p = params_now(b);
            p.mode = 0;
            p.write_energy = 0;
            p.write_forces = 0;
            p.warn = b->warn_scratch;
            for (int r = 0; r < 2; r++) BCHK(b, enqueue_batch_step(b, p, b->stream));
            BCHK(b, hipEventRecord(e0, b->stream));
            for (int r = 0; r < 4; r++) BCHK(b, enqueue_batch_step(b, p, b->stream));
            BCHK(b, hipEventRecord(e1, b->stream));
            int st = batch_sync(b);
            if (st) return st;
            float ms = 0.f;
            BCHK(b, hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) { best = ms; pick_b = c; pick_l = l; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->blocks = pick_b;
    b->lanes = pick_l;
    return BROMOC_OK;
}

int batch_check_flags(bromoc_batch *b) {
    int32_t q[2] = {0, 0};
    BCHK(b, hipMemcpy(q, b->base.queue, sizeof(q), hipMemcpyDeviceToHost));
    if (q[1] & 1) return bfail(b, BROMOC_E_CAPACITY, "a replica exceeded its capacity in GRAND");
    if (q[1] & 2) return bfail(b, BROMOC_E_STATE, "nat does not match ibuffer in GRAND");
    return BROMOC_OK;
}

}  // namespace

extern "C" {

const char *bromoc_batch_last_error(const bromoc_batch *b) { return b ? b->err.c_str() : "null batch"; }

int bromoc_batch_destroy(bromoc_batch *b) {
    if (!b) return BROMOC_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto &g : b->graph)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (void *p : b->allocs) (void)hipFree(p);
    if (b->shared) bromoc_destroy(b->shared);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return BROMOC_OK;
}

int bromoc_batch_create(const bromoc_system *s, int32_t nrep, const bromoc_replica *reps,
                        int device, bromoc_batch **out) {
    if (!s || !reps || !out || nrep < 1 || nrep > (1 << 16)) return BROMOC_E_ARG;
    *out = nullptr;
    // the replica kernels implement pairs, the trilinear grids and DYNAMICS0/1/2 only
    const uint32_t unsupported = BROMOC_QNUCL | BROMOC_QRFPAR | BROMOC_QRFPSIN | BROMOC_QDNAFREE |
                                 BROMOC_QMEMB | BROMOC_QBSPLINE | BROMOC_QMMIJ | BROMOC_QNMCDEN |
                                 BROMOC_QSPHERE | BROMOC_QECYL;
    if (s->flags & unsupported) {
        fprintf(stderr, "bromoc_batch_create: replicas are ion-only boxes (no CG-DNA, RFPAR, "
                        "MMIJ, membrane, B-spline, per-species map or curved system)\n");
        return BROMOC_E_UNSUPPORTED;
    }
    if (s->nsites != 0 || !(s->flags & BROMOC_QPAR)) return BROMOC_E_ARG;
    const bool qphix = (s->flags & BROMOC_QPHIX) != 0;
    int64_t total = 0;
    int gmax = 0;
    for (int r = 0; r < nrep; r++) {
        const bromoc_replica &R = reps[r];
        if (R.ntot < 0 || R.nfix < 0 || R.nfix > R.ntot || R.capacity < R.ntot ||
            R.capacity < 1 || R.capacity > (1 << 22))
            return BROMOC_E_ARG;
        if (R.ntot > 0 && (!R.x || !R.y || !R.z || !R.typei)) return BROMOC_E_ARG;
        if (qphix && !R.phix) return BROMOC_E_ARG;
        for (int i = 0; i < R.ntot; i++) {
            const int t = std::abs(R.typei[i]);
            if (t <= s->nold || t > s->ntype) return BROMOC_E_ARG;
        }
        if (R.gcmc) {
            const bromoc_gcmc &G = *R.gcmc;
            if (G.nbuffer < 0 || G.nbuffer > 64) return BROMOC_E_ARG;
            if (G.nbuffer > 0 && (!G.ibfftyp || !G.avnum || !G.mu || !G.kb || !G.lzmin ||
                                  !G.lzmax || !G.bufferbias || !G.ntotat || !R.nat ||
                                  (R.ntot > 0 && !R.ibuffer)))
                return BROMOC_E_ARG;
            std::vector<int> cnt(G.nbuffer + 1, 0);
            for (int b = 1; b <= G.nbuffer; b++)
                if (G.ibfftyp[b] <= s->nold || G.ibfftyp[b] > s->ntype) return BROMOC_E_ARG;
            for (int i = 0; i < R.ntot; i++) {
                const int v = R.ibuffer[i];
                if (v < 0 || v > G.nbuffer || (v > 0 && i < R.nfix)) return BROMOC_E_ARG;
                cnt[v]++;
            }
            for (int b = 1; b <= G.nbuffer; b++)
                if (R.nat[b] != cnt[b]) return BROMOC_E_ARG;
            gmax = std::max(gmax, G.nbuffer);
        }
        total += R.capacity;
    }
    if (total > ((int64_t)1 << 27)) return BROMOC_E_ARG;
    // the shared tables through a particle-less single-system context
    bromoc_system sh = *s;
    sh.ntot = 0;
    sh.nfix = 0;
    sh.capacity = 1;
    const double zero = 0.0;
    const int32_t izero = 0;
    sh.x = sh.y = sh.z = &zero;
    sh.typei = &izero;
    if (qphix) sh.phix.data = reps[0].phix;
    bromoc_batch *b = new bromoc_batch();
    b->device = device;
    b->sys = *s;
    b->nrep = nrep;
    b->total = total;
    StepParams sp{};
    int st = shared_context(&sh, device, &b->shared, &sp);
    if (st) {
        delete b;
        return st;
    }
    auto cleanup = [&](int code) { bromoc_batch_destroy(b); return code; };
    if (hipSetDevice(device) != hipSuccess) return cleanup(BROMOC_E_HIP);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(BROMOC_E_HIP);
#define ACHK(expr) do { if ((expr) != hipSuccess) return cleanup(BROMOC_E_HIP); } while (0)
    auto &al = b->allocs;
    b->h_off.resize(nrep);
    b->h_cap.resize(nrep);
    b->h_nfix.resize(nrep);
    b->h_nbuf.resize(nrep);
    b->h_probofs.resize(nrep);
    std::vector<int32_t> ntot(nrep);
    int64_t o = 0, po = 0;
    b->gstride = gmax + 1;
    for (int r = 0; r < nrep; r++) {
        b->h_off[r] = (int32_t)o;
        b->h_cap[r] = reps[r].capacity;
        b->h_nfix[r] = reps[r].nfix;
        b->h_nbuf[r] = reps[r].gcmc ? reps[r].gcmc->nbuffer : 0;
        b->h_probofs[r] = po;
        ntot[r] = reps[r].ntot;
        o += reps[r].capacity;
        po += (int64_t)b->h_nbuf[r] * (reps[r].capacity + 1);
    }
    std::vector<int32_t> order(nrep);
    for (int r = 0; r < nrep; r++) order[r] = r;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return b->h_cap[a] > b->h_cap[c]; });
    // particles
    std::vector<double4> p4(total, make_double4(0.0, 0.0, 0.0, 0.0));
    std::vector<int32_t> ty(total, 0), ibuf(total, 0);
    for (int r = 0; r < nrep; r++) {
        const bromoc_replica &R = reps[r];
        for (int i = 0; i < R.ntot; i++) {
            const int64_t k = b->h_off[r] + i;
            const int t2 = s->nwtype[std::abs(R.typei[i])];
            p4[k] = make_double4(R.x[i], R.y[i], R.z[i], pack_t2(t2));
            ty[k] = R.typei[i];
            if (R.gcmc && R.ibuffer) ibuf[k] = R.ibuffer[i];
        }
    }
    for (int q = 0; q < 2; q++) ACHK(bupload(al, &b->pos[q], p4.data(), total));
    BatchParams &B = b->base;
    B.sp = sp;
    B.sp.ntargets = 0;
    for (int r = 0; r < nrep; r++) B.sp.ntargets += reps[r].capacity - reps[r].nfix;
    B.nrep = nrep;
    int32_t *d_i = nullptr;
    ACHK(bupload(al, &d_i, order.data(), nrep)); B.order = d_i;
    ACHK(bupload(al, &d_i, b->h_off.data(), nrep)); B.off = d_i;
    ACHK(bupload(al, &d_i, b->h_cap.data(), nrep)); B.cap = d_i;
    ACHK(bupload(al, &d_i, b->h_nfix.data(), nrep)); B.nfix = d_i;
    ACHK(bupload(al, &B.ntot, ntot.data(), nrep));
    ACHK(balloc(al, &B.wstart, nrep + 1));
    ACHK(balloc(al, &B.queue, 2));
    ACHK(bupload(al, &B.typei, ty.data(), total));
    ACHK(bupload(al, &B.ibuffer, ibuf.data(), total));
    ACHK(balloc(al, &B.fx, total));
    ACHK(balloc(al, &B.fy, total));
    ACHK(balloc(al, &B.fz, total));
    ACHK(balloc(al, &B.epart, (size_t)NE * total));
    B.ecap = total;
    // static fields: one allocation per replica, a device table of their addresses
    {
        std::vector<const float *> tab(nrep, nullptr);
        if (qphix) {
            const size_t n = (size_t)s->phix.nclx * s->phix.ncly * s->phix.nclz;
            for (int r = 0; r < nrep; r++) {
                float *d = nullptr;
                ACHK(bupload(al, &d, reps[r].phix, n));
                tab[r] = d;
            }
        }
        const float **d_tab = nullptr;
        ACHK(bupload(al, (const float ***)&d_tab, (const float **)tab.data(), nrep));
        B.phix = d_tab;
    }
    const int nt = std::max(1, s->ntype - s->nold), np = std::max(1, s->cntpts);
    b->flux_stride = nt * np;
    b->warn_stride = s->nttyp + 1;
    B.flux_stride = b->flux_stride;
    B.warn_stride = b->warn_stride;
    ACHK(balloc(al, &B.nfwd, (size_t)nrep * b->flux_stride));
    ACHK(balloc(al, &B.nbwd, (size_t)nrep * b->flux_stride));
    ACHK(balloc(al, &B.warn, (size_t)nrep * b->warn_stride));
    ACHK(balloc(al, &b->warn_scratch, (size_t)nrep * b->warn_stride));
    // GCMC buffers, [nrep][gstride] with entry 0 unused as in grandmod
    {
        const int gs = b->gstride;
        std::vector<int32_t> ibf((size_t)nrep * gs, 0), tot((size_t)nrep * gs, 0), nat((size_t)nrep * gs, 0);
        std::vector<double> av((size_t)nrep * gs, 0.0), mu_((size_t)nrep * gs, 0.0),
            kb_((size_t)nrep * gs, 0.0), zl((size_t)nrep * gs, 0.0), zh((size_t)nrep * gs, 0.0);
        std::vector<int8_t> bi((size_t)nrep * gs, 0);
        std::vector<double> box((size_t)nrep * 6, 0.0);
        for (int r = 0; r < nrep; r++) {
            const bromoc_gcmc *G = reps[r].gcmc;
            if (!G) continue;
            for (int q = 1; q <= G->nbuffer; q++) {
                const size_t k = (size_t)r * gs + q;
                ibf[k] = G->ibfftyp[q]; tot[k] = G->ntotat[q]; av[k] = G->avnum[q];
                mu_[k] = G->mu[q]; kb_[k] = G->kb[q]; zl[k] = G->lzmin[q]; zh[k] = G->lzmax[q];
                bi[k] = G->bufferbias[q]; nat[k] = reps[r].nat[q];
            }
            double *bx = box.data() + 6 * (size_t)r;
            bx[0] = G->lx; bx[1] = G->ly; bx[2] = G->cx; bx[3] = G->cy; bx[4] = G->cz; bx[5] = G->ikbt;
        }
        ACHK(bupload(al, &d_i, b->h_nbuf.data(), nrep)); B.nbuf = d_i;
        ACHK(bupload(al, &d_i, ibf.data(), ibf.size())); B.ibfftyp = d_i;
        ACHK(bupload(al, &d_i, tot.data(), tot.size())); B.ntotat = d_i;
        double *d_d = nullptr;
        ACHK(bupload(al, &d_d, av.data(), av.size())); B.avnum = d_d;
        ACHK(bupload(al, &d_d, mu_.data(), mu_.size())); B.mu = d_d;
        ACHK(bupload(al, &d_d, kb_.data(), kb_.size())); B.kb = d_d;
        ACHK(bupload(al, &d_d, zl.data(), zl.size())); B.lzmin = d_d;
        ACHK(bupload(al, &d_d, zh.data(), zh.size())); B.lzmax = d_d;
        int8_t *d_b = nullptr;
        ACHK(bupload(al, &d_b, bi.data(), bi.size())); B.bias = d_b;
        ACHK(bupload(al, &d_d, box.data(), box.size())); B.gbox = d_d;
        ACHK(bupload(al, &B.nat, nat.data(), nat.size()));
        B.gstride = gs;
        ACHK(balloc(al, &B.prob, (size_t)std::max<int64_t>(po, 1)));
        int64_t *d_l = nullptr;
        ACHK(bupload(al, &d_l, b->h_probofs.data(), nrep)); B.probofs = d_l;
        ACHK(balloc(al, &B.gener, nrep));
    }
    ACHK(balloc(al, &B.ctl, 8));
    ACHK(balloc(al, &b->ered, (size_t)nrep * NE));
#undef ACHK
    *out = b;
    return BROMOC_OK;
}

int bromoc_batch_run(bromoc_batch *b, int32_t nsteps, uint64_t seed, uint64_t counter0,
                     int32_t gcmc_every, int32_t ngcmc, uint64_t cycle0,
                     bromoc_energies *energies) {
    if (!b || nsteps < 0 || gcmc_every < 0 || ngcmc < 0) return BROMOC_E_ARG;
    (void)hipSetDevice(b->device);
    int st = batch_autotune(b);
    if (st) return st;
    b->h_ctl[0] = 0; b->h_ctl[1] = (int64_t)counter0; b->h_ctl[2] = 0; b->h_ctl[3] = (int64_t)seed;
    b->h_ctl[4] = 0; b->h_ctl[5] = 0; b->h_ctl[6] = (int64_t)cycle0; b->h_ctl[7] = 0;
    BCHK(b, hipMemcpyAsync(b->base.ctl, b->h_ctl, sizeof(b->h_ctl), hipMemcpyHostToDevice, b->stream));
    const int32_t qz[2] = {0, 0};
    BCHK(b, hipMemcpyAsync(b->base.queue, qz, sizeof(qz), hipMemcpyHostToDevice, b->stream));
    const bool gc = gcmc_every > 0 && ngcmc > 0;
    // one unit: [GRAND] + len steps; whole units replay a captured graph
    const int len = gc ? gcmc_every : BATCH_GRAPH_K;
    int done = 0;
    auto unit_direct = [&](int nst, bool with_gcmc, bool last_energy) -> int {
        if (with_gcmc) {
            BatchParams p = params_now(b);
            p.ngcmc = ngcmc;
            BCHK(b, enqueue_batch_gcmc(p, b->stream));
        }
        for (int k = 0; k < nst; k++) {
            BatchParams p = params_now(b);
            p.mode = 1;
            const bool wr = last_energy && k == nst - 1;
            p.write_energy = wr ? 1 : 0;
            p.write_forces = wr ? 1 : 0;
            BCHK(b, enqueue_batch_step(b, p, b->stream));
            b->cur ^= 1;
        }
        return BROMOC_OK;
    };
    static const bool graphs_on = [] {
        const char *e = std::getenv("BROMOC_GRAPH");
        return !(e && e[0] == '0');
    }();
    while (done < nsteps) {
        const int nst = std::min(len, nsteps - done);
        const bool last = done + nst >= nsteps;
        const bool with_gcmc = gc && (done % gcmc_every == 0);
        if (!graphs_on || nst < len || (last && energies)) {
            st = unit_direct(nst, with_gcmc, last && energies != nullptr);
            if (st) return st;
            done += nst;
            continue;
        }
        bromoc_batch::Graph &g = b->graph[b->cur];
        const int gflag = with_gcmc ? 1 : 0;
        if (!g.exec || g.len != len || g.gcmc != gflag || g.ngcmc != ngcmc ||
            g.blocks != b->blocks || g.lanes != b->lanes) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
            const int cur0 = b->cur;
            BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int cst = unit_direct(len, with_gcmc, false);
            hipGraph_t graph = nullptr;
            hipError_t ce = hipStreamEndCapture(b->stream, &graph);
            b->cur = cur0;
            if (cst) return cst;
            BCHK(b, ce);
            hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            BCHK(b, ie);
            g.len = len; g.gcmc = gflag; g.ngcmc = ngcmc; g.blocks = b->blocks;
            g.lanes = b->lanes;
        }
        BCHK(b, hipGraphLaunch(g.exec, b->stream));
        if (len & 1) b->cur ^= 1;
        done += nst;
    }
    st = batch_sync(b);
    if (st) return st;
    st = batch_check_flags(b);
    if (st) return st;
    if (energies) {
        BCHK(b, launch_batch_reduce(params_now(b), b->ered, b->stream));
        std::vector<double> r((size_t)b->nrep * NE);
        BCHK(b, hipMemcpyAsync(r.data(), b->ered, r.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        st = batch_sync(b);
        if (st) return st;
        for (int k = 0; k < b->nrep; k++) batch_energies(r.data() + (size_t)k * NE, &energies[k]);
    }
    return BROMOC_OK;
}

int bromoc_batch_energy(bromoc_batch *b, bromoc_energies *energies) {
    if (!b || !energies) return BROMOC_E_ARG;
    (void)hipSetDevice(b->device);
    int st = batch_autotune(b);
    if (st) return st;
    BatchParams p = params_now(b);
    p.mode = 0;
    p.write_energy = 1;
    p.write_forces = 1;
    BCHK(b, enqueue_batch_step(b, p, b->stream));
    BCHK(b, launch_batch_reduce(p, b->ered, b->stream));
    std::vector<double> r((size_t)b->nrep * NE);
    BCHK(b, hipMemcpyAsync(r.data(), b->ered, r.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    st = batch_sync(b);
    if (st) return st;
    for (int k = 0; k < b->nrep; k++) batch_energies(r.data() + (size_t)k * NE, &energies[k]);
    return BROMOC_OK;
}

int bromoc_batch_get(bromoc_batch *b, int32_t r, double *x, double *y, double *z, int32_t *typei,
                     int32_t *ntot, int32_t *ibuffer, int32_t *nat, double *gcmc_ener) {
    if (!b || r < 0 || r >= b->nrep) return BROMOC_E_ARG;
    (void)hipSetDevice(b->device);
    int32_t n = 0;
    BCHK(b, hipMemcpy(&n, b->base.ntot + r, sizeof(n), hipMemcpyDeviceToHost));
    const int64_t off = b->h_off[r];
    if (x || y || z) {
        std::vector<double4> p4(std::max(n, 1));
        if (n) BCHK(b, hipMemcpy(p4.data(), b->pos[b->cur] + off, sizeof(double4) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; i++) {
            if (x) x[i] = p4[i].x;
            if (y) y[i] = p4[i].y;
            if (z) z[i] = p4[i].z;
        }
    }
    if (typei && n) BCHK(b, hipMemcpy(typei, b->base.typei + off, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (ibuffer && n) BCHK(b, hipMemcpy(ibuffer, b->base.ibuffer + off, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (nat) {
        nat[0] = 0;
        if (b->h_nbuf[r] > 0)
            BCHK(b, hipMemcpy(nat + 1, b->base.nat + (size_t)r * b->gstride + 1,
                              sizeof(int32_t) * b->h_nbuf[r], hipMemcpyDeviceToHost));
    }
    if (gcmc_ener) BCHK(b, hipMemcpy(gcmc_ener, b->base.gener + r, sizeof(double), hipMemcpyDeviceToHost));
    if (ntot) *ntot = n;
    return BROMOC_OK;
}

int bromoc_batch_get_forces(bromoc_batch *b, int32_t r, double *fx, double *fy, double *fz) {
    if (!b || r < 0 || r >= b->nrep || !fx || !fy || !fz) return BROMOC_E_ARG;
    (void)hipSetDevice(b->device);
    int32_t n = 0;
    BCHK(b, hipMemcpy(&n, b->base.ntot + r, sizeof(n), hipMemcpyDeviceToHost));
    const int64_t off = b->h_off[r];
    if (n) {
        BCHK(b, hipMemcpy(fx, b->base.fx + off, sizeof(double) * n, hipMemcpyDeviceToHost));
        BCHK(b, hipMemcpy(fy, b->base.fy + off, sizeof(double) * n, hipMemcpyDeviceToHost));
        BCHK(b, hipMemcpy(fz, b->base.fz + off, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
    return BROMOC_OK;
}

int bromoc_batch_get_flux(bromoc_batch *b, int32_t r, int32_t *nforward, int32_t *nbackward) {
    if (!b || r < 0 || r >= b->nrep || !nforward || !nbackward) return BROMOC_E_ARG;
    (void)hipSetDevice(b->device);
    const int nt = b->sys.ntype - b->sys.nold, np = b->sys.cntpts;
    if (nt <= 0 || np <= 0) return BROMOC_OK;
    const size_t o = (size_t)r * b->flux_stride;
    BCHK(b, hipMemcpy(nforward, b->base.nfwd + o, sizeof(int32_t) * nt * np, hipMemcpyDeviceToHost));
    BCHK(b, hipMemcpy(nbackward, b->base.nbwd + o, sizeof(int32_t) * nt * np, hipMemcpyDeviceToHost));
    return BROMOC_OK;
}

int bromoc_batch_get_prob(bromoc_batch *b, int32_t r, int32_t *prob) {
    if (!b || r < 0 || r >= b->nrep || !prob) return BROMOC_E_ARG;
    (void)hipSetDevice(b->device);
    const size_t n = (size_t)b->h_nbuf[r] * (b->h_cap[r] + 1);
    if (n) BCHK(b, hipMemcpy(prob, b->base.prob + b->h_probofs[r], sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    return BROMOC_OK;
}

int bromoc_batch_time_steps(bromoc_batch *b, int32_t nsteps, uint64_t seed, uint64_t counter0,
                            double *total_ms, double *kernel_ms) {
    if (!b || nsteps <= 0 || !total_ms || !kernel_ms) return BROMOC_E_ARG;
    (void)hipSetDevice(b->device);
    int st = batch_autotune(b);
    if (st) return st;
    b->h_ctl[0] = 0; b->h_ctl[1] = (int64_t)counter0; b->h_ctl[3] = (int64_t)seed;
    BCHK(b, hipMemcpyAsync(b->base.ctl, b->h_ctl, sizeof(b->h_ctl), hipMemcpyHostToDevice, b->stream));
    std::vector<hipEvent_t> ev(2 * nsteps);
    for (auto &e : ev) BCHK(b, hipEventCreate(&e));
    for (int k = 0; k < nsteps; k++) {
        BatchParams p = params_now(b);
        p.mode = 1;
        BCHK(b, launch_batch_prep(p, 0, b->stream));
        BCHK(b, hipEventRecord(ev[2 * k], b->stream));
        BCHK(b, launch_batch_step(p, b->blocks, b->lanes, b->stream));
        BCHK(b, hipEventRecord(ev[2 * k + 1], b->stream));
        b->cur ^= 1;
    }
    st = batch_sync(b);
    if (st) return st;
    double sum = 0.0;
    for (int k = 0; k < nsteps; k++) {
        float ms = 0.f;
        BCHK(b, hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
        sum += ms;
    }
    float all = 0.f;
    BCHK(b, hipEventElapsedTime(&all, ev[0], ev[2 * nsteps - 1]));
    for (auto &e : ev) (void)hipEventDestroy(e);
    *total_ms = all;
    *kernel_ms = sum / nsteps;
    return BROMOC_OK;
}

int bromoc_batch_get_tuning(bromoc_batch *b, int32_t *blocks, int32_t *lanes) {
    if (!b || !blocks || !lanes) return BROMOC_E_ARG;
    (void)hipSetDevice(b->device);
    int st = batch_autotune(b);
    if (st) return st;
    *blocks = b->blocks;
    *lanes = b->lanes;
    return BROMOC_OK;
}

}  // extern "C"
