// adjoint.hip -- reverse-mode derivative of a Brownian-dynamics trajectory with respect to
// the tabulated pair potentials, for a loss on the trajectory's smoothed radial
// distribution functions.  No counterpart in the reference (its potentials are fitted by
// the separate inverse-Monte-Carlo program); the forward model is the reference's BD step
// (ENERGY pair forces through getyd, polynom.f90:197-230, and the trilinear grids of STATICF1
// and VDWGD1TRLN, whose Jacobian J enters as J^T mu and J dx; DYNAMICS0, dynamics0.mono.f90:
// 31-46 with its bdmax clamp; FIXCOOR's box clamp, fixcoor.f90:49-64) with injected noise.
//
// Step map x' = fixcoor(x + clamp(F(x) f1) + f2 xi).  With D_box, D_bd the 0/1 masks of
// the components the two clamps leave alone, the adjoint of one step is
//   mu      = f1 * D_bd * D_box * lambda'
//   lambda  = D_box * lambda' + H(x) mu + dl/dx           (H = dF/dx, symmetric)
//   dL/dth += (dF/dth)^T mu
// For a pair with separation d = x_i - x_j, r = |d|, F_i = phi(r) d, so
//   (H mu)_i   = sum_j phi (mu_i - mu_j) + (phi'/r) ((mu_i - mu_j).d) d
//   dL/dth    += sum_pairs (dphi/dth) (mu_i - mu_j).d
// with phi = -ep2/r - 2 ep3 inside the table (knot k), 12 sc1/r^14 below it and
// 6 sc3/r^8 + fct/r^3 above it.
//
// The tangent-linear map of the same step (k_tlm) carries a parameter direction v forward:
//   dx' = D_box dx + f1 D_bd D_box (H(x) dx + (dF/dth) v)
// and the derivative of every RDF histogram bin; <dL/dth, v> from the reverse pass must
// equal the tangent's dL for every v (the two are transposes of one linear map).
//
// Every sum a kernel gathers for one particle runs in a fixed order; sums that many
// threads add into one place (histogram bins, knot and pair-type gradients) are integer
// fixed-point sums (fix_add), so the results are bit-reproducible run to run.
#include <hip/hip_runtime.h>

#include "device.h"

namespace bromoc {

// region of a table pair at squared distance d2 (getyd, polynom.f90:197-230): phi and
// dphi/dr, and which parameters it depends on (kind 0 wall: sc1, 1 table: knot k, 2 tail)
__device__ __forceinline__ void pair_phi(const DevPair &p, const double4 *__restrict__ ep,
                                         double ires, double d2, double r, double &phi,
                                         double &dphi, int &kind, int &k) {
    const double ir = 1.0 / r, ix2 = ir * ir;
    if (d2 < p.dm2lo) {
        const double c = p.sc1 * pow6(ix2);
        phi = 12.0 * c * ix2;
        dphi = -14.0 * phi * ir;
        kind = 0;
        k = -1;
    } else if (d2 > p.dm2hi) {
        const double c = p.sc3 * pow3(ix2);
        const double f = p.fct * ir;          // zero for uncharged pairs
        phi = (6.0 * c + f) * ix2;
        dphi = (-48.0 * c - 3.0 * f) * ix2 * ir;
        kind = 2;
        k = -1;
    } else {
        k = (int)((r - p.dmi) * ires) + p.nxi;
        k = min(max(k, p.nxi), p.nlast);
        const double4 e = ep[k];
        phi = -e.y * ir - 2.0 * e.z;
        dphi = e.y * ix2;
        kind = 1;
    }
}

// smoothed RDF kernel exp(-u^2 / (2 sigma^2)) within |u| <= cut
__device__ __forceinline__ double rdf_k(double u, double sigma) {
    return exp(-0.5 * u * u / (sigma * sigma));
}

// k_rdf: raw smoothed histograms of the pairs each moving ion heads (a moving partner of
// higher index, or any fixed partner), added to the ion's own row.
__global__ __launch_bounds__(256) void k_rdf(AdjParams A) {
    extern __shared__ unsigned long long s_h[];   // [nh][2] fixed-point bins
    const StepParams &P = A.sp;
    const int nf0 = P.nsites + P.nfix;
    const int i = nf0 + blockIdx.x;
    const int tid = threadIdx.x;
    const int nh = A.ngroup * A.nbins, nt1 = P.nttyp + 1;
    for (int k = tid; k < 2 * nh; k += blockDim.x) s_h[k] = 0ull;
    __syncthreads();
    const double4 xi = A.x[i];
    const int ti = __double2loint(xi.w);
    for (int j = P.nsites + tid; j < P.ntot; j += blockDim.x) {
        if (j == i || (j >= nf0 && j < i)) continue;
        const double4 xj = A.x[j];
        const int g = A.group[ti * nt1 + __double2loint(xj.w)];
        if (g < 0) continue;
        const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const int b0 = max(0, (int)ceil((r - A.cut - A.r0) / A.dr - 0.5));
        const int b1 = min(A.nbins - 1, (int)floor((r + A.cut - A.r0) / A.dr - 0.5));
        for (int b = b0; b <= b1; b++) {
            const double u = r - (A.r0 + (b + 0.5) * A.dr);
            if (fabs(u) <= A.cut) fix_add<2>(&s_h[2 * (g * A.nbins + b)], rdf_k(u, A.sigma));
        }
    }
    __syncthreads();
    double *row = A.hist + (size_t)i * nh;
    for (int k = tid; k < nh; k += blockDim.x) row[k] += fix_value(&s_h[2 * k], 2);
}

// k_adj_force: F(x_s) of each moving ion, pairs in index order
__global__ __launch_bounds__(256) void k_adj_force(AdjParams A) {
    __shared__ double s_r[3][4];
    const StepParams &P = A.sp;
    const int i = P.nsites + P.nfix + blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nt1 = P.nttyp + 1;
    const double4 xi = A.x[i];
    const int ti = __double2loint(xi.w);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int j = tid; j < P.ntot; j += blockDim.x) {
        if (j == i) continue;
        const double4 xj = A.x[j];
        const DevPair &p = P.rows[ti * nt1 + __double2loint(xj.w)];
        const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const double d2 = dx * dx + dy * dy + dz * dz;
        double y, yp;
        getyd(p, P.ep, P.ires, d2, y, yp);
        fx += yp * dx;
        fy += yp * dy;
        fz += yp * dz;
    }
    fx = wave_sum(fx); fy = wave_sum(fy); fz = wave_sum(fz);
    if (lane == 0) { s_r[0][wv] = fx; s_r[1][wv] = fy; s_r[2][wv] = fz; }
    __syncthreads();
    if (tid == 0) {
        double G[3] = {0.0, 0.0, 0.0}, J[6];
        if (P.flags & (QPHIX | QPHIV)) grid_force_jac(P, i, xi.x, xi.y, xi.z, ti, G, J);
        A.f[i] = make_double4(s_r[0][0] + s_r[0][1] + s_r[0][2] + s_r[0][3] + G[0],
                              s_r[1][0] + s_r[1][1] + s_r[1][2] + s_r[1][3] + G[1],
                              s_r[2][0] + s_r[2][1] + s_r[2][2] + s_r[2][3] + G[2], 0.0);
    }
}

// k_adj_mu: the clamp masks of DYNAMICS0 + FIXCOOR at x_s and mu = f1 D_bd D_box lambda'
__global__ void k_adj_mu(AdjParams A) {
    const StepParams &P = A.sp;
    const int i = P.nsites + P.nfix + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.ntot) return;
    const int itype = abs(P.typei[i]);
    const double f1 = P.fact1a[itype], f2 = P.fact2a[itype];
    const double4 x = A.x[i], F = A.f[i], l = A.lam1[i];
    const double xv[3] = {x.x, x.y, x.z}, fv[3] = {F.x, F.y, F.z}, lv[3] = {l.x, l.y, l.z};
    const double lo[3] = {P.lx2m, P.ly2m, P.lz2m}, hi[3] = {P.lx2p, P.ly2p, P.lz2p};
    double m[3], b[3];
    for (int c = 0; c < 3; c++) {
        double del = fv[c] * f1;
        const bool free_bd = !(fabs(del) > P.bdmax);
        if (!free_bd) del = fsign(P.bdmax, del);
        const double xt = xv[c] + del + f2 * A.xi[3 * (size_t)i + c];
        const bool free_box = xt >= lo[c] && xt <= hi[c];
        b[c] = free_box ? 1.0 : 0.0;
        m[c] = (free_bd && free_box) ? f1 * lv[c] : 0.0;
    }
    A.mu[i] = make_double4(m[0], m[1], m[2], 0.0);
    A.dbox[i] = make_double4(b[0], b[1], b[2], 0.0);
}

// k_adj_pairs: lambda_s of each moving ion (Hessian-vector product, box mask, RDF-loss
// term) and the parameter gradient of the pairs it heads (a moving partner of lower index,
// or a fixed partner), one block per moving ion
__global__ __launch_bounds__(256) void k_adj_pairs(AdjParams A) {
    __shared__ double s_r[6][4];
    const StepParams &P = A.sp;
    const int nf0 = P.nsites + P.nfix;
    const int i = nf0 + blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nt1 = P.nttyp + 1;
    __shared__ unsigned long long s_sc[MAX_LDS_TYPES][2][ADJ_W];   // wall / tail by partner type
    const double4 xi = A.x[i], mi = A.mu[i];
    const int ti = __double2loint(xi.w);
    for (int k = tid; k < MAX_LDS_TYPES * 2 * ADJ_W; k += blockDim.x) (&s_sc[0][0][0])[k] = 0ull;
    __syncthreads();
    double hx = 0.0, hy = 0.0, hz = 0.0, ox = 0.0, oy = 0.0, oz = 0.0;
    const double is2 = 1.0 / (A.sigma * A.sigma);
    for (int j = tid; j < P.ntot; j += blockDim.x) {
        if (j == i) continue;
        const double4 xj = A.x[j];
        const int tj = __double2loint(xj.w);
        const bool jmov = j >= nf0;
        const double4 mj = jmov ? A.mu[j] : make_double4(0.0, 0.0, 0.0, 0.0);
        const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const double d2 = dx * dx + dy * dy + dz * dz;
        const double r = sqrt(d2);
        const DevPair &p = P.rows[ti * nt1 + tj];
        double phi, dphi;
        int kind, k;
        pair_phi(p, P.ep, P.ires, d2, r, phi, dphi, kind, k);
        const double ux = mi.x - mj.x, uy = mi.y - mj.y, uz = mi.z - mj.z;
        const double s = ux * dx + uy * dy + uz * dz;
        const double q = dphi / r * s;
        hx += phi * ux + q * dx;
        hy += phi * uy + q * dy;
        hz += phi * uz + q * dz;
        // parameter gradient, once per pair
        if (!jmov || j < i) {
            if (kind == 1) {
                fix_add<ADJ_W>(&A.gep[2 * ADJ_W * (size_t)k], -s / r);
                fix_add<ADJ_W>(&A.gep[2 * ADJ_W * (size_t)k + ADJ_W], -2.0 * s);
            } else {
                // wall and tail terms of the block's pairs gather by partner type in LDS
                // first: one global sum per pair type and block instead of one per pair
                const double ix2 = 1.0 / d2;
                if (kind == 0) fix_add<ADJ_W>(&s_sc[tj][0][0], 12.0 * pow6(ix2) * ix2 * s);
                else fix_add<ADJ_W>(&s_sc[tj][1][0], 6.0 * pow3(ix2) * ix2 * s);
            }
        }
        // RDF-loss term: d/dx_i of the group's smoothed histogram, weighted
        if (A.obs) {
            const int g = A.group[ti * nt1 + tj];
            if (g >= 0) {
                const int b0 = max(0, (int)ceil((r - A.cut - A.r0) / A.dr - 0.5));
                const int b1 = min(A.nbins - 1, (int)floor((r + A.cut - A.r0) / A.dr - 0.5));
                double om = 0.0;
                for (int b = b0; b <= b1; b++) {
                    const double u = r - (A.r0 + (b + 0.5) * A.dr);
                    if (fabs(u) <= A.cut) om += A.w[g * A.nbins + b] * (-u * is2) * rdf_k(u, A.sigma);
                }
                const double c = om / r;
                ox += c * dx;
                oy += c * dy;
                oz += c * dz;
            }
        }
    }
    double v[6] = {hx, hy, hz, ox, oy, oz};
    for (int c = 0; c < 6; c++) {
        v[c] = wave_sum(v[c]);
        if (lane == 0) s_r[c][wv] = v[c];
    }
    __syncthreads();
    if (tid < nt1 * 2) {
        // this block's wall (q = 0) and tail (q = 1) sums of partner type tj, into the pair
        // type's fixed-point gradient (integer sums: the order of blocks does not matter)
        const int tj = tid >> 1, q = tid & 1;
        unsigned long long v[ADJ_W];
        bool any = false;
        for (int w = 0; w < ADJ_W; w++) { v[w] = s_sc[tj][q][w]; any |= v[w] != 0ull; }
        if (any) {
            fix_carry<ADJ_W>(v);
            const int is = P.pairidx[ti * nt1 + tj];
            for (int w = 0; w < ADJ_W; w++)
                atomicAdd(&A.gsc[ADJ_W * (2 * (size_t)is + q) + w], v[w]);
        }
    }
    if (tid == 0) {
        double t[6];
        for (int c = 0; c < 6; c++) t[c] = s_r[c][0] + s_r[c][1] + s_r[c][2] + s_r[c][3];
        const double4 l1 = A.lam1[i], db = A.dbox[i];
        // flux loss at the last state: d(charge displacement)/dz_i = q_i
        const double zf = A.zflux != 0.0 ? A.zflux * P.cg[abs(P.typei[i])] : 0.0;
        // the grids' share: J^T mu (J = dF/dx of STATICF1 and VDWGD1TRLN at x_i)
        double g[3] = {0.0, 0.0, 0.0};
        if (P.flags & (QPHIX | QPHIV)) {
            double G[3], J[6];
            grid_force_jac(P, i, xi.x, xi.y, xi.z, ti, G, J);
            g[0] = J[2] * mi.y + J[4] * mi.z;
            g[1] = J[0] * mi.x + J[5] * mi.z;
            g[2] = J[1] * mi.x + J[3] * mi.y;
        }
        A.lam0[i] = make_double4(db.x * l1.x + t[0] + t[3] + g[0], db.y * l1.y + t[1] + t[4] + g[1],
                                 db.z * l1.z + t[2] + t[5] + g[2] + zf, 0.0);
    }
}

// k_tlm: the tangent-linear step of moving ion i.  Gathers dF_i = (H dx)_i + (dF/dth v)_i
// over every partner, and the derivative of the smoothed histograms along dx for the
// pairs ion i heads (as k_rdf counts them) when obs is set; writes dx' = D_box dx +
// mu (.) dF, with mu = f1 D_bd D_box from k_adj_mu (run with lambda' = 1).
__global__ __launch_bounds__(256) void k_tlm(AdjParams A) {
    extern __shared__ unsigned long long s_dh[];   // [nh][ADJ_W]
    __shared__ double s_r[3][4];
    const StepParams &P = A.sp;
    const int nf0 = P.nsites + P.nfix;
    const int i = nf0 + blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nt1 = P.nttyp + 1, nh = A.ngroup * A.nbins;
    if (A.obs)
        for (int k = tid; k < ADJ_W * nh; k += blockDim.x) s_dh[k] = 0ull;
    __syncthreads();
    const double4 xi = A.x[i], di = A.dx[i];
    const int ti = __double2loint(xi.w);
    const double is2 = 1.0 / (A.sigma * A.sigma);
    double hx = 0.0, hy = 0.0, hz = 0.0;
    for (int j = tid; j < P.ntot; j += blockDim.x) {
        if (j == i) continue;
        const double4 xj = A.x[j];
        const int tj = __double2loint(xj.w);
        const bool jmov = j >= nf0;
        const double4 dj = jmov ? A.dx[j] : make_double4(0.0, 0.0, 0.0, 0.0);
        const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const double d2 = dx * dx + dy * dy + dz * dz;
        const double r = sqrt(d2);
        const DevPair &p = P.rows[ti * nt1 + tj];
        double phi, dphi;
        int kind, k;
        pair_phi(p, P.ep, P.ires, d2, r, phi, dphi, kind, k);
        // the parameters' share: d phi = (d phi / d th) . v
        double vphi;
        if (kind == 1) {
            vphi = -A.vep[2 * (size_t)k] / r - 2.0 * A.vep[2 * (size_t)k + 1];
        } else {
            const int is = P.pairidx[ti * nt1 + tj];
            const double ix2 = 1.0 / d2;
            vphi = kind == 0 ? 12.0 * pow6(ix2) * ix2 * A.vsc[2 * is]
                             : 6.0 * pow3(ix2) * ix2 * A.vsc[2 * is + 1];
        }
        const double ux = di.x - dj.x, uy = di.y - dj.y, uz = di.z - dj.z;
        const double s = ux * dx + uy * dy + uz * dz;
        const double q = dphi / r * s;
        hx += phi * ux + q * dx + vphi * dx;
        hy += phi * uy + q * dy + vphi * dy;
        hz += phi * uz + q * dz + vphi * dz;
        if (A.obs && !(jmov && j < i)) {
            const int g = A.group[ti * nt1 + tj];
            if (g >= 0) {
                const int b0 = max(0, (int)ceil((r - A.cut - A.r0) / A.dr - 0.5));
                const int b1 = min(A.nbins - 1, (int)floor((r + A.cut - A.r0) / A.dr - 0.5));
                const double sr = s / r;
                for (int b = b0; b <= b1; b++) {
                    const double u = r - (A.r0 + (b + 0.5) * A.dr);
                    if (fabs(u) <= A.cut)
                        fix_add<ADJ_W>(&s_dh[ADJ_W * (g * A.nbins + b)],
                                       -u * is2 * rdf_k(u, A.sigma) * sr);
                }
            }
        }
    }
    hx = wave_sum(hx); hy = wave_sum(hy); hz = wave_sum(hz);
    if (lane == 0) { s_r[0][wv] = hx; s_r[1][wv] = hy; s_r[2][wv] = hz; }
    __syncthreads();
    if (A.obs) {
        for (int k = tid; k < nh; k += blockDim.x) {
            unsigned long long v[ADJ_W];
            bool any = false;
            for (int w = 0; w < ADJ_W; w++) { v[w] = s_dh[ADJ_W * k + w]; any |= v[w] != 0ull; }
            if (any) {
                fix_carry<ADJ_W>(v);
                for (int w = 0; w < ADJ_W; w++) atomicAdd(&A.dh[ADJ_W * (size_t)k + w], v[w]);
            }
        }
    }
    if (tid == 0) {
        double fx = s_r[0][0] + s_r[0][1] + s_r[0][2] + s_r[0][3];
        double fy = s_r[1][0] + s_r[1][1] + s_r[1][2] + s_r[1][3];
        double fz = s_r[2][0] + s_r[2][1] + s_r[2][2] + s_r[2][3];
        if (P.flags & (QPHIX | QPHIV)) {   // the grids' share: J dx (J = dF/dx at x_i)
            double G[3], J[6];
            grid_force_jac(P, i, xi.x, xi.y, xi.z, ti, G, J);
            fx += J[0] * di.y + J[1] * di.z;
            fy += J[2] * di.x + J[3] * di.z;
            fz += J[4] * di.x + J[5] * di.y;
        }
        const double4 m = A.mu[i], db = A.dbox[i];
        A.dxo[i] = make_double4(db.x * di.x + m.x * fx, db.y * di.y + m.y * fy,
                                db.z * di.z + m.z * fz, 0.0);
    }
}

hipError_t launch_adj(int which, const AdjParams &a, hipStream_t s) {
    const int nmov = a.sp.ntot - a.sp.nsites - a.sp.nfix;
    if (nmov <= 0) return hipSuccess;
    switch (which) {
        case 0: {
            const size_t lds = 2 * sizeof(unsigned long long) * (size_t)a.ngroup * a.nbins;
            hipLaunchKernelGGL(k_rdf, dim3(nmov), dim3(256), lds, s, a);
            break;
        }
        case 1: hipLaunchKernelGGL(k_adj_force, dim3(nmov), dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_adj_mu, dim3((nmov + 255) / 256), dim3(256), 0, s, a); break;
        case 3: hipLaunchKernelGGL(k_adj_pairs, dim3(nmov), dim3(256), 0, s, a); break;
        default: {
            const size_t lds = a.obs ? ADJ_W * sizeof(unsigned long long) * (size_t)a.ngroup * a.nbins : 0;
            hipLaunchKernelGGL(k_tlm, dim3(nmov), dim3(256), lds, s, a);
            break;
        }
    }
    return hipGetLastError();
}

}  // namespace bromoc
