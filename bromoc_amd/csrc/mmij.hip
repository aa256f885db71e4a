// mmij.hip -- the MMIJ reaction field of a rectangular region (RECTBOX; rect_rf1.f90 for
// ENERGY, rect_rf0.f90 for INTERACT).  The field is a quadratic form in ntpol coefficients
// q_m = sum over the moving ions inside the region of charge * bnorm_m * P_lx(x xscal)
// P_ly(y yscal) P_lz(z zscal):  egsbpb = celec/2 q.(M q rfscal).  Per step:
//   k_mm_leg    one thread per ion: its Legendre values (and charge, 0 outside)
//   k_mm_coef   one workgroup per basis function: q_m, fixed-order tree over the ions
//   k_mm_mq     one wavefront per row: (M q rfscal)_m and the row's energy term
//   k_mm_energy one workgroup: the energy (only when energies are written)
//   k_mm_force  one wavefront per ion: -charge celec sum_m grad b_m (M q rfscal)_m into frf
// Every sum runs in a fixed order: results are deterministic.
#include "device.h"

namespace bromoc {

namespace {

// P_0..P_{n-1}(xs) by the three-term recurrence as rect_rf1.f90:48-63 writes it (n >= 3)
__device__ inline void legendre(double xs, int n, double *lp) {
#pragma clang fp contract(off)
    lp[0] = 1.0;
    lp[1] = xs;
    lp[2] = 0.5 * (3.0 * xs * xs - 1.0);
    for (int k = 3; k <= n - 1; k++)
        lp[k] = ((2.0 * k - 1) * xs * lp[k - 1] - (k - 1.0) * lp[k - 2]) / k;
}

// the region |x| <= 1/xscal, ... (shell_simul.f90:2793-2798), bounds as the reference forms them
__device__ inline bool mm_inside(const StepParams &P, double x, double y, double z) {
    const double xmax = 1.0 / P.xscal, ymax = 1.0 / P.yscal, zmax = 1.0 / P.zscal;
    const double xmin = -1.0 / P.xscal, ymin = -1.0 / P.yscal, zmin = -1.0 / P.zscal;
    return x >= xmin && x <= xmax && y >= ymin && y <= ymax && z >= zmin && z <= zmax;
}

// ion slots ion0..n-1; slot j (when 0 <= j) takes the trial (xj, yj, zj, jtype)
__global__ void k_mm_leg(StepParams P, int n, int j, double xj, double yj, double zj, int jtype) {
#pragma clang fp contract(off)
    const int i = P.nsites + P.nfix + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x, y, z;
    int it;
    if (i == j) {
        x = xj; y = yj; z = zj; it = jtype;
    } else {
        const double4 p = P.pos[i];
        x = p.x; y = p.y; z = p.z; it = abs(P.typei[i]);
    }
    double *l = P.mleg + (size_t)i * P.npl;
    if (!mm_inside(P, x, y, z)) {
        for (int k = 0; k < P.npl; k++) l[k] = 0.0;
        P.mchg[i] = 0.0;
        return;
    }
    P.mchg[i] = P.cg[it];
    legendre(P.xscal * x, P.xnpol, l);
    legendre(P.yscal * y, P.ynpol, l + P.xnpol);
    legendre(P.zscal * z, P.znpol, l + P.xnpol + P.ynpol);
}

// charge * b_m at the polynomials l, in the reference's product order
__device__ inline double charge_basis(const StepParams &P, double q, const double *l, int m) {
#pragma clang fp contract(off)
    const int d = P.lstp[m];
    return q * P.bnorm[m] * l[d & 255] * l[P.xnpol + ((d >> 8) & 255)] *
           l[P.xnpol + P.ynpol + (d >> 16)];
}

// q_m over ion slots ion0..n-1, in the product order charge*norm*Px*Py*Pz (rect_rf1.f90:69)
__global__ __launch_bounds__(256) void k_mm_coef(StepParams P, int n) {
#pragma clang fp contract(off)
    __shared__ double s[256];
    const int m = blockIdx.x, t = threadIdx.x;
    const int d = P.lstp[m];
    const int ox = d & 255, oy = P.xnpol + ((d >> 8) & 255), oz = P.xnpol + P.ynpol + (d >> 16);
    const double norm = P.bnorm[m];
    double acc = 0.0;
    for (int i = P.nsites + P.nfix + t; i < n; i += 256) {
        const double *l = P.mleg + (size_t)i * P.npl;
        acc += P.mchg[i] * norm * l[ox] * l[oy] * l[oz];
    }
    s[t] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) s[t] += s[t + h];
        __syncthreads();
    }
    if (t == 0) P.mcoef[m] = s[0];
}

// row m of M q rfscal (rect_rf1.f90:73-79) and its energy term 0.5 q_m mq_m
__global__ __launch_bounds__(256) void k_mm_mq(StepParams P) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & (WAVE - 1);
    const int m = blockIdx.x * WPB + (threadIdx.x >> 6);
    if (m >= P.ntpol) return;
    const double *row = P.mmij + (size_t)m * P.ntpol;
    double acc = 0.0;
    for (int k = lane; k < P.ntpol; k += WAVE) acc += row[k] * P.mcoef[k] * P.rfscal;
    acc = wave_sum(acc);
    if (lane == 0) {
        P.mq[m] = acc;
        P.mer[m] = 0.5 * P.mcoef[m] * acc;
    }
}

// egsbpb = celec * sum of the rows' terms, into slot 0 of its energy category
__global__ __launch_bounds__(256) void k_mm_energy(StepParams P) {
    __shared__ double s[256];
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int m = t; m < P.ntpol; m += 256) acc += P.mer[m];
    s[t] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) s[t] += s[t + h];
        __syncthreads();
    }
    if (t == 0) P.epart[(size_t)E_GSBPB * P.ecap] = s[0] * P.celec;
}

// one wavefront per moving ion: the derivative tables of x, y, z (rect_rf1.f90:101-127) in
// LDS, then lanes over the basis functions (rect_rf1.f90:128-144)
__global__ __launch_bounds__(256) void k_mm_force(StepParams P) {
#pragma clang fp contract(off)
    __shared__ double s_dl[WPB][3 * 64];
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x >> 6;
    const int i0 = P.nsites + P.nfix + blockIdx.x * WPB + wv;
    const bool valid = i0 < P.ntot;   // the last block may hold fewer ions than wavefronts
    const int i = valid ? i0 : P.nsites + P.nfix;
    const double chg = valid ? P.mchg[i] : 0.0;
    const double *l = P.mleg + (size_t)i * P.npl;
    double *dl = s_dl[wv];
    if (valid && lane < 3) {
        const int n = lane == 0 ? P.xnpol : (lane == 1 ? P.ynpol : P.znpol);
        const int o = lane == 0 ? 0 : (lane == 1 ? P.xnpol : P.xnpol + P.ynpol);
        const double sc = lane == 0 ? P.xscal : (lane == 1 ? P.yscal : P.zscal);
        const double xs = l[o + 1];   // P_1 is the scaled coordinate itself
        double *d = dl + o;
        d[0] = 0.0;
        d[1] = 1.0;
        d[2] = 3.0 * xs;
        for (int k = 3; k <= n - 1; k++) d[k] = xs * d[k - 1] + k * l[o + k - 1];
        for (int k = 0; k < n; k++) d[k] = d[k] * sc;
    }
    __syncthreads();   // the tables written by lanes 0..2, read by every lane
    double rx = 0.0, ry = 0.0, rz = 0.0;
    if (chg != 0.0) {
        const int ny0 = P.xnpol, nz0 = P.xnpol + P.ynpol;
        for (int m = lane; m < P.ntpol; m += WAVE) {
            const int d = P.lstp[m];
            const int px = d & 255, py = (d >> 8) & 255, pz = d >> 16;
            const double norm = P.bnorm[m];
            const double xl = l[px], yl = l[ny0 + py], zl = l[nz0 + pz];
            const double dx = norm * dl[px] * yl * zl;
            const double dy = norm * xl * dl[ny0 + py] * zl;
            const double dz = norm * xl * yl * dl[nz0 + pz];
            const double q = P.mq[m];
            rx = rx - dx * q;
            ry = ry - dy * q;
            rz = rz - dz * q;
        }
    }
    rx = wave_sum(rx);
    ry = wave_sum(ry);
    rz = wave_sum(rz);
    if (valid && lane == 0) {
        const double ccc = chg * P.celec;
        P.frf[i] = make_double4(rx * ccc, ry * ccc, rz * ccc, 0.0);
    }
}

// RECT_RF0 row ii: 0.5 (ci^2 - cj^2) M_ii + sum over jj > ii of (ci q_jj - cj q2_jj) M_ij,
// q2 = q without ion j (rect_rf0.f90:84-110)
__global__ __launch_bounds__(256) void k_mm_rf0(StepParams P, int j) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & (WAVE - 1);
    const int ii = blockIdx.x * WPB + (threadIdx.x >> 6);
    if (ii >= P.ntpol) return;
    const double qj = P.mchg[j];
    const double *lj = P.mleg + (size_t)j * P.npl;
    const double *row = P.mmij + (size_t)ii * P.ntpol;
    const double ci = P.mcoef[ii];
    const double cj = ci - charge_basis(P, qj, lj, ii);
    double acc = 0.0;
    for (int jj = ii + 1 + lane; jj < P.ntpol; jj += WAVE) {
        const double c2 = P.mcoef[jj] - charge_basis(P, qj, lj, jj);
        acc += (ci * P.mcoef[jj] - cj * c2) * row[jj];
    }
    acc = wave_sum(acc);
    if (lane == 0) P.mer[ii] = 0.5 * (ci * ci - cj * cj) * row[ii] + acc;
}

__global__ __launch_bounds__(256) void k_mm_rf0_sum(StepParams P, double *out) {
    __shared__ double s[256];
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int m = t; m < P.ntpol; m += 256) acc += P.mer[m];
    s[t] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) s[t] += s[t + h];
        __syncthreads();
    }
    if (t == 0) out[0] += s[0] * P.celec * P.rfscal;
}

}  // namespace

hipError_t launch_mmij(const StepParams &p, hipStream_t s) {
    if (!p.mmon) return hipSuccess;
    const int nion = p.ntot - p.nsites - p.nfix;
    if (nion > 0)
        hipLaunchKernelGGL(k_mm_leg, dim3((nion + 255) / 256), dim3(256), 0, s, p, p.ntot, -1,
                           0.0, 0.0, 0.0, 0);
    hipLaunchKernelGGL(k_mm_coef, dim3(p.ntpol), dim3(256), 0, s, p, p.ntot);
    hipLaunchKernelGGL(k_mm_mq, dim3((p.ntpol + WPB - 1) / WPB), dim3(256), 0, s, p);
    if (p.write_energy) hipLaunchKernelGGL(k_mm_energy, dim3(1), dim3(256), 0, s, p);
    if (nion > 0)
        hipLaunchKernelGGL(k_mm_force, dim3((nion + WPB - 1) / WPB), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_mmij_interact(const StepParams &p, double xj, double yj, double zj, int jtype,
                                int j, double *out, hipStream_t s) {
    const int ion0 = p.nsites + p.nfix;
    // a site or fixed ion has no term (the reference reads polynomials it never set)
    if (!p.mmon || j < ion0) return hipSuccess;
    const int n = j == p.ntot ? p.ntot + 1 : p.ntot;   // an insertion adds slot ntot
    hipLaunchKernelGGL(k_mm_leg, dim3((n - ion0 + 255) / 256), dim3(256), 0, s, p, n, j, xj, yj,
                       zj, jtype);
    hipLaunchKernelGGL(k_mm_coef, dim3(p.ntpol), dim3(256), 0, s, p, n);
    hipLaunchKernelGGL(k_mm_rf0, dim3((p.ntpol + WPB - 1) / WPB), dim3(256), 0, s, p, j);
    hipLaunchKernelGGL(k_mm_rf0_sum, dim3(1), dim3(256), 0, s, p, out);
    return hipGetLastError();
}

}  // namespace bromoc
