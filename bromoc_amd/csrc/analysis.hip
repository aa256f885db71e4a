// analysis.hip -- the statistics of the reference's cycle loop (simul1.f90:477-538) and the
// fold of the virial-pressure sums (simul1.f90:365-368), on the device.  The histograms
// are integer counters bumped with atomics: counts are exact, so the order in which the
// threads arrive does not matter.  Everything here is a few bytes per particle per cycle.
#include "engine.h"

namespace bromoc {

namespace {

// One thread per particle.  Bins as the reference computes them, int() truncating toward
// zero: a sample with -1 < (z - zmini) * idelz < 0 lands in bin 1; samples the reference
// would write below its arrays (iz < 1) are left out.
__global__ void k_hist(AnParams a) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.ntot) return;
    const double4 p = a.pos[i];
    const int ion0 = a.nsites + a.nfix;
    if (i < a.nsites) {
        // rDNA(typtyp, iz), simul1.f90:487-495
        if (a.what & BROMOC_AN_RDNA) {
            const double v = (p.z - a.zmini) * a.idelz;
            if (v > -1.0 && v < (double)a.nzmax) {
                const int iz = (int)v + 1;
                const int t = __double2loint(p.w);
                atomicAdd(&a.rdna[(size_t)(t - 1) * a.nzmax + (iz - 1)], 1ull);
            }
        }
        return;
    }
    if (i < ion0) return;   // fixed ions take part in none of the statistics
    const int it = abs(a.typei[i]);
    // rho(itype, iz), simul1.f90:477-485
    if (a.what & BROMOC_AN_RHO) {
        const double v = (p.z - a.zmini) * a.idelz;
        if (v > -1.0 && v < (double)a.nzmax) {
            const int iz = (int)v + 1;
            atomicAdd(&a.rho[(size_t)(it - 1) * a.nzmax + (iz - 1)], 1ull);
        }
    }
    // g(r) around particle igr, simul1.f90:497-509
    if (a.what & BROMOC_AN_GR) {
        const double4 c = a.pos[a.igr];
        double dist = (p.x - c.x) * (p.x - c.x) + (p.y - c.y) * (p.y - c.y) +
                      (p.z - c.z) * (p.z - c.z);
        dist = sqrt(dist);
        const double q = dist / a.dr;
        if (q < (double)a.npoints) {
            const int ir = (int)q + 1;
            atomicMax(a.irmax, ir);
            atomicAdd(&a.gr[(size_t)(it - 1) * a.npoints + (ir - 1)], 1ull);
        } else {
            atomicMax(a.irmax, a.npoints);   // irmax is capped at npoints
        }
    }
    // ions of each type in the channel and in the system, simul1.f90:511-533
    if (a.what & BROMOC_AN_PROB) {
        if (a.chan != 0) {
            bool ok;
            if (a.chan == 1) {   // cylindrical pore of the membrane
                const double r = sqrt(p.x * p.x + p.y * p.y);
                ok = p.z >= a.zmemb1 && p.z <= a.zmemb2;
                const double rc = a.rcylinder[a.nwtype[it]];
                if (ok) ok = rc > 0.0 && r <= rc;
            } else {
                ok = p.z >= a.czmin && p.z <= a.czmax;
            }
            if (ok) atomicAdd(&a.ncnt[it], 1);
        }
        atomicAdd(&a.ncnt[(a.ntype + 1) + it], 1);
    }
}

// One thread per ion type: prob2(itype, count) and prob3(itype, count) of this cycle
// (simul1.f90:527-538, every type nold+1..ntype, also those with no ion), then the cycle's
// counts are cleared for the next one.
__global__ void k_prob(AnParams a) {
    const int it = a.nold + 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (it > a.ntype) return;
    const size_t row = (size_t)(it - 1) * (a.datom + 1);
    if (a.chan != 0) {
        const int n2 = a.ncnt[it];
        if (n2 <= a.datom) a.prob2[row + n2] += 1ull;
    }
    const int n3 = a.ncnt[(a.ntype + 1) + it];
    if (n3 <= a.datom) a.prob3[row + n3] += 1ull;
    a.ncnt[it] = 0;
    a.ncnt[(a.ntype + 1) + it] = 0;
}

// vsum[0] += sum of vacc[0..n): per-thread strided sums, then a fixed tree in LDS.
__global__ __launch_bounds__(256) void k_vir_fold(double *vacc, int n, double *vsum) {
    __shared__ double s[256];
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int i = t; i < n; i += 256) {
        acc += vacc[i];
        vacc[i] = 0.0;
    }
    s[t] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) s[t] += s[t + h];
        __syncthreads();
    }
    if (t == 0) vsum[0] += s[0];
}

}  // namespace

hipError_t launch_analyze(const AnParams &a, hipStream_t s) {
    if (a.ntot > 0)
        hipLaunchKernelGGL(k_hist, dim3((a.ntot + 255) / 256), dim3(256), 0, s, a);
    const int nion = a.ntype - a.nold;
    if ((a.what & BROMOC_AN_PROB) && nion > 0)
        hipLaunchKernelGGL(k_prob, dim3((nion + 63) / 64), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_vir_fold(double *vacc, int n, double *vsum, hipStream_t s) {
    hipLaunchKernelGGL(k_vir_fold, dim3(1), dim3(256), 0, s, vacc, n, vsum);
    return hipGetLastError();
}

}  // namespace bromoc
