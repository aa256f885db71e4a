// tables.hip -- EFPOT tables rebuilt on the device from new samples (the knot values a
// potential fit changes), into the context's own knot and pair arrays, and the
// vector-Jacobian product of the table construction for the adjoint.  The construction
// is the host's (tables.h): a table built here is bit-identical to bromoc_build_table's.
// Each table is a sequential recurrence over its knots (squarespline's integration and
// end correction), so a table is one thread, and so is the reverse sweep of its transpose.
#include <mutex>
#include <vector>

#include "engine.h"
#include "tables.h"

namespace bromoc {

// Workspace of bromoc_table_vjp_device_noalloc, a file-scope device array: each GPU the
// library runs on holds its own copy until the process ends.  `scratch` names a host-side
// stand-in; the device address of each GPU's copy comes from hipGetSymbolAddress, taken once
// per device (vjp_scratch_init) and used for every copy and kernel argument.
constexpr size_t VJP_SCRATCH_DOUBLES = size_t(2) << 20;   // 16 MiB
static __device__ __align__(256) double scratch[VJP_SCRATCH_DOUBLES];

static std::mutex g_scratch_mu;
static std::vector<double *> g_scratch_ptr;   // by device; null until taken

// Takes the address of this device's copy of scratch (the current device must be
// `device`); kept for later calls.
hipError_t vjp_scratch_init(int device, double **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0) return hipErrorInvalidDevice;
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        double *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = ptr;
    }
    if (out) *out = g_scratch_ptr[device];
    return hipSuccess;
}

size_t vjp_scratch_doubles() { return VJP_SCRATCH_DOUBLES; }

namespace {

// one thread per table: SPLINEPOT into a staging copy, then the knots as (ep1, ep2, ep3, 0)
__global__ void k_tab_build(TabJob j) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= j.ntab) return;
    const int64_t o = j.off[t];
    const int n = (int)(j.off[t + 1] - o);
    double *stage = j.scratch + 6 * o, *scr = stage + 3 * (int64_t)n;
    TabEnds e{};
    table_splinepot(n, j.r + o, j.u + o, j.kf[t], stage, e, scr);
    for (int k = 0; k < n; k++)
        j.ep[j.nxi[t] + k] = make_double4(stage[3 * k], stage[3 * k + 1], stage[3 * k + 2], 0.0);
    j.ends[t] = e;
}

__device__ inline void set_ends(DevPair &d, const TabEnds &e) {
    d.dmi = e.dmi; d.dm2lo = e.dm2lo; d.dm2hi = e.dm2hi;
    d.sc1 = e.sc1; d.sc2 = e.sc2; d.sc3 = e.sc3;
}

// the wall, range and tail of each rebuilt table into its pair type and every type-pair row
__global__ void k_tab_rows(TabJob j) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < j.nrows) {
        const int t = j.tab_of[j.pairidx[k]];
        if (t >= 0) set_ends(j.rows[k], j.ends[t]);
    }
    if (k <= j.npair) {
        const int t = j.tab_of[k];
        if (t >= 0) set_ends(j.pairs[k], j.ends[t]);
    }
}

// knot k = k0 + thread: the table of the unit samples e_k dotted with the gradient, in
// the order of the host's bromoc_table_vjp
// the transposed table build (table_splinepot_vjp), one thread: the sweep is sequential
__global__ void k_tab_vjp(int n, const double *r, const double *gep, const double *gsc,
                          double *out, double *scratch) {
    if (blockIdx.x * blockDim.x + threadIdx.x != 0) return;
    table_splinepot_vjp(n, r, gep, gsc, out, scratch);
}

}  // namespace

hipError_t launch_table_update(const TabJob &j, hipStream_t s) {
    hipLaunchKernelGGL(k_tab_build, dim3((j.ntab + 63) / 64), dim3(64), 0, s, j);
    const int m = j.nrows > j.npair + 1 ? j.nrows : j.npair + 1;
    hipLaunchKernelGGL(k_tab_rows, dim3((m + 255) / 256), dim3(256), 0, s, j);
    return hipGetLastError();
}

hipError_t launch_table_vjp(int n, const double *r, const double *gep, const double *gsc,
                            double *out, double *scratch, hipStream_t s) {
    hipLaunchKernelGGL(k_tab_vjp, dim3(1), dim3(64), 0, s, n, r, gep, gsc, out, scratch);
    return hipGetLastError();
}

}  // namespace bromoc
