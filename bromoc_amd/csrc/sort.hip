// sort.hip -- spatial order of the partner array (cell list as an integer key sort).
//
// Every particle gets the Morton code of its cell on a regular lattice over the box
// (10 bits per axis); a stable radix sort of (key, particle index) gives the partner order,
// and a gather writes the partner array spos (x, y, z, packed type and original index) in
// that order plus rank[i], the slot of particle i.  The step kernel keeps spos current
// between sorts by writing each updated particle to its slot, so a sort is only needed
// every few steps; the order only decides which partners share a wavefront (and thus
// whether the wavefront can take the analytic-tail path), never which pairs are summed.
#include <hipcub/hipcub.hpp>

#include "engine.h"

namespace bromoc {

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t cell_coord(double x, double lo, double inv, int n) {
    double c = (x - lo) * inv;
    int k = c > 0.0 ? (int)c : 0;   // also catches NaN
    return (uint32_t)(k < n ? k : n - 1);
}

__global__ void k_keys(const double4 *__restrict__ pos, int n, SortGeom g, uint32_t *keys,
                       int32_t *vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 p = pos[i];
    const uint32_t cx = cell_coord(p.x, g.lo[0], g.inv, g.n[0]);
    const uint32_t cy = cell_coord(p.y, g.lo[1], g.inv, g.n[1]);
    const uint32_t cz = cell_coord(p.z, g.lo[2], g.inv, g.n[2]);
    keys[i] = spread10(cx) | (spread10(cy) << 1) | (spread10(cz) << 2);
    vals[i] = i;
}

__global__ void k_gather(const double4 *__restrict__ pos, const int32_t *__restrict__ order, int n,
                         double4 *spos0, double4 *spos1, int32_t *rank) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int j = order[k];
    const double4 p = pos[j];
    spos0[k] = p;
    spos1[k] = p;
    rank[j] = k;
}

size_t sort_temp_bytes(int n) {
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t *)nullptr,
                                             (uint32_t *)nullptr, (const int32_t *)nullptr,
                                             (int32_t *)nullptr, n, 0, 30);
    return bytes;
}

hipError_t sort_partners(const SortBufs &b, const double4 *pos, int n, double4 *spos0,
                         double4 *spos1, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int blk = 256, grid = (n + blk - 1) / blk;
    hipLaunchKernelGGL(k_keys, dim3(grid), dim3(blk), 0, s, pos, n, b.geom, b.keys, b.vals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = b.temp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(b.temp, bytes, b.keys, b.keys_out, b.vals, b.order, n,
                                           0, 30, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gather, dim3(grid), dim3(blk), 0, s, pos, b.order, n, spos0, spos1,
                       b.rank);
    return hipGetLastError();
}

}  // namespace bromoc
